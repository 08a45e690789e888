"""Deterministic synthetic workloads (libish_workload.so).

Each workload is reproducible from (kind, seed, n): the same bytes come out
on any machine and with any thread count. The parameters and the choices the
task leaves open are documented in csrc/workload.cpp.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import LIB_DIR, IshError, Targets

_wl = None

KIND_QUERY_LEN = {1: 150, 2: 350, 4: 1000, 5: 350}
KIND_MATRIX = {1: "actgn", 2: "blosum62", 3: "blosum62", 4: "actgn0", 5: "blosum62"}


def _lib():
    global _wl
    if _wl is None:
        path = os.path.join(LIB_DIR, "libish_workload.so")
        if not os.path.exists(path):
            raise IshError(f"{path} is missing: run `python -m ish_amd.build` first")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER
        L.ish_wl_query.argtypes = [ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32, P(ctypes.c_uint8)]
        L.ish_wl_query.restype = ctypes.c_int32
        L.ish_wl_lengths.argtypes = [ctypes.c_int32, ctypes.c_uint64, ctypes.c_int64, P(ctypes.c_int64)]
        L.ish_wl_lengths.restype = ctypes.c_int64
        L.ish_wl_targets.argtypes = [ctypes.c_int32, ctypes.c_uint64, P(ctypes.c_uint8), ctypes.c_int32,
                                     ctypes.c_int64, P(ctypes.c_int64), P(ctypes.c_uint64),
                                     P(ctypes.c_uint8), P(ctypes.c_uint8), ctypes.c_int32]
        L.ish_wl_targets.restype = ctypes.c_int32
        L.ish_wl_hash.argtypes = [P(ctypes.c_uint8), ctypes.c_int64]
        L.ish_wl_hash.restype = ctypes.c_uint64
        _wl = L
    return _wl


@dataclass
class Workload:
    kind: int
    seed: int
    query: np.ndarray
    targets: Targets
    planted: np.ndarray
    matrix: str

    @property
    def cells(self) -> int:
        return int(len(self.query)) * int(self.targets.offsets[-1])

    def content_hash(self) -> str:
        L = _lib()
        h1 = L.ish_wl_hash(self.targets.residues.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                           int(self.targets.offsets[-1]))
        h2 = L.ish_wl_hash(self.query.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), len(self.query))
        return f"{h2:016x}-{h1:016x}"


def write_fastx(path: str, targets: Targets, matrix: str, fastq: bool = False, width: int = 60) -> None:
    """Writes targets as FASTA (or FASTQ) text, residues decoded with the
    matrix's decoding table (records named "t<index> synthetic")."""
    from . import decoding
    L = _lib()
    dec = np.zeros(256, np.uint8)
    d = decoding(matrix)
    dec[: len(d)] = d
    P = ctypes.POINTER
    L.ish_wl_write_fastx.argtypes = [ctypes.c_char_p, P(ctypes.c_uint8), P(ctypes.c_uint64), ctypes.c_int64,
                                     P(ctypes.c_uint8), ctypes.c_int32, ctypes.c_int32]
    L.ish_wl_write_fastx.restype = ctypes.c_int32
    rc = L.ish_wl_write_fastx(os.fsencode(path), targets.residues.ctypes.data_as(P(ctypes.c_uint8)),
                              targets.offsets.ctypes.data_as(P(ctypes.c_uint64)), targets.n,
                              dec.ctypes.data_as(P(ctypes.c_uint8)), int(fastq), width)
    if rc != 0:
        raise IshError(f"cannot write {path}")


def query(kind: int, seed: int, qlen: int | None = None) -> np.ndarray:
    qlen = KIND_QUERY_LEN.get(kind, 350) if qlen is None else qlen
    out = np.zeros(max(qlen, 1), np.uint8)
    _lib().ish_wl_query(kind, seed, qlen, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return out[:qlen]


def generate(kind: int, n: int, seed: int = 1, qlen: int | None = None, threads: int = 0) -> Workload:
    """kind 1: nucleotide lines, 2: protein FASTA, 3: the records of 1 as
    BLOSUM62 letters (A C G T N) with protein queries, 4: long reads, 5:
    serving database (see workload.cpp)."""
    L = _lib()
    q = query(kind, seed, qlen)
    lens = np.zeros(max(n, 1), np.int64)
    L.ish_wl_lengths(kind, seed, n, lens.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    lens = lens[:n]
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(lens, out=off[1:])
    res = np.zeros(max(int(off[-1]), 1), np.uint8)
    planted = np.zeros(max(n, 1), np.uint8)
    L.ish_wl_targets(kind, seed, q.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), len(q), n,
                     lens.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                     off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                     res.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                     planted.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), threads)
    return Workload(kind, seed, q, Targets(res, off), planted[:n].astype(bool), KIND_MATRIX[kind])


@dataclass
class Request:
    arrival: float      # seconds after the start of the stream
    query: np.ndarray   # BLOSUM62 codes at background frequencies
    ends_free: str      # "FFTT" or "TTTT"


def serving_requests(count: int, seed: int = 1, rate: float = 200.0, median: float = 300.0,
                     sigma: float = 0.5, lo: int = 20, hi: int = 2000,
                     ttt_fraction: float = 0.5, sources: Targets | None = None,
                     hit_fraction: float = 0.0) -> list[Request]:
    """Open-loop query stream of the serving workload: Poisson arrivals at
    `rate` per second, ln Q ~ N(ln median, sigma^2) rounded and clipped to
    [lo, hi], ends-free TTTT with probability ttt_fraction, else FFTT.
    Query residues come from the workload generator (background
    frequencies), one stream per request. With `sources`, a hit_fraction
    of the requests are instead a window of a random target of sources
    (clipped to that target's length), so that they have at least that
    target as a hit at any threshold."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, count)
    arrivals = np.cumsum(gaps) - gaps[0]
    lens = np.clip(np.rint(np.exp(np.log(median) + sigma * rng.standard_normal(count))), lo, hi).astype(int)
    ttt = rng.random(count) < ttt_fraction
    hit = rng.random(count) < hit_fraction if sources is not None and sources.n > 0 else np.zeros(count, bool)
    out = []
    for k in range(count):
        if hit[k]:
            t = int(rng.integers(0, sources.n))
            a, e = int(sources.offsets[t]), int(sources.offsets[t + 1])
            L = min(int(lens[k]), e - a)
            o = a + int(rng.integers(0, e - a - L + 1))
            q = np.array(sources.residues[o:o + L], np.uint8)
        else:
            q = query(5, (seed << 20) + k + 1, int(lens[k]))
        out.append(Request(float(arrivals[k]), q, "TTTT" if ttt[k] else "FFTT"))
    return out
