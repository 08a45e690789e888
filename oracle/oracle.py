"""ctypes binding of the CPU oracle (oracle/lib/liboracle*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product (ish_amd/) never
imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MATRIX_KINDS = {"ascii": 0, "actgn": 1, "actgn0": 2, "blosum62": 3}


class End(ctypes.Structure):
    _fields_ = [("score", ctypes.c_int32), ("end_ref", ctypes.c_int32),
                ("end_query", ctypes.c_int32)]


class StartEnd(ctypes.Structure):
    _fields_ = [("score", ctypes.c_int32), ("query_start", ctypes.c_int32),
                ("query_end", ctypes.c_int32), ("target_start", ctypes.c_int32),
                ("target_end", ctypes.c_int32)]


class Local(ctypes.Structure):
    _fields_ = [("score1", ctypes.c_int32), ("ref_begin1", ctypes.c_int32),
                ("ref_end1", ctypes.c_int32), ("read_begin1", ctypes.c_int32),
                ("read_end1", ctypes.c_int32)]


def host_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
        return all(x in flags for x in ("avx512f", "avx512bw", "avx512vl", "avx512dq", "avx512cd"))
    except OSError:
        return False


def build() -> None:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    name = "liboracle_v4.so" if host_has_avx512() else "liboracle.so"
    path = os.path.join(_HERE, "lib", name)
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    i8p = ctypes.POINTER(ctypes.c_int8)
    i32p = ctypes.POINTER(ctypes.c_int32)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    i32, u32, i64 = ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64
    L.oracle_alphabet_size.argtypes = [i32]
    L.oracle_alphabet_size.restype = i32
    L.oracle_matrix.argtypes = [i32, i8p]
    L.oracle_matrix.restype = i32
    L.oracle_encoding.argtypes = [i32, u8p]
    L.oracle_encoding.restype = i32
    L.oracle_decoding.argtypes = [i32, u8p]
    L.oracle_decoding.restype = i32
    L.oracle_encode_and_score.argtypes = [i32, u8p, i64, u8p]
    L.oracle_encode_and_score.restype = i64
    sig = [u8p, i32, u8p, i32, i8p, i32, i32, i32]
    L.oracle_sg.argtypes = sig + [u32]
    L.oracle_sg.restype = End
    L.oracle_sg_start_end.argtypes = sig + [u32, ctypes.c_float, ctypes.POINTER(StartEnd)]
    L.oracle_sg_start_end.restype = i32
    L.oracle_sw.argtypes = sig
    L.oracle_sw.restype = End
    L.oracle_sw_align.argtypes = sig + [ctypes.c_float, ctypes.POINTER(Local)]
    L.oracle_sw_align.restype = i32
    L.oracle_sg_emit.argtypes = [i32, i64, ctypes.c_float]
    L.oracle_sg_emit.restype = i32
    L.oracle_sw_emit.argtypes = [i32, i32, i32, i64, ctypes.c_float]
    L.oracle_sw_emit.restype = i32
    bsig = [i32, u8p, i32, u8p, u64p, i64, i8p, i32, i32, i32, u32, i32p, i32p, i32p, i32]
    L.oracle_batch.argtypes = bsig
    L.oracle_batch.restype = i32
    L.oracle_batch_striped.argtypes = bsig
    L.oracle_batch_striped.restype = i32
    L.oracle_sg_striped.argtypes = sig + [u32, i32]
    L.oracle_sg_striped.restype = End
    L.oracle_sw_striped.argtypes = sig + [i32]
    L.oracle_sw_striped.restype = End
    _LIB = L
    return L


def _u8(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def flags_from_str(s: str) -> int:
    """--sg-ends-free letters (qs, qe, ts, te) -> flag bits."""
    s = s.upper()
    assert len(s) == 4 and set(s) <= {"T", "F"}
    return sum(1 << k for k, c in enumerate(s) if c == "T")


class Matrix:
    def __init__(self, kind: str):
        self.kind = MATRIX_KINDS[kind]
        L = lib()
        self.A = L.oracle_alphabet_size(self.kind)
        self.values = np.zeros(65536, dtype=np.int8)
        L.oracle_matrix(self.kind, self.values.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)))
        self.values = self.values[: self.A * self.A].copy()
        self.enc = np.zeros(256, dtype=np.uint8)
        L.oracle_encoding(self.kind, self.enc.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        dec = np.zeros(256, dtype=np.uint8)
        n = L.oracle_decoding(self.kind, dec.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        self.dec = dec[:n].copy()

    def encode(self, s) -> np.ndarray:
        b = np.frombuffer(s.encode() if isinstance(s, str) else bytes(s), dtype=np.uint8)
        return self.enc[b]

    def max_score(self, s) -> int:
        e = self.encode(s)
        return int(self.values.reshape(self.A, self.A)[e, e].sum())

    def ptr(self):
        return self.values.ctypes.data_as(ctypes.POINTER(ctypes.c_int8))


def sg(q, r, m: Matrix, go, ge, flags):
    q, qp = _u8(q)
    r, rp = _u8(r)
    e = lib().oracle_sg(qp, len(q), rp, len(r), m.ptr(), m.A, go, ge, flags)
    return (e.score, e.end_ref, e.end_query)


def sg_striped(q, r, m: Matrix, go, ge, flags, width=16):
    q, qp = _u8(q)
    r, rp = _u8(r)
    e = lib().oracle_sg_striped(qp, len(q), rp, len(r), m.ptr(), m.A, go, ge, flags, width)
    return (e.score, e.end_ref, e.end_query)


def sg_start_end(q, r, m: Matrix, go, ge, flags, cutoff=-1e30):
    q, qp = _u8(q)
    r, rp = _u8(r)
    out = StartEnd()
    ok = lib().oracle_sg_start_end(qp, len(q), rp, len(r), m.ptr(), m.A, go, ge, flags,
                                   ctypes.c_float(cutoff), ctypes.byref(out))
    if not ok:
        return None
    return dict(score=out.score, query_start=out.query_start, query_end=out.query_end,
                target_start=out.target_start, target_end=out.target_end)


def sw(q, r, m: Matrix, go, ge):
    q, qp = _u8(q)
    r, rp = _u8(r)
    e = lib().oracle_sw(qp, len(q), rp, len(r), m.ptr(), m.A, go, ge)
    return (e.score, e.end_ref, e.end_query)


def sw_striped(q, r, m: Matrix, go, ge, width=16):
    q, qp = _u8(q)
    r, rp = _u8(r)
    e = lib().oracle_sw_striped(qp, len(q), rp, len(r), m.ptr(), m.A, go, ge, width)
    return (e.score, e.end_ref, e.end_query)


def sw_align(q, r, m: Matrix, go, ge, cutoff=0.0):
    q, qp = _u8(q)
    r, rp = _u8(r)
    out = Local()
    ok = lib().oracle_sw_align(qp, len(q), rp, len(r), m.ptr(), m.A, go, ge,
                               ctypes.c_float(cutoff), ctypes.byref(out))
    if not ok:
        return None
    return dict(score1=out.score1, ref_begin1=out.ref_begin1, ref_end1=out.ref_end1,
                read_begin1=out.read_begin1, read_end1=out.read_end1)


def sg_emit(score, max_score, thr) -> bool:
    return bool(lib().oracle_sg_emit(int(score), int(max_score), ctypes.c_float(thr)))


def sw_emit(score, end_ref, end_q, max_score, thr) -> bool:
    return bool(lib().oracle_sw_emit(int(score), int(end_ref), int(end_q), int(max_score),
                                     ctypes.c_float(thr)))


def sat_lanes(type_: str, width: int, op: str, a, b):
    """The striped restatement's saturating add / sub on `width` lanes of
    type "u8", "i16" or "u16" (the overloads its kernels use)."""
    t = {"u8": 0, "i16": 1, "u16": 2}[type_]
    a = np.ascontiguousarray(a, np.int32)
    b = np.ascontiguousarray(b, np.int32)
    r = np.zeros(width, np.int32)
    P = ctypes.POINTER(ctypes.c_int32)
    L = lib()
    L.oracle_sat_lanes.argtypes = [ctypes.c_int32] * 3 + [P, P, P]
    L.oracle_sat_lanes.restype = ctypes.c_int32
    rc = L.oracle_sat_lanes(t, width, 0 if op == "add" else 1, a.ctypes.data_as(P), b.ctypes.data_as(P),
                            r.ctypes.data_as(P))
    if rc != 0:
        raise ValueError(f"no {type_} lanes at width {width}")
    return r


def local_profile(q, values, A, bias, type_: str, width: int):
    """The biased ssw profile of the striped restatement: (A * seg, width)."""
    q, qp = _u8(q)
    S = np.ascontiguousarray(values, np.int8)
    seg = (len(q) + width - 1) // width
    out = np.zeros(A * seg * width, np.int32)
    L = lib()
    L.oracle_local_profile.argtypes = [ctypes.POINTER(ctypes.c_uint8), ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_int8), ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.oracle_local_profile.restype = ctypes.c_int32
    got = L.oracle_local_profile(qp, len(q), S.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)), A, bias,
                                 {"u8": 0, "u16": 2}[type_], width, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    if got != seg:
        raise ValueError(f"no {type_} profile at width {width}")
    return out.reshape(A * seg, width)


def batch(algo, q, residues, offsets, m: Matrix, go, ge, flags, nthreads=0, striped=False):
    """Scores every target; returns (score, end_ref, end_query, threads_used)."""
    q, qp = _u8(q)
    residues, rp = _u8(residues)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    n = len(offsets) - 1
    out = [np.zeros(n, dtype=np.int32) for _ in range(3)]
    i32p = ctypes.POINTER(ctypes.c_int32)
    fn = lib().oracle_batch_striped if striped else lib().oracle_batch
    used = fn(algo, qp, len(q), rp, offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
              n, m.ptr(), m.A, go, ge, flags, *(o.ctypes.data_as(i32p) for o in out), nthreads)
    return out[0], out[1], out[2], used
