"""ish_amd: MI355X-native alignment-scoring engine for ish.

Python binding of the C ABI in include/ish_hip.h (libish_hip.so). The engine
itself is the HIP library; this module only marshals arrays. There is no
CPU fallback: every scoring call goes to the GPU and fails loudly when the
library or the device is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")

ISH_OK = 0
ERRORS = {-1: "ISH_E_INVALID", -2: "ISH_E_HIP", -3: "ISH_E_NOMEM", -4: "ISH_E_NO_QUERY",
          -5: "ISH_E_RANGE", -6: "ISH_E_NO_DEVICE", -7: "ISH_E_TIMEOUT"}
ISH_E_TIMEOUT = -7
MATRIX = {"ascii": 0, "actgn": 1, "actgn0": 2, "blosum62": 3}
ALGO = {"striped-semi-global": 0, "semi-global": 0, "striped-local": 1, "local": 1}
ENGINE = {"batched": 0, "plain": 1}
SCHEDULE = {"auto": 0, "wave": 1, "workgroup": 2}
FREE_QUERY_START, FREE_QUERY_END, FREE_TARGET_START, FREE_TARGET_END = 1, 2, 4, 8
STATUS_I16, STATUS_I32, STATUS_F16, STATUS_WIDENED = 16, 32, 0x40, 0x80

# every symbol the public header declares
ABI_SYMBOLS = [
    "ish_hip_version", "ish_hip_device_count", "ish_hip_create", "ish_hip_destroy",
    "ish_hip_last_error", "ish_hip_matrix", "ish_hip_encoding", "ish_hip_decoding",
    "ish_hip_encode_and_score", "ish_hip_set_query", "ish_hip_set_engine",
    "ish_hip_set_tile", "ish_hip_set_schedule", "ish_hip_set_f16", "ish_hip_autotune", "ish_hip_tuning",
    "ish_hip_set_graphs", "ish_hip_align_batch", "ish_hip_align_batch_starts",
    "ish_hip_align_batch_submit", "ish_hip_align_batch_wait", "ish_hip_set_slice_bytes", "ish_hip_db_load", "ish_hip_db_load_ascii", "ish_hip_db_free", "ish_hip_db_size",
    "ish_hip_align_db", "ish_hip_align_db_noalloc", "ish_hip_align_db_async", "ish_hip_sync", "ish_hip_db_fetch",
    "ish_hip_last_timing", "ish_hip_emit", "ish_hip_find_starts", "ish_hip_align_db_multi",
    "ish_hip_server_create", "ish_hip_server_destroy", "ish_hip_server_last_error",
    "ish_hip_server_clock", "ish_hip_server_submit", "ish_hip_server_wait", "ish_hip_server_hits",
    "ish_hip_server_scores", "ish_hip_server_release", "ish_hip_server_get_stats",
    "ish_hip_search_file", "ish_hip_records_read", "ish_hip_records_read_ex", "ish_hip_records_count", "ish_hip_records_flags",
    "ish_hip_records_status", "ish_hip_records_get", "ish_hip_records_free",
]
RECORD = {"line": 0, "fastx": 1}


class SearchConfig(ctypes.Structure):
    _fields_ = [("matrix_kind", ctypes.c_int32), ("algo", ctypes.c_int32), ("gap_open", ctypes.c_int32),
                ("gap_extend", ctypes.c_int32), ("ends_free", ctypes.c_uint32), ("threshold", ctypes.c_float),
                ("record_type", ctypes.c_int32), ("batch_bytes", ctypes.c_int64),
                ("reader_threads", ctypes.c_int32), ("reader_segment_bytes", ctypes.c_int32),
                ("tty", ctypes.c_int32)]


class SearchStats(ctypes.Structure):
    _fields_ = [("records", ctypes.c_int64), ("residues", ctypes.c_int64), ("matches", ctypes.c_int64),
                ("batches", ctypes.c_int64), ("cells", ctypes.c_double), ("seconds_total", ctypes.c_double),
                ("seconds_parse", ctypes.c_double), ("seconds_gpu", ctypes.c_double),
                ("seconds_write", ctypes.c_double), ("reader_status", ctypes.c_int32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("matrix_kind", ctypes.c_int32), ("matrix", ctypes.POINTER(ctypes.c_int8)),
                ("matrix_len", ctypes.c_int32), ("gap_open", ctypes.c_int32),
                ("gap_extend", ctypes.c_int32), ("threshold", ctypes.c_float),
                ("max_batch", ctypes.c_int32), ("slots", ctypes.c_int32),
                ("batch_cells", ctypes.c_int64), ("hits_per_query", ctypes.c_int64),
                ("keep_scores", ctypes.c_int32), ("aging", ctypes.c_double),
                ("batch_seconds", ctypes.c_double)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("n_hits", ctypes.c_int64), ("cells", ctypes.c_int64), ("t_submit", ctypes.c_double),
                ("t_launch", ctypes.c_double), ("t_done", ctypes.c_double),
                ("batch_id", ctypes.c_int32), ("batch_size", ctypes.c_int32),
                ("status", ctypes.c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_int64), ("queries", ctypes.c_int64), ("cells", ctypes.c_int64),
                ("kernel_ms", ctypes.c_double), ("pending", ctypes.c_int64),
                ("busy_slots", ctypes.c_int64)]


class IshError(RuntimeError):
    pass


_lib = None


def lib_path() -> str:
    return os.path.join(LIB_DIR, "libish_hip.so")


def lib():
    """Loads libish_hip.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise IshError(f"{path} is missing: run `python -m ish_amd.build` first")
    L = ctypes.CDLL(path)
    P = ctypes.POINTER
    u8p, i8p, i32p, u64p = P(ctypes.c_uint8), P(ctypes.c_int8), P(ctypes.c_int32), P(ctypes.c_uint64)
    i32, u32, i64, vp = ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64, ctypes.c_void_p
    sigs = {
        "ish_hip_version": ([], i32),
        "ish_hip_device_count": ([], i32),
        "ish_hip_create": ([i32, P(vp)], i32),
        "ish_hip_destroy": ([vp], None),
        "ish_hip_last_error": ([vp], ctypes.c_char_p),
        "ish_hip_matrix": ([i32, i8p, i32], i32),
        "ish_hip_encoding": ([i32, u8p], i32),
        "ish_hip_decoding": ([i32, u8p, i32], i32),
        "ish_hip_encode_and_score": ([i32, u8p, i64, u8p], i64),
        "ish_hip_set_query": ([vp, u8p, i32, i8p, i32, i32, i32, i32, u32, i32], i32),
        "ish_hip_set_engine": ([vp, i32], i32),
        "ish_hip_set_tile": ([vp, i32], i32),
        "ish_hip_set_schedule": ([vp, i32], i32),
        "ish_hip_set_f16": ([vp, i32], i32),
        "ish_hip_autotune": ([vp, vp, i32], i32),
        "ish_hip_tuning": ([vp, vp, i32p, i32p], i32),
        "ish_hip_set_graphs": ([vp, i32], i32),
        "ish_hip_set_slice_bytes": ([vp, ctypes.c_int64], i32),
        "ish_hip_align_batch": ([vp, u8p, u64p, i64, i32p, i32p, i32p, u8p], i32),
        "ish_hip_align_batch_starts": ([vp, u8p, u64p, i64, i64, ctypes.c_float, i32p, i32p, i32p, u8p,
                                        u8p, i32p, i32p], i32),
        "ish_hip_align_batch_submit": ([vp, u8p, u64p, i64, i64, ctypes.c_float, i32p, i32p, i32p, u8p,
                                        u8p, i32p, i32p], i32),
        "ish_hip_align_batch_wait": ([vp], i32),
        "ish_hip_db_load": ([vp, u8p, u64p, i64, P(vp)], i32),
        "ish_hip_db_load_ascii": ([vp, u8p, u64p, i64, i32, P(vp)], i32),
        "ish_hip_db_free": ([vp], None),
        "ish_hip_db_size": ([vp, P(i64), P(i64)], i32),
        "ish_hip_align_db": ([vp, vp, i32p, i32p, i32p, u8p], i32),
        "ish_hip_align_db_noalloc": ([vp, vp, vp, vp, vp, vp], i32),
        "ish_hip_align_db_async": ([vp, vp], i32),
        "ish_hip_sync": ([vp], i32),
        "ish_hip_db_fetch": ([vp, vp, i32p, i32p, i32p, u8p], i32),
        "ish_hip_last_timing": ([vp, P(ctypes.c_float), i32p], i32),
        "ish_hip_emit": ([i32, i32p, i32p, i32p, i64, i64, ctypes.c_float, u8p], i32),
        "ish_hip_find_starts": ([vp, vp, P(i64), i32p, i32p, i64, i32p, i32p], i32),
        "ish_hip_align_db_multi": ([vp, vp, u8p, u64p, i32, i8p, i32, i32, i32, i32, u32,
                                    i32p, i32p, i32p, u8p], i32),
        "ish_hip_server_create": ([vp, P(ServerConfig), P(vp)], i32),
        "ish_hip_server_destroy": ([vp], None),
        "ish_hip_server_last_error": ([vp], ctypes.c_char_p),
        "ish_hip_server_clock": ([], ctypes.c_double),
        "ish_hip_server_submit": ([vp, u8p, i32, u32, P(i64)], i32),
        "ish_hip_server_wait": ([vp, i64, i32, P(RequestInfo)], i32),
        "ish_hip_server_hits": ([vp, i64, P(i64), i32p, i32p, i32p, i64], i32),
        "ish_hip_server_scores": ([vp, i64, i32p, i32p, i32p], i32),
        "ish_hip_server_release": ([vp, i64], i32),
        "ish_hip_server_get_stats": ([vp, P(ServerStats)], i32),
        "ish_hip_search_file": ([vp, ctypes.c_char_p, u8p, i32, P(SearchConfig), i32, P(SearchStats)], i32),
        "ish_hip_records_read": ([ctypes.c_char_p, i32, P(vp)], i32),
        "ish_hip_records_read_ex": ([ctypes.c_char_p, i32, i32, ctypes.c_int64, P(vp)], i32),
        "ish_hip_records_count": ([vp], i64),
        "ish_hip_records_flags": ([vp], i32),
        "ish_hip_records_status": ([vp], i32),
        "ish_hip_records_get": ([vp, i64, P(u8p), P(i64), P(u8p), P(i64), P(u8p), P(i64)], i32),
        "ish_hip_records_free": ([vp], None),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _matrix_arg(values):
    """(pointer, matrix_len) of a caller's scoring matrix, passed the way
    GpuMatcher.matrix_bytes() / matrix_len() hand it over: the flat int8
    values and their count; (None, 0) selects the built-in matrix."""
    if values is None:
        return None, 0, None
    flat = np.ascontiguousarray(np.asarray(values, np.int8).reshape(-1))
    return _p(flat, ctypes.c_int8), len(flat), flat


def device_count() -> int:
    return int(lib().ish_hip_device_count())


def matrix(kind: str) -> np.ndarray:
    out = np.zeros(65536, np.int8)
    A = lib().ish_hip_matrix(MATRIX[kind], _p(out, ctypes.c_int8), 65536)
    if A < 0:
        raise IshError(f"unknown matrix {kind}")
    return out[: A * A].reshape(A, A).copy()


def encoding(kind: str) -> np.ndarray:
    out = np.zeros(256, np.uint8)
    if lib().ish_hip_encoding(MATRIX[kind], _p(out, ctypes.c_uint8)) != 0:
        raise IshError(f"unknown matrix {kind}")
    return out


def decoding(kind: str) -> np.ndarray:
    out = np.zeros(256, np.uint8)
    A = lib().ish_hip_decoding(MATRIX[kind], _p(out, ctypes.c_uint8), 256)
    if A < 0:
        raise IshError(f"unknown matrix {kind}")
    return out[:A].copy()


def encode_and_score(kind: str, text) -> tuple[np.ndarray, int]:
    b = np.frombuffer(text.encode() if isinstance(text, str) else bytes(text), np.uint8).copy()
    out = np.zeros(max(len(b), 1), np.uint8)
    s = lib().ish_hip_encode_and_score(MATRIX[kind], _p(b, ctypes.c_uint8), len(b), _p(out, ctypes.c_uint8))
    return out[: len(b)], int(s)


def emit(algo, score, end_ref, end_query, max_score: int, threshold: float) -> np.ndarray:
    score = np.ascontiguousarray(score, np.int32)
    end_ref = np.ascontiguousarray(end_ref, np.int32)
    end_query = np.ascontiguousarray(end_query, np.int32)
    out = np.zeros(max(len(score), 1), np.uint8)
    rc = lib().ish_hip_emit(ALGO.get(algo, algo) if isinstance(algo, str) else algo,
                            _p(score, ctypes.c_int32), _p(end_ref, ctypes.c_int32),
                            _p(end_query, ctypes.c_int32), len(score), int(max_score),
                            ctypes.c_float(threshold), _p(out, ctypes.c_uint8))
    if rc != 0:
        raise IshError(f"ish_hip_emit failed ({rc})")
    return out[: len(score)].astype(bool)


def flags_from_str(s: str) -> int:
    """--sg-ends-free letters (query start, query end, target start, target end)."""
    s = s.upper()
    if len(s) != 4 or set(s) - {"T", "F"}:
        raise ValueError("ends-free must be 4 letters of T/F, like FFTT")
    return sum(1 << k for k, c in enumerate(s) if c == "T")


def read_records(path, record_type: str = "line", threads: int = 0, segment_bytes: int = 0):
    """Every record of a file as the runner reads it (host only): returns
    (records, flags, status); records are (seq, name, qual) bytes, flags bit 0
    = binary, bit 1 = FASTQ. threads / segment_bytes set the parallel parse of
    plain files (0 = defaults); the records never depend on them."""
    L = lib()
    h = ctypes.c_void_p()
    rc = L.ish_hip_records_read_ex(os.fsencode(path), RECORD[record_type], threads, segment_bytes,
                                   ctypes.byref(h))
    if rc != 0:
        raise IshError(f"ish_hip_records_read failed: {ERRORS.get(rc, rc)}")
    try:
        n = L.ish_hip_records_count(h)
        out = []
        P = ctypes.POINTER(ctypes.c_uint8)
        ptrs = [P() for _ in range(3)]
        lens = [ctypes.c_int64() for _ in range(3)]
        for i in range(n):
            L.ish_hip_records_get(h, i, ctypes.byref(ptrs[0]), ctypes.byref(lens[0]), ctypes.byref(ptrs[1]),
                                  ctypes.byref(lens[1]), ctypes.byref(ptrs[2]), ctypes.byref(lens[2]))
            out.append(tuple(ctypes.string_at(p, n_.value) if n_.value else b"" for p, n_ in zip(ptrs, lens)))
        return out, int(L.ish_hip_records_flags(h)), int(L.ish_hip_records_status(h))
    finally:
        L.ish_hip_records_free(h)


class Targets:
    """Encoded targets as one residue buffer plus offsets[n+1]."""

    def __init__(self, residues: np.ndarray, offsets: np.ndarray):
        self.residues = np.ascontiguousarray(residues, np.uint8)
        self.offsets = np.ascontiguousarray(offsets, np.uint64)
        if len(self.residues) == 0:
            self.residues = np.zeros(1, np.uint8)

    @classmethod
    def from_list(cls, seqs):
        lens = np.array([len(s) for s in seqs], np.uint64)
        off = np.zeros(len(seqs) + 1, np.uint64)
        np.cumsum(lens, out=off[1:])
        res = np.concatenate([np.asarray(s, np.uint8) for s in seqs]) if len(seqs) and off[-1] else np.zeros(1, np.uint8)
        return cls(res, off)

    @property
    def n(self) -> int:
        return len(self.offsets) - 1


class Engine:
    """One ish_hip context on one device."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.ish_hip_create(device, ctypes.byref(h))
        if rc != 0:
            raise IshError(f"ish_hip_create failed: {ERRORS.get(rc, rc)}")
        self._h = h
        self.algo = 0
        self._pending = None

    def close(self):
        if getattr(self, "_h", None):
            self._L.ish_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != 0:
            msg = self._L.ish_hip_last_error(self._h)
            raise IshError(f"{ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def set_engine(self, engine: str):
        self._check(self._L.ish_hip_set_engine(self._h, ENGINE[engine]))

    def set_tile(self, rows: int):
        self._check(self._L.ish_hip_set_tile(self._h, rows))

    def set_schedule(self, schedule: str):
        self._check(self._L.ish_hip_set_schedule(self._h, SCHEDULE[schedule]))

    def set_f16(self, on: bool):
        """Binary16 lanes for the groups within the query's range bound
        (default on); off keeps every group in int16. Results are the same."""
        self._check(self._L.ish_hip_set_f16(self._h, int(bool(on))))

    def set_graphs(self, on: bool):
        self._check(self._L.ish_hip_set_graphs(self._h, int(bool(on))))

    def set_slice_bytes(self, nbytes: int):
        """Slice size of align() in residue bytes (0: automatic, the
        default; < 0: never slice): a larger batch is uploaded and scored
        slice by slice, each upload overlapping the scoring of the slice
        before."""
        self._check(self._L.ish_hip_set_slice_bytes(self._h, int(nbytes)))

    def autotune(self, db: "Database", reps: int = 2):
        """Times whole passes over db with every launch configuration for
        the current query and keeps the fastest; returns tuning(db)."""
        self._check(self._L.ish_hip_autotune(self._h, db._h, reps))
        return self.tuning(db)

    def tuning(self, db: "Database"):
        """(tile, schedule) tuned for the current query and db, or None."""
        t, sc = ctypes.c_int32(), ctypes.c_int32()
        n = self._L.ish_hip_tuning(self._h, db._h, ctypes.byref(t), ctypes.byref(sc))
        if n < 0:
            self._check(n)
        if n == 0:
            return None
        inv = {v: k for k, v in SCHEDULE.items()}
        return int(t.value), inv.get(int(sc.value), "auto")

    def set_query(self, query: np.ndarray, matrix_kind: str, gap_open: int = 3,
                  gap_extend: int = 1, ends_free="FFTT", algo="striped-semi-global",
                  matrix=None):
        """matrix: the kind's score values (A x A or flat, int8), passed as
        pointer and entry count; None selects the built-in matrix."""
        q = np.ascontiguousarray(query, np.uint8)
        qb = q if len(q) else np.zeros(1, np.uint8)
        fl = flags_from_str(ends_free) if isinstance(ends_free, str) else int(ends_free)
        self.algo = ALGO[algo] if isinstance(algo, str) else int(algo)
        mp, mlen, _keep = _matrix_arg(matrix)
        self._check(self._L.ish_hip_set_query(self._h, _p(qb, ctypes.c_uint8), len(q), mp, mlen,
                                              MATRIX[matrix_kind], gap_open, gap_extend, fl,
                                              self.algo))

    def align(self, targets: Targets):
        n = targets.n
        out = [np.zeros(max(n, 1), np.int32) for _ in range(3)]
        st = np.zeros(max(n, 1), np.uint8)
        self._check(self._L.ish_hip_align_batch(
            self._h, _p(targets.residues, ctypes.c_uint8), _p(targets.offsets, ctypes.c_uint64), n,
            _p(out[0], ctypes.c_int32), _p(out[1], ctypes.c_int32), _p(out[2], ctypes.c_int32),
            _p(st, ctypes.c_uint8)))
        return out[0][:n], out[1][:n], out[2][:n], st[:n]

    def align_starts(self, targets: Targets, max_score: int, threshold: float):
        """One batch with start coordinates (ish_hip_align_batch_starts):
        returns score, end_ref, end_query, status, emit (bool), target_start,
        query_start (-1 where not emitted)."""
        n = targets.n
        out = [np.zeros(max(n, 1), np.int32) for _ in range(5)]
        st = np.zeros(max(n, 1), np.uint8)
        em = np.zeros(max(n, 1), np.uint8)
        self._check(self._L.ish_hip_align_batch_starts(
            self._h, _p(targets.residues, ctypes.c_uint8), _p(targets.offsets, ctypes.c_uint64), n,
            int(max_score), ctypes.c_float(threshold), *(_p(o, ctypes.c_int32) for o in out[:3]),
            _p(st, ctypes.c_uint8), _p(em, ctypes.c_uint8), _p(out[3], ctypes.c_int32),
            _p(out[4], ctypes.c_int32)))
        return (out[0][:n], out[1][:n], out[2][:n], st[:n], em[:n].astype(bool), out[3][:n], out[4][:n])

    def submit(self, targets: Targets, max_score=None, threshold: float = 0.8):
        """Starts one batch on the context's host thread and returns a handle
        for wait(); with max_score the batch also emits and finds starts. The
        targets and the result arrays stay referenced until wait()."""
        n = targets.n
        res = {"targets": targets, "n": n, "out": [np.zeros(max(n, 1), np.int32) for _ in range(3)],
               "st": np.zeros(max(n, 1), np.uint8), "starts": max_score is not None}
        P = ctypes.POINTER
        nul32, nul8 = ctypes.cast(None, P(ctypes.c_int32)), ctypes.cast(None, P(ctypes.c_uint8))
        if res["starts"]:
            res["em"] = np.zeros(max(n, 1), np.uint8)
            res["ts"] = np.zeros(max(n, 1), np.int32)
            res["qs"] = np.zeros(max(n, 1), np.int32)
            extra = (_p(res["em"], ctypes.c_uint8), _p(res["ts"], ctypes.c_int32), _p(res["qs"], ctypes.c_int32))
        else:
            extra = (nul8, nul32, nul32)
        self._check(self._L.ish_hip_align_batch_submit(
            self._h, _p(targets.residues, ctypes.c_uint8), _p(targets.offsets, ctypes.c_uint64), n,
            int(max_score or 0), ctypes.c_float(threshold), *(_p(o, ctypes.c_int32) for o in res["out"]),
            _p(res["st"], ctypes.c_uint8), *extra))
        self._pending = res
        return res

    def wait(self):
        """Waits for the submitted batch; returns align()'s tuple, or
        align_starts()'s when it was submitted with max_score."""
        res, self._pending = self._pending, None
        self._check(self._L.ish_hip_align_batch_wait(self._h))
        n = res["n"]
        base = (res["out"][0][:n], res["out"][1][:n], res["out"][2][:n], res["st"][:n])
        if not res["starts"]:
            return base
        return base + (res["em"][:n].astype(bool), res["ts"][:n], res["qs"][:n])

    def load_db(self, targets: Targets) -> "Database":
        h = ctypes.c_void_p()
        self._check(self._L.ish_hip_db_load(self._h, _p(targets.residues, ctypes.c_uint8),
                                            _p(targets.offsets, ctypes.c_uint64), targets.n,
                                            ctypes.byref(h)))
        return Database(self, h, targets.n)

    def load_db_ascii(self, targets: Targets, matrix_kind: str) -> "Database":
        """Device-resident database from raw record bytes, encoded on the
        device with the matrix's encoding."""
        h = ctypes.c_void_p()
        self._check(self._L.ish_hip_db_load_ascii(self._h, _p(targets.residues, ctypes.c_uint8),
                                                  _p(targets.offsets, ctypes.c_uint64), targets.n,
                                                  MATRIX[matrix_kind], ctypes.byref(h)))
        return Database(self, h, targets.n)

    def align_db(self, db: "Database"):
        self.align_db_async(db)
        return self.fetch(db)

    def align_db_noalloc(self, db: "Database", d_score: int, d_end_ref: int, d_end_query: int,
                         d_status: int):
        """ish_hip_align_db into caller-owned device arrays (raw device
        addresses, n entries each), with the kernels' scratch in the
        library's static device array: allocates nothing."""
        self._check(self._L.ish_hip_align_db_noalloc(self._h, db._h, d_score, d_end_ref, d_end_query,
                                                     d_status))

    def align_db_async(self, db: "Database"):
        self._check(self._L.ish_hip_align_db_async(self._h, db._h))

    def sync(self):
        self._check(self._L.ish_hip_sync(self._h))

    def fetch(self, db: "Database"):
        n = db.n
        out = [np.zeros(max(n, 1), np.int32) for _ in range(3)]
        st = np.zeros(max(n, 1), np.uint8)
        self._check(self._L.ish_hip_db_fetch(self._h, db._h, _p(out[0], ctypes.c_int32),
                                             _p(out[1], ctypes.c_int32), _p(out[2], ctypes.c_int32),
                                             _p(st, ctypes.c_uint8)))
        return out[0][:n], out[1][:n], out[2][:n], st[:n]

    def find_starts(self, db: "Database", sel, end_ref, end_query):
        """Start coordinates (target_start, query_start) of the selected
        targets from their forward end coordinates (reverse pass on the GPU)."""
        sel = np.ascontiguousarray(sel, np.int64)
        er = np.ascontiguousarray(end_ref, np.int32)
        eq = np.ascontiguousarray(end_query, np.int32)
        n = len(sel)
        ts = np.zeros(max(n, 1), np.int32)
        qs = np.zeros(max(n, 1), np.int32)
        self._check(self._L.ish_hip_find_starts(
            self._h, db._h, _p(sel if n else np.zeros(1, np.int64), ctypes.c_int64),
            _p(er if n else np.zeros(1, np.int32), ctypes.c_int32),
            _p(eq if n else np.zeros(1, np.int32), ctypes.c_int32), n,
            _p(ts, ctypes.c_int32), _p(qs, ctypes.c_int32)))
        return ts[:n], qs[:n]

    def align_db_multi(self, db: "Database", queries, matrix_kind: str, gap_open: int = 3,
                       gap_extend: int = 1, ends_free="FFTT", matrix=None):
        """Scores every target of db against every query (semi-global) in one
        persistent launch per cell width; returns [nq, n] arrays."""
        qs = [np.ascontiguousarray(q, np.uint8) for q in queries]
        qoff = np.zeros(len(qs) + 1, np.uint64)
        np.cumsum([len(q) for q in qs], out=qoff[1:])
        qbuf = np.concatenate(qs) if qoff[-1] else np.zeros(1, np.uint8)
        fl = flags_from_str(ends_free) if isinstance(ends_free, str) else int(ends_free)
        nq, n = len(qs), db.n
        out = [np.zeros(max(nq * n, 1), np.int32) for _ in range(3)]
        st = np.zeros(max(nq * n, 1), np.uint8)
        mp, mlen, _keep = _matrix_arg(matrix)
        self._check(self._L.ish_hip_align_db_multi(
            self._h, db._h, _p(qbuf, ctypes.c_uint8), _p(qoff, ctypes.c_uint64), nq, mp, mlen,
            MATRIX[matrix_kind], gap_open, gap_extend, fl, _p(out[0], ctypes.c_int32),
            _p(out[1], ctypes.c_int32), _p(out[2], ctypes.c_int32), _p(st, ctypes.c_uint8)))
        return tuple(a[: nq * n].reshape(nq, n) for a in (*out, st))

    def search_file(self, path, pattern: bytes, matrix_kind: str, algo="striped-semi-global",
                    gap_open: int = 3, gap_extend: int = 1, ends_free="FFTT", threshold: float = 0.8,
                    record_type: str = "line", out_fd: int = 1, batch_bytes: int = 0,
                    reader_threads: int = 0, reader_segment_bytes: int = 0, tty: bool = False):
        """Searches one file (plain or gzip) and writes the emitted records to
        out_fd in the reference runners' format; returns the run's stats."""
        pat = np.frombuffer(bytes(pattern), np.uint8).copy()
        pb = pat if len(pat) else np.zeros(1, np.uint8)
        fl = flags_from_str(ends_free) if isinstance(ends_free, str) else int(ends_free)
        cfg = SearchConfig(MATRIX[matrix_kind], ALGO[algo] if isinstance(algo, str) else int(algo), gap_open,
                           gap_extend, fl, threshold, RECORD[record_type], batch_bytes, reader_threads,
                           reader_segment_bytes, int(bool(tty)))
        st = SearchStats()
        self._check(self._L.ish_hip_search_file(self._h, os.fsencode(path), _p(pb, ctypes.c_uint8), len(pat),
                                                ctypes.byref(cfg), out_fd, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in SearchStats._fields_}

    def last_timing(self):
        ms = ctypes.c_float()
        k = ctypes.c_int32()
        self._check(self._L.ish_hip_last_timing(self._h, ctypes.byref(ms), ctypes.byref(k)))
        return float(ms.value), int(k.value)


class Database:
    def __init__(self, engine: Engine, handle, n: int):
        self._engine = engine
        self._h = handle
        self.n = n

    def close(self):
        if getattr(self, "_h", None):
            self._engine._L.ish_hip_db_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Server:
    """Serving loop over a device-resident database (semi-global): queries
    are submitted as they arrive and batched onto the GPU continuously.
    submit() returns a ticket; result() waits and returns the emitted
    targets (ascending target index) with scores and end coordinates."""

    def __init__(self, db: Database, matrix_kind: str, gap_open: int = 3, gap_extend: int = 1,
                 threshold: float = 0.8, max_batch: int = 16, slots: int = 2,
                 batch_cells: int = 0, hits_per_query: int = 0, keep_scores: bool = False,
                 matrix=None, aging: float = 0.0, batch_seconds: float = 0.0):
        self._L = lib()
        self._db = db  # keeps the database alive while the server runs
        mp, mlen, _keep = _matrix_arg(matrix)
        cfg = ServerConfig(MATRIX[matrix_kind], mp, mlen, gap_open, gap_extend, threshold,
                           max_batch, slots, batch_cells, hits_per_query, int(keep_scores),
                           aging, batch_seconds)
        h = ctypes.c_void_p()
        rc = self._L.ish_hip_server_create(db._h, ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise IshError(f"ish_hip_server_create failed: {ERRORS.get(rc, rc)}")
        self._h = h
        self.n = db.n

    @staticmethod
    def clock() -> float:
        return float(lib().ish_hip_server_clock())

    def _check(self, rc):
        if rc != 0:
            msg = self._L.ish_hip_server_last_error(self._h)
            raise IshError(f"{ERRORS.get(rc, rc)}: {msg.decode() if msg else ''}")

    def submit(self, query, ends_free="FFTT") -> int:
        q = np.ascontiguousarray(query, np.uint8)
        qb = q if len(q) else np.zeros(1, np.uint8)
        fl = flags_from_str(ends_free) if isinstance(ends_free, str) else int(ends_free)
        t = ctypes.c_int64()
        self._check(self._L.ish_hip_server_submit(self._h, _p(qb, ctypes.c_uint8), len(q), fl,
                                                  ctypes.byref(t)))
        return int(t.value)

    def wait(self, ticket: int, timeout_ms: int = -1):
        """Request info as a dict, or None on timeout."""
        info = RequestInfo()
        rc = self._L.ish_hip_server_wait(self._h, ticket, timeout_ms, ctypes.byref(info))
        if rc == ISH_E_TIMEOUT:
            return None
        self._check(rc)
        return {k: getattr(info, k) for k, _ in RequestInfo._fields_}

    def hits(self, ticket: int, n_hits: int):
        m = max(n_hits, 1)
        t = np.zeros(m, np.int64)
        a = [np.zeros(m, np.int32) for _ in range(3)]
        self._check(self._L.ish_hip_server_hits(self._h, ticket, _p(t, ctypes.c_int64),
                                                _p(a[0], ctypes.c_int32), _p(a[1], ctypes.c_int32),
                                                _p(a[2], ctypes.c_int32), n_hits))
        return t[:n_hits], a[0][:n_hits], a[1][:n_hits], a[2][:n_hits]

    def scores(self, ticket: int):
        n = max(self.n, 1)
        a = [np.zeros(n, np.int32) for _ in range(3)]
        self._check(self._L.ish_hip_server_scores(self._h, ticket, _p(a[0], ctypes.c_int32),
                                                  _p(a[1], ctypes.c_int32), _p(a[2], ctypes.c_int32)))
        return tuple(x[: self.n] for x in a)

    def release(self, ticket: int):
        self._check(self._L.ish_hip_server_release(self._h, ticket))

    def result(self, ticket: int, release: bool = True):
        """Waits for a request; returns (info, (target, score, end_ref, end_query))."""
        info = self.wait(ticket)
        hits = self.hits(ticket, info["n_hits"])
        if release:
            self.release(ticket)
        return info, hits

    def stats(self):
        st = ServerStats()
        self._check(self._L.ish_hip_server_get_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            self._L.ish_hip_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
