"""ish_hip_align_db_noalloc through the C ABI: the database pass whose
kernels keep their scratch in the library's static device array and whose
results go to caller-owned device arrays.

Each call is checked bit-exact (score, end_ref, end_query, status on every
target; the status without ish_hip_align_db's binary16 flag, since this pass
always runs int16 lanes) against ish_hip_align_db on the same context, on two
calls in a row
whose second database is smaller, for semi-global and local queries; the
second call must leave the device's free memory unchanged; a target longer
than the array can hold boundary rows for is rejected before any launch.
"""
import ctypes
import os

import numpy as np
import pytest

import ish_amd
from ish_amd import workloads

pytestmark = pytest.mark.gpu

_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        _HIP = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    return _HIP


class DeviceOut:
    """score / end_ref / end_query / status arrays in device memory the
    test allocates (hipMalloc), read back with hipMemcpy."""

    def __init__(self, n):
        self.n = n
        self.ptrs = []
        for nbytes in (4 * n, 4 * n, 4 * n, n):
            p = ctypes.c_void_p()
            assert _hip().hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(nbytes, 4))) == 0
            self.ptrs.append(p.value)

    def read(self):
        out = []
        for p, dt in zip(self.ptrs, (np.int32, np.int32, np.int32, np.uint8)):
            a = np.zeros(max(self.n, 1), dt)
            assert _hip().hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(p),
                                    ctypes.c_size_t(a.itemsize * self.n), 2) == 0  # device to host
            out.append(a[: self.n])
        return out

    def free(self):
        for p in self.ptrs:
            _hip().hipFree(ctypes.c_void_p(p))
        self.ptrs = []


def _free_bytes():
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert _hip().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


@pytest.fixture(scope="module")
def eng():
    if ish_amd.device_count() == 0:
        pytest.fail("no gfx950 device visible to the HIP library")
    e = ish_amd.Engine(0)
    yield e
    e.close()


def _check_pair(eng, db, n, check_memory=False):
    ref = eng.align_db(db)
    out = DeviceOut(n)
    try:
        before = _free_bytes()
        eng.align_db_noalloc(db, *out.ptrs)
        after = _free_bytes()
        got = out.read()
    finally:
        out.free()
    keep = 0xFF ^ ish_amd.STATUS_F16
    ref = (*ref[:3], ref[3] & keep)
    for name, a, b in zip(("score", "end_ref", "end_query", "status"), got, ref):
        assert np.array_equal(a, b), f"{name} differs on {int((a != b).sum())} of {n} targets"
    if check_memory:
        assert after == before, f"free device memory changed by {before - after} bytes"


@pytest.mark.parametrize("case", ["semi-global", "local"])
def test_noalloc_matches_align_db_two_calls(eng, case):
    if case == "semi-global":
        big = workloads.generate(2, 20000, seed=5)
        small = workloads.generate(2, 3000, seed=6)
        eng.set_query(big.query, "blosum62", 3, 1, "FFTT")
    else:
        big = workloads.generate(4, 60, seed=5)
        small = workloads.generate(4, 15, seed=6)
        eng.set_query(big.query, "actgn0", 3, 1, "FFFF", "striped-local")
    print(f"noalloc {case}: {big.targets.n} then {small.targets.n} targets", flush=True)
    for w, mem in ((big, False), (small, True)):
        db = eng.load_db(w.targets)
        try:
            # a first call loads the kernels' code objects; the second is
            # measured for device allocations
            _check_pair(eng, db, w.targets.n)
            _check_pair(eng, db, w.targets.n, check_memory=mem)
        finally:
            db.close()


def test_noalloc_free_ends_and_int32_groups(eng):
    """Every free-end setting, with targets long enough that some groups run
    in int32 lanes (status 0x80|32), on one database."""
    rng = np.random.default_rng(17)
    lens = np.concatenate([rng.integers(0, 400, 700), [3000, 2500, 0, 1]]).astype(np.int64)
    seqs = [rng.integers(0, 24, int(L)).astype(np.uint8) for L in lens]
    t = ish_amd.Targets.from_list(seqs)
    q = rng.integers(0, 24, 1500).astype(np.uint8)
    db = eng.load_db(t)
    try:
        for f in range(16):
            eng.set_query(q, "blosum62", 5, 2, f)
            _check_pair(eng, db, t.n)
    finally:
        db.close()


def test_noalloc_rejects_targets_beyond_the_scratch(eng):
    rng = np.random.default_rng(3)
    t = ish_amd.Targets.from_list([rng.integers(0, 4, 270000).astype(np.uint8),
                                   rng.integers(0, 4, 50).astype(np.uint8)])
    eng.set_query(rng.integers(0, 4, 40).astype(np.uint8), "actgn", 3, 1, "FFTT")
    db = eng.load_db(t)
    out = DeviceOut(t.n)
    try:
        with pytest.raises(ish_amd.IshError, match="ISH_E_RANGE"):
            eng.align_db_noalloc(db, *out.ptrs)
    finally:
        out.free()
        db.close()
    # the context and the scratch still serve a database that fits
    small = ish_amd.Targets.from_list([rng.integers(0, 4, int(L)).astype(np.uint8) for L in (0, 7, 300, 41)])
    db = eng.load_db(small)
    try:
        _check_pair(eng, db, small.n)
    finally:
        db.close()
