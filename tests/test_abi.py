"""CPU-side checks of the C ABI library: it loads, exports every symbol the
public header declares, and its host-only functions (tables, encoding,
emission rule) agree with the oracle. No GPU compute is called here."""
import ctypes
import os
import re

import numpy as np
import pytest

import ish_amd
from ish_amd import workloads
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "ish_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ish_hip_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    L = ish_amd.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(ish_amd.ABI_SYMBOLS) == syms


def test_version():
    assert ish_amd.lib().ish_hip_version() == 10000


@pytest.mark.parametrize("kind", ["ascii", "actgn", "actgn0", "blosum62"])
def test_tables_match_oracle(kind):
    m = O.Matrix(kind)
    assert np.array_equal(ish_amd.matrix(kind).ravel(), m.values)
    assert np.array_equal(ish_amd.encoding(kind), m.enc)
    assert np.array_equal(ish_amd.decoding(kind), m.dec)
    text = bytes(range(256)) * 3
    enc, score = ish_amd.encode_and_score(kind, text)
    assert np.array_equal(enc, m.enc[np.frombuffer(text, np.uint8)])
    assert score == m.max_score(text)


def test_emit_matches_oracle():
    rng = np.random.default_rng(5)
    n = 4000
    score = rng.integers(-50, 400, n).astype(np.int32)
    er = rng.integers(-1, 5, n).astype(np.int32)
    eq = rng.integers(-1, 5, n).astype(np.int32)
    for ms in (1, 3, 299, 300, 301, 1000):
        for thr in (0.0, 0.5, 0.8, 0.95, 1.0):
            sg = ish_amd.emit("striped-semi-global", score, er, eq, ms, thr)
            lo = ish_amd.emit("striped-local", score, er, eq, ms, thr)
            for t in range(0, n, 7):
                assert sg[t] == O.sg_emit(score[t], ms, thr)
                assert lo[t] == O.sw_emit(score[t], er[t], eq[t], ms, thr)


def test_no_device_is_reported_not_faked():
    if ish_amd.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(ish_amd.IshError):
        ish_amd.Engine(0)


def test_workload_determinism():
    a = workloads.generate(2, 3000, seed=11, threads=1)
    b = workloads.generate(2, 3000, seed=11, threads=8)
    assert a.content_hash() == b.content_hash()
    assert np.array_equal(a.planted, b.planted)
    c = workloads.generate(2, 3000, seed=12)
    assert c.content_hash() != a.content_hash()
    w1 = workloads.generate(1, 20000, seed=3)
    lens = np.diff(w1.targets.offsets)
    assert lens.min() >= 100 and lens.max() <= 300
    assert w1.targets.residues.max() <= 4
    w2 = workloads.generate(2, 20000, seed=3)
    lens = np.diff(w2.targets.offsets)
    assert lens.min() >= 30 and lens.max() <= 5000
    assert 300 < np.median(lens) < 400
    assert w2.targets.residues.max() < 20
