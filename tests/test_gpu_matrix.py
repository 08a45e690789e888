"""Caller-supplied scoring matrices through the C ABI, passed the way the
Mojo binding passes them: GpuMatcher.matrix_bytes() and matrix_len(), i.e.
the flat int8 values of the kind's ScoringMatrix and their count A*A
(ishlib/matcher/__init__.mojo:79-86, striped_semi_global_matcher.mojo:
139-144).

For every matrix kind the results equal the built-in-table run and the CPU
oracle; a modified matrix is honoured (checked against the oracle with the
same values) by set_query, align_db_multi and the serving loop; a count
other than A*A is rejected.
"""
import numpy as np
import pytest

import ish_amd
from oracle import oracle as O

pytestmark = pytest.mark.gpu

KINDS = ["ascii", "actgn", "actgn0", "blosum62"]


@pytest.fixture(scope="module")
def eng():
    if ish_amd.device_count() == 0:
        pytest.fail("no gfx950 device visible to the HIP library")
    e = ish_amd.Engine(0)
    yield e
    e.close()


def _targets(kind, n, seed):
    """Random targets over codes the kind's encoding produces (ascii: a
    handful of printable bytes plus a few others)."""
    rng = np.random.default_rng(seed)
    codes = {"ascii": np.frombuffer(b"ACGTNacgtn-*xyz", np.uint8), "actgn": np.arange(5),
             "actgn0": np.arange(5), "blosum62": np.arange(24)}[kind]
    lens = rng.integers(0, 300, n)
    seqs = [rng.choice(codes, int(L)).astype(np.uint8) for L in lens]
    return ish_amd.Targets.from_list(seqs), rng.choice(codes, 90).astype(np.uint8)


def _oracle(algo, q, t, m, go, ge, flags):
    s, er, eq, _ = O.batch(algo, q, t.residues, t.offsets, m, go, ge, flags)
    return s, er, eq


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("algo", ["striped-semi-global", "striped-local"])
def test_matrix_as_the_binding_passes_it(eng, kind, algo):
    t, q = _targets(kind, 1500, 7 + KINDS.index(kind))
    flags = "FFTT" if algo == "striped-semi-global" else "FFFF"
    m = ish_amd.matrix(kind)
    A = m.shape[0]
    assert m.size == A * A and m.size == {"ascii": 65536, "actgn": 25, "actgn0": 25, "blosum62": 576}[kind]
    eng.set_query(q, kind, 3, 1, flags, algo, matrix=m.reshape(-1))
    got = eng.align(t)[:3]
    eng.set_query(q, kind, 3, 1, flags, algo)  # built-in table
    builtin = eng.align(t)[:3]
    om = O.Matrix(kind)
    ref = _oracle(1 if algo == "striped-local" else 0, q, t, om, 3, 1, O.flags_from_str(flags))
    for a, b, c in zip(got, builtin, ref):
        assert np.array_equal(a, b)
        assert np.array_equal(a, c)


def _modified_blosum62():
    m = ish_amd.matrix("blosum62").astype(np.int16)
    m = np.clip(m * 2 - 1, -20, 20)  # a different matrix with the same shape
    np.fill_diagonal(m, np.clip(np.diag(m) + 3, -20, 24))
    return m.astype(np.int8)


def test_modified_matrix_is_used(eng):
    t, q = _targets("blosum62", 2000, 21)
    mm = _modified_blosum62()
    om = O.Matrix("blosum62")
    om.values = mm.reshape(-1).copy()
    got = {}
    for flags in ("FFTT", "TTTT", "FFFF"):
        eng.set_query(q, "blosum62", 4, 2, flags, matrix=mm)
        s, er, eq, _ = eng.align(t)
        got[flags] = s
        rs, rer, req = _oracle(0, q, t, om, 4, 2, O.flags_from_str(flags))
        assert np.array_equal(s, rs) and np.array_equal(er, rer) and np.array_equal(eq, req)
    # the built-in matrix gives other scores under the same settings
    eng.set_query(q, "blosum62", 4, 2, "FFTT")
    assert not np.array_equal(eng.align(t)[0], got["FFTT"])


def test_modified_matrix_multi_query(eng):
    t, _ = _targets("blosum62", 1200, 22)
    rng = np.random.default_rng(5)
    qs = [rng.integers(0, 24, int(L)).astype(np.uint8) for L in (30, 200, 450)]
    mm = _modified_blosum62()
    om = O.Matrix("blosum62")
    om.values = mm.reshape(-1).copy()
    db = eng.load_db(t)
    try:
        s, er, eq, _ = eng.align_db_multi(db, qs, "blosum62", 3, 1, "FFTT", matrix=mm)
        for k, q in enumerate(qs):
            rs, rer, req = _oracle(0, q, t, om, 3, 1, O.flags_from_str("FFTT"))
            assert np.array_equal(s[k], rs) and np.array_equal(er[k], rer) and np.array_equal(eq[k], req)
    finally:
        db.close()


def test_modified_matrix_serving(eng):
    t, _ = _targets("blosum62", 1200, 23)
    rng = np.random.default_rng(6)
    qs = [rng.integers(0, 24, int(L)).astype(np.uint8) for L in (25, 180)]
    mm = _modified_blosum62()
    om = O.Matrix("blosum62")
    om.values = mm.reshape(-1).copy()
    db = eng.load_db(t)
    srv = ish_amd.Server(db, "blosum62", 3, 1, 0.8, keep_scores=True, matrix=mm)
    try:
        tickets = [srv.submit(q, "TTTT") for q in qs]
        for q, tk in zip(qs, tickets):
            srv.wait(tk)
            s, er, eq = srv.scores(tk)
            rs, rer, req = _oracle(0, q, t, om, 3, 1, O.flags_from_str("TTTT"))
            assert np.array_equal(s, rs) and np.array_equal(er, rer) and np.array_equal(eq, req)
            srv.release(tk)
    finally:
        srv.close()
        db.close()


@pytest.mark.parametrize("kind,count", [("blosum62", 575), ("blosum62", 24), ("actgn", 5),
                                        ("actgn0", 0), ("ascii", 256)])
def test_wrong_matrix_len_is_rejected(eng, kind, count):
    q = np.zeros(10, np.uint8)
    flat = np.zeros(max(count, 1), np.int8)[:count]
    with pytest.raises(ish_amd.IshError, match="ISH_E_INVALID.*matrix_len"):
        eng.set_query(q, kind, 3, 1, "FFTT", matrix=flat)
