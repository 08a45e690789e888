"""Binary16 lanes of the semi-global lockstep tiles (ish_hip_set_f16).

A group runs in binary16 halves when the query's range bound keeps every
value of its pass within [-2040, 2040], where binary16 integer arithmetic is
exact; wider groups run in int16 in the same launch. The checks:
  - results with the lanes on equal the results with them off (int16 only)
    on every target, and the oracle on a sample;
  - the lanes ran (status flag) for targets within the bound and not for
    targets past it;
  - targets built to reach the top of the bound: tandem copies of a
    high-scoring query, at lengths just inside and just past the widest
    group the bound allows, where the pass's largest values lie within a
    few units of 2040.
"""
import os

import numpy as np
import pytest

import ish_amd
from ish_amd import Targets
from oracle import oracle as O

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)
GO, GE = 3, 1
# BLOSUM62 codes with a large diagonal: W 11, C 9, H 8, Y 7, P 7, N 6, D 6, F 6
HIGH = np.array([17, 4, 8, 18, 14, 2, 3, 13], np.uint8)


@pytest.fixture(scope="module")
def eng():
    if ish_amd.device_count() == 0:
        pytest.fail("no gfx950 device visible to the HIP library")
    e = ish_amd.Engine(0)
    yield e
    e.set_f16(True)
    e.close()


def widest_group(q, m):
    """Columns of the widest group the bound admits for query q (BLOSUM62,
    gap open 3, extend 1): sizes the targets of the edge cases."""
    S = m.values.reshape(m.A, m.A).astype(np.int64)
    P = int(np.maximum(S[q.astype(np.int64), :20].max(axis=1), 0).sum())
    L = -3 * GO - 2 * GE - 4 - 2
    return (L + 4080 - P) // GE - len(q) + 2


def edge_targets(q, w16, rng, n_edge=1200, n_beyond=128, n_rand=3792):
    """Tandem copies of q cut at lengths just inside the widest group
    (w16 rounded down to whole 4-column blocks), a group's worth past it,
    and random targets of the 20 standard residues below it."""
    top = w16 & ~3
    rep = np.tile(q, top // len(q) + 4)
    seqs = []
    for k in range(n_edge):
        L = int(top - rng.integers(0, 40))
        r = rep[:L].copy()
        if k % 3 == 1:  # a few substitutions
            pos = rng.integers(0, L, 5)
            r[pos] = rng.integers(0, 20, 5)
        seqs.append(r)
    for _ in range(n_beyond):
        L = int(top + 8 + rng.integers(0, 200))
        seqs.append(rep[:L].copy())
    for _ in range(n_rand):
        L = int(rng.integers(0, top))
        seqs.append(rng.integers(0, 20, L).astype(np.uint8))
    order = rng.permutation(len(seqs))
    return [seqs[i] for i in order], order


def run(eng, q, T, flags, f16, tile, schedule):
    eng.set_f16(f16)
    eng.set_engine("batched")
    eng.set_tile(tile)
    eng.set_schedule(schedule)
    eng.set_query(q, "blosum62", GO, GE, O.flags_from_str(flags))
    try:
        return eng.align(T)
    finally:
        eng.set_schedule("auto")
        eng.set_f16(True)


def same(label, a, b):
    for name, x, y in zip(("score", "end_ref", "end_query"), a, b):
        bad = np.flatnonzero(np.asarray(x) != np.asarray(y))
        assert len(bad) == 0, f"{label}: {name} differs on {len(bad)} targets, first {bad[0]}"


@pytest.mark.parametrize("tile,schedule", [(32, "wave"), (32, "auto"), (16, "wave"), (24, "wave"), (24, "auto")])
@pytest.mark.parametrize("flags", ["FFTT", "TTTT", "FFFF", "TFTF", "FTFT"])
def test_f16_equals_int16_at_the_range_edge(eng, flags, tile, schedule):
    rng = np.random.default_rng(7)
    q = HIGH[rng.integers(0, len(HIGH), 150)]
    m = O.Matrix("blosum62")
    w16 = widest_group(q, m)
    assert 1000 < w16 < 4000, w16
    seqs, order = edge_targets(q, w16, rng)
    T = Targets.from_list(seqs)
    on = run(eng, q, T, flags, True, tile, schedule)
    off = run(eng, q, T, flags, False, tile, schedule)
    same(f"{flags} binary16 vs int16", on[:3], off[:3])
    lens = np.diff(T.offsets).astype(np.int64)
    st = on[3]
    assert not (off[3] & ish_amd.STATUS_F16).any()
    # every target past the bound ran in int16; the tandem copies at the
    # edge ran in binary16
    assert not (st[lens > (w16 & ~3)] & ish_amd.STATUS_F16).any()
    if schedule == "wave":
        # (AUTO pipelines most groups of a database this small, in int16)
        edge = order < 1200
        assert (st[edge] & ish_amd.STATUS_F16).all(), (st[edge] & ish_amd.STATUS_F16).mean()
    # the oracle on the edge and past-the-bound targets and a sample of the rest
    idx = np.flatnonzero((order < 1200 + 128) | (np.arange(T.n) % 9 == 0))
    sub = [T.residues[int(T.offsets[i]):int(T.offsets[i + 1])] for i in idx]
    S = Targets.from_list(sub)
    ref = O.batch(0, q, S.residues, S.offsets, m, GO, GE, O.flags_from_str(flags),
                  nthreads=THREADS, striped=True)[:3]
    same(f"{flags} binary16 vs oracle", [a[idx] for a in on[:3]], ref)


def test_f16_off_when_most_groups_exceed_the_bound(eng):
    """A long high-scoring query leaves a small bound: when more than a few
    groups exceed it, the whole launch stays in int16."""
    rng = np.random.default_rng(3)
    q = HIGH[rng.integers(0, len(HIGH), 450)]
    m = O.Matrix("blosum62")
    assert widest_group(q, m) < 600
    seqs = [rng.integers(0, 20, int(L)).astype(np.uint8) for L in rng.integers(200, 1500, 3000)]
    T = Targets.from_list(seqs)
    for flags in ("FFTT", "TTTT"):
        on = run(eng, q, T, flags, True, 32, "auto")
        assert not (on[3] & ish_amd.STATUS_F16).any()
        ref = O.batch(0, q, T.residues, T.offsets, m, GO, GE, O.flags_from_str(flags),
                      nthreads=THREADS, striped=True)[:3]
        same(f"{flags} int16 vs oracle", on[:3], ref)


def test_f16_protein_workload_reuse(eng):
    """The protein workload's generator (log-normal lengths, background
    frequencies) through a resident database with several queries in turn:
    binary16 lanes on and off agree on every target and the lanes ran for
    almost all of them (one group per wave: AUTO would pipeline most groups
    of a database this small, in int16)."""
    from ish_amd import workloads
    w = workloads.generate(2, 60_000, seed=5)
    db = eng.load_db(w.targets)
    eng.set_engine("batched")
    eng.set_tile(32)
    eng.set_schedule("wave")
    try:
        for k, (qlen, flags) in enumerate(((350, "FFTT"), (120, "TTTT"), (300, "FTFT"))):
            q = workloads.query(2, 40 + k, qlen)
            eng.set_query(q, "blosum62", GO, GE, O.flags_from_str(flags))
            on = eng.align_db(db)
            eng.set_f16(False)
            off = eng.align_db(db)
            eng.set_f16(True)
            same(f"query {qlen} {flags}", on[:3], off[:3])
            assert (on[3] & ish_amd.STATUS_F16).mean() > 0.9, (qlen, (on[3] & ish_amd.STATUS_F16).mean())
            idx = np.arange(0, w.targets.n, 13)
            sub = [w.targets.residues[int(w.targets.offsets[i]):int(w.targets.offsets[i + 1])] for i in idx]
            S = Targets.from_list(sub)
            ref = O.batch(0, q, S.residues, S.offsets, O.Matrix("blosum62"), GO, GE,
                          O.flags_from_str(flags), nthreads=THREADS, striped=True)[:3]
            same(f"query {qlen} {flags} vs oracle", [a[idx] for a in on[:3]], ref)
    finally:
        eng.set_schedule("auto")
        eng.set_f16(True)
        db.close()


@pytest.mark.parametrize("flags", ["FFTT", "TTTT", "TFTF"])
def test_f16_multi_query_per_query_lanes(eng, flags):
    """One multi-query launch over the workgroup-shared paired profiles:
    each query whose bound admits every group of the database runs in
    binary16 lanes, the high-scoring long query in int16; every query's
    results equal those with the lanes off and the oracle's on a sample."""
    rng = np.random.default_rng(11)
    seqs = [rng.integers(0, 20, int(L)).astype(np.uint8) for L in rng.integers(0, 600, 6000)]
    for k in range(0, len(seqs), 50):  # tandem copies of a query: high scores
        seqs[k] = np.tile(HIGH, 80)[:len(seqs[k])].copy()
    T = Targets.from_list(seqs)
    queries = [rng.integers(0, 20, int(L)).astype(np.uint8) for L in (30, 150, 350, 500)]
    queries.append(HIGH[rng.integers(0, len(HIGH), 600)])
    m = O.Matrix("blosum62")
    widest = (int(np.diff(T.offsets).max()) + 3) & ~3
    expect = [widest_group(q, m) >= widest for q in queries]
    assert expect == [True, True, True, True, False], [widest_group(q, m) for q in queries]
    db = eng.load_db(T)
    try:
        eng.set_engine("batched")
        eng.set_tile(32)
        eng.set_f16(True)
        on = eng.align_db_multi(db, queries, "blosum62", GO, GE, flags)
        eng.set_f16(False)
        off = eng.align_db_multi(db, queries, "blosum62", GO, GE, flags)
    finally:
        eng.set_f16(True)
        db.close()
    idx = np.arange(0, T.n, 11)
    sub = Targets.from_list([T.residues[int(T.offsets[i]):int(T.offsets[i + 1])] for i in idx])
    for k, q in enumerate(queries):
        same(f"query {k} {flags} binary16 vs int16", [a[k] for a in on[:3]], [a[k] for a in off[:3]])
        assert not (off[3][k] & ish_amd.STATUS_F16).any()
        assert bool((on[3][k] & ish_amd.STATUS_F16).all()) == expect[k], k
        ref = O.batch(0, q, sub.residues, sub.offsets, m, GO, GE, O.flags_from_str(flags),
                      nthreads=THREADS, striped=True)[:3]
        same(f"query {k} {flags} vs oracle", [a[k][idx] for a in on[:3]], ref)


def test_f16_multi_query_nucleotide_codes(eng):
    """The multi-query workload's database holds only A, C, G, T and N codes
    scored with BLOSUM62: the bound over those codes admits queries of up to
    1,000 residues, which run in binary16 lanes."""
    from ish_amd import workloads
    w = workloads.generate(3, 30_000, seed=2)
    queries = [workloads.query(3, 300 + k, L) for k, L in enumerate((60, 400, 1000))]
    db = eng.load_db(w.targets)
    try:
        eng.set_engine("batched")
        eng.set_tile(32)
        on = eng.align_db_multi(db, queries, "blosum62", GO, GE, "FFTT")
        eng.set_f16(False)
        off = eng.align_db_multi(db, queries, "blosum62", GO, GE, "FFTT")
    finally:
        eng.set_f16(True)
        db.close()
    idx = np.arange(0, w.targets.n, 17)
    sub = Targets.from_list([w.targets.residues[int(w.targets.offsets[i]):int(w.targets.offsets[i + 1])]
                             for i in idx])
    for k, q in enumerate(queries):
        same(f"query {k} binary16 vs int16", [a[k] for a in on[:3]], [a[k] for a in off[:3]])
        assert (on[3][k] & ish_amd.STATUS_F16).all(), k
        ref = O.batch(0, q, sub.residues, sub.offsets, O.Matrix("blosum62"), GO, GE,
                      O.flags_from_str("FFTT"), nthreads=THREADS, striped=True)[:3]
        same(f"query {k} vs oracle", [a[k][idx] for a in on[:3]], ref)


def run_local(eng, q, T, kind, f16, tile=32, schedule="workgroup"):
    eng.set_f16(f16)
    eng.set_engine("batched")
    eng.set_tile(tile)
    eng.set_schedule(schedule)
    eng.set_query(q, kind, GO, GE, 0, "striped-local")
    try:
        return eng.align(T)
    finally:
        eng.set_schedule("auto")
        eng.set_f16(True)


@pytest.mark.parametrize("qlen,lanes", [(1020, True), (1021, False), (300, True)])
def test_f16_local_at_the_range_edge(eng, qlen, lanes):
    """Local alignment of reads in the workgroup pipeline: an actgn0 query
    of 1,020 nt scores at most 2,040, the largest plain value binary16
    lanes admit; reads holding exact copies of it among random flanks
    (several 256-column offset windows) reach that score. One more query
    residue leaves the bound and the pass stays in int16. Both lane types
    give the same results, equal to the oracle's."""
    rng = np.random.default_rng(qlen)
    q = rng.integers(0, 4, qlen).astype(np.uint8)
    seqs = []
    for k in range(1500):
        L = int(rng.integers(qlen + 50, 4000))
        r = rng.integers(0, 5, L).astype(np.uint8)
        if k % 3 == 0:  # an exact copy at a random offset
            o = int(rng.integers(0, L - qlen + 1))
            r[o:o + qlen] = q
        elif k % 3 == 1:  # a copy with a few substitutions and an N run
            o = int(rng.integers(0, L - qlen + 1))
            c = q.copy()
            c[rng.integers(0, qlen, 8)] = rng.integers(0, 5, 8)
            c[qlen // 2:qlen // 2 + 5] = 4
            r[o:o + qlen] = c
        seqs.append(r)
    T = Targets.from_list(seqs)
    for tile in (32, 16):
        on = run_local(eng, q, T, "actgn0", True, tile)
        off = run_local(eng, q, T, "actgn0", False, tile)
        same(f"local {qlen} tile {tile} binary16 vs int16", on[:3], off[:3])
        assert bool((on[3] & ish_amd.STATUS_F16).all()) == lanes, (on[3] & ish_amd.STATUS_F16).mean()
        assert not (off[3] & ish_amd.STATUS_F16).any()
    if qlen == 1020:
        assert on[0].max() == 2040
    idx = np.arange(0, T.n, 7)
    S = Targets.from_list([T.residues[int(T.offsets[i]):int(T.offsets[i + 1])] for i in idx])
    ref = O.batch(1, q, S.residues, S.offsets, O.Matrix("actgn0"), GO, GE, 0, nthreads=THREADS, striped=True)[:3]
    same(f"local {qlen} vs oracle", [a[idx] for a in on[:3]], ref)


def test_f16_local_long_read_workload(eng):
    """The long-read workload's generator (log-normal read lengths around
    8 kb, errors and N, planted query copies) through the automatic
    schedule: binary16 and int16 lanes agree on every read."""
    from ish_amd import workloads
    w = workloads.generate(4, 3000, seed=4)
    on = run_local(eng, w.query, w.targets, "actgn0", True, 32, "auto")
    off = run_local(eng, w.query, w.targets, "actgn0", False, 32, "auto")
    same("long reads binary16 vs int16", on[:3], off[:3])
    assert (on[3] & ish_amd.STATUS_F16).all()
    idx = np.arange(0, w.targets.n, 19)
    S = Targets.from_list([w.targets.residues[int(w.targets.offsets[i]):int(w.targets.offsets[i + 1])] for i in idx])
    ref = O.batch(1, w.query, S.residues, S.offsets, O.Matrix("actgn0"), GO, GE, 0, nthreads=THREADS,
                  striped=True)[:3]
    same("long reads vs oracle", [a[idx] for a in on[:3]], ref)


@pytest.mark.parametrize("go,ge,lanes", [(3, 0, True), (11, 1, True), (700, 1, False), (2, 3, False)])
def test_f16_gap_penalties(eng, go, ge, lanes):
    """Gap penalties at the bound's limits: a zero extension admits groups
    of any width (no column offset), a large open moves the lower bound
    past the binary16 range, and go < ge has no top padding rows; the
    last two stay in int16. Every case equals the int16 lanes and the
    oracle, on targets up to 5,000 residues."""
    rng = np.random.default_rng(go * 10 + ge)
    q = rng.integers(0, 20, 120).astype(np.uint8)
    lens = [int(x) for x in rng.integers(2000, 5000, 40)] + [int(x) for x in rng.integers(0, 400, 1500)]
    seqs = [rng.integers(0, 20, L).astype(np.uint8) for L in lens]
    for k in range(0, len(seqs), 4):
        r = seqs[k]
        if len(r) > 120:
            o = int(rng.integers(0, len(r) - 119))
            r[o:o + 120] = q
    T = Targets.from_list(seqs)
    m = O.Matrix("blosum62")
    for flags in ("FFTT", "TTTT", "FTFT"):
        fl = O.flags_from_str(flags)
        res = []
        for f16 in (True, False):
            eng.set_f16(f16)
            eng.set_engine("batched")
            eng.set_tile(32)
            eng.set_schedule("wave")
            eng.set_query(q, "blosum62", go, ge, fl)
            try:
                res.append(eng.align(T))
            finally:
                eng.set_schedule("auto")
                eng.set_f16(True)
        on, off = res
        same(f"go {go} ge {ge} {flags} binary16 vs int16", on[:3], off[:3])
        assert bool((on[3] & ish_amd.STATUS_F16).any()) == lanes
        # (with go <= ge the reference's striped answer depends on its SIMD
        # width; the engine's contract there is textbook Gotoh, the scalar
        # oracle)
        ref = O.batch(0, q, T.residues, T.offsets, m, go, ge, fl, nthreads=THREADS, striped=go > ge)[:3]
        same(f"go {go} ge {ge} {flags} vs oracle", on[:3], ref)
