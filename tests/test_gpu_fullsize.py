"""Parity at the full sizes of the benchmark workloads: the targets and
queries that the bench lines run. Each check uses a property that does not
shrink with the size: two independent GPU paths equal on every target, and
the oracle on every target or on a large even sample.

- 200k nucleotide line records, 150 nt query, actgn, FFTT.
  - Every target's score and end coordinates against the oracle.
  - Tiles of 40 and 16 rows equal to tile 32 on every target.
  - The plain int32 engine equal to the batched one.
  - The emitted set at 0.8 equal to the oracle's emission rule.
- 2M protein targets, 350 aa query, blosum62, FFTT.
  - The batched launch (packed int16, hybrid) equal to the plain int32
    engine on every target.
  - Every target against the oracle's striped restatement.
- 64 protein queries x 4M records.
  - The multi-query launch equal, on every target, to single-query passes
    for every one of the 64 queries.
  - Each query against the oracle's striped restatement on an even sample.
- 500k long reads, 1,000 nt query, local.
  - The workgroup pipeline over the paired profile equal to the one-wave
    kernel over the int16 profile on every read.
  - Both against the oracle on every 25th read.
"""
import os

import numpy as np
import pytest

import ish_amd
from ish_amd import Targets, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def eng():
    if ish_amd.device_count() == 0:
        pytest.fail("no gfx950 device visible to the HIP library")
    e = ish_amd.Engine(0)
    yield e
    e.close()


def _sample(t: Targets, idx):
    lens = (t.offsets[idx + 1] - t.offsets[idx]).astype(np.int64)
    off = np.zeros(len(idx) + 1, np.uint64)
    np.cumsum(lens, out=off[1:])
    res = np.concatenate([t.residues[int(t.offsets[i]):int(t.offsets[i + 1])] for i in idx])
    return res, off


def _same(label, a, b):
    for name, x, y in zip(("score", "end_ref", "end_query"), a, b):
        bad = np.flatnonzero(np.asarray(x) != np.asarray(y))
        assert len(bad) == 0, f"{label}: {name} differs on {len(bad)} targets, first {bad[0]}: {x[bad[0]]} vs {y[bad[0]]}"


def _scores(eng, db):
    eng.set_engine("batched")
    eng.set_schedule("auto")
    s, er, eq, _ = eng.align_db(db)
    return s, er, eq


def test_nucleotide_200k_full(eng):
    w = workloads.generate(1, 200_000, seed=1)
    flags = O.flags_from_str("FFTT")
    eng.set_query(w.query, "actgn", 3, 1, "FFTT")
    db = eng.load_db(w.targets)
    try:
        got = _scores(eng, db)
        for tile in (40, 16):  # four 40-row chunks: one round of the four-wave pipeline
            eng.set_tile(tile)
            _same(f"tile {tile} vs tile 32", _scores(eng, db), got)
    finally:
        eng.set_tile(32)
        db.close()
    eng.set_engine("plain")
    try:
        plain = eng.align(w.targets)[:3]
    finally:
        eng.set_engine("batched")
    _same("batched vs plain", got, plain)
    m = O.Matrix("actgn")
    ref = O.batch(0, w.query, w.targets.residues, w.targets.offsets, m, 3, 1, flags, nthreads=THREADS)[:3]
    _same("batched vs oracle", got, ref)
    ms = int(m.values.reshape(m.A, m.A)[w.query, w.query].sum())
    emit = ish_amd.emit("striped-semi-global", *got, ms, 0.8)
    keys, inv = np.unique(ref[0], return_inverse=True)
    emit_ref = np.array([O.sg_emit(int(k), ms, 0.8) for k in keys], bool)[inv]
    assert np.array_equal(emit, emit_ref)
    assert emit.sum() > 0  # the planted copies are found


def test_protein_2M_full(eng):
    w = workloads.generate(2, 2_000_000, seed=1)
    eng.set_query(w.query, "blosum62", 3, 1, "FFTT")
    db = eng.load_db(w.targets)
    try:
        got = _scores(eng, db)
    finally:
        db.close()
    eng.set_engine("plain")
    try:
        plain = eng.align(w.targets)[:3]
    finally:
        eng.set_engine("batched")
    _same("batched vs plain", got, plain)
    # the one-shot call cuts this 0.84 GB batch into two automatic slices
    _same("one-shot call vs resident pass", eng.align(w.targets)[:3], got)
    ref = O.batch(0, w.query, w.targets.residues, w.targets.offsets, O.Matrix("blosum62"), 3, 1,
                  O.flags_from_str("FFTT"), nthreads=THREADS, striped=True)[:3]
    _same("batched vs oracle", got, ref)


def test_protein_10M_one_shot_slices(eng):
    """A 4.19 GB protein batch (10 M targets, near ish's 4 GiB default
    batch) through one-shot calls: six automatic slices, five fixed 860 MiB
    slices and no slicing give the same results; a sample of 20 k targets
    matches the oracle."""
    w = workloads.generate(2, 10_000_000, seed=1)
    eng.set_engine("batched")
    eng.set_query(w.query, "blosum62", 3, 1, "FFTT")
    try:
        auto = eng.align(w.targets)[:3]
        eng.set_slice_bytes(860 << 20)
        _same("860 MiB slices vs automatic", eng.align(w.targets)[:3], auto)
        eng.set_slice_bytes(-1)
        _same("unsliced vs automatic", eng.align(w.targets)[:3], auto)
    finally:
        eng.set_slice_bytes(0)
    idx = np.linspace(0, w.targets.n - 1, 20_000).astype(np.int64)
    res, off = _sample(w.targets, idx)
    ref = O.batch(0, w.query, res, off, O.Matrix("blosum62"), 3, 1, O.flags_from_str("FFTT"),
                  nthreads=THREADS, striped=True)[:3]
    _same("automatic slices vs oracle on a sample", [a[idx] for a in auto], ref)


def test_multiquery_64x4M_full(eng):
    w = workloads.generate(3, 4_000_000, seed=1)
    rng = np.random.default_rng(1)
    qlens = rng.integers(50, 1001, 64)
    queries = [workloads.query(3, 1000 + k, int(L)) for k, L in enumerate(qlens)]
    db = eng.load_db(w.targets)
    try:
        ms, mer, meq, _ = eng.align_db_multi(db, queries, "blosum62", 3, 1, "FFTT")
        idx = np.linspace(0, w.targets.n - 1, 20_000).astype(np.int64)
        res, off = _sample(w.targets, idx)
        for k in range(64):
            eng.set_query(queries[k], "blosum62", 3, 1, "FFTT")
            one = _scores(eng, db)
            _same(f"multi vs single, query {k}", (ms[k], mer[k], meq[k]), one)
            ref = O.batch(0, queries[k], res, off, O.Matrix("blosum62"), 3, 1, O.flags_from_str("FFTT"),
                          nthreads=THREADS, striped=True)[:3]
            _same(f"query {k} vs oracle", tuple(a[idx] for a in one), ref)
    finally:
        db.close()


def test_longread_500k_full(eng):
    w = workloads.generate(4, 500_000, seed=1)
    eng.set_query(w.query, "actgn0", 3, 1, "FFFF", "striped-local")
    db = eng.load_db(w.targets)
    try:
        eng.set_tile(32)
        pipe = _scores(eng, db)
        eng.set_schedule("wave")
        wave = eng.align_db(db)[:3]
    finally:
        eng.set_schedule("auto")
        db.close()
    _same("workgroup pipeline vs one wave per group", pipe, wave)
    idx = np.arange(0, w.targets.n, 25, dtype=np.int64)
    res, off = _sample(w.targets, idx)
    ref = O.batch(1, w.query, res, off, O.Matrix("actgn0"), 3, 1, 0, nthreads=THREADS, striped=True)[:3]
    _same("pipeline vs oracle", tuple(a[idx] for a in pipe), ref)


def test_file_runner_protein_2M_default_batches(eng, tmp_path):
    """The 2M protein targets as a 0.89 GB FASTA file through the file runner
    with its default batching (batches grow with the longest record, each
    loaded while the one before is scored): the output equals, byte for
    byte, the records the in-memory pass emits, in input order."""
    w = workloads.generate(2, 2_000_000, seed=1)
    path = tmp_path / "targets.fa"
    workloads.write_fastx(str(path), w.targets, "blosum62", fastq=False, width=60)
    dec = np.asarray(ish_amd.decoding("blosum62"), np.uint8)
    pattern = bytes(dec[w.query])
    out = tmp_path / "out.fa"
    fd = os.open(str(out), os.O_WRONLY | os.O_CREAT | os.O_TRUNC)
    try:
        st = eng.search_file(str(path), pattern, "blosum62", "striped-semi-global", 3, 1, "FFTT", 0.25,
                             "fastx", fd)
    finally:
        os.close(fd)
    assert st["batches"] >= 2 and st["records"] == w.targets.n
    eng.set_query(w.query, "blosum62", 3, 1, "FFTT")
    db = eng.load_db(w.targets)
    s, er, eq = _scores(eng, db)
    db.close()
    ms = int(ish_amd.matrix("blosum62")[w.query, w.query].sum())
    emit = ish_amd.emit("striped-semi-global", s, er, eq, ms, 0.25)
    assert emit.sum() > 100
    off = w.targets.offsets
    expected = b"".join(b">t%d synthetic\n" % t + bytes(dec[w.targets.residues[int(off[t]):int(off[t + 1])]]) + b"\n"
                        for t in np.flatnonzero(emit))
    assert out.read_bytes() == expected
