"""Parity of the HIP engine with the CPU oracle, through the C ABI.

Bit-exact on score, end_ref and end_query for every target; identical
emitted record sets at --score 0.8. Both engines are checked: the batched
persistent kernel (packed int16, int32 for wide groups) and the plain
one-launch-per-batch int32 kernel.
"""
import json
import os
import zlib

import numpy as np
import pytest

import ish_amd
from ish_amd import Targets, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
VEC = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))
ENGINES = ["batched", "plain"]


@pytest.fixture(scope="module")
def eng():
    if ish_amd.device_count() == 0:
        pytest.fail("no gfx950 device visible to the HIP library")
    e = ish_amd.Engine(0)
    yield e
    e.close()


def oracle_batch(q, targets, kind, go, ge, flags):
    m = O.Matrix(kind)
    s, er, eq, _ = O.batch(0, q, targets.residues, targets.offsets, m, go, ge, flags)
    return s, er, eq


def check(eng, engine, q, targets, kind, go, ge, flags, label="", tile=32, schedule="auto"):
    eng.set_engine(engine)
    eng.set_tile(tile)
    eng.set_schedule(schedule)
    eng.set_query(q, kind, go, ge, flags, "striped-semi-global")
    try:
        s, er, eq, st = eng.align(targets)
    finally:
        eng.set_schedule("auto")
    os_, oer, oeq = oracle_batch(q, targets, kind, go, ge, flags)
    bad = np.nonzero((s != os_) | (er != oer) | (eq != oeq))[0]
    if len(bad):
        t = bad[0]
        raise AssertionError(
            f"{label} engine={engine} schedule={schedule} {len(bad)}/{targets.n} differ; first t={t} "
            f"R={targets.offsets[t + 1] - targets.offsets[t]} gpu=({s[t]},{er[t]},{eq[t]}) "
            f"oracle=({os_[t]},{oer[t]},{oeq[t]}) status={st[t]}")
    return s, er, eq, st


@pytest.mark.parametrize("engine", ENGINES)
def test_golden_vectors(eng, engine):
    eng.set_engine(engine)
    for c in VEC["semi_global"]:
        m = O.Matrix(c["matrix"])
        q, t = m.encode(c["query"]), m.encode(c["target"])
        eng.set_query(q, c["matrix"], c["go"], c["ge"], c["flags"])
        s, er, eq, _ = eng.align(Targets.from_list([t]))
        assert s[0] == c["score"], c
        if "end_ref" in c:
            assert er[0] == c["end_ref"], c
        if "end_query" in c:
            assert eq[0] == c["end_query"], c
        assert (s[0], er[0], eq[0]) == O.sg(q, t, m, c["go"], c["ge"], O.flags_from_str(c["flags"]))


@pytest.mark.parametrize("engine,tile,schedule", [("batched", 32, "auto"), ("batched", 16, "auto"),
                                                  ("batched", 32, "wave"), ("batched", 32, "workgroup"),
                                                  ("batched", 16, "workgroup"), ("batched", 40, "auto"),
                                                  ("batched", 40, "wave"), ("batched", 40, "workgroup"),
                                                  ("plain", 32, "auto")])
@pytest.mark.parametrize("kind", ["actgn", "actgn0", "blosum62", "ascii"])
def test_random_all_flags(eng, engine, tile, kind, schedule):
    rng = np.random.default_rng(zlib.crc32(f"{engine}-{kind}".encode()))
    alpha = {"actgn": 5, "actgn0": 5, "blosum62": 24, "ascii": 256}[kind]
    for flags in range(16):
        Q = int(rng.integers(1, 140))
        q = rng.integers(0, alpha if kind != "ascii" else 60, Q).astype(np.uint8) + (0 if kind != "ascii" else 40)
        seqs = []
        for t in range(300):
            R = int(rng.integers(0, 260))
            r = rng.integers(0, alpha, R).astype(np.uint8)
            if kind == "ascii":
                r = (rng.integers(0, 60, R) + 40).astype(np.uint8)
            if R and t % 5 == 0:  # plant part of the query
                k = min(Q, R)
                off = int(rng.integers(0, R - k + 1))
                r[off:off + k] = q[:k]
            seqs.append(r)
        go = int(rng.integers(1, 12))
        ge = int(rng.integers(0, go))
        check(eng, engine, q, Targets.from_list(seqs), kind, go, ge, flags, f"flags={flags}", tile, schedule)


@pytest.mark.parametrize("schedule", ["auto", "wave", "workgroup"])
@pytest.mark.parametrize("tile", [32, 24, 16])
@pytest.mark.parametrize("ncodes", [9, 14, 20, 21, 22])
def test_shared_paired_profile_class_counts(eng, tile, schedule, ncodes):
    """Databases whose codes fall into 9..21 profile classes run over the
    workgroup-shared paired profile (tiles of four groups in lockstep, long
    groups pipelined over the int16 profile in the same launch); 22 classes
    fall back to the int16 profile. A BLOSUM62 query holding every residue
    gives each database code its own class."""
    rng = np.random.default_rng(ncodes * 100 + tile)
    q = np.concatenate([np.arange(24), rng.integers(0, 24, 176)]).astype(np.uint8)
    rng.shuffle(q)
    lens = [int(x) for x in rng.integers(3000, 6000, 24)] + [int(x) for x in rng.integers(0, 400, 700)]
    seqs = [rng.integers(0, ncodes, L).astype(np.uint8) for L in lens]
    for k in range(0, len(seqs), 9):  # plant part of the query (its codes below ncodes)
        r = seqs[k]
        if len(r) > 60:
            piece = q[q < ncodes][:60]
            o = int(rng.integers(0, len(r) - len(piece) + 1))
            r[o:o + len(piece)] = piece
    T = Targets.from_list(seqs)
    for flags in ("FFTT", "TTTT", "FFFF", "TFTF"):
        check(eng, "batched", q, T, "blosum62", 3, 1, O.flags_from_str(flags), f"codes={ncodes} {flags}",
              tile, schedule)


@pytest.mark.parametrize("flags", ["FFTT", "TTTT", "FFFF", "TFTF", "FTFT"])
def test_hybrid_long_groups_among_short(eng, flags):
    """A few very long targets among many short ones: the automatic
    schedule scores the long groups with whole workgroups and the rest one
    group per wave, in one launch; workgroups take several items."""
    rng = np.random.default_rng(zlib.crc32(flags.encode()))
    q = rng.integers(0, 24, 200).astype(np.uint8)
    seqs = [rng.integers(0, 24, int(rng.integers(0, 120))).astype(np.uint8) for _ in range(60000)]
    for k in range(0, 60000, 211):
        seqs[k] = rng.integers(0, 24, int(rng.integers(2000, 6000))).astype(np.uint8)
        if k % 3 == 0:
            seqs[k][100:300] = q
    fl = O.flags_from_str(flags)
    check(eng, "batched", q, Targets.from_list(seqs), "blosum62", 3, 1, fl, "hybrid", 32, "auto")
    check(eng, "batched", q, Targets.from_list(seqs[:3000]), "blosum62", 3, 1, fl, "all pipelined", 16,
          "workgroup")


@pytest.mark.parametrize("tile", [16, 32, 40])
def test_two_wave_pipelines(eng, tile):
    """Chunk counts below 8 that are 1 or 2 modulo 4 run pipelined groups
    over two-wave workgroups (the 40-row tile over four), 9 and 10 over four
    (the workgroup schedule pipelines every group); one or two chunks stream
    the groups through the pipeline without draining (nucleotide codes and
    no query-end rule). Each must equal the oracle under every end-free
    setting the kernel distinguishes."""
    rng = np.random.default_rng(31 + tile)
    seqs = [rng.integers(0, 5, int(rng.integers(1, 700))).astype(np.uint8) for _ in range(700)]
    targets = Targets.from_list(seqs)
    for nch in (1, 2, 5, 6, 9, 10):
        Q = max(1, nch * tile - int(rng.integers(0, tile)))
        q = rng.integers(0, 5, Q).astype(np.uint8)
        for flags in ("FFFF", "FFTT", "TTFF", "FTFT", "TFTF", "TTTT"):
            check(eng, "batched", q, targets, "actgn", 3, 1, O.flags_from_str(flags),
                  f"nch={nch} Q={Q} {flags}", tile, "workgroup")


@pytest.mark.parametrize("engine", ENGINES)
def test_gap_open_not_above_extend(eng, engine):
    """go <= ge: the engine computes textbook Gotoh (the oracle)."""
    rng = np.random.default_rng(9)
    q = rng.integers(0, 5, 77).astype(np.uint8)
    seqs = [rng.integers(0, 5, int(rng.integers(1, 200))).astype(np.uint8) for _ in range(400)]
    for go, ge in ((1, 1), (2, 3), (0, 2), (0, 0)):
        for flags in (0, 5, 10, 15):
            check(eng, engine, q, Targets.from_list(seqs), "actgn", go, ge, flags, f"go={go} ge={ge}")


@pytest.mark.parametrize("tile", [16, 32, 40])
@pytest.mark.parametrize("schedule", ["wave", "workgroup"])
def test_top_padding_every_row_offset(eng, tile, schedule):
    """With go >= ge the kernels put padding rows above query row 0 so that
    the query fills whole chunks (0 .. tile-1 of them). Query lengths at
    every offset modulo the tile height, and a few multi-chunk ones, rotate
    through target start free and not (the padding rows' two forms), go == ge
    and go > ge, the other free ends, and nucleotide and protein codes (per
    wave and workgroup-shared paired profiles): every result equals the
    oracle."""
    rng = np.random.default_rng(zlib.crc32(f"toppad-{tile}-{schedule}".encode()))
    nuc = Targets.from_list([rng.integers(0, 5, int(rng.integers(0, 260))).astype(np.uint8) for _ in range(300)])
    prot = Targets.from_list([rng.integers(0, 24, int(rng.integers(0, 260))).astype(np.uint8) for _ in range(300)])
    combos = [(ts, gaps, fl, kind) for ts in (True, False) for gaps in ((3, 1), (2, 2))
              for fl in ("TT", "FF", "TF", "FT") for kind in ("actgn", "blosum62")]
    lengths = list(range(1, tile + 2)) + [3 * tile - 5, 3 * tile, 3 * tile + 3, 5 * tile + 7]
    for n, Q in enumerate(lengths):
        ts, (go, ge), fl, kind = combos[(n * 5) % len(combos)]
        # flags string: qs qe ts te
        flags = fl[0] + fl[1][0] + ("T" if ts else "F") + fl[1][0]
        if n % 3 == 1:
            flags = flags[:3] + ("F" if flags[3] == "T" else "T")
        A = 5 if kind == "actgn" else 24
        q = rng.integers(0, A, Q).astype(np.uint8)
        check(eng, "batched", q, nuc if kind == "actgn" else prot, kind, go, ge, O.flags_from_str(flags),
              f"Q={Q} {flags} go={go} ge={ge} {kind}", tile, schedule)


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("n", [1, 63, 127, 128, 129, 1000])
def test_batch_sizes_and_empty_targets(eng, engine, n):
    rng = np.random.default_rng(n)
    q = rng.integers(0, 24, 61).astype(np.uint8)
    seqs = [rng.integers(0, 24, int(rng.integers(0, 9)) if t % 3 == 0 else int(rng.integers(0, 400))).astype(np.uint8)
            for t in range(n)]
    for flags in (0, 12, 3, 15):
        check(eng, engine, q, Targets.from_list(seqs), "blosum62", 3, 1, flags)


@pytest.mark.parametrize("engine", ENGINES)
def test_empty_query(eng, engine):
    seqs = [np.array([1, 2, 3], np.uint8), np.zeros(0, np.uint8)]
    eng.set_engine(engine)
    eng.set_query(np.zeros(0, np.uint8), "actgn", 3, 1, "FFTT")
    s, er, eq, _ = eng.align(Targets.from_list(seqs))
    assert list(s) == [0, 0] and list(er) == [0, 0] and list(eq) == [0, 0]


def test_int32_widening(eng):
    """Groups whose proven range exceeds int16 run in int32 and stay exact."""
    rng = np.random.default_rng(4)
    q = rng.integers(0, 24, 300).astype(np.uint8)
    lens = [20000, 12000, 9000, 6000, 3000, 800, 50] + [int(x) for x in rng.integers(1, 600, 300)]
    seqs = [rng.integers(0, 24, L).astype(np.uint8) for L in lens]
    for flags in (12, 0, 15, 2):
        for tile in (16, 32):
            s, er, eq, st = check(eng, "batched", q, Targets.from_list(seqs), "blosum62", 5, 4, flags, "", tile)
            assert (st[:3] & ish_amd.STATUS_WIDENED).all()
            assert ((st & (0xFF ^ ish_amd.STATUS_F16)) == ish_amd.STATUS_I16).sum() > 100


def test_gap_open_below_extend_runs_int32(eng):
    """go < ge: the padding rows above the query would not repeat the top
    boundary, so every group runs in the int32 kernels (which read row Q-1
    from any register row) and stays exact."""
    rng = np.random.default_rng(12)
    q = rng.integers(0, 24, 150).astype(np.uint8)
    seqs = [rng.integers(0, 24, int(L)).astype(np.uint8) for L in rng.integers(1, 500, 600)]
    for flags in (0, 4, 10, 15):
        for tile in (16, 32, 40):
            s, er, eq, st = check(eng, "batched", q, Targets.from_list(seqs), "blosum62", 2, 3, flags,
                                  f"flags={flags}", tile)
            assert (st & ish_amd.STATUS_WIDENED).all()


@pytest.mark.parametrize("engine", ENGINES)
def test_nucleotide_workload(eng, engine):
    w = workloads.generate(1, 20000, seed=5)
    s, er, eq, _ = check(eng, engine, w.query, w.targets, "actgn", 3, 1, O.flags_from_str("FFTT"), "wl1")
    ms = int(O.Matrix("actgn").values.reshape(5, 5)[w.query, w.query].sum())
    emitted = ish_amd.emit("striped-semi-global", s, er, eq, ms, 0.8)
    os_, _, _ = oracle_batch(w.query, w.targets, "actgn", 3, 1, O.flags_from_str("FFTT"))
    ref = np.array([O.sg_emit(v, ms, 0.8) for v in os_])
    assert np.array_equal(emitted, ref)
    assert emitted.sum() > 0


@pytest.mark.parametrize("engine,tile", [("batched", 32), ("batched", 16), ("batched", 40), ("plain", 32)])
def test_protein_workload(eng, engine, tile):
    w = workloads.generate(2, 6000, seed=6)
    for flags in ("FFTT", "TTTT"):
        check(eng, engine, w.query, w.targets, "blosum62", 3, 1, O.flags_from_str(flags), "wl2", tile)


@pytest.mark.parametrize("slice_bytes", [1, 40_000, 1_000_000])
def test_sliced_batches(eng, slice_bytes):
    """A batch cut into input-order slices (each uploaded while the one
    before is scored) gives the unsliced results, semi-global and local,
    with empty targets at slice edges."""
    w = workloads.generate(2, 3000 if slice_bytes > 1 else 40, seed=9)
    seqs = [w.targets.residues[w.targets.offsets[t]:w.targets.offsets[t + 1]] for t in range(w.targets.n)]
    for t in range(0, len(seqs), 7):
        seqs[t] = seqs[t][:0]
    targets = Targets.from_list(seqs)
    try:
        for algo, flags in (("striped-semi-global", "FFTT"), ("striped-semi-global", "TTTT"),
                            ("striped-local", "FFFF")):
            eng.set_engine("batched")
            eng.set_tile(32)
            eng.set_query(w.query, "blosum62", 3, 1, flags, algo)
            eng.set_slice_bytes(-1)
            whole = eng.align(targets)
            eng.set_slice_bytes(slice_bytes)
            sliced = eng.align(targets)
            for a, b in zip(whole[:3], sliced[:3]):
                assert np.array_equal(a, b), (algo, flags)
            # the lanes of a launch follow its own groups, so a slice may run
            # in binary16 where the whole batch did not: same cell width
            keep = 0xFF ^ ish_amd.STATUS_F16
            assert np.array_equal(whole[3] & keep, sliced[3] & keep), (algo, flags)
            m = O.Matrix("blosum62")
            if algo == "striped-local":
                ref = O.batch(1, w.query, targets.residues, targets.offsets, m, 3, 1, 0)
            else:
                ref = O.batch(0, w.query, targets.residues, targets.offsets, m, 3, 1, O.flags_from_str(flags))
            for a, b in zip(sliced[:3], ref[:3]):
                assert np.array_equal(a, b), (algo, flags)
    finally:
        eng.set_slice_bytes(0)


def test_resident_loads_after_one_shot_calls():
    """A one-shot call's copy stream ends with the call; the staging it used
    must still serve later resident loads, in the same context and in a new
    one after the first is destroyed (a wait on an event left on the ended
    stream made the next load fail)."""
    w = workloads.generate(2, 4000, seed=12)
    m = O.Matrix("blosum62")
    ref = O.batch(0, w.query, w.targets.residues, w.targets.offsets, m, 3, 1, O.flags_from_str("FFTT"))
    for round_ in range(2):
        e = ish_amd.Engine(0)
        try:
            e.set_engine("batched")
            e.set_query(w.query, "blosum62", 3, 1, "FFTT")
            e.set_slice_bytes(100_000)  # several slices: the copy stream is used
            one = e.align(w.targets)
            db = e.load_db(w.targets)
            res = e.align_db(db)
            db.close()
            for got in (one, res):
                for a, b in zip(got[:3], ref[:3]):
                    assert np.array_equal(a, b), round_
        finally:
            e.close()


def test_bad_offsets_rejected_before_any_read(eng):
    """Decreasing offsets and over-long targets are errors of the call, found
    before any residue is read (the residue buffer here is 8 bytes)."""
    res = np.arange(8, dtype=np.uint8)
    eng.set_engine("batched")
    eng.set_query(np.array([1, 2, 3], np.uint8), "blosum62", 3, 1, "FFTT")
    cases = [(np.array([0, 5, 3, 8], np.uint64), "INVALID"),
             (np.array([0, 4, 1 << 31], np.uint64), "RANGE")]
    for off, code in cases:
        t = Targets(res, off)
        with pytest.raises(ish_amd.IshError, match=code):
            eng.align(t)
        with pytest.raises(ish_amd.IshError, match=code):
            eng.load_db(t)
    # the context still works afterwards
    s, _, _, _ = eng.align(Targets(res, np.array([0, 3, 8], np.uint64)))
    assert len(s) == 2


def test_database_reuse_across_queries(eng):
    w = workloads.generate(2, 5000, seed=8)
    eng.set_engine("batched")
    eng.set_tile(32)
    db = eng.load_db(w.targets)
    m = O.Matrix("blosum62")
    for qlen, flags in ((350, "FFTT"), (57, "TTTT"), (1000, "FFFF"), (1, "TFTF")):
        q = workloads.query(2, qlen + 17, qlen)
        eng.set_query(q, "blosum62", 3, 1, flags)
        s, er, eq, st = eng.align_db(db)
        os_, oer, oeq, _ = O.batch(0, q, w.targets.residues, w.targets.offsets, m, 3, 1,
                                   O.flags_from_str(flags))
        assert np.array_equal(s, os_) and np.array_equal(er, oer) and np.array_equal(eq, oeq), (qlen, flags)
        assert ((st & (0xFF ^ ish_amd.STATUS_F16)) == ish_amd.STATUS_I16).all()
    db.close()


# ---- local (Smith-Waterman) ------------------------------------------------
def check_local(eng, q, targets, kind, go, ge, tile=32, label="", schedule="auto"):
    eng.set_engine("batched")
    eng.set_tile(tile)
    eng.set_schedule(schedule)
    eng.set_query(q, kind, go, ge, "FFFF", "striped-local")
    try:
        s, er, eq, st = eng.align(targets)
    finally:
        eng.set_schedule("auto")
    m = O.Matrix(kind)
    os_, oer, oeq, _ = O.batch(1, q, targets.residues, targets.offsets, m, go, ge, 0)
    bad = np.nonzero((s != os_) | (er != oer) | (eq != oeq))[0]
    if len(bad):
        t = bad[0]
        raise AssertionError(
            f"{label} local tile={tile} schedule={schedule} {len(bad)}/{targets.n} differ; first t={t} "
            f"R={targets.offsets[t + 1] - targets.offsets[t]} gpu=({s[t]},{er[t]},{eq[t]}) "
            f"oracle=({os_[t]},{oer[t]},{oeq[t]}) status={st[t]}")
    return s, er, eq, st


@pytest.mark.parametrize("tile", [32, 16])
def test_local_known_answers(eng, tile):
    eng.set_tile(tile)
    for c in VEC["local"]:
        m = O.Matrix(c["matrix"])
        q, t = m.encode(c["read"]), m.encode(c["reference"])
        eng.set_query(q, c["matrix"], c["go"], c["ge"], "FFFF", "striped-local")
        s, er, eq, _ = eng.align(Targets.from_list([t]))
        assert (s[0], er[0], eq[0]) == (c["score"], c["ref_end"], c["read_end"]), c


@pytest.mark.parametrize("schedule", ["wave", "workgroup"])
@pytest.mark.parametrize("tile", [32, 16])
@pytest.mark.parametrize("kind", ["actgn", "actgn0", "blosum62", "ascii"])
def test_local_random(eng, tile, kind, schedule):
    rng = np.random.default_rng(zlib.crc32(f"local-{kind}-{tile}".encode()))
    alpha = {"actgn": 5, "actgn0": 5, "blosum62": 24, "ascii": 256}[kind]
    for trial in range(6):
        Q = int(rng.integers(1, 200))
        base = 40 if kind == "ascii" else 0
        span = 60 if kind == "ascii" else alpha
        q = (rng.integers(0, span, Q) + base).astype(np.uint8)
        seqs = []
        for t in range(250):
            R = int(rng.integers(0, 700 if t % 7 else 2600))
            r = (rng.integers(0, span, R) + base).astype(np.uint8)
            if R and t % 4 == 0:
                k = min(Q, R)
                off = int(rng.integers(0, R - k + 1))
                r[off:off + k] = q[:k]
            seqs.append(r)
        ge = int(rng.integers(0, 3))
        go = ge + int(rng.integers(1, 6))
        check_local(eng, q, Targets.from_list(seqs), kind, go, ge, tile, f"trial={trial}", schedule)


@pytest.mark.parametrize("schedule", ["wave", "workgroup"])
def test_local_long_reads_and_windows(eng, schedule):
    """Reads far longer than the 256-column offset window, hits near the end."""
    rng = np.random.default_rng(21)
    q = rng.integers(0, 4, 500).astype(np.uint8)
    seqs = []
    for L in (20000, 9000, 2049, 1023, 257, 256, 255, 3):
        r = rng.integers(0, 4, L).astype(np.uint8)
        if L > 600:
            r[L - 520:L - 20] = q
        seqs.append(r)
    for tile in (32, 16):
        s, er, eq, st = check_local(eng, q, Targets.from_list(seqs), "actgn0", 3, 1, tile,
                                    schedule=schedule)
        assert s[0] == 2 * 500 and er[0] == 20000 - 21
        assert ((st & (0xFF ^ ish_amd.STATUS_F16)) == ish_amd.STATUS_I16).all()


@pytest.mark.parametrize("schedule", ["wave", "workgroup"])
@pytest.mark.parametrize("tile", [32, 16])
def test_local_degenerate_inputs(eng, tile, schedule):
    """Queries of 0, 1 and 2 residues against empty, one-residue and short
    reads and one long read; with the workgroup schedule the (one-class or
    few-class) paired profile is used."""
    rng = np.random.default_rng(21 + tile)
    seqs = [np.zeros(0, np.uint8), np.array([0], np.uint8), np.array([3, 3], np.uint8)]
    seqs += [rng.integers(0, 5, int(L)).astype(np.uint8) for L in rng.integers(1, 12, 200)]
    seqs += [rng.integers(0, 5, 5000).astype(np.uint8)]
    for Q in (0, 1, 2):
        q = (np.arange(Q) % 4).astype(np.uint8)
        check_local(eng, q, Targets.from_list(seqs), "actgn0", 3, 1, tile, f"Q={Q}", schedule)


@pytest.mark.parametrize("tile", [32, 16])
def test_local_paired_profile_wide_alphabet(eng, tile):
    """A query over a few bytes of the 256-code ASCII matrix has a profile of
    few classes, which the local pipeline reads as a paired profile. The
    targets hold every byte value (most of them in the query's absent
    class), stretches of the query's letters and copies of query pieces."""
    rng = np.random.default_rng(770 + tile)
    letters = np.frombuffer(b"ACGT", np.uint8)
    q = rng.choice(letters, 300)
    seqs = []
    for k in range(600):
        L = int(rng.integers(200, 3000))
        t = rng.integers(0, 256, L).astype(np.uint8)
        if k % 3 == 0:
            o = int(rng.integers(0, L - 100))
            t[o:o + 100] = rng.choice(letters, 100)
        if k % 7 == 0 and L > 400:
            a = int(rng.integers(0, 150))
            o = int(rng.integers(0, L - 150))
            t[o:o + 150] = q[a:a + 150]
        seqs.append(t)
    check_local(eng, q, Targets.from_list(seqs), "ascii", 3, 1, tile, "paired ascii", "workgroup")


def test_local_int32_path(eng):
    rng = np.random.default_rng(22)
    q = rng.integers(0, 24, 600).astype(np.uint8)
    seqs = [rng.integers(0, 24, int(rng.integers(1, 900))).astype(np.uint8) for _ in range(300)]
    s, er, eq, st = check_local(eng, q, Targets.from_list(seqs), "blosum62", 60, 50)
    assert (st & ish_amd.STATUS_WIDENED).all()


@pytest.mark.parametrize("schedule", ["auto", "wave", "workgroup"])
def test_local_long_read_workload(eng, schedule):
    w = workloads.generate(4, 160, seed=4)
    check_local(eng, w.query, w.targets, "actgn0", 3, 1, 32, "wl4", schedule)


@pytest.mark.parametrize("Q", [130, 31])
def test_local_workgroup_many_groups(eng, Q):
    """More groups than workgroups, so every workgroup scores several groups
    (LDS ring and combine area reused); Q=130 leaves waves 1-3 idle in the
    last round, Q=31 has a single chunk (waves 1-3 idle throughout)."""
    rng = np.random.default_rng(Q)
    q = rng.integers(0, 4, Q).astype(np.uint8)
    seqs = [rng.integers(0, 5, int(rng.integers(0, 150))).astype(np.uint8) for _ in range(70000)]
    for k in range(0, 70000, 97):
        if len(seqs[k]) >= Q // 2:
            seqs[k][: Q // 2] = q[: Q // 2]
    check_local(eng, q, Targets.from_list(seqs), "actgn", 3, 1, 32, f"many Q={Q}", "workgroup")


def _records(text, rtype):
    if rtype == "line":
        lines = text.split("\n")
        if text.endswith("\n"):
            lines = lines[:-1]
        return [(l, l) for l in lines]
    recs, name, seq = [], None, []
    for l in text.split("\n"):
        if l.startswith(">"):
            if name is not None:
                recs.append((name, "".join(seq)))
            name, seq = l, []
        elif name is not None and l:
            seq.append(l)
    if name is not None:
        recs.append((name, "".join(seq)))
    return recs


def test_searcher_fixtures_on_gpu(eng):
    """tests/test_searcher.mojo:17-200 through the GPU engine: 5 inputs x
    {striped-local, striped-semi-global} x {TTTT, FFTT}, output compared
    exactly (line) or case-insensitively (fastx)."""
    cases = json.load(open(os.path.join(HERE, "golden", "searcher_cases.json")))
    eng.set_engine("batched")
    eng.set_tile(32)
    for c in cases["cases"]:
        kind = c["matrix"]
        recs = _records(cases["inputs"][c["input"]], c["record_type"])
        q, ms = ish_amd.encode_and_score(kind, c["pattern"])
        dec = ish_amd.decoding(kind)
        targets = Targets.from_list([ish_amd.encode_and_score(kind, seq.encode())[0] for _, seq in recs])
        for algo in ("striped-local", "striped-semi-global"):
            for ends in ("TTTT", "FFTT"):
                eng.set_query(q, kind, 3, 1, ends, algo)
                s, er, eq, _ = eng.align(targets)
                emit = ish_amd.emit(algo, s, er, eq, ms, 0.8)
                out = []
                for k, (name, seq) in enumerate(recs):
                    if emit[k]:
                        if c["record_type"] == "line":
                            out.append(name + "\n")
                        else:
                            enc = ish_amd.encode_and_score(kind, seq.encode())[0]
                            out.append(name + "\n" + bytes(dec[enc]).decode() + "\n")
                got = "".join(out)
                exp = c["expected"]
                assert (got == exp) if c["record_type"] == "line" else (got.upper() == exp.upper()), (kind, algo, ends)


# ---- start coordinates (reverse pass on the GPU) ---------------------------
def test_starts_known_answers(eng):
    eng.set_engine("batched")
    eng.set_tile(32)
    for c in VEC["semi_global"]:
        if "target_start" not in c:
            continue
        m = O.Matrix(c["matrix"])
        q, t = m.encode(c["query"]), m.encode(c["target"])
        eng.set_query(q, c["matrix"], c["go"], c["ge"], c["flags"])
        db = eng.load_db(Targets.from_list([t]))
        s, er, eq, _ = eng.align_db(db)
        ts, qs = eng.find_starts(db, [0], er, eq)
        assert (ts[0], er[0]) == (c["target_start"], c["target_end"]), c
        db.close()
    for c in VEC["local"]:
        m = O.Matrix(c["matrix"])
        q, t = m.encode(c["read"]), m.encode(c["reference"])
        eng.set_query(q, c["matrix"], c["go"], c["ge"], "FFFF", "striped-local")
        db = eng.load_db(Targets.from_list([t]))
        s, er, eq, _ = eng.align_db(db)
        ts, qs = eng.find_starts(db, [0], er, eq)
        assert (ts[0], qs[0]) == (c["ref_begin"], c["read_begin"]), c
        db.close()


@pytest.mark.parametrize("flags", range(16))
def test_semi_global_starts_random(eng, flags):
    rng = np.random.default_rng(300 + flags)
    m = O.Matrix("blosum62")
    q = rng.integers(0, 20, 90).astype(np.uint8)
    seqs = []
    for t in range(200):
        R = int(rng.integers(1, 400))
        r = rng.integers(0, 20, R).astype(np.uint8)
        if t % 3 == 0:
            k = min(len(q), R)
            off = int(rng.integers(0, R - k + 1))
            r[off:off + k] = q[:k]
        seqs.append(r)
    T = Targets.from_list(seqs)
    eng.set_engine("batched")
    eng.set_tile(32)
    eng.set_query(q, "blosum62", 3, 1, flags)
    db = eng.load_db(T)
    s, er, eq, _ = eng.align_db(db)
    sel = np.arange(T.n)
    ts, qs = eng.find_starts(db, sel, er, eq)
    for k in sel:
        se = O.sg_start_end(q, seqs[k], m, 3, 1, flags)
        assert (ts[k], qs[k]) == (se["target_start"], se["query_start"]), (k, flags)
    db.close()


def test_local_starts_random(eng):
    rng = np.random.default_rng(77)
    m = O.Matrix("actgn")
    q = rng.integers(0, 5, 120).astype(np.uint8)
    seqs = []
    for t in range(300):
        R = int(rng.integers(1, 900))
        r = rng.integers(0, 5, R).astype(np.uint8)
        if t % 2 == 0:
            k = min(len(q) - 10, R)
            off = int(rng.integers(0, R - k + 1))
            r[off:off + k] = q[5:5 + k]
        seqs.append(r)
    T = Targets.from_list(seqs)
    eng.set_engine("batched")
    eng.set_query(q, "actgn", 3, 1, "FFFF", "striped-local")
    db = eng.load_db(T)
    s, er, eq, _ = eng.align_db(db)
    sel = np.arange(T.n)
    ts, qs = eng.find_starts(db, sel, er, eq)
    for k in sel:
        a = O.sw_align(q, seqs[k], m, 3, 1, cutoff=-1.0)
        assert a is not None
        assert (ts[k], qs[k]) == (a["ref_begin1"], a["read_begin1"]), k
    db.close()


# ---- many queries in one launch --------------------------------------------
@pytest.mark.parametrize("flags", ["FFTT", "TTTT", "FFFF", "TFTF"])
def test_multi_query(eng, flags):
    rng = np.random.default_rng(zlib.crc32(flags.encode()))
    m = O.Matrix("blosum62")
    seqs = [rng.integers(0, 24, int(rng.integers(0, 700))).astype(np.uint8) for _ in range(1500)]
    seqs += [rng.integers(0, 24, L).astype(np.uint8) for L in (15000, 12000)]  # int32 for the longest query
    T = Targets.from_list(seqs)
    queries = [rng.integers(0, 20, L).astype(np.uint8) for L in (1, 7, 33, 64, 100, 257, 350, 600, 1000)]
    queries[5][:50] = seqs[3][:50] if len(seqs[3]) >= 50 else queries[5][:50]
    eng.set_engine("batched")
    eng.set_tile(32)
    db = eng.load_db(T)
    s, er, eq, st = eng.align_db_multi(db, queries, "blosum62", 3, 2, flags)
    fl = O.flags_from_str(flags)
    for k, q in enumerate(queries):
        os_, oer, oeq, _ = O.batch(0, q, T.residues, T.offsets, m, 3, 2, fl)
        assert np.array_equal(s[k], os_) and np.array_equal(er[k], oer) and np.array_equal(eq[k], oeq), (k, len(q))
    assert (st[-1, -2:] & ish_amd.STATUS_WIDENED).all()
    db.close()


@pytest.mark.parametrize("tile", [32, 16])
@pytest.mark.parametrize("ncodes", [5, 12, 20, 21, 22])
def test_multi_query_paired_class_counts(eng, tile, ncodes):
    """Multi-query launches over databases of 5 codes (per-wave paired
    profiles), 12 to 21 (workgroup-shared paired profiles, tiles in
    lockstep) and 22 (int16 profile), with queries of every length class,
    int32 groups and per-query free ends (FFTT and TTTT in one launch
    through the serving-style per-query flags are covered by the server
    tests; here one set per launch)."""
    rng = np.random.default_rng(ncodes * 10 + tile)
    m = O.Matrix("blosum62")
    seqs = [rng.integers(0, ncodes, int(rng.integers(0, 500))).astype(np.uint8) for _ in range(900)]
    seqs += [rng.integers(0, ncodes, L).astype(np.uint8) for L in (9000, 4000)]
    T = Targets.from_list(seqs)
    queries = [np.concatenate([np.arange(24), rng.integers(0, 24, L)]).astype(np.uint8)
               for L in (0, 9, 100, 333, 700)]
    eng.set_engine("batched")
    eng.set_tile(tile)
    db = eng.load_db(T)
    try:
        for flags in ("FFTT", "TTTT", "FFFF"):
            s, er, eq, st = eng.align_db_multi(db, queries, "blosum62", 3, 1, flags)
            fl = O.flags_from_str(flags)
            for k, q in enumerate(queries):
                os_, oer, oeq, _ = O.batch(0, q, T.residues, T.offsets, m, 3, 1, fl)
                bad = np.flatnonzero((s[k] != os_) | (er[k] != oer) | (eq[k] != oeq))
                assert len(bad) == 0, (ncodes, tile, flags, k, len(q), bad[:5])
    finally:
        db.close()


# ---- autotuner and HIP graphs ------------------------------------------------
def test_autotune_keeps_results(eng):
    """The tuned launch plan gives the same results as the defaults: a
    protein and a nucleotide semi-global query (the latter with pipelined
    groups, whose workgroup width is tuned too) and a local query."""
    w = workloads.generate(2, 30000, seed=8)
    eng.set_engine("batched")
    eng.set_tile(32)
    db = eng.load_db(w.targets)
    try:
        eng.set_query(w.query, "blosum62", 3, 1, "FFTT")
        ref = eng.align_db(db)
        tuned = eng.autotune(db)
        assert tuned[0] in (16, 24, 32) and tuned[1] in ("wave", "auto", "workgroup")
        got = eng.align_db(db)
        for a, b in zip(ref[:3], got[:3]):
            assert np.array_equal(a, b)
        # the tuned plan may run other groups in binary16 lanes
        keep = 0xFF ^ ish_amd.STATUS_F16
        assert np.array_equal(ref[3] & keep, got[3] & keep)
        lr = workloads.generate(4, 400, seed=9)
        db2 = eng.load_db(lr.targets)
        eng.set_tile(32)
        eng.set_query(lr.query, "actgn0", 3, 1, "FFFF", "striped-local")
        ref = eng.align_db(db2)
        tuned = eng.autotune(db2, 1)
        assert tuned is not None and eng.tuning(db2) == tuned
        got = eng.align_db(db2)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)
        m = O.Matrix("actgn0")
        os_, oer, oeq, _ = O.batch(1, lr.query, lr.targets.residues, lr.targets.offsets, m, 3, 1, 0)
        assert np.array_equal(got[0], os_) and np.array_equal(got[1], oer) and np.array_equal(got[2], oeq)
        db2.close()
        nt = workloads.generate(1, 20000, seed=10)
        db3 = eng.load_db(nt.targets)
        eng.set_tile(32)
        eng.set_query(nt.query, "actgn", 3, 1, "FFTT")
        ref = eng.align_db(db3)
        assert eng.autotune(db3, 1) is not None
        got = eng.align_db(db3)
        db3.close()
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)
    finally:
        eng.set_tile(32)  # switches the tuning table off for later tests
        db.close()


def test_graph_replays_match_eager(eng):
    """Captured passes replayed several times, with eager work and
    synchronisations in between; every replay equals the eager result."""
    w = workloads.generate(2, 20000, seed=12)
    rng = np.random.default_rng(3)
    qa, qb = w.query, rng.integers(0, 20, 120).astype(np.uint8)
    eng.set_engine("batched")
    eng.set_tile(32)
    db = eng.load_db(w.targets)
    other = Targets.from_list([rng.integers(0, 20, 300).astype(np.uint8) for _ in range(500)])
    try:
        eager = {}
        for name, q, fl in (("a", qa, "FFTT"), ("b", qb, "TTTT")):
            eng.set_query(q, "blosum62", 3, 1, fl)
            eager[name] = eng.align_db(db)
        eng.set_graphs(True)
        for rnd in range(3):
            for name, q, fl in (("a", qa, "FFTT"), ("b", qb, "TTTT")):
                eng.set_query(q, "blosum62", 3, 1, fl)
                for rep in range(3):
                    got = eng.align_db(db)
                    for x, y in zip(got, eager[name]):
                        assert np.array_equal(x, y), (rnd, name, rep)
                    # eager work between replays
                    eng.set_engine("plain")
                    eng.align(other)
                    eng.set_engine("batched")
                    eng.sync()
    finally:
        eng.set_graphs(False)
        db.close()


@pytest.mark.parametrize("kind", ["actgn", "blosum62", "ascii"])
def test_db_load_ascii_encodes_on_device(eng, kind):
    """Raw bytes (every byte value, lengths that leave 16-byte tails)
    encoded on the device score exactly like host-encoded residues."""
    rng = np.random.default_rng(zlib.crc32(kind.encode()))
    raw = [rng.integers(0, 256, int(rng.integers(0, 300))).astype(np.uint8) for _ in range(777)]
    raw[5] = np.arange(256, dtype=np.uint8)
    enc = ish_amd.encoding(kind)
    A = len(ish_amd.decoding(kind))
    q = rng.integers(0, A if kind != "ascii" else 256, 90).astype(np.uint8)
    q = enc[q] if kind == "ascii" else q
    eng.set_engine("batched")
    eng.set_tile(32)
    eng.set_query(q, kind, 3, 1, "TTTT")
    db_a = eng.load_db_ascii(Targets.from_list(raw), kind)
    db_e = eng.load_db(Targets.from_list([enc[r] for r in raw]))
    try:
        for x, y in zip(eng.align_db(db_a), eng.align_db(db_e)):
            assert np.array_equal(x, y)
    finally:
        db_a.close()
        db_e.close()
