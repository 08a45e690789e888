"""ish_hip_align_batch_starts and the async submit/wait pair through the C
ABI: one call scores a batch, applies the matchers' emission rule and finds
the start coordinates of the emitted targets on the GPU, from the residues
already uploaded (the reference's per-batch GPU step, ishlib/gpu/
process_batch.mojo:129-160, with find_start from cpu/process_batch.mojo:
36-68).

Checked against the oracle: the forward results on every target, emission
by the oracle's rule, and for every emitted target the start coordinates of
the oracle's semi_global_aln_start_end / ssw_align. Cases: the reference's
searcher fixtures (tests/test_searcher.mojo:17-128) for local and
semi-global, TTTT and FFTT; 20,000-target random batches; batches cut into
many slices; an empty batch and one with no hit; the async pair.
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


@pytest.fixture(scope="module")
def eng():
    if ish_amd.device_count() == 0:
        pytest.fail("no gfx950 device visible to the HIP library")
    e = ish_amd.Engine(0)
    e.set_engine("batched")
    yield e
    e.close()


def _check(label, q, seqs, T, kind, algo, flags, thr, got, go=3, ge=1):
    s, er, eq, st, emit, ts, qs = got
    m = O.Matrix(kind)
    local = algo == "striped-local"
    fl = 0 if local else O.flags_from_str(flags)
    os_, oer, oeq, _ = O.batch(1 if local else 0, q, T.residues, T.offsets, m, go, ge, fl)
    assert np.array_equal(s, os_) and np.array_equal(er, oer) and np.array_equal(eq, oeq), label
    ms = int(m.values.reshape(m.A, m.A)[q, q].sum())
    for k in range(T.n):
        want = O.sw_emit(s[k], er[k], eq[k], ms, thr) if local else O.sg_emit(s[k], ms, thr)
        assert bool(emit[k]) == want, (label, k)
        if not emit[k]:
            assert (ts[k], qs[k]) == (-1, -1), (label, k)
            continue
        if local:
            a = O.sw_align(q, seqs[k], m, go, ge, cutoff=-1.0)
            assert (ts[k], qs[k]) == (a["ref_begin1"], a["read_begin1"]), (label, k)
        else:
            se = O.sg_start_end(q, seqs[k], m, go, ge, fl)
            assert (ts[k], qs[k], er[k], eq[k]) == (se["target_start"], se["query_start"], se["target_end"],
                                                    se["query_end"]), (label, k)
    return int(emit.sum())


def _records(text, record_type):
    lines = text.split("\n")
    if record_type == "line":
        return [(ln, ln) for ln in lines if ln]
    recs, name, seq = [], None, []
    for ln in lines:
        if ln.startswith(">"):
            if name is not None:
                recs.append((name, "".join(seq)))
            name, seq = ln, []
        elif name is not None:
            seq.append(ln)
    if name is not None:
        recs.append((name, "".join(seq)))
    return recs


def test_starts_searcher_fixtures(eng):
    cases = json.load(open(os.path.join(HERE, "golden", "searcher_cases.json")))
    total = 0
    for c in cases["cases"]:
        kind = c["matrix"]
        recs = _records(cases["inputs"][c["input"]], c["record_type"])
        q, ms = ish_amd.encode_and_score(kind, c["pattern"])
        seqs = [ish_amd.encode_and_score(kind, seq.encode())[0] for _, seq in recs]
        T = Targets.from_list(seqs)
        for algo in ("striped-local", "striped-semi-global"):
            for ends in ("TTTT", "FFTT"):
                eng.set_query(q, kind, 3, 1, ends, algo)
                got = eng.align_starts(T, ms, 0.8)
                total += _check(f"{c['input']} {algo} {ends}", q, seqs, T, kind, algo, ends, 0.8, got)
    assert total > 0


@pytest.mark.parametrize("algo,kind,flags", [("striped-semi-global", "blosum62", "FFTT"),
                                             ("striped-semi-global", "blosum62", "TTTT"),
                                             ("striped-semi-global", "actgn", "FTFT"),
                                             ("striped-local", "actgn0", "FFFF"),
                                             ("striped-local", "blosum62", "FFFF")])
def test_starts_random_20k(eng, algo, kind, flags):
    rng = np.random.default_rng(zlib.crc32(f"{algo}-{kind}-{flags}".encode()))
    alpha = 24 if kind == "blosum62" else 5
    q = rng.integers(0, alpha, 120).astype(np.uint8)
    seqs = []
    for t in range(20000):
        R = int(rng.integers(0, 400))
        r = rng.integers(0, alpha, R).astype(np.uint8)
        if t % 7 == 0 and R >= 60:  # a stretch of the query, so that it emits at 0.3
            k = min(len(q), R)
            o = int(rng.integers(0, R - k + 1))
            r[o:o + k] = q[:k]
        seqs.append(r)
    T = Targets.from_list(seqs)
    eng.set_query(q, kind, 3, 1, flags, algo)
    m = O.Matrix(kind)
    ms = int(m.values.reshape(m.A, m.A)[q, q].sum())
    got = eng.align_starts(T, ms, 0.3)
    n = _check(f"{algo} {kind} {flags}", q, seqs, T, kind, algo, flags, 0.3, got)
    assert n > 1000
    # the same through many input-order slices (each slice's prefixes are
    # gathered before the slice's buffers are reused)
    eng.set_slice_bytes(200_000)
    try:
        sliced = eng.align_starts(T, ms, 0.3)
    finally:
        eng.set_slice_bytes(0)
    for a, b in zip(sliced, got):
        assert np.array_equal(a, b)


def test_starts_empty_and_no_hit(eng):
    w = workloads.generate(2, 500, seed=4)
    eng.set_query(w.query, "blosum62", 3, 1, "FFTT")
    empty = eng.align_starts(Targets.from_list([]), 100, 0.8)
    assert all(len(a) == 0 for a in empty)
    got = eng.align_starts(w.targets, 10 ** 6, 0.99)  # nothing reaches this
    assert not got[4].any() and (got[5] == -1).all() and (got[6] == -1).all()
    plain = eng.align(w.targets)
    for a, b in zip(got[:4], plain):
        assert np.array_equal(a, b)


def test_submit_wait_matches_sync(eng):
    w = workloads.generate(2, 30000, seed=9)
    T = w.targets
    eng.set_query(w.query, "blosum62", 3, 1, "FFTT")
    m = O.Matrix("blosum62")
    ms = int(m.values.reshape(m.A, m.A)[w.query, w.query].sum())
    sync = eng.align_starts(T, ms, 0.5)
    plain = eng.align(T)
    eng.submit(T, ms, 0.5)
    host_work = sum(int(x) for x in range(20000))  # the host is free meanwhile
    got = eng.wait()
    assert host_work > 0
    for a, b in zip(got, sync):
        assert np.array_equal(a, b)
    eng.submit(T)
    got = eng.wait()
    for a, b in zip(got, plain):
        assert np.array_equal(a, b)
    assert sync[4].sum() > 0
    # a second submit before wait is refused
    eng.submit(T)
    with pytest.raises(ish_amd.IshError, match="in flight"):
        eng._check(eng._L.ish_hip_align_batch_submit(eng._h, None, None, 0, 0, 0.0, None, None, None,
                                                      None, None, None, None))
    eng.wait()
