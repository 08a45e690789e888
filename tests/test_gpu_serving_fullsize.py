"""The serving loop at the serving workload's own size: a device-resident
database of 10,000,000 protein targets (generator kind 5: lengths uniform in
[100, 300] at background amino-acid frequencies, 2.0 G residues, blosum62),
answering serving requests (ends-free FFTT / TTTT, lengths log-normal with
median 300 clipped to [20, 2000], a quarter of them windows of database
targets so that they emit at 0.8).

Twelve requests are submitted at once: the first ones take the two launch
slots, the others wait and are admitted together, so batches hold several
requests. For every request:
  - the full per-target results (keep_scores) equal a single-query
    ish_hip_align_db pass on every one of the 10M targets;
  - the emitted hits equal that pass's emission at 0.8;
  - the results equal the oracle on an even sample of 20,000 targets.
"""
import os

import numpy as np
import pytest

import ish_amd
from ish_amd import workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def _same(label, a, b):
    for name, x, y in zip(("score", "end_ref", "end_query"), a, b):
        bad = np.flatnonzero(np.asarray(x) != np.asarray(y))
        assert len(bad) == 0, f"{label}: {name} differs on {len(bad)} targets, first {bad[0]}"


def test_serving_10M_batches_match_single_passes_and_oracle():
    if ish_amd.device_count() == 0:
        pytest.fail("no gfx950 device visible to the HIP library")
    w = workloads.generate(5, 10_000_000, seed=1)
    T = w.targets
    assert T.n == 10_000_000
    lens = T.offsets[1:] - T.offsets[:-1]
    assert lens.min() >= 100 and lens.max() <= 300
    reqs = workloads.serving_requests(12, seed=7, rate=200.0, sources=T, hit_fraction=0.25)
    print(f"serving 10M: {int(T.offsets[-1])} residues, query lengths {[len(r.query) for r in reqs]}", flush=True)
    eng = ish_amd.Engine(0)
    db = eng.load_db(T)
    srv = ish_amd.Server(db, "blosum62", 3, 1, 0.8, keep_scores=True)
    idx = np.linspace(0, T.n - 1, 20_000).astype(np.int64)
    slens = (T.offsets[idx + 1] - T.offsets[idx]).astype(np.int64)
    soff = np.zeros(len(idx) + 1, np.uint64)
    np.cumsum(slens, out=soff[1:])
    sres = np.concatenate([T.residues[int(T.offsets[i]):int(T.offsets[i + 1])] for i in idx])
    mat = O.Matrix("blosum62")
    m62 = ish_amd.matrix("blosum62")
    try:
        tickets = [srv.submit(r.query, r.ends_free) for r in reqs]
        sizes, total_hits = [], 0
        for k, (r, t) in enumerate(zip(reqs, tickets)):
            info = srv.wait(t)
            assert info["status"] == 0
            sizes.append(info["batch_size"])
            got = srv.scores(t)
            hits = srv.hits(t, info["n_hits"])
            srv.release(t)
            eng.set_query(r.query, "blosum62", 3, 1, r.ends_free)
            one = eng.align_db(db)[:3]
            _same(f"request {k} ({len(r.query)} aa, {r.ends_free}) vs single-query pass", got, one)
            ms = int(m62[r.query, r.query].sum())
            emit = np.flatnonzero(ish_amd.emit("semi-global", *one, ms, 0.8))
            assert np.array_equal(hits[0], emit), k
            assert np.array_equal(hits[1], one[0][emit]) and np.array_equal(hits[2], one[1][emit]), k
            total_hits += len(emit)
            ref = O.batch(0, r.query, sres, soff, mat, 3, 1, O.flags_from_str(r.ends_free),
                          nthreads=THREADS)[:3]
            _same(f"request {k} vs oracle on the sample", tuple(a[idx] for a in got), ref)
        print(f"batch sizes {sizes}, hits {total_hits}", flush=True)
        assert max(sizes) > 1, sizes  # requests shared a batch
        assert total_hits > 0
    finally:
        srv.close()
        db.close()
        eng.close()
