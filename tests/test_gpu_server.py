"""Serving loop (ish_hip_server_*) through the C ABI.

Every request's results must be the same whether it runs alone or batched
with others (bit-exact on score, end_ref, end_query for every target, and
the same emitted set), and equal to the CPU oracle. Batches mix all 16
free-end settings, so both end-rule variants of the multi-query kernel and
the per-query flags run in one batch.
"""
import threading
import time
import zlib

import numpy as np
import pytest

import ish_amd
from ish_amd import workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FLAGS = ["".join("T" if (f >> k) & 1 else "F" for k in range(4)) for f in range(16)]


@pytest.fixture(scope="module")
def setup():
    if ish_amd.device_count() == 0:
        pytest.fail("no gfx950 device visible to the HIP library")
    eng = ish_amd.Engine(0)
    w = workloads.generate(2, 6000, seed=11)
    db = eng.load_db(w.targets)
    yield eng, w, db
    db.close()
    eng.close()


def _queries(w):
    rng = np.random.default_rng(zlib.crc32(b"server"))
    qs = [w.query]  # the planted homologs' query
    for k in range(15):
        L = int(rng.integers(1, 700))
        qs.append(workloads.query(5, 1000 + k, L))
    # exact windows of targets for the settings with both target ends free
    # (FFTT, TFTT, FTTT, TTTT), so that those targets emit at 0.8
    T = w.targets
    lens = (T.offsets[1:] - T.offsets[:-1]).astype(np.int64)
    long_ = np.flatnonzero(lens >= 260)
    for k in range(12, 16):
        t = int(long_[k * 37 % len(long_)])
        o = int(T.offsets[t]) + 20
        qs[k] = T.residues[o:o + 40 + 50 * (k - 12)].copy()
    return qs


def _max_score(q):
    m = ish_amd.matrix("blosum62")
    return int(sum(int(m[c, c]) for c in q))


def _expected_hits(score, er, eq, q, thr=0.8):
    e = ish_amd.emit("semi-global", score, er, eq, _max_score(q), thr)
    t = np.flatnonzero(e)
    return t, score[t], er[t], eq[t]


def test_server_alone_vs_mix_vs_oracle(setup):
    eng, w, db = setup
    qs = _queries(w)
    T = w.targets
    mat = O.Matrix("blosum62")
    srv = ish_amd.Server(db, "blosum62", 3, 1, 0.8, max_batch=8, slots=2, keep_scores=True)
    try:
        alone = []
        for q, f in zip(qs, FLAGS):
            t = srv.submit(q, f)
            info = srv.wait(t)
            assert info["batch_size"] == 1 and info["status"] == 0
            alone.append((srv.scores(t), srv.hits(t, info["n_hits"])))
            srv.release(t)
        # the same requests arriving together from several threads
        tickets = [None] * len(qs)

        def client(idx):
            for k in idx:
                time.sleep(0.001 * (k % 3))
                tickets[k] = srv.submit(qs[k], FLAGS[k])

        th = [threading.Thread(target=client, args=(list(range(i, len(qs), 4)),)) for i in range(4)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        sizes = []
        for k, t in enumerate(tickets):
            info = srv.wait(t)
            sizes.append(info["batch_size"])
            s, er, eq = srv.scores(t)
            hits = srv.hits(t, info["n_hits"])
            a_s, a_er, a_eq = alone[k][0]
            assert np.array_equal(s, a_s) and np.array_equal(er, a_er) and np.array_equal(eq, a_eq), k
            for x, y in zip(hits, alone[k][1]):
                assert np.array_equal(x, y), k
            # oracle, every target
            fl = O.flags_from_str(FLAGS[k])
            os_, oer, oeq, _ = O.batch(0, qs[k], T.residues, T.offsets, mat, 3, 1, fl)
            bad = np.flatnonzero((s != os_) | (er != oer) | (eq != oeq))
            assert len(bad) == 0, (k, FLAGS[k], len(qs[k]), bad[:5])
            for x, y in zip(hits, _expected_hits(os_, oer, oeq, qs[k])):
                assert np.array_equal(x, y), k
            srv.release(t)
        assert max(sizes) > 1, sizes  # requests really shared batches
        assert sum(len(a[1][0]) for a in alone) > 0  # something was emitted
        st = srv.stats()
        assert st["queries"] == 2 * len(qs) and st["pending"] == 0
    finally:
        srv.close()


def test_server_matches_single_query_engine(setup):
    eng, w, db = setup
    srv = ish_amd.Server(db, "blosum62", 3, 1, 0.8, max_batch=4, slots=3)
    try:
        reqs = workloads.serving_requests(12, seed=5)
        tickets = [srv.submit(r.query, r.ends_free) for r in reqs]
        for r, t in zip(reqs, tickets):
            info, hits = srv.result(t)
            eng.set_engine("batched")
            eng.set_query(r.query, "blosum62", 3, 1, r.ends_free)
            s, er, eq, _ = eng.align_db(db)
            for x, y in zip(hits, _expected_hits(s, er, eq, r.query)):
                assert np.array_equal(x, y)
    finally:
        srv.close()


def test_server_hit_list_overflow(setup):
    eng, w, db = setup
    # threshold 0 with all ends free: nearly every target emits, far beyond
    # the 3-entry device hit list, so the host emits from the full results
    srv = ish_amd.Server(db, "blosum62", 3, 1, 0.0, max_batch=4, hits_per_query=3)
    try:
        qs = [w.query[:120], workloads.query(5, 77, 200)]
        tickets = [srv.submit(q, "TTTT") for q in qs]
        for q, t in zip(qs, tickets):
            info, hits = srv.result(t)
            assert info["n_hits"] > 3
            eng.set_query(q, "blosum62", 3, 1, "TTTT")
            s, er, eq, _ = eng.align_db(db)
            for x, y in zip(hits, _expected_hits(s, er, eq, q, 0.0)):
                assert np.array_equal(x, y)
    finally:
        srv.close()


def test_server_rejects_profile_beyond_lds(setup):
    eng, w, db = setup
    srv = ish_amd.Server(db, "blosum62", 3, 1, 0.8)
    try:
        with pytest.raises(ish_amd.IshError, match="RANGE"):
            srv.submit(workloads.query(5, 3, 5000), "FFTT")
        # the server keeps serving after a rejected request
        info, _ = srv.result(srv.submit(workloads.query(5, 4, 100), "FFTT"))
        assert info["status"] == 0
    finally:
        srv.close()
