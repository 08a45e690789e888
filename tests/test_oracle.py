"""Pins the CPU oracle before it is trusted as the checker.

- every known-answer vector transcribed from the reference's own tests;
- the end-to-end emitted record sets of tests/test_searcher.mojo:17-128;
- brute-force enumeration of all alignment paths on tiny inputs (the DP must
  equal the maximum over paths, for every free-end flag combination);
- the scalar textbook DP against the step-by-step striped restatement
  (Farrar + lazy-F) at 8, 16 and 32 lanes, for both algorithms;
- closed forms (exact copy = diagonal sum, empty input = (0,0,0));
- the striped restatement's saturating uint8 / int16 / uint16 lanes against
  the saturating_add / saturating_sub vectors of
  tests/matcher/alignment/test_striped_utils.mojo:9-66 (and against clipped
  integer arithmetic on random pairs), and its biased ssw profile against
  the layout vectors of tests/test_local_and_profile.mojo:19-152.
"""
import json
import os
import random

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
VEC = json.load(open(os.path.join(HERE, "golden", "reference_vectors.json")))
SEARCH = json.load(open(os.path.join(HERE, "golden", "searcher_cases.json")))
ALL_FLAGS = ["".join("T" if (k >> b) & 1 else "F" for b in range(4)) for k in range(16)]


@pytest.mark.parametrize("case", VEC["semi_global"], ids=lambda c: f"{c['src'].split(':')[-1]}-{c['flags']}")
@pytest.mark.parametrize("impl", ["scalar", "w8", "w16", "w32"])
def test_semi_global_known_answers(case, impl):
    m = O.Matrix(case["matrix"])
    q, t = m.encode(case["query"]), m.encode(case["target"])
    f = O.flags_from_str(case["flags"])
    if impl == "scalar":
        s, er, eq = O.sg(q, t, m, case["go"], case["ge"], f)
    else:
        s, er, eq = O.sg_striped(q, t, m, case["go"], case["ge"], f, int(impl[1:]))
    assert s == case["score"]
    if "end_ref" in case:
        assert er == case["end_ref"]
    if "end_query" in case:
        assert eq == case["end_query"]
    if "target_start" in case:
        se = O.sg_start_end(q, t, m, case["go"], case["ge"], f)
        assert (se["target_start"], se["target_end"]) == (case["target_start"], case["target_end"])


@pytest.mark.parametrize("case", VEC["local"], ids=lambda c: c["src"].split(":")[-1])
def test_local_known_answers(case):
    m = O.Matrix(case["matrix"])
    q, t = m.encode(case["read"]), m.encode(case["reference"])
    a = O.sw_align(q, t, m, case["go"], case["ge"])
    assert a == dict(score1=case["score"], ref_begin1=case["ref_begin"], ref_end1=case["ref_end"],
                     read_begin1=case["read_begin"], read_end1=case["read_end"])
    for w in (8, 16, 32):
        assert O.sw_striped(q, t, m, case["go"], case["ge"], w) == (case["score"], case["ref_end"], case["read_end"])


def parse_records(text, rtype):
    """Line records, or FASTA records as (header line, concatenated sequence)."""
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


@pytest.mark.parametrize("ci", range(len(SEARCH["cases"])))
@pytest.mark.parametrize("algo", ["striped-local", "striped-semi-global"])
@pytest.mark.parametrize("ends", ["TTTT", "FFTT"])
def test_searcher_fixtures(ci, algo, ends):
    c = SEARCH["cases"][ci]
    m = O.Matrix(c["matrix"])
    q = m.encode(c["pattern"])
    ms = m.max_score(c["pattern"])
    out = []
    for name, seq in parse_records(SEARCH["inputs"][c["input"]], c["record_type"]):
        t = m.encode(seq.encode("utf-8"))
        if algo == "striped-semi-global":
            s, er, eq = O.sg(q, t, m, 3, 1, O.flags_from_str(ends))
            emit = O.sg_emit(s, ms, 0.8)
        else:
            s, er, eq = O.sw(q, t, m, 3, 1)
            emit = O.sw_emit(s, er, eq, ms, 0.8)
        if emit:
            out.append(name + "\n" if c["record_type"] == "line"
                       else name + "\n" + bytes(m.dec[t]).decode() + "\n")
    got = "".join(out)
    if c["record_type"] == "line":
        assert got == c["expected"]
    else:
        assert got.upper() == c["expected"].upper()


# ---- brute force over alignment paths -------------------------------------
def _paths_best(q, r, S, A, go, ge, qs, ts):
    """H(j,i) = max over every move sequence from (-1,-1) that ends by entering
    (j,i). An interior gap move costs go, or min(go, ge) when it continues a
    gap of the same direction (Gotoh's recurrence may re-open a gap from the
    gap state itself). Runs on the left edge (qs) and top edge (ts) are free
    when flagged, else cost exactly go+(L-1)*ge (the boundary formula)."""
    Q, R = len(q), len(r)
    best = {}

    def walk(j, i, score, last, run):
        # last: 'D','V','H' or None; run: length of the current gap run
        if j >= 0 and i >= 0:
            key = (j, i)
            if score > best.get(key, -10**9):
                best[key] = score
        if j + 1 < Q and i + 1 < R:
            walk(j + 1, i + 1, score + S[r[i + 1] * A + q[j + 1]], "D", 0)
        if j + 1 < Q:  # vertical move
            if i == -1:
                cost = 0 if qs else (ge if last == "V" else go)
            else:
                cost = min(go, ge) if last == "V" else go
            walk(j + 1, i, score - cost, "V", run + 1)
        if i + 1 < R:  # horizontal move
            if j == -1:
                cost = 0 if ts else (ge if last == "H" else go)
            else:
                cost = min(go, ge) if last == "H" else go
            walk(j, i + 1, score - cost, "H", run + 1)

    walk(-1, -1, 0, None, 0)
    return best


def _end_rules(H, Q, R, flags, neg_limit):
    qs, qe, ts, te = (bool(flags >> k & 1) for k in range(4))
    vmax, end_ref = neg_limit, -1
    for i in range(R):
        if H[(Q - 1, i)] > vmax:
            vmax, end_ref = H[(Q - 1, i)], i
    score, end_q = neg_limit, Q - 1
    if te:
        score = vmax
    if qe:
        for j in range(Q):
            v = H[(j, R - 1)]
            if v > score:
                score, end_q, end_ref = v, j, R - 1
            elif v == score and end_ref == R - 1 and j < end_q:
                end_q = j
    if not te and not qe:
        score, end_ref, end_q = H[(Q - 1, R - 1)], R - 1, Q - 1
    return score, end_ref, end_q


@pytest.mark.parametrize("seed", range(6))
def test_semi_global_brute_force(seed):
    rng = random.Random(seed)
    for kind in ("actgn", "blosum62"):
        m = O.Matrix(kind)
        S = m.values.tolist()
        mn = min(0, min(S))
        for _ in range(8):
            Q, R = rng.randint(1, 4), rng.randint(1, 4)
            q = np.array([rng.randrange(min(m.A, 6)) for _ in range(Q)], np.uint8)
            r = np.array([rng.randrange(min(m.A, 6)) for _ in range(R)], np.uint8)
            go, ge = rng.randint(1, 5), rng.randint(1, 3)
            neg_limit = (-32768 + go + 1) if -go < mn else (-32768 - mn + 1)
            for fl in range(16):
                H = _paths_best(q.tolist(), r.tolist(), S, m.A, go, ge, fl & 1, fl >> 2 & 1)
                assert O.sg(q, r, m, go, ge, fl) == _end_rules(H, Q, R, fl, neg_limit), (q, r, go, ge, fl)


@pytest.mark.parametrize("seed", range(4))
def test_local_brute_force(seed):
    """SW H(j,i) = max(0, max over substring starts of the path maximum)."""
    rng = random.Random(100 + seed)
    m = O.Matrix("actgn")
    S = m.values.tolist()
    for _ in range(12):
        Q, R = rng.randint(1, 4), rng.randint(1, 5)
        q = [rng.randrange(5) for _ in range(Q)]
        r = [rng.randrange(5) for _ in range(R)]
        go, ge = rng.randint(1, 4), rng.randint(1, 2)
        H = {(j, i): 0 for j in range(Q) for i in range(R)}
        for a in range(Q):
            for b in range(R):
                sub = _paths_best(q[a:], r[b:], S, m.A, go, ge, 0, 0)
                for (j, i), v in sub.items():
                    H[(j + a, i + b)] = max(H[(j + a, i + b)], v)
        best, end_ref = 0, -1
        for i in range(R):
            cm = max(H[(j, i)] for j in range(Q))
            if cm > best:
                best, end_ref = cm, i
        end_q = 0 if best == 0 else min(j for j in range(Q) if H[(j, end_ref)] == best)
        assert O.sw(np.array(q, np.uint8), np.array(r, np.uint8), m, go, ge) == (best, end_ref, end_q)


# ---- scalar oracle vs striped restatement ----------------------------------
@pytest.mark.parametrize("kind", ["actgn", "actgn0", "blosum62", "ascii"])
def test_scalar_equals_striped(kind):
    rng = np.random.default_rng(7)
    m = O.Matrix(kind)
    alpha = min(m.A, 24) if kind != "ascii" else 256
    for trial in range(60):
        Q = int(rng.integers(1, 90))
        R = int(rng.integers(1, 140))
        q = rng.integers(0, alpha, Q).astype(np.uint8)
        r = rng.integers(0, alpha, R).astype(np.uint8)
        if trial % 3 == 0:  # plant a copy of the query so scores go high
            k = min(Q, R)
            off = int(rng.integers(0, R - k + 1))
            r[off:off + k] = q[:k]
        ge = int(rng.integers(0, 4))
        go = ge + 1 + int(rng.integers(0, 10))  # go > ge: see the next test
        fl = int(rng.integers(0, 16))
        ref = O.sg(q, r, m, go, ge, fl)
        for w in (8, 16, 32):
            assert O.sg_striped(q, r, m, go, ge, fl, w) == ref, (trial, w)
        ref = O.sw(q, r, m, go, ge)
        for w in (8, 16, 32):
            assert O.sw_striped(q, r, m, go, ge, w) == ref, (trial, w)


def test_striped_is_width_dependent_when_open_le_extend():
    """With gap_open <= gap_extend the reference's lazy-F early exit
    (striped.mojo:428-432, local_aln/striped.mojo:413-416) compares F-ge with
    H-go after H was raised by F, and stops before F-go reaches the next lane.
    The reference's answer then depends on its SIMD width, i.e. on the host
    CPU. The engine's contract is textbook Gotoh (the brute-force-pinned
    oracle); bit-exact parity with the reference holds for go > ge."""
    m = O.Matrix("actgn")
    q = np.array([3, 2, 0], np.uint8)
    r = np.array([3], np.uint8)
    # go == ge == 3, FFFT: textbook -4; one row per lane loses F-go at W=8
    assert O.sg(q, r, m, 3, 3, O.flags_from_str("FFFT")) == (-4, 0, 2)
    assert O.sg_striped(q, r, m, 3, 3, O.flags_from_str("FFFT"), 8) == (-8, 0, 2)
    rng = np.random.default_rng(3)
    differ = 0
    for _ in range(400):
        Q, R = int(rng.integers(1, 60)), int(rng.integers(1, 60))
        q = rng.integers(0, 5, Q).astype(np.uint8)
        r = rng.integers(0, 5, R).astype(np.uint8)
        go = int(rng.integers(0, 3))
        ge = go + int(rng.integers(0, 3))
        fl = int(rng.integers(0, 16))
        res = {O.sg_striped(q, r, m, go, ge, fl, w) for w in (8, 16, 32)}
        differ += len(res) > 1
    assert differ > 0


def test_closed_forms():
    m = O.Matrix("blosum62")
    q = m.encode("MAFSAEDVLKEYDRRRRMEALLLSLYYPNDRKLLDYKEWSPPRVQVECPKAPVEWNNPPS")
    diag = int(m.values.reshape(24, 24)[q, q].sum())
    for fl in range(16):
        assert O.sg(q, q, m, 3, 1, fl)[0] == diag
    assert O.sw(q, q, m, 3, 1) == (diag, len(q) - 1, len(q) - 1)
    for fl in range(16):
        assert O.sg(q[:0], q, m, 3, 1, fl) == (0, 0, 0)
        assert O.sg(q, q[:0], m, 3, 1, fl) == (0, 0, 0)
    assert O.sw(q, q[:0], m, 3, 1) == (0, -1, 0)


def test_emission_float32():
    # 0.8 is not exact in float32: f32(300)*f32(0.8) = 240.00000357... rounds to 240.0
    assert O.sg_emit(240, 300, 0.8) is True
    assert O.sg_emit(239, 300, 0.8) is False
    assert O.sw_emit(240, 5, 5, 300, 0.8) is False  # local cutoff is strict
    assert O.sw_emit(241, 5, 5, 300, 0.8) is True
    assert O.sw_emit(241, 0, 5, 300, 0.8) is False  # no begin at ref 0
    assert O.sw_emit(241, 5, 0, 300, 0.8) is False


# ---- the striped restatement's lanes and profile, against the reference's vectors
SAT = VEC["saturating"]
LIMITS = {"u8": (0, 255), "i16": (-32768, 32767), "u16": (0, 65535)}


@pytest.mark.parametrize("type_", ["u8", "i16", "u16"])
@pytest.mark.parametrize("width", [8, 16, 32, 64])
def test_saturating_lanes_reference_vectors(type_, width):
    """saturating_add / saturating_sub (test_striped_utils.mojo:9-66) on the
    lane types the striped restatement uses, at every lane count it can be
    built with (the AVX2 / AVX-512 overloads at their native widths)."""
    lo, hi = LIMITS[type_]
    word = {"MAX": hi, "MIN": lo}
    for c in SAT["cases"]:
        a = np.full(width, word.get(c["lhs"], c["lhs"]), np.int32)
        b = np.full(width, c["rhs"], np.int32)
        r = O.sat_lanes(type_, width, c["op"], a, b)
        assert (r == word.get(c["expected"], c["expected"])).all(), (type_, width, c)


@pytest.mark.parametrize("type_", ["u8", "i16", "u16"])
def test_saturating_lanes_exhaustive_pairs(type_):
    """Beyond the reference's four vectors: every lane clamps exactly like
    integer arithmetic clipped to the type, on random pairs with the limits
    mixed in (a dropped clamp or a wrap fails here)."""
    lo, hi = LIMITS[type_]
    rng = np.random.default_rng(5)
    for width in (8, 16, 32, 64):
        for _ in range(50):
            a = rng.integers(lo, hi + 1, width)
            b = rng.integers(lo, hi + 1, width)
            a[:4] = [lo, hi, lo, hi]
            b[:4] = [hi, hi, lo, lo]
            for op, f in (("add", np.add), ("sub", np.subtract)):
                r = O.sat_lanes(type_, width, op, a, b)
                assert np.array_equal(r, np.clip(f(a, b), lo, hi)), (type_, width, op)


def test_ssw_profile_reference_layout():
    """The biased u8 (16 lanes) and u16 (8 lanes) profiles of
    test_local_and_profile.mojo:19-152."""
    P = VEC["ssw_profile"]
    A = P["alphabet"]
    q = np.array(P["query"], np.uint8)
    S = np.full((A, A), P["mismatch"], np.int8)
    np.fill_diagonal(S, P["match"])
    for type_, lanes, nvec in (("u8", P["u8_lanes"], P["u8_vectors"]), ("u16", P["u16_lanes"], P["u16_vectors"])):
        prof = O.local_profile(q, S.reshape(-1), A, P["bias"], type_, lanes)
        assert prof.shape == (nvec, lanes)
        seg = nvec // A
        for c in range(A):
            want = np.zeros((seg, lanes), np.int32)
            want[c % seg, c // seg] = 3
            assert np.array_equal(prof[c * seg:(c + 1) * seg], want), (type_, c)
        if type_ == "u8":
            for c, vecs in enumerate(P["u8_first_chars"]):
                assert np.array_equal(prof[c * seg:(c + 1) * seg], np.array(vecs)), c


@pytest.mark.parametrize("type_,lanes", [("u8", 64), ("u8", 32), ("u16", 32), ("u16", 16)])
def test_ssw_profile_native_widths(type_, lanes):
    """The same layout rule at the lane counts the timed baseline uses:
    lane k of vector s of char c holds S[c][q[s + k*seg]] + bias, and bias
    past the query."""
    rng = np.random.default_rng(lanes)
    A = 5
    S = rng.integers(-3, 4, (A, A)).astype(np.int8)
    bias = int(-S.min())
    for Q in (1, lanes - 1, lanes, 3 * lanes + 7):
        q = rng.integers(0, A, Q).astype(np.uint8)
        prof = O.local_profile(q, S.reshape(-1), A, bias, type_, lanes)
        seg = (Q + lanes - 1) // lanes
        for c in range(A):
            for s in range(seg):
                for k in range(lanes):
                    j = s + k * seg
                    want = bias if j >= Q else int(S[c, q[j]]) + bias
                    assert prof[c * seg + s, k] == want, (Q, c, s, k)
