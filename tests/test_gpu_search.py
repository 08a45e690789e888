"""File search runner (ish_hip_search_file) end to end on the GPU.

- The reference's searcher fixtures (tests/test_searcher.mojo:17-200): 4
  inputs x {striped-local, striped-semi-global} x {TTTT, FFTT}, from plain
  and gzip files, output compared exactly (line) or case-insensitively
  (fastx, as the reference's test does).
- Seeded random FASTA / FASTQ / line files with planted matches, read in
  many small batches: the output must equal, byte for byte, the records the
  restated reader yields, emitted by the CPU oracle's scores, in the
  runners' format.
"""
import gzip
import json
import os
import zlib

import numpy as np
import pytest

import ish_amd
from oracle import oracle as O
from test_ingest import expected_records

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def eng():
    if ish_amd.device_count() == 0:
        pytest.fail("no gfx950 device visible to the HIP library")
    e = ish_amd.Engine(0)
    yield e
    e.close()


def _write(path, data, gz):
    if gz:
        with gzip.open(path, "wb") as f:
            f.write(data)
    else:
        with open(path, "wb") as f:
            f.write(data)


def _search(eng, tmp_path, path, **kw):
    out = tmp_path / "out.txt"
    fd = os.open(out, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
    try:
        st = eng.search_file(path, out_fd=fd, **kw)
    finally:
        os.close(fd)
    return out.read_bytes(), st


@pytest.mark.parametrize("gz", [False, True])
def test_searcher_fixtures_from_files(eng, tmp_path, gz):
    cases = json.load(open(os.path.join(HERE, "golden", "searcher_cases.json")))
    for c in cases["cases"]:
        path = str(tmp_path / ("in" + (".gz" if gz else "")))
        _write(path, cases["inputs"][c["input"]].encode(), gz)
        for algo in ("striped-local", "striped-semi-global"):
            for ends in ("TTTT", "FFTT"):
                got, st = _search(eng, tmp_path, path, pattern=c["pattern"].encode(), matrix_kind=c["matrix"],
                                  algo=algo, ends_free=ends, threshold=0.8, record_type=c["record_type"])
                exp = c["expected"]
                g = got.decode()
                assert (g == exp) if c["record_type"] == "line" else (g.upper() == exp.upper()), \
                    (c["input"], algo, ends, g[:200])


PURPLE, GREEN, RED, RESET, BOLD_RED = b"\x1b[35m", b"\x1b[32m", b"\x1b[31m", b"\x1b[0m", b"\x1b[1;31m"


def _expected_output(data, rtype, pattern, kind, algo, ends, thr, tty_path=None):
    """The runners' output for the oracle's results; with tty_path, their
    terminal format (ishlib/parallel_line_search_runner.mojo:123-150,
    parallel_fastx_search_runner.mojo:150-167) with the match span
    [start, end_ref + 1) from the oracle's start/end functions."""
    recs, flags, _ = expected_records(data, rtype)
    if rtype == "line" and flags & 1:
        return b""
    m = O.Matrix(kind)
    q = m.encode(pattern)
    enc = ish_amd.encoding(kind)
    dec = ish_amd.decoding(kind)
    seqs = [enc[np.frombuffer(s, np.uint8)] if s else np.zeros(0, np.uint8) for s, _, _ in recs]
    T = ish_amd.Targets.from_list(seqs)
    fl = O.flags_from_str(ends)
    s, er, eq, _ = O.batch(1 if algo == "striped-local" else 0, q, T.residues, T.offsets, m, 3, 1, fl)
    ms = int(sum(int(m.values.reshape(m.A, m.A)[c, c]) for c in q))
    emit = ish_amd.emit(algo, s, er, eq, ms, thr)
    out = []
    fastq = bool(flags & 2)
    for k, (seq, name, qual) in enumerate(recs):
        if not emit[k]:
            continue
        body = bytes(dec[seqs[k]])
        if tty_path is not None:
            if algo == "striped-local":
                a = O.sw_align(q, seqs[k], m, 3, 1, cutoff=-1.0)
                m0, m1 = a["ref_begin1"], a["ref_end1"] + 1
            else:
                se = O.sg_start_end(q, seqs[k], m, 3, 1, fl)
                m0, m1 = se["target_start"], se["target_end"] + 1
            span = (body[:m0], body[m0:m1], body[m1:])
            if rtype == "line":
                out.append(PURPLE + tty_path + RESET + b":" + GREEN + str(k + 1).encode() + RESET + b": " +
                           span[0] + RED + span[1] + RESET + span[2] + b"\n")
            else:
                out.append(b">" + name + b"\n" + span[0] + BOLD_RED + RED + span[1] + RESET + span[2] +
                           ((b"\n+\n" + qual) if fastq else b"") + b"\n")
            continue
        if rtype == "line":
            out.append(body + b"\n")
        else:
            out.append(b">" + name + b"\n" + body + ((b"\n+\n" + qual) if fastq else b"") + b"\n")
    return b"".join(out)


def _b(rng, alphabet, n):
    return bytes(np.asarray(rng.choice(list(alphabet), n), np.uint8)) if n else b""


def _random_records(rng, kind, n, pattern):
    parts = []
    for k in range(n):
        L = int(rng.integers(20, 400))
        seq = bytearray(_b(rng, b"ACGTacgtN", L))
        if k % 5 == 0 and L > len(pattern):
            o = int(rng.integers(0, L - len(pattern) + 1))
            seq[o:o + len(pattern)] = pattern
            for _ in range(int(rng.integers(0, 3))):  # a few substitutions
                seq[o + int(rng.integers(0, len(pattern)))] = ord("ACGT"[int(rng.integers(0, 4))])
        if kind == "fastq":
            parts.append(b"@read%d extra\n" % k + bytes(seq) + b"\n+\n" + _b(rng, b"!5?I", L) + b"\n")
        elif kind == "fasta":
            cut = int(rng.integers(1, L))
            parts.append(b">seq%d desc\n" % k + bytes(seq[:cut]) + b"\n" + bytes(seq[cut:]) + b"\n")
        else:
            parts.append(bytes(seq) + (b"\r\n" if k % 9 == 0 else b"\n"))
    return b"".join(parts)


@pytest.mark.parametrize("kind,rtype", [("fastq", "fastx"), ("fasta", "fastx"), ("line", "line")])
def test_random_files_match_oracle(eng, tmp_path, kind, rtype):
    rng = np.random.default_rng(zlib.crc32(kind.encode()))
    pattern = _b(rng, b"ACGT", 40)
    data = _random_records(rng, kind, 3000, pattern)
    for gz in (False, True):
        path = str(tmp_path / ("in" + (".gz" if gz else "")))
        _write(path, data, gz)
        for algo, ends in (("striped-semi-global", "FFTT"), ("striped-local", "FFFF"), ("striped-semi-global", "TTTT")):
            got, st = _search(eng, tmp_path, path, pattern=pattern, matrix_kind="actgn", algo=algo, ends_free=ends,
                              threshold=0.7, record_type=rtype, batch_bytes=50000)
            exp = _expected_output(data, rtype, pattern, "actgn", algo, ends, 0.7)
            assert st["batches"] > 5 and st["records"] == 3000
            assert got == exp, (kind, gz, algo, ends, len(got), len(exp))
            assert st["matches"] > 0


@pytest.mark.parametrize("kind,rtype", [("fastq", "fastx"), ("fasta", "fastx"), ("line", "line")])
def test_tty_format_matches_oracle_spans(eng, tmp_path, kind, rtype):
    """The terminal format: line numbers across small batches, and the match
    span of every emitted record from the GPU start pass equal to the
    oracle's start/end coordinates."""
    rng = np.random.default_rng(zlib.crc32(b"tty" + kind.encode()))
    pattern = _b(rng, b"ACGT", 40)
    data = _random_records(rng, kind, 1500, pattern)
    path = str(tmp_path / "in.txt")
    _write(path, data, False)
    for algo, ends in (("striped-semi-global", "FFTT"), ("striped-local", "FFFF"), ("striped-semi-global", "TTTT")):
        got, st = _search(eng, tmp_path, path, pattern=pattern, matrix_kind="actgn", algo=algo, ends_free=ends,
                          threshold=0.7, record_type=rtype, batch_bytes=40000, tty=True)
        exp = _expected_output(data, rtype, pattern, "actgn", algo, ends, 0.7, tty_path=path.encode())
        assert st["batches"] > 3 and st["matches"] > 0
        assert got == exp, (kind, algo, ends, got[:300], exp[:300])


def test_tty_format_searcher_fixtures(eng, tmp_path):
    cases = json.load(open(os.path.join(HERE, "golden", "searcher_cases.json")))
    for c in cases["cases"]:
        path = str(tmp_path / "in")
        data = cases["inputs"][c["input"]].encode()
        _write(path, data, False)
        for algo in ("striped-local", "striped-semi-global"):
            for ends in ("TTTT", "FFTT"):
                got, _ = _search(eng, tmp_path, path, pattern=c["pattern"].encode(), matrix_kind=c["matrix"],
                                 algo=algo, ends_free=ends, threshold=0.8, record_type=c["record_type"], tty=True)
                exp = _expected_output(data, c["record_type"], c["pattern"].encode(), c["matrix"], algo, ends, 0.8,
                                       tty_path=path.encode())
                assert got == exp, (c["input"], algo, ends, got[:300], exp[:300])


def test_binary_file_skipped_in_line_mode(eng, tmp_path):
    path = str(tmp_path / "bin.txt")
    _write(path, b"ACGT\nAC\x00GT\nACGTACGT\n", False)
    got, st = _search(eng, tmp_path, path, pattern=b"ACGT", matrix_kind="actgn", record_type="line")
    assert got == b"" and st["records"] == 0
