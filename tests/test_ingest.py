"""Record reader of the file runner (ish_hip_records_*, host only).

The expected records come from a Python restatement of the reference's
kseq port (ishlib/vendor/kseq.mojo:265-461: BufferedReader.read_until with
its trailing-'\\r' rule, FastxReader.read with read_comment=False) and of
peek_file (ishlib/peek_file.mojo:37-98), driven the way the runners drive
them (ishlib/parallel_line_search_runner.mojo:83-103: records until
read_until < 0; ishlib/parallel_fastx_search_runner.mojo:93-123: records
while read() > 0). Hand-written edge cases plus seeded random files, plain
and gzip.
"""
import gzip
import os
import zlib

import numpy as np
import pytest

import ish_amd


class KseqStream:
    """BufferedReader over a byte string (buffer boundaries do not change
    what read_until / read_byte return)."""

    def __init__(self, data: bytes):
        self.d = data
        self.pos = 0

    def read_byte(self):
        if self.pos >= len(self.d):
            return -1
        c = self.d[self.pos]
        self.pos += 1
        return c

    def read_until_newline(self, out: bytearray):
        # kseq.mojo:270-325: nothing left -> -1; else append up to '\n'
        # (consumed, not kept); drop a trailing '\r' when len(out) > 1
        if self.pos >= len(self.d):
            return -1
        j = self.d.find(b"\n", self.pos)
        if j < 0:
            out += self.d[self.pos:]
            self.pos = len(self.d)
        else:
            out += self.d[self.pos:j]
            self.pos = j + 1
        if len(out) > 1 and out[-1] == 13:
            del out[-1]
        return len(out)


class KseqFastx:
    """FastxReader.read, read_comment=False (kseq.mojo:370-461)."""

    def __init__(self, stream):
        self.s = stream
        self.last = 0
        self.name = bytearray()
        self.seq = bytearray()
        self.qual = bytearray()

    def read(self):
        s = self.s
        if self.last == 0:
            c = s.read_byte()
            while c >= 0 and c not in (62, 64):
                c = s.read_byte()
            if c < 0:
                return -1
        self.seq = bytearray()
        self.qual = bytearray()
        self.name = bytearray()
        if s.read_until_newline(self.name) < 0:
            return -1
        c = s.read_byte()
        while c >= 0 and c not in (62, 43, 64):
            if c == 13:
                c = s.read_byte()
                continue
            self.seq.append(c)
            s.read_until_newline(self.seq)
            c = s.read_byte()
        if c in (62, 64):
            self.last = c
        if c != 43:
            return len(self.seq)
        c = s.read_byte()
        while c >= 0 and c != 10:
            c = s.read_byte()
        if c < 0:
            return -2
        r = s.read_until_newline(self.qual)
        while r >= 0 and len(self.qual) < len(self.seq):
            r = s.read_until_newline(self.qual)
        if r < 0:
            return -3
        self.last = 0
        if len(self.qual) != len(self.seq):
            return -2
        return len(self.seq)


def expected_records(data: bytes, record_type: str):
    head = data[:8192]
    binary = 0 in head
    rd = KseqFastx(KseqStream(head))
    fastq = False
    while rd.read() > 0:
        fastq = fastq or len(rd.qual) > 0
    flags = (1 if binary else 0) | (2 if fastq else 0)
    recs, status = [], 0
    if record_type == "line":
        st = KseqStream(data)
        while True:
            buf = bytearray()
            if st.read_until_newline(buf) < 0:
                break
            recs.append((bytes(buf), b"", b""))
    else:
        rd = KseqFastx(KseqStream(data))
        while True:
            r = rd.read()
            if r <= 0:
                status = r if r < -1 else 0
                break
            recs.append((bytes(rd.seq), bytes(rd.name), bytes(rd.qual)))
    return recs, flags, status


EDGE = [
    b"",
    b"\n",
    b"\r\n",
    b"\r",
    b"a\r\n\r\nb",
    b"one\ntwo\nthree",
    b"one\ntwo\nthree\n",
    b"\n\n\n",
    b">r1 comment here\nACGT\nAC\r\n>r2\nGG\n",
    b"junk before\n>r1\nACGT\n",
    b">r1\r\nAC\r\nGT\r\n>r2\r\nTT\r\n",
    b">r1\nAC\n\nGT\n>r2\nA\n",
    b">empty\n>r2\nACGT\n",
    b"@q1\nACGT\n+\nIIII\n@q2\nAC\n+q2\nII\n",
    b"@q1\nACGTACGT\n+\nIIII\nIIII\n@q2\nA\n+\nI\n",
    b"@q1\nACGT\n+\nII\n",
    b"@q1\nACGT\n+",
    b"@q1\nACGT\n+\n",
    b"@q1\nAC\n+\nII\n>r2\nGGG\n",
    b">r1\nAC\x00GT\n",
    b">r1\nACGT",
    b"  >not a header\n>r1\nA\n",
]


def _b(rng, alphabet, n):
    """n random bytes drawn from alphabet."""
    return bytes(np.asarray(rng.choice(list(alphabet), n), np.uint8)) if n else b""


def _write(tmp_path, name, data, gz):
    p = tmp_path / (name + (".gz" if gz else ""))
    if gz:
        with gzip.open(p, "wb") as f:
            f.write(data)
    else:
        p.write_bytes(data)
    return str(p)


@pytest.mark.parametrize("gz", [False, True])
@pytest.mark.parametrize("rtype", ["line", "fastx"])
def test_reader_edge_cases(tmp_path, rtype, gz):
    for k, data in enumerate(EDGE):
        path = _write(tmp_path, f"e{k}", data, gz)
        got = ish_amd.read_records(path, rtype)
        exp = expected_records(data, rtype)
        assert got == exp, (k, data)


def _random_file(rng, kind):
    parts = []
    alphabet = b"ACGTNacgtn"
    for _ in range(int(rng.integers(0, 40))):
        r = rng.random()
        if kind == "fastq" and r < 0.8:
            L = int(rng.integers(0, 30))
            seq = _b(rng, alphabet, L)
            qlen = L if rng.random() < 0.9 else int(rng.integers(0, 30))
            q = _b(rng, b"!#5I", qlen)
            split = int(rng.integers(0, qlen + 1))
            qtxt = q[:split] + (b"\n" + q[split:] if split < qlen and rng.random() < 0.3 else q[split:])
            parts.append(b"@r%d x\n" % len(parts) + seq + b"\n+" + (b"r" if rng.random() < 0.5 else b"") + b"\n" + qtxt + b"\n")
        elif r < 0.9:
            lines = []
            for _ in range(int(rng.integers(0, 4))):
                L = int(rng.integers(0, 25))
                ln = _b(rng, alphabet, L)
                if rng.random() < 0.15:
                    ln += b"\r"
                lines.append(ln)
            hdr = b">s%d" % len(parts) + (b" some comment" if rng.random() < 0.5 else b"")
            if rng.random() < 0.1:
                hdr += b"\r"
            parts.append(hdr + b"\n" + b"\n".join(lines) + (b"\n" if rng.random() < 0.9 else b""))
        else:
            parts.append(_b(rng, b"xy \r\n>@+\x00" if rng.random() < 0.2 else b"xy \r\n>@+", int(rng.integers(0, 12))))
    return b"".join(parts)


@pytest.mark.parametrize("kind", ["fasta", "fastq"])
def test_reader_random_files(tmp_path, kind):
    rng = np.random.default_rng(zlib.crc32(kind.encode()))
    for k in range(150):
        data = _random_file(rng, kind)
        gz = k % 3 == 0
        path = _write(tmp_path, f"r{k}", data, gz)
        for rtype in ("fastx", "line"):
            assert ish_amd.read_records(path, rtype) == expected_records(data, rtype), (k, rtype, data[:200])


def test_reader_large_file_buffer_boundaries(tmp_path):
    """Lines and records spanning the reader's 64 KiB refills and the 8 KiB
    peek window."""
    rng = np.random.default_rng(5)
    parts = []
    for k in range(3000):
        L = int(rng.integers(0, 300))
        parts.append(b">r%d\n" % k + _b(rng, b"ACGT", L) + (b"\r\n" if k % 7 == 0 else b"\n"))
    data = b"".join(parts)
    for gz in (False, True):
        path = _write(tmp_path, "big", data, gz)
        for rtype in ("fastx", "line"):
            assert ish_amd.read_records(path, rtype) == expected_records(data, rtype)


def test_reader_missing_file(tmp_path):
    with pytest.raises(ish_amd.IshError):
        ish_amd.read_records(str(tmp_path / "nope.fa"), "fastx")


# The parallel reader of plain files (segments parsed by worker threads and
# stitched at the serial parse's record starts) must give the records of the
# serial restatement whatever the segment size and thread count; tiny
# segments put a boundary inside every kind of line.
PARALLEL = [(3, 1), (4, 7), (2, 64), (8, 4096)]


def _adversarial(rng, kind):
    """Candidate record starts that are not record starts: '@' at the start
    of quality lines, '>' lines swallowed after an empty or '\\r' line, '+'
    and '>' inside qualities, truncated records."""
    parts = []
    for k in range(int(rng.integers(1, 60))):
        if kind == "fastq":
            L = int(rng.integers(0, 20))
            seq = _b(rng, b"ACGT", L)
            q = _b(rng, b"@>+I", L)
            if rng.random() < 0.5 and L > 2:
                q = q[:L // 2] + b"\n" + q[L // 2:]  # quality over two lines
            parts.append(b"@q%d\n" % k + seq + b"\n+\n" + q + b"\n")
            if rng.random() < 0.1:
                parts.append(b"@" + _b(rng, b"I@", int(rng.integers(0, 6))) + b"\n")
        else:
            body = b"\n".join(_b(rng, b"ACGT", int(rng.integers(0, 15))) for _ in range(int(rng.integers(0, 3))))
            sep = [b"\n", b"\n\n", b"\n\r\n", b"\r\n", b"\n\r\r\n"][int(rng.integers(0, 5))]
            parts.append(b">f%d c\n" % k + body + sep)
    data = b"".join(parts)
    if rng.random() < 0.3:
        data = data[:int(rng.integers(0, len(data) + 1))]
    return data


@pytest.mark.parametrize("threads,segment", PARALLEL)
def test_parallel_reader_edge_and_random(tmp_path, threads, segment):
    rng = np.random.default_rng(threads * 1000 + segment)
    files = list(EDGE)
    files += [_random_file(rng, kind) for kind in ("fasta", "fastq") for _ in range(40)]
    files += [_adversarial(rng, kind) for kind in ("fasta", "fastq") for _ in range(40)]
    for k, data in enumerate(files):
        path = _write(tmp_path, f"p{k}", data, False)
        for rtype in ("fastx", "line"):
            got = ish_amd.read_records(path, rtype, threads=threads, segment_bytes=segment)
            assert got == expected_records(data, rtype), (k, rtype, threads, segment, data[:200])


def test_parallel_reader_many_segments(tmp_path):
    """A file of a few thousand records over hundreds of segments and
    several threads, FASTA and FASTQ, against the serial restatement."""
    rng = np.random.default_rng(11)
    for kind in ("fasta", "fastq"):
        parts = []
        for k in range(4000):
            L = int(rng.integers(0, 200))
            s = _b(rng, b"ACGT", L)
            if kind == "fastq":
                parts.append(b"@r%d\n" % k + s + b"\n+\n" + _b(rng, b"@I#", L) + b"\n")
            else:
                parts.append(b">r%d\n" % k + b"\n".join(s[i:i + 60] for i in range(0, L, 60)) + b"\n")
        data = b"".join(parts)
        path = _write(tmp_path, kind, data, False)
        exp = {rt: expected_records(data, rt) for rt in ("fastx", "line")}
        for threads, segment in ((8, 1000), (3, 777), (16, 50000)):
            for rtype in ("fastx", "line"):
                got = ish_amd.read_records(path, rtype, threads=threads, segment_bytes=segment)
                assert got == exp[rtype], (kind, rtype, threads, segment)
