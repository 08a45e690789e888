// File search runner: reads line or FASTA/FASTQ records (plain or gzip),
// encodes them, scores each batch on the GPU and writes the emitted records
// in input order, with the output bytes of the reference's parallel runners
// in non-tty mode (and, with ish_hip_search_config.tty, in their terminal
// format with the match span coloured):
//   line  (ishlib/parallel_line_search_runner.mojo:75-161):
//         decode(encode(line)) "\n"
//   fastx (ishlib/parallel_fastx_search_runner.mojo:79-182):
//         ">" name "\n" decode(encode(seq)) ["\n+\n" qual] "\n"
// Record parsing follows the reference's kseq port
// (ishlib/vendor/kseq.mojo:265-461, read_comment=False as the runners use
// it), including its edge cases: a trailing '\r' is dropped from a line of
// two or more bytes, '\r' at the start of a sequence line is skipped, an
// empty line inside a sequence contributes a '\n', and a FASTA/FASTQ record
// with an empty sequence ends the file. Binary files (a NUL byte in the
// first 8192 bytes) are skipped in line mode (ishlib/peek_file.mojo:37-58);
// whether the records are FASTQ is decided by the first 8192 bytes of the
// file, as peek_file does for the first file.
//
// Pipeline: a parser thread reads and splits records into batches of raw
// bytes (two in flight) while the calling thread uploads the previous batch
// (the device translates the bytes to residue codes), scores it on the GPU
// and writes its output.
//
// Plain (uncompressed) files are parsed in parallel: the file is mapped,
// cut into segments, and each segment is parsed by a worker thread from its
// first candidate record start (a line start; for FASTA/FASTQ a '>' or '@'
// at a line start). A candidate need not be a record start of the serial
// parse (a '>' line swallowed after an empty line, an '@' that begins a
// quality line), so segment k's records are used only from the position
// where the serial parse, continued from segment k-1, starts its next
// record: records from that position on are the serial parse's own, since a
// parse that has just read a record's header character is in the same
// state however it got there. When segment k has no record there, it is
// parsed again from that position. gzip streams are read serially.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unistd.h>
#include <vector>

#include "../../include/ish_hip.h"
#include "pieces.h"

namespace {

constexpr int kPeekSize = 8192;
constexpr size_t kReadBuf = 1 << 16;

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Buffered reader with the reference's read_until semantics, over zlib or
// over bytes in memory (a mapped file: positions are then file offsets).
struct Reader {
    gzFile f = nullptr;
    std::vector<uint8_t> own; // zlib reads land here
    const uint8_t *base = nullptr;
    size_t start = 0, end = 0;
    bool eof = false;
    bool error = false;

    explicit Reader(gzFile fh) : f(fh), own(kReadBuf) { base = own.data(); }
    Reader(const uint8_t *data, size_t from, size_t to)
        : base(data), start(from), end(to), eof(true) {}

    bool fill() {
        if (!f) {
            eof = true;
            return false;
        }
        start = 0;
        base = own.data();
        const int n = gzread(f, own.data(), (unsigned)own.size());
        if (n < 0) {
            error = true;
            end = 0;
            eof = true;
            return false;
        }
        end = (size_t)n;
        if (end < own.size()) eof = true;
        return end > 0;
    }

    // Next byte, or -1 at the end of the stream.
    int read_byte() {
        if (start >= end) {
            if (eof) return -1;
            if (!fill()) return -1;
        }
        return base[start++];
    }

    // Appends bytes up to (not including) the next '\n' to out, consuming the
    // '\n'. Returns -1 when nothing was left to read, else out's length.
    // A trailing '\r' is dropped when out holds more than one byte.
    int64_t read_until_newline(std::vector<uint8_t> &out) {
        bool got_any = false;
        for (;;) {
            if (start >= end) {
                if (eof) break;
                if (!fill()) break;
            }
            got_any = true;
            const uint8_t *p = base + start;
            const uint8_t *q = (const uint8_t *)memchr(p, '\n', end - start);
            const size_t r = q ? (size_t)(q - base) : end;
            out.insert(out.end(), base + start, base + r);
            start = r + 1;
            if (r < end) break;
        }
        if (!got_any && eof) return -1;
        if (out.size() > 1 && out.back() == '\r') out.pop_back();
        return (int64_t)out.size();
    }
};

// One record of a FASTA/FASTQ stream (FastxReader.read, read_comment=False).
struct FastxParser {
    Reader &rd;
    int last_char = 0;
    std::vector<uint8_t> name, seq, qual;
    // memory readers: position of the last record's header character, and
    // a position at or after which read() stops before the record's body
    int64_t rec_pos = -1;
    int64_t limit = INT64_MAX;

    explicit FastxParser(Reader &r) : rd(r) {}

    // >= 0: sequence length; -1 end of file; -2 truncated quality; -3 error;
    // -4 the next record starts at or after `limit` (at rec_pos)
    int64_t read() {
        if (last_char == 0) {
            int c = rd.read_byte();
            while (c >= 0 && c != '>' && c != '@') c = rd.read_byte();
            if (c < 0) return -1;
        }
        rec_pos = (int64_t)rd.start - 1;
        if (rec_pos >= limit) return -4;
        seq.clear();
        qual.clear();
        name.clear();
        if (rd.read_until_newline(name) < 0) return -1;
        int c = rd.read_byte();
        while (c >= 0 && c != '>' && c != '+' && c != '@') {
            if (c == '\r') {
                c = rd.read_byte();
                continue;
            }
            seq.push_back((uint8_t)c);
            rd.read_until_newline(seq);
            c = rd.read_byte();
        }
        if (c == '>' || c == '@') last_char = c;
        if (c != '+') return (int64_t)seq.size();
        // FASTQ: skip the rest of the '+' line, then read the quality
        c = rd.read_byte();
        while (c >= 0 && c != '\n') c = rd.read_byte();
        if (c < 0) return -2;
        int64_t r = rd.read_until_newline(qual);
        while (r >= 0 && qual.size() < seq.size()) r = rd.read_until_newline(qual);
        if (r < 0) return -3;
        last_char = 0;
        if (qual.size() != seq.size()) return -2;
        return (int64_t)seq.size();
    }
};

// Records as the parser produced them.
struct Recs {
    std::vector<uint8_t> res;       // record bytes (the device encodes them)
    std::vector<uint64_t> off{0};   // record offsets into res
    std::vector<uint8_t> names;     // fastx names
    std::vector<uint64_t> noff{0};
    std::vector<uint8_t> quals;     // fastq qualities
    std::vector<uint64_t> qoff{0};
    size_t count() const { return off.size() - 1; }
    uint64_t bytes(size_t i) const {
        return off[i + 1] - off[i] + (noff.size() > i + 1 ? noff[i + 1] - noff[i] : 0) +
               (qoff.size() > i + 1 ? qoff[i + 1] - qoff[i] : 0);
    }
};

// Records parsed from one segment of a mapped file (or one run of a
// serial read).
struct Segment {
    Recs recs;
    std::vector<int64_t> start; // file offset where each record starts
    int64_t stop = -1;          // start of the first record at or after the limit
    bool ended = false;         // the file ends inside this segment
    int32_t status = 0;         // the reader's code when it ended
};

// A batch: records [first, end) of each part's segment, in order. The
// records are not copied into the batch; the upload gathers them.
struct Batch {
    struct Part {
        std::shared_ptr<const Segment> seg;
        size_t first, end;
    };
    std::vector<Part> parts;
    size_t nrec = 0;
    int64_t bytes = 0;
    bool last = false;
    int32_t error = 0;
    size_t records() const { return nrec; }
};

struct Queue {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::unique_ptr<Batch>> q;
    size_t cap = 2;
    void push(std::unique_ptr<Batch> b) {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return q.size() < cap; });
        q.push_back(std::move(b));
        cv.notify_all();
    }
    std::unique_ptr<Batch> pop() {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return !q.empty(); });
        auto b = std::move(q.front());
        q.pop_front();
        cv.notify_all();
        return b;
    }
};

struct Peek {
    bool binary = false, fastq = false, ok = true;
};

Peek peek_file(const char *path) {
    Peek pk;
    gzFile f = gzopen(path, "rb");
    if (!f) {
        pk.ok = false;
        return pk;
    }
    std::vector<uint8_t> head(kPeekSize);
    const int n = gzread(f, head.data(), kPeekSize);
    gzclose(f);
    if (n < 0) {
        pk.ok = false;
        return pk;
    }
    head.resize((size_t)n);
    pk.binary = memchr(head.data(), 0, head.size()) != nullptr;
    // FASTQ if a record of the first bytes has a quality string (peek_file
    // parses only what the first read returned, with the same record rules)
    Reader rd(head.data(), 0, head.size());
    FastxParser fp(rd);
    while (fp.read() > 0) pk.fastq = pk.fastq || !fp.qual.empty();
    return pk;
}

// Reads the next record into (seq, name, qual). Returns false at the end
// of the records; *status then holds the reader's code (0 at a clean end).
struct RecordSource {
    Reader rd;
    FastxParser fp;
    int32_t type;
    std::vector<uint8_t> line;
    int32_t status = 0;
    RecordSource(gzFile f, int32_t record_type) : rd(f), fp(rd), type(record_type) {}
    bool next(const std::vector<uint8_t> *&seq, const std::vector<uint8_t> *&name,
              const std::vector<uint8_t> *&qual) {
        if (type == ISH_RECORD_LINE) {
            line.clear();
            if (rd.read_until_newline(line) < 0) {
                if (rd.error) status = ISH_E_INVALID;
                return false;
            }
            seq = &line;
            name = qual = nullptr;
            return true;
        }
        const int64_t r = fp.read();
        if (r <= 0) { // end of file, error, or an empty sequence
            status = r < -1 ? (int32_t)r : 0;
            if (rd.error) status = -3;
            return false;
        }
        seq = &fp.seq;
        name = &fp.name;
        qual = &fp.qual;
        return true;
    }
};

// ---------------------------------------------------------------------------
// Batches: the records of a file, in order, cut into batches of about
// batch_bytes (a batch is handed over once its bytes reach batch_bytes; the
// last one, possibly empty, has last = true and the reader's status).
// ---------------------------------------------------------------------------
using BatchSink = std::function<void(std::unique_ptr<Batch>)>;

// Parses records of m[from, n) as the serial reader does from a record
// start, until the next record starts at or after `limit`.
void parse_segment(const uint8_t *m, size_t n, int32_t type, size_t from, int64_t limit,
                   Segment &sg) {
    Reader rd(m, from, n);
    Recs &b = sg.recs;
    if (type == ISH_RECORD_LINE) {
        std::vector<uint8_t> line;
        for (;;) {
            if ((int64_t)rd.start >= limit) {
                sg.stop = (int64_t)rd.start;
                return;
            }
            const int64_t at = (int64_t)rd.start;
            line.clear();
            if (rd.read_until_newline(line) < 0) {
                sg.ended = true;
                return;
            }
            sg.start.push_back(at);
            b.res.insert(b.res.end(), line.begin(), line.end());
            b.off.push_back(b.res.size());
        }
    }
    FastxParser fp(rd);
    fp.limit = limit;
    for (;;) {
        const int64_t r = fp.read();
        if (r == -4) {
            sg.stop = fp.rec_pos;
            return;
        }
        if (r <= 0) { // end of file, an empty sequence, or a reader error
            sg.ended = true;
            sg.status = r < -1 ? (int32_t)r : 0;
            return;
        }
        sg.start.push_back(fp.rec_pos);
        b.res.insert(b.res.end(), fp.seq.begin(), fp.seq.end());
        b.off.push_back(b.res.size());
        b.names.insert(b.names.end(), fp.name.begin(), fp.name.end());
        b.noff.push_back(b.names.size());
        b.quals.insert(b.quals.end(), fp.qual.begin(), fp.qual.end());
        b.qoff.push_back(b.quals.size());
    }
}

// Cuts records into batches of about batch_bytes.
// Largest batch the adaptive rule asks for (the reference's default batch
// size, searcher_settings.mojo:138).
constexpr int64_t kMaxAdaptiveBatch = (int64_t)4 << 30;

struct Batcher {
    const BatchSink &sink;
    int64_t batch_bytes;
    // > 0: a batch also holds at least this many residues per residue of
    // the longest record seen so far (up to kMaxAdaptiveBatch)
    int64_t per_longest;
    int64_t longest = 0;
    std::unique_ptr<Batch> b = std::make_unique<Batch>();
    Batcher(const BatchSink &s, int64_t bb, int64_t pl) : sink(s), batch_bytes(bb), per_longest(pl) {}
    int64_t limit() const {
        return std::max(batch_bytes, std::min(per_longest * longest, kMaxAdaptiveBatch));
    }
    // records [i, count) of sg
    void add(std::shared_ptr<const Segment> sg, size_t i) {
        const Recs &r = sg->recs;
        const size_t j = r.count();
        while (i < j) {
            // records up to the one that brings the batch to its limit
            size_t e = i;
            while (e < j && b->bytes < limit()) {
                const int64_t L = (int64_t)r.bytes(e++);
                b->bytes += L;
                longest = std::max(longest, L);
            }
            b->parts.push_back({sg, i, e});
            b->nrec += e - i;
            i = e;
            if (b->bytes >= limit()) {
                sink(std::move(b));
                b = std::make_unique<Batch>();
            }
        }
    }
    void finish(int32_t status) {
        b->last = true;
        b->error = status;
        sink(std::move(b));
    }
};

// Serial producer (gzip streams): records go into segments of about
// batch_bytes, each handed over as it fills.
void produce_serial(gzFile f, int32_t type, int64_t batch_bytes, int64_t per_longest,
                    const BatchSink &sink) {
    RecordSource src(f, type);
    const bool fastx = type == ISH_RECORD_FASTX;
    Batcher bt(sink, batch_bytes, per_longest);
    auto sg = std::make_shared<Segment>();
    int64_t bytes = 0;
    const std::vector<uint8_t> *seq = nullptr, *name = nullptr, *qual = nullptr;
    while (src.next(seq, name, qual)) {
        Recs &r = sg->recs;
        r.res.insert(r.res.end(), seq->begin(), seq->end());
        r.off.push_back(r.res.size());
        bytes += (int64_t)seq->size();
        if (fastx) {
            r.names.insert(r.names.end(), name->begin(), name->end());
            r.noff.push_back(r.names.size());
            r.quals.insert(r.quals.end(), qual->begin(), qual->end());
            r.qoff.push_back(r.quals.size());
            bytes += (int64_t)(name->size() + qual->size());
        }
        if (bytes >= batch_bytes) {
            bt.add(std::move(sg), 0);
            sg = std::make_shared<Segment>();
            bytes = 0;
        }
    }
    bt.add(std::move(sg), 0);
    bt.finish(src.status);
}

struct Mapped {
    int fd = -1;
    const uint8_t *p = nullptr;
    size_t n = 0;
    bool open(const char *path) {
        fd = ::open(path, O_RDONLY);
        if (fd < 0) return false;
        struct stat sb;
        if (fstat(fd, &sb) != 0) return false;
        n = (size_t)sb.st_size;
        if (n == 0) return true;
        void *m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) return false;
        p = (const uint8_t *)m;
        (void)madvise(m, n, MADV_SEQUENTIAL);
        return true;
    }
    ~Mapped() {
        if (p) munmap((void *)p, n);
        if (fd >= 0) ::close(fd);
    }
};

// First candidate record start at or after s (0 at the start of the file,
// n when there is none).
size_t candidate(const uint8_t *m, size_t n, int32_t type, size_t s) {
    if (s == 0) return 0;
    size_t pos = s - 1; // a line starts at s when m[s-1] is '\n'
    while (pos < n) {
        const uint8_t *q = (const uint8_t *)memchr(m + pos, '\n', n - pos);
        if (!q) return n;
        const size_t y = (size_t)(q - m) + 1;
        if (y >= n) return n;
        if (type == ISH_RECORD_LINE || m[y] == '>' || m[y] == '@') return y;
        pos = y;
    }
    return n;
}

// Parallel producer over a mapped plain file (see the top of the file).
void produce_parallel(const Mapped &mf, int32_t type, int64_t batch_bytes, int64_t per_longest, int threads,
                      int64_t seg_bytes, const BatchSink &sink) {
    const uint8_t *m = mf.p;
    const size_t n = mf.n;
    const size_t nseg = n == 0 ? 1 : (size_t)((n + (size_t)seg_bytes - 1) / (size_t)seg_bytes);
    // candidate starts, each at or after the previous one (segments may be
    // empty); cand[nseg] = n
    std::vector<size_t> cand(nseg + 1, n);
    cand[0] = 0;
    for (size_t k = 1; k < nseg; ++k) {
        const size_t s = k * (size_t)seg_bytes;
        cand[k] = cand[k - 1] >= s ? cand[k - 1] : candidate(m, n, type, s);
    }
    auto limit_of = [&](size_t k) -> int64_t { return k + 1 < nseg ? (int64_t)cand[k + 1] : INT64_MAX; };

    const size_t window = 2 * (size_t)threads; // segments parsed ahead of the stitcher
    std::vector<std::unique_ptr<Segment>> seg(nseg);
    std::vector<char> ready(nseg, 0);
    std::mutex mu;
    std::condition_variable cv;
    size_t consumed = 0;
    bool stop = false;
    std::atomic<size_t> next{0};
    auto worker = [&] {
        for (;;) {
            const size_t k = next.fetch_add(1);
            if (k >= nseg) return;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || k < consumed + window; });
                if (stop) return;
            }
            auto sg = std::make_unique<Segment>();
            parse_segment(m, n, type, cand[k], limit_of(k), *sg);
            std::lock_guard<std::mutex> lk(mu);
            seg[k] = std::move(sg);
            ready[k] = 1;
            cv.notify_all();
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(worker);

    Batcher bt(sink, batch_bytes, per_longest);
    int32_t status = 0;
    int64_t expected = 0; // where the serial parse starts its next record
    for (size_t k = 0; k < nseg; ++k) {
        std::unique_ptr<Segment> sg;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return ready[k] != 0; });
            sg = std::move(seg[k]);
        }
        size_t first = 0;
        if (expected != (int64_t)cand[k]) {
            // the serial parse continued past this candidate: use the
            // segment from its record at `expected`, or parse it again
            // from there
            const auto it = std::lower_bound(sg->start.begin(), sg->start.end(), expected);
            if (it != sg->start.end() && *it == expected) {
                first = (size_t)(it - sg->start.begin());
            } else {
                sg = std::make_unique<Segment>();
                parse_segment(m, n, type, (size_t)expected, limit_of(k), *sg);
            }
        }
        const bool ended = sg->ended;
        status = sg->status;
        expected = sg->stop;
        bt.add(std::shared_ptr<const Segment>(std::move(sg)), first);
        {
            std::lock_guard<std::mutex> lk(mu);
            consumed = k + 1;
            cv.notify_all();
        }
        if (ended) break;
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
        cv.notify_all();
    }
    for (auto &t : pool) t.join();
    bt.finish(status);
}

bool is_gzip(const char *path) {
    uint8_t h[2] = {0, 0};
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    const ssize_t r = ::read(fd, h, 2);
    ::close(fd);
    return r == 2 && h[0] == 0x1f && h[1] == 0x8b;
}

// Every record of path, in batches. Returns 0 or ISH_E_INVALID when the
// file cannot be opened.
int32_t produce_batches(const char *path, int32_t type, int64_t batch_bytes, int64_t per_longest, int threads,
                        int64_t seg_bytes, const BatchSink &sink) {
    if (threads <= 0) threads = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (seg_bytes <= 0) seg_bytes = (int64_t)8 << 20;
    if (!is_gzip(path)) {
        Mapped mf;
        if (mf.open(path)) {
            produce_parallel(mf, type, batch_bytes, per_longest, threads, seg_bytes, sink);
            return ISH_OK;
        }
    }
    gzFile f = gzopen(path, "rb");
    if (!f) return ISH_E_INVALID;
    gzbuffer(f, 1 << 20);
    produce_serial(f, type, batch_bytes, per_longest, sink);
    gzclose(f);
    return ISH_OK;
}

struct Writer {
    int fd;
    std::vector<uint8_t> buf;
    bool failed = false;
    explicit Writer(int f) : fd(f) { buf.reserve(1 << 20); }
    void put(const uint8_t *p, size_t n) {
        buf.insert(buf.end(), p, p + n);
        if (buf.size() >= (1 << 20)) flush();
    }
    void put(const char *s) { put((const uint8_t *)s, strlen(s)); }
    void flush() {
        size_t o = 0;
        while (o < buf.size() && !failed) {
            const ssize_t w = ::write(fd, buf.data() + o, buf.size() - o);
            if (w <= 0) failed = true;
            else o += (size_t)w;
        }
        buf.clear();
    }
};

} // namespace

extern "C" int32_t ish_hip_search_file(ish_hip_ctx *ctx, const char *path, const uint8_t *pattern,
                                       int32_t plen, const ish_hip_search_config *cfg,
                                       int32_t out_fd, ish_hip_search_stats *stats) {
    if (!ctx || !path || !cfg || plen < 0 || (plen > 0 && !pattern)) return ISH_E_INVALID;
    if (cfg->record_type != ISH_RECORD_LINE && cfg->record_type != ISH_RECORD_FASTX) return ISH_E_INVALID;
    const double t_start = now_s();
    ish_hip_search_stats st{};
    uint8_t enc[256], dec[256];
    if (ish_hip_encoding(cfg->matrix_kind, enc) != ISH_OK) return ISH_E_INVALID;
    const int32_t A = ish_hip_decoding(cfg->matrix_kind, dec, 256);
    if (A < 0) return ISH_E_INVALID;
    std::vector<uint8_t> q((size_t)std::max(plen, 1));
    const int64_t max_score = ish_hip_encode_and_score(cfg->matrix_kind, pattern, plen, q.data());
    int32_t rc = ish_hip_set_query(ctx, q.data(), plen, nullptr, 0, cfg->matrix_kind, cfg->gap_open,
                                   cfg->gap_extend, cfg->ends_free, cfg->algo);
    if (rc) return rc;

    const Peek pk = peek_file(path);
    if (!pk.ok) return ISH_E_INVALID;
    if (cfg->record_type == ISH_RECORD_LINE && pk.binary) {
        if (stats) *stats = st;
        return ISH_OK; // binary files are skipped
    }
    const bool fastq = pk.fastq;
    {
        const int fd = ::open(path, O_RDONLY);
        if (fd < 0) return ISH_E_INVALID;
        ::close(fd);
    }
    const int64_t batch_bytes = cfg->batch_bytes > 0 ? cfg->batch_bytes : ((int64_t)256 << 20);
    // With the default size, a batch grows to hold at least cus x 256
    // residues per residue of the longest record seen: each batch's pass
    // ends with its longest group pipelined over four waves, and a smaller
    // batch would be mostly that tail.
    const int64_t per_longest = cfg->batch_bytes > 0 ? 0 : (int64_t)ish::ctx_cus(ctx) * 256;

    Queue queue;
    double t_parse = 0;
    std::thread parser([&] {
        const double t0 = now_s();
        const BatchSink sink = [&](std::unique_ptr<Batch> b) { queue.push(std::move(b)); };
        if (produce_batches(path, cfg->record_type, batch_bytes, per_longest, cfg->reader_threads,
                            cfg->reader_segment_bytes, sink) != ISH_OK) {
            auto b = std::make_unique<Batch>();
            b->last = true;
            b->error = -3;
            queue.push(std::move(b));
        }
        t_parse = now_s() - t0;
    });

    Writer out(out_fd);
    rc = ISH_OK;
    // Batch k+1 is uploaded and built (copy stream) while batch k is
    // scored, and queued behind it; batch k's results are then fetched on
    // the copy stream as its pass ends, and its matches written while
    // batch k+1 scores. Batches use the context's two one-shot buffer
    // slots in turn, so nothing is allocated or freed between passes.
    std::unique_ptr<Batch> cur; // the batch whose database is being scored
    ish_hip_db *cur_db = nullptr;
    int cur_slot = 0, next_slot = 0;
    int64_t cur_first = 0; // records of the file before cur
    std::vector<int32_t> score, er, eq;
    std::vector<uint8_t> emit;
    // tty output: start coordinates of cur's emitted records (per record;
    // -1 elsewhere)
    const bool tty = cfg->tty != 0;
    std::vector<int32_t> tstart;
    // fetches cur's results (and, for tty output, the start coordinates of
    // its emitted records, from the residues still on the device) and drops
    // its database; false when no matches are to be written
    auto fetch = [&]() {
        if (!cur_db) return false;
        int64_t n = 0;
        (void)ish_hip_db_size(cur_db, &n, nullptr);
        score.resize((size_t)n);
        er.resize((size_t)n);
        eq.resize((size_t)n);
        emit.resize((size_t)n);
        const double tg = now_s();
        int32_t r = ish::db_fetch_after(ctx, cur_db, cur_slot, score.data(), er.data(), eq.data(), nullptr);
        if (r == ISH_OK)
            r = ish_hip_emit(cfg->algo, score.data(), er.data(), eq.data(), n, max_score,
                             cfg->threshold, emit.data());
        if (r == ISH_OK && tty) {
            std::vector<int64_t> sel;
            std::vector<int32_t> ser, seq;
            for (int64_t t = 0; t < n; ++t)
                if (emit[(size_t)t]) {
                    sel.push_back(t);
                    ser.push_back(er[(size_t)t]);
                    seq.push_back(eq[(size_t)t]);
                }
            std::vector<int32_t> ts(sel.size()), qs(sel.size());
            r = ish_hip_find_starts(ctx, cur_db, sel.data(), ser.data(), seq.data(), (int64_t)sel.size(),
                                    ts.data(), qs.data());
            tstart.assign((size_t)n, -1);
            for (size_t k = 0; k < sel.size(); ++k) tstart[(size_t)sel[k]] = ts[k];
        }
        ish_hip_db_free(cur_db);
        cur_db = nullptr;
        st.seconds_gpu += now_s() - tg;
        if (r != ISH_OK && rc == ISH_OK) rc = r;
        return r == ISH_OK;
    };
    // the runners' tty colours (ishlib/__init__.mojo:5-10)
    static const char kPurple[] = "\x1b[35m", kGreen[] = "\x1b[32m", kRed[] = "\x1b[31m",
                      kReset[] = "\x1b[0m", kBoldRed[] = "\x1b[1;31m";
    const size_t path_len = strlen(path);
    auto write = [&](const Batch &b) {
        const double tw = now_s();
        std::vector<uint8_t> tmp;
        size_t k = 0;
        for (const Batch::Part &pt : b.parts) {
            const Recs &r = pt.seg->recs;
            for (size_t x = pt.first; x < pt.end; ++x, ++k) {
                if (!emit[k]) continue;
                st.matches++;
                const uint64_t a = r.off[x], e = r.off[x + 1];
                tmp.resize(e - a);
                for (uint64_t y = a; y < e; ++y) tmp[y - a] = dec[enc[r.res[y]]];
                // the match span [start, end_ref + 1) of MatchResult
                size_t m0 = 0, m1 = 0;
                if (tty) {
                    m1 = (size_t)std::min<int64_t>(std::max<int64_t>((int64_t)er[k] + 1, 0), (int64_t)tmp.size());
                    m0 = (size_t)std::min<int64_t>(std::max<int64_t>(tstart[k], 0), (int64_t)m1);
                }
                if (cfg->record_type == ISH_RECORD_LINE) {
                    if (tty) {
                        // PURPLE file RESET ":" GREEN line RESET ": " pre RED match RESET post
                        // (parallel_line_search_runner.mojo:123-150)
                        char num[32];
                        snprintf(num, sizeof num, "%lld", (long long)(cur_first + (int64_t)k + 1));
                        out.put(kPurple);
                        out.put((const uint8_t *)path, path_len);
                        out.put(kReset);
                        out.put(":");
                        out.put(kGreen);
                        out.put(num);
                        out.put(kReset);
                        out.put(": ");
                        out.put(tmp.data(), m0);
                        out.put(kRed);
                        out.put(tmp.data() + m0, m1 - m0);
                        out.put(kReset);
                        out.put(tmp.data() + m1, tmp.size() - m1);
                    } else {
                        out.put(tmp.data(), tmp.size());
                    }
                    out.put("\n");
                } else {
                    out.put(">");
                    out.put(r.names.data() + r.noff[x], r.noff[x + 1] - r.noff[x]);
                    out.put("\n");
                    if (tty) {
                        // pre BOLD-RED RED match RESET post
                        // (parallel_fastx_search_runner.mojo:150-167)
                        out.put(tmp.data(), m0);
                        out.put(kBoldRed);
                        out.put(kRed);
                        out.put(tmp.data() + m0, m1 - m0);
                        out.put(kReset);
                        out.put(tmp.data() + m1, tmp.size() - m1);
                    } else {
                        out.put(tmp.data(), tmp.size());
                    }
                    if (fastq) {
                        out.put("\n+\n");
                        out.put(r.quals.data() + r.qoff[x], r.qoff[x + 1] - r.qoff[x]);
                    }
                    out.put("\n");
                }
            }
        }
        st.seconds_write += now_s() - tw;
    };
    rc = ish::one_shot_begin(ctx);
    for (;;) {
        std::unique_ptr<Batch> b = queue.pop();
        const int64_t n = (int64_t)b->records();
        ish_hip_db *db = nullptr;
        if (rc == ISH_OK && n > 0) {
            const double tg = now_s();
            // the batch's record bytes, piece by piece, and their offsets
            std::vector<ish::HostPiece> pieces;
            std::vector<uint64_t> off((size_t)n + 1, 0);
            uint64_t acc = 0;
            size_t k = 0;
            for (const Batch::Part &pt : b->parts) {
                const Recs &r = pt.seg->recs;
                const uint64_t a = r.off[pt.first];
                for (size_t x = pt.first; x < pt.end; ++x) off[++k] = acc + (r.off[x + 1] - a);
                pieces.push_back({r.res.data() + a, r.off[pt.end] - a});
                acc += r.off[pt.end] - a;
            }
            rc = ish::db_load_ascii_pieces(ctx, pieces.data(), pieces.size(), off.data(), n,
                                           cfg->matrix_kind, &db, next_slot);
            st.seconds_gpu += now_s() - tg;
            st.records += n;
            st.residues += (int64_t)acc;
            st.batches++;
        }
        if (db && rc == ISH_OK) {
            const double tg = now_s();
            rc = ish::db_score_once_async(ctx, db, next_slot);
            st.seconds_gpu += now_s() - tg;
        }
        const bool have = fetch(); // the pass before this one
        if (rc != ISH_OK && db) {
            (void)ish_hip_sync(ctx);
            ish_hip_db_free(db);
            db = nullptr;
        }
        if (have && rc == ISH_OK) write(*cur);
        if (cur) cur_first += (int64_t)cur->records();
        cur = std::move(b);
        cur_db = db;
        if (db) {
            cur_slot = next_slot;
            next_slot ^= 1;
        }
        // like the reference's runners, a reader error ends the file
        // without failing the search; its code is reported in the stats
        if (cur->error) st.reader_status = cur->error;
        if (cur->last) break;
    }
    if (fetch() && rc == ISH_OK) write(*cur);
    ish::one_shot_end(ctx);
    parser.join();
    out.flush();
    if (out.failed && rc == ISH_OK) rc = ISH_E_INVALID;
    st.cells = (double)st.residues * (double)plen;
    st.seconds_parse = t_parse;
    st.seconds_total = now_s() - t_start;
    if (stats) *stats = st;
    return rc;
}

struct ish_hip_records {
    std::vector<uint8_t> seqs, names, quals;
    std::vector<uint64_t> soff{0}, noff{0}, qoff{0};
    int32_t flags = 0, status = 0;
};

extern "C" int32_t ish_hip_records_read_ex(const char *path, int32_t record_type, int32_t threads,
                                           int64_t segment_bytes, ish_hip_records **out) {
    if (!path || !out || (record_type != ISH_RECORD_LINE && record_type != ISH_RECORD_FASTX))
        return ISH_E_INVALID;
    *out = nullptr;
    const Peek pk = peek_file(path);
    if (!pk.ok) return ISH_E_INVALID;
    auto r = std::make_unique<ish_hip_records>();
    r->flags = (pk.binary ? 1 : 0) | (pk.fastq ? 2 : 0);
    const bool fastx = record_type == ISH_RECORD_FASTX;
    const BatchSink sink = [&](std::unique_ptr<Batch> b) {
        for (const Batch::Part &pt : b->parts) {
            const Recs &s = pt.seg->recs;
            for (size_t x = pt.first; x < pt.end; ++x) {
                r->seqs.insert(r->seqs.end(), s.res.begin() + (ptrdiff_t)s.off[x],
                               s.res.begin() + (ptrdiff_t)s.off[x + 1]);
                r->soff.push_back(r->seqs.size());
                if (fastx) {
                    r->names.insert(r->names.end(), s.names.begin() + (ptrdiff_t)s.noff[x],
                                    s.names.begin() + (ptrdiff_t)s.noff[x + 1]);
                    r->quals.insert(r->quals.end(), s.quals.begin() + (ptrdiff_t)s.qoff[x],
                                    s.quals.begin() + (ptrdiff_t)s.qoff[x + 1]);
                }
                r->noff.push_back(r->names.size());
                r->qoff.push_back(r->quals.size());
            }
        }
        if (b->last) r->status = b->error;
    };
    if (produce_batches(path, record_type, (int64_t)64 << 20, 0, threads, segment_bytes, sink) != ISH_OK)
        return ISH_E_INVALID;
    *out = r.release();
    return ISH_OK;
}

extern "C" int32_t ish_hip_records_read(const char *path, int32_t record_type, ish_hip_records **out) {
    return ish_hip_records_read_ex(path, record_type, 0, 0, out);
}

extern "C" int64_t ish_hip_records_count(const ish_hip_records *r) {
    return r ? (int64_t)r->soff.size() - 1 : ISH_E_INVALID;
}
extern "C" int32_t ish_hip_records_flags(const ish_hip_records *r) { return r ? r->flags : ISH_E_INVALID; }
extern "C" int32_t ish_hip_records_status(const ish_hip_records *r) { return r ? r->status : ISH_E_INVALID; }

extern "C" int32_t ish_hip_records_get(const ish_hip_records *r, int64_t i, const uint8_t **seq,
                                       int64_t *seq_len, const uint8_t **name, int64_t *name_len,
                                       const uint8_t **qual, int64_t *qual_len) {
    if (!r || i < 0 || i >= (int64_t)r->soff.size() - 1) return ISH_E_INVALID;
    const size_t k = (size_t)i;
    if (seq) *seq = r->seqs.data() + r->soff[k];
    if (seq_len) *seq_len = (int64_t)(r->soff[k + 1] - r->soff[k]);
    if (name) *name = r->names.data() + r->noff[k];
    if (name_len) *name_len = (int64_t)(r->noff[k + 1] - r->noff[k]);
    if (qual) *qual = r->quals.data() + r->qoff[k];
    if (qual_len) *qual_len = (int64_t)(r->qoff[k + 1] - r->qoff[k]);
    return ISH_OK;
}

extern "C" void ish_hip_records_free(ish_hip_records *r) { delete r; }
