// File search runner: reads line or FASTA/FASTQ records (plain or gzip),
// encodes them, scores each batch on the GPU and writes the emitted records
// in input order, with the output bytes of the reference's parallel runners
// in non-tty mode:
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
#include <zlib.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unistd.h>
#include <vector>

#include "../../include/ish_hip.h"

namespace {

constexpr int kPeekSize = 8192;
constexpr size_t kReadBuf = 1 << 16;

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Buffered reader over zlib with the reference's read_until semantics.
struct Reader {
    gzFile f = nullptr;
    std::vector<uint8_t> buf;
    size_t start = 0, end = 0;
    bool eof = false;
    bool error = false;

    explicit Reader(gzFile fh) : f(fh), buf(kReadBuf) {}

    bool fill() {
        start = 0;
        const int n = gzread(f, buf.data(), (unsigned)buf.size());
        if (n < 0) {
            error = true;
            end = 0;
            eof = true;
            return false;
        }
        end = (size_t)n;
        if (end < buf.size()) eof = true;
        return end > 0;
    }

    // Next byte, or -1 at the end of the stream.
    int read_byte() {
        if (start >= end) {
            if (eof) return -1;
            if (!fill()) return -1;
        }
        return buf[start++];
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
            const uint8_t *p = buf.data() + start;
            const uint8_t *q = (const uint8_t *)memchr(p, '\n', end - start);
            const size_t r = q ? (size_t)(q - buf.data()) : end;
            out.insert(out.end(), buf.data() + start, buf.data() + r);
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

    explicit FastxParser(Reader &r) : rd(r) {}

    // >= 0: sequence length; -1 end of file; -2 truncated quality; -3 error
    int64_t read() {
        if (last_char == 0) {
            int c = rd.read_byte();
            while (c >= 0 && c != '>' && c != '@') c = rd.read_byte();
            if (c < 0) return -1;
        }
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

struct Batch {
    std::vector<uint8_t> res;       // encoded residues
    std::vector<uint64_t> off{0};   // record offsets into res
    std::vector<uint8_t> names;     // fastx names
    std::vector<uint64_t> noff{0};
    std::vector<uint8_t> quals;     // fastq qualities
    std::vector<uint64_t> qoff{0};
    bool last = false;
    int32_t error = 0;
    size_t records() const { return off.size() - 1; }
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
    Reader rd(nullptr);
    rd.buf = head;
    rd.start = 0;
    rd.end = head.size();
    rd.eof = true;
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
    gzFile f = gzopen(path, "rb");
    if (!f) return ISH_E_INVALID;
    gzbuffer(f, 1 << 20);
    const int64_t batch_bytes = cfg->batch_bytes > 0 ? cfg->batch_bytes : ((int64_t)256 << 20);

    Queue queue;
    double t_parse = 0;
    std::thread parser([&] {
        const double t0 = now_s();
        RecordSource src(f, cfg->record_type);
        auto b = std::make_unique<Batch>();
        bool more = true;
        while (more) {
            const std::vector<uint8_t> *seq = nullptr, *name = nullptr, *qual = nullptr;
            if (!src.next(seq, name, qual)) {
                more = false;
                b->error = src.status;
            } else {
                // raw bytes: the device encodes them after the upload
                b->res.insert(b->res.end(), seq->begin(), seq->end());
                b->off.push_back(b->res.size());
                if (name) {
                    b->names.insert(b->names.end(), name->begin(), name->end());
                    b->noff.push_back(b->names.size());
                    b->quals.insert(b->quals.end(), qual->begin(), qual->end());
                    b->qoff.push_back(b->quals.size());
                }
            }
            const int64_t bytes = (int64_t)(b->res.size() + b->names.size() + b->quals.size());
            if (!more || bytes >= batch_bytes) {
                b->last = !more;
                queue.push(std::move(b));
                b = std::make_unique<Batch>();
            }
        }
        t_parse = now_s() - t0;
    });

    Writer out(out_fd);
    rc = ISH_OK;
    for (;;) {
        std::unique_ptr<Batch> b = queue.pop();
        const int64_t n = (int64_t)b->records();
        if (rc == ISH_OK && n > 0) {
            const double tg = now_s();
            std::vector<int32_t> score((size_t)n), er((size_t)n), eq((size_t)n);
            std::vector<uint8_t> emit((size_t)n);
            ish_hip_db *db = nullptr;
            if (b->res.empty()) b->res.push_back(0);
            rc = ish_hip_db_load_ascii(ctx, b->res.data(), b->off.data(), n, cfg->matrix_kind, &db);
            if (rc == ISH_OK) rc = ish_hip_align_db(ctx, db, score.data(), er.data(), eq.data(), nullptr);
            if (db) ish_hip_db_free(db);
            if (rc == ISH_OK)
                rc = ish_hip_emit(cfg->algo, score.data(), er.data(), eq.data(), n, max_score,
                                  cfg->threshold, emit.data());
            const double tw = now_s();
            st.seconds_gpu += tw - tg;
            if (rc == ISH_OK) {
                std::vector<uint8_t> tmp;
                for (int64_t k = 0; k < n; ++k) {
                    if (!emit[(size_t)k]) continue;
                    st.matches++;
                    const uint64_t a = b->off[(size_t)k], e = b->off[(size_t)k + 1];
                    tmp.resize(e - a);
                    for (uint64_t x = a; x < e; ++x) tmp[x - a] = dec[enc[b->res[x]]];
                    if (cfg->record_type == ISH_RECORD_LINE) {
                        out.put(tmp.data(), tmp.size());
                        out.put("\n");
                    } else {
                        out.put(">");
                        out.put(b->names.data() + b->noff[(size_t)k], b->noff[(size_t)k + 1] - b->noff[(size_t)k]);
                        out.put("\n");
                        out.put(tmp.data(), tmp.size());
                        if (fastq) {
                            out.put("\n+\n");
                            out.put(b->quals.data() + b->qoff[(size_t)k], b->qoff[(size_t)k + 1] - b->qoff[(size_t)k]);
                        }
                        out.put("\n");
                    }
                }
            }
            st.seconds_write += now_s() - tw;
            st.records += n;
            st.residues += (int64_t)(b->off.back());
            st.batches++;
        }
        // like the reference's runners, a reader error ends the file
        // without failing the search; its code is reported in the stats
        if (b->error) st.reader_status = b->error;
        if (b->last) break;
    }
    parser.join();
    gzclose(f);
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

extern "C" int32_t ish_hip_records_read(const char *path, int32_t record_type, ish_hip_records **out) {
    if (!path || !out || (record_type != ISH_RECORD_LINE && record_type != ISH_RECORD_FASTX))
        return ISH_E_INVALID;
    *out = nullptr;
    const Peek pk = peek_file(path);
    if (!pk.ok) return ISH_E_INVALID;
    gzFile f = gzopen(path, "rb");
    if (!f) return ISH_E_INVALID;
    auto r = std::make_unique<ish_hip_records>();
    r->flags = (pk.binary ? 1 : 0) | (pk.fastq ? 2 : 0);
    RecordSource src(f, record_type);
    const std::vector<uint8_t> *seq = nullptr, *name = nullptr, *qual = nullptr;
    while (src.next(seq, name, qual)) {
        r->seqs.insert(r->seqs.end(), seq->begin(), seq->end());
        r->soff.push_back(r->seqs.size());
        if (name) r->names.insert(r->names.end(), name->begin(), name->end());
        r->noff.push_back(r->names.size());
        if (qual) r->quals.insert(r->quals.end(), qual->begin(), qual->end());
        r->qoff.push_back(r->quals.size());
    }
    r->status = src.status;
    gzclose(f);
    *out = r.release();
    return ISH_OK;
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
