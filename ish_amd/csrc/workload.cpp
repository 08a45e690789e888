// Deterministic synthetic workloads for the benchmark and the tests, in
// encoded residue codes (ACTGN: A0 C1 G2 T3 N4; BLOSUM62: index in
// ARNDCQEGHILKMFPSTWYVBZX*). Every target draws from its own counter-based
// stream (splitmix64 of seed and target index), so the output does not
// depend on the number of threads.
//
// Workloads (parameters the task leaves open are fixed here and recorded):
//  1 nucleotide line records: query 150 nt iid ACGT; targets U[100,300]
//    iid ACGT with N at 0.5 %; 1 % of targets carry a mutated copy of the
//    query (5 % substitutions to a different base, indel events at 1 % per
//    position with length U{1,2,3}, insertion/deletion 50/50, inserted bases
//    uniform), overwritten at an offset U[0, R-|copy|] (a copy longer than
//    the target contributes its first R residues).
//  2 protein FASTA: query 350 aa at Robinson-Robinson background
//    frequencies; targets with ln R ~ N(ln 350, 0.6^2) rounded and clipped
//    to [30, 5000], same frequencies; 0.1 % carry a homolog of the query
//    (each position substituted with probability 0.30 by b != a drawn with
//    weight p_b * 2^(S62(a,b)/2)) at an offset U[0, R-Q] (first R residues
//    of the homolog when R < Q).
//  3 multi-query targets: the records of 1 (including its planted copies of
//    the 150 nt query of 1), with the letters A, C, G, T, N encoded as the
//    BLOSUM62 residues of those letters (A0 C4 G7 T16 N2); the queries are
//    protein (background frequencies).
//  4 long reads: query 1000 nt iid ACGT; ln R ~ N(ln 8000, 0.5^2) clipped to
//    [1000, 50000]; per base 3 % substitution, 3 % insertion, 2 % deletion,
//    2 % N; 2 % of reads carry the query passed through the same error
//    model.
//  5 serving database: protein targets as in 2 but with the length
//    distribution of 1, R ~ U[100, 300]; 0.1 % carry a homolog of the
//    350 aa base query (its first R residues).
#include <math.h>
#include <stdio.h>
#include <pthread.h>
#include <stdint.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s;
    Rng(uint64_t seed, uint64_t stream) {
        s = seed ^ (stream * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull);
        splitmix64(s);
    }
    uint64_t next() { return splitmix64(s); }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint32_t below(uint32_t n) { return (uint32_t)((next() >> 32) * (uint64_t)n >> 32); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
};

// Robinson & Robinson (1991) amino-acid background frequencies, per mille,
// in BLOSUM62 code order A R N D C Q E G H I L K M F P S T W Y V.
const double kAAFreq[20] = {78.05, 51.29, 44.87, 53.64, 19.25, 42.64, 62.95,
                            73.77, 21.99, 51.42, 90.19, 57.44, 22.43, 38.56,
                            52.03, 71.20, 58.41, 13.30, 32.16, 64.41};

const int8_t kB62[20][20] = {
    {4, -1, -2, -2, 0, -1, -1, 0, -2, -1, -1, -1, -1, -2, -1, 1, 0, -3, -2, 0},
    {-1, 5, 0, -2, -3, 1, 0, -2, 0, -3, -2, 2, -1, -3, -2, -1, -1, -3, -2, -3},
    {-2, 0, 6, 1, -3, 0, 0, 0, 1, -3, -3, 0, -2, -3, -2, 1, 0, -4, -2, -3},
    {-2, -2, 1, 6, -3, 0, 2, -1, -1, -3, -4, -1, -3, -3, -1, 0, -1, -4, -3, -3},
    {0, -3, -3, -3, 9, -3, -4, -3, -3, -1, -1, -3, -1, -2, -3, -1, -1, -2, -2, -1},
    {-1, 1, 0, 0, -3, 5, 2, -2, 0, -3, -2, 1, 0, -3, -1, 0, -1, -2, -1, -2},
    {-1, 0, 0, 2, -4, 2, 5, -2, 0, -3, -3, 1, -2, -3, -1, 0, -1, -3, -2, -2},
    {0, -2, 0, -1, -3, -2, -2, 6, -2, -4, -4, -2, -3, -3, -2, 0, -2, -2, -3, -3},
    {-2, 0, 1, -1, -3, 0, 0, -2, 8, -3, -3, -1, -2, -1, -2, -1, -2, -2, 2, -3},
    {-1, -3, -3, -3, -1, -3, -3, -4, -3, 4, 2, -3, 1, 0, -3, -2, -1, -3, -1, 3},
    {-1, -2, -3, -4, -1, -2, -3, -4, -3, 2, 4, -2, 2, 0, -3, -2, -1, -2, -1, 1},
    {-1, 2, 0, -1, -3, 1, 1, -2, -1, -3, -2, 5, -1, -3, -1, 0, -1, -3, -2, -2},
    {-1, -1, -2, -3, -1, 0, -2, -3, -2, 1, 2, -1, 5, 0, -2, -1, -1, -1, -1, 1},
    {-2, -3, -3, -3, -2, -3, -3, -3, -1, 0, 0, -3, 0, 6, -4, -2, -2, 1, 3, -1},
    {-1, -2, -2, -1, -3, -1, -1, -2, -2, -3, -3, -1, -2, -4, 7, -1, -1, -4, -3, -2},
    {1, -1, 1, 0, -1, 0, 0, 0, -1, -2, -2, 0, -1, -2, -1, 4, 1, -3, -2, -2},
    {0, -1, 0, -1, -1, -1, -1, -2, -2, -1, -1, -1, -1, -2, -1, 1, 5, -2, -2, 0},
    {-3, -3, -4, -4, -2, -2, -3, -2, -2, -3, -2, -3, -1, 1, -4, -3, -2, 11, 2, -3},
    {-2, -2, -2, -3, -2, -1, -2, -3, 2, -1, -1, -2, -1, 3, -3, -2, -2, 2, 7, -1},
    {0, -3, -3, -3, -1, -2, -2, -3, -3, 3, 1, -2, 1, -1, -2, -2, 0, -3, -1, 4},
};

struct AATables {
    double cdf[20];
    double sub_cdf[20][20]; // P(b | a), b != a, weight p_b 2^(S(a,b)/2)
    AATables() {
        double tot = 0;
        for (int k = 0; k < 20; ++k) tot += kAAFreq[k];
        double acc = 0;
        for (int k = 0; k < 20; ++k) {
            acc += kAAFreq[k] / tot;
            cdf[k] = acc;
        }
        cdf[19] = 1.0;
        for (int a = 0; a < 20; ++a) {
            double w[20], s = 0;
            for (int b = 0; b < 20; ++b) {
                w[b] = b == a ? 0.0 : kAAFreq[b] * pow(2.0, kB62[a][b] / 2.0);
                s += w[b];
            }
            double c = 0;
            for (int b = 0; b < 20; ++b) {
                c += w[b] / s;
                sub_cdf[a][b] = c;
            }
            sub_cdf[a][19] = 1.0;
            if (a == 19) sub_cdf[a][18] = 1.0;
        }
    }
};
const AATables &aa() {
    static AATables t;
    return t;
}

inline uint8_t draw_cdf(Rng &r, const double *cdf, int n) {
    const double u = r.uniform();
    int k = 0;
    while (k < n - 1 && u >= cdf[k]) ++k;
    return (uint8_t)k;
}

int64_t lognormal_len(Rng &r, double median, double sigma, int64_t lo, int64_t hi) {
    double v = exp(log(median) + sigma * r.normal());
    int64_t L = (int64_t)llround(v);
    return std::min(hi, std::max(lo, L));
}

// Copy of `src` through substitution / indel errors (nucleotides).
void mutate_nt(Rng &r, const uint8_t *src, int64_t n, double p_sub, double p_ins, double p_del,
               double p_indel_event, std::vector<uint8_t> &out) {
    out.clear();
    for (int64_t k = 0; k < n; ++k) {
        if (p_indel_event > 0 && r.uniform() < p_indel_event) {
            const int len = 1 + (int)r.below(3);
            if (r.below(2) == 0) { // insertion
                for (int z = 0; z < len; ++z) out.push_back((uint8_t)r.below(4));
            } else { // deletion
                k += len - 1;
                continue;
            }
        }
        if (p_ins > 0 && r.uniform() < p_ins) out.push_back((uint8_t)r.below(4));
        if (p_del > 0 && r.uniform() < p_del) continue;
        uint8_t c = src[k];
        if (r.uniform() < p_sub) c = (uint8_t)((c + 1 + r.below(3)) & 3);
        out.push_back(c);
    }
}

// Overwrites the target with the planted copy at an offset U[0, R-|copy|];
// a copy longer than the target contributes its window copy[0..R).
void plant(Rng &r, const std::vector<uint8_t> &copy, uint8_t *dst, int64_t R) {
    const int64_t L = std::min<int64_t>((int64_t)copy.size(), R);
    const int64_t off = r.below((uint32_t)(R - L + 1));
    memcpy(dst + off, copy.data(), (size_t)L);
}

struct Job {
    int kind;
    uint64_t seed;
    int64_t n;
    const uint8_t *query;
    int32_t qlen;
    const int64_t *lens;
    const uint64_t *offsets;
    uint8_t *res;
    uint8_t *planted;
    const uint8_t *recode; // code translation applied to every target, or null
    int64_t next;
    pthread_mutex_t mu;
};

void gen_target(const Job &jb, int64_t t) {
    Rng r(jb.seed, (uint64_t)t + 1);
    const int64_t R = jb.lens[t];
    uint8_t *dst = jb.res + jb.offsets[t];
    // advance past the length draw so residues use a fresh part of the stream
    r.next();
    r.next();
    std::vector<uint8_t> copy;
    if (jb.kind == 1) {
        for (int64_t k = 0; k < R; ++k) dst[k] = r.uniform() < 0.005 ? 4 : (uint8_t)r.below(4);
        if (r.uniform() < 0.01) {
            mutate_nt(r, jb.query, jb.qlen, 0.05, 0.0, 0.0, 0.01, copy);
            plant(r, copy, dst, R);
            jb.planted[t] = 1;
        }
    } else if (jb.kind == 2 || jb.kind == 5) {
        const AATables &T = aa();
        for (int64_t k = 0; k < R; ++k) dst[k] = draw_cdf(r, T.cdf, 20);
        if (r.uniform() < 0.001) {
            copy.resize((size_t)jb.qlen);
            for (int32_t k = 0; k < jb.qlen; ++k) {
                uint8_t a = jb.query[k];
                if (r.uniform() < 0.30) a = draw_cdf(r, T.sub_cdf[a], 20);
                copy[(size_t)k] = a;
            }
            plant(r, copy, dst, R);
            jb.planted[t] = 1;
        }
    } else { // 4: long reads
        for (int64_t k = 0; k < R; ++k) dst[k] = (uint8_t)r.below(4);
        if (r.uniform() < 0.02) {
            mutate_nt(r, jb.query, jb.qlen, 0.03, 0.03, 0.02, 0.0, copy);
            plant(r, copy, dst, R);
            jb.planted[t] = 1;
        }
        // read-level errors over the whole read: substitutions already in the
        // background are iid; add N at 2 %
        for (int64_t k = 0; k < R; ++k)
            if (r.uniform() < 0.02) dst[k] = 4;
    }
}

void *worker(void *arg) {
    Job *jb = (Job *)arg;
    for (;;) {
        pthread_mutex_lock(&jb->mu);
        const int64_t lo = jb->next;
        jb->next += 256;
        pthread_mutex_unlock(&jb->mu);
        if (lo >= jb->n) break;
        const int64_t hi = std::min(jb->n, lo + 256);
        for (int64_t t = lo; t < hi; ++t) {
            gen_target(*jb, t);
            if (jb->recode) {
                uint8_t *d = jb->res + jb->offsets[t];
                for (int64_t k = 0; k < jb->lens[t]; ++k) d[k] = jb->recode[d[k]];
            }
        }
    }
    return nullptr;
}

} // namespace

extern "C" {

// Query of a workload (codes). Returns its length.
int32_t ish_wl_query(int32_t kind, uint64_t seed, int32_t qlen, uint8_t *out) {
    Rng r(seed, 0);
    if (kind == 2 || kind == 3 || kind == 5) {
        for (int32_t k = 0; k < qlen; ++k) out[k] = draw_cdf(r, aa().cdf, 20);
    } else {
        for (int32_t k = 0; k < qlen; ++k) out[k] = (uint8_t)r.below(4);
    }
    return qlen;
}

// Target lengths of a workload; returns the total number of residues.
int64_t ish_wl_lengths(int32_t kind, uint64_t seed, int64_t n, int64_t *lens) {
    int64_t tot = 0;
    for (int64_t t = 0; t < n; ++t) {
        Rng r(seed, (uint64_t)t + 1);
        int64_t L;
        if (kind == 1 || kind == 3 || kind == 5) L = 100 + (int64_t)r.below(201);
        else if (kind == 2) L = lognormal_len(r, 350.0, 0.6, 30, 5000);
        else L = lognormal_len(r, 8000.0, 0.5, 1000, 50000);
        lens[t] = L;
        tot += L;
    }
    return tot;
}

// Fills residues (offsets[n] bytes) and the planted flags.
int32_t ish_wl_targets(int32_t kind, uint64_t seed, const uint8_t *query, int32_t qlen,
                       int64_t n, const int64_t *lens, const uint64_t *offsets,
                       uint8_t *residues, uint8_t *planted, int32_t nthreads) {
    Job jb;
    jb.kind = kind == 3 ? 1 : kind;
    jb.seed = seed;
    jb.n = n;
    jb.query = query;
    jb.qlen = qlen;
    // kind 3: the records of kind 1 with its own query planted, letters
    // recoded to BLOSUM62 (ACTGN A0 C1 G2 T3 N4 -> A0 C4 G7 T16 N2)
    uint8_t nt_query[150];
    static const uint8_t kNtToBlosum[256] = {0, 4, 7, 16, 2};
    jb.recode = nullptr;
    if (kind == 3) {
        ish_wl_query(1, seed, 150, nt_query);
        jb.query = nt_query;
        jb.qlen = 150;
        jb.recode = kNtToBlosum;
    }
    jb.lens = lens;
    jb.offsets = offsets;
    jb.res = residues;
    jb.planted = planted;
    jb.next = 0;
    memset(planted, 0, (size_t)n);
    pthread_mutex_init(&jb.mu, nullptr);
    if (nthreads <= 0) nthreads = (int32_t)sysconf(_SC_NPROCESSORS_ONLN);
    nthreads = std::max(1, std::min(nthreads, 64));
    std::vector<pthread_t> th((size_t)nthreads);
    for (auto &t : th) pthread_create(&t, nullptr, worker, &jb);
    for (auto &t : th) pthread_join(t, nullptr);
    pthread_mutex_destroy(&jb.mu);
    return nthreads;
}

// Writes the targets as a FASTA (fastq = 0) or FASTQ (fastq = 1) file:
// record t is named "t<t> synthetic", its residues decoded through
// decode[code], sequence lines of at most `width` residues (0 = one line),
// FASTQ qualities all 'I'. Returns 0, or -1 when the file cannot be written.
int32_t ish_wl_write_fastx(const char *path, const uint8_t *res, const uint64_t *off, int64_t n,
                           const uint8_t *decode, int32_t fastq, int32_t width) {
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    std::vector<char> buf;
    buf.reserve(1 << 22);
    auto flush = [&]() {
        if (!buf.empty() && fwrite(buf.data(), 1, buf.size(), f) != buf.size()) return false;
        buf.clear();
        return true;
    };
    bool ok = true;
    char hdr[64];
    for (int64_t t = 0; t < n && ok; ++t) {
        const int k = snprintf(hdr, sizeof hdr, "%ct%lld synthetic\n", fastq ? '@' : '>', (long long)t);
        buf.insert(buf.end(), hdr, hdr + k);
        const uint64_t a = off[t], e = off[t + 1];
        uint64_t col = 0;
        for (uint64_t x = a; x < e; ++x) {
            buf.push_back((char)decode[res[x]]);
            if (width > 0 && !fastq && ++col == (uint64_t)width && x + 1 < e) {
                buf.push_back('\n');
                col = 0;
            }
        }
        buf.push_back('\n');
        if (fastq) {
            buf.push_back('+');
            buf.push_back('\n');
            buf.insert(buf.end(), (size_t)(e - a), 'I');
            buf.push_back('\n');
        }
        if (buf.size() >= (1 << 22)) ok = flush();
    }
    ok = ok && flush();
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : -1;
}

// FNV-1a 64 of a buffer (content hash recorded with each workload).
uint64_t ish_wl_hash(const uint8_t *p, int64_t n) {
    uint64_t h = 1469598103934665603ull;
    for (int64_t k = 0; k < n; ++k) {
        h ^= p[k];
        h *= 1099511628211ull;
    }
    return h;
}
}
