// Host side of the ish_hip C ABI (include/ish_hip.h): contexts, query
// profiles, the device-resident target database and kernel dispatch.
#include "../../include/ish_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "engine.h"
#include "host.h"
#include "pieces.h"
#include "tables.h"

using namespace ish;

// Autotuning table: (matrix kind, algorithm, ceil(log2 Q), length bucket
// of the database's longest target) -> the launch configuration of a pass.
constexpr int kLenBuckets = 4;
constexpr int32_t kLenBucketLo[kLenBuckets] = {4096, 1024, 256, 0};
struct TuneChoice {
    int8_t chunk, schedule;
    int8_t waves; // waves per workgroup of pipelined semi-global groups (0: the default rule)
};
struct TuneKey {
    int32_t kind, algo, qbucket, dbucket;
    bool operator<(const TuneKey &o) const {
        if (kind != o.kind) return kind < o.kind;
        if (algo != o.algo) return algo < o.algo;
        if (qbucket != o.qbucket) return qbucket < o.qbucket;
        return dbucket < o.dbucket;
    }
};

struct ish_hip_ctx {
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // query
    bool have_query = false;
    int32_t Q = 0, algo = 0, go = 0, ge = 0, kind = 0, nchunks = 0, stride = 0, rows = 0;
    int32_t front = 0; // profile entries before query row 0 (top padding)
    bool top = false;  // semi-global top padding rows
    uint32_t flags = 0;
    int32_t neg_limit = 0, min_score = 0, max_score = 0;
    int64_t r16max = -1; // longest target routed to the int16 kernel
    // binary16 lanes (semi-global lockstep tiles): groups of at most w16
    // columns may run in them, every value held minus fshift; f16 = 0
    // keeps every group in int16
    int64_t w16 = -1;
    int32_t fshift = 0;
    bool f16 = true;
    int32_t prof_words = 0;
    size_t lds = 0;
    DevBuf prof_t, prof_rt, prof_p, lut; // forward, reversed-query, plain profiles
    // host copy of the query profile (paired profiles are built from it per
    // database) and a counter of set_query calls
    Profile hprof;
    uint64_t query_gen = 0;
    // paired profile of the current query for the code set pair_mask
    // (pair_rows classes; 0: the codes need more than kPairMaxRows)
    DevBuf pprof, plut;
    PairProfile pairs;
    uint64_t pair_gen = 0; // query_gen it was built for (0: none)
    uint32_t pair_mask[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int32_t pair_rows = 0, pqpad = 0;
    int engine = ISH_ENGINE_BATCHED;
    int chunk = kChunkDefault; // query rows per lane in the batched kernels
    int schedule = ISH_SCHEDULE_AUTO; // one wave or one workgroup per group
    // autotuning: (matrix, algorithm, query-length bucket) -> choice per
    // target-length bucket
    std::map<TuneKey, TuneChoice> tune;
    bool tuned = false;
    // HIP graph of the last db pass, replayed while its launches match
    bool graphs = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint8_t> graph_sig;
    // work buffers
    DevBuf scratch, counter, in_res, in_off, out_score, out_er, out_eq, out_st;
    // pinned staging of database uploads: two buffers, each with the event
    // of its last copy to the device
    uint8_t *stage[2] = {nullptr, nullptr};
    hipEvent_t stage_done[2] = {nullptr, nullptr};
    // One-shot calls (sliced batches, file batches) load databases on a
    // stream of their own, beside the scoring queued on `stream`: passes
    // record pass_done[slot] after their launches, and their results are
    // fetched on `copy` once it fires; span brackets a sliced batch's
    // launches. `copy` exists only during such a call: a process gets
    // few hardware queues (4 by default), and a stream held between calls
    // would make a server's launch streams share one, which serializes
    // the batches they overlap.
    hipStream_t copy = nullptr;
    hipEvent_t pass_done[2] = {nullptr, nullptr};
    hipEvent_t span[2] = {nullptr, nullptr};
    // host workers for the parallel host copies (created at first use)
    std::unique_ptr<ish::HostPool> pool;
    // pinned landing buffer of fetched results (grow-only)
    uint8_t *res_host = nullptr;
    size_t res_host_bytes = 0;
    // pinned staging of a build's offsets and group tables (grow-only), apart
    // from the residue staging so that they need not wait for its last
    // copies; tab_done marks the end of the last table copies
    uint8_t *tab_host = nullptr;
    size_t tab_bytes = 0;
    hipEvent_t tab_done = nullptr;
    // pinned landing of a build's residue-code mask (a copy into pageable
    // memory would hold the host until the upload before it ends);
    // mask_done marks its arrival and is recorded on `stream` only
    uint32_t *mask_host = nullptr;
    hipEvent_t mask_done = nullptr;
    // buffers of one-shot databases (sliced batches, file batches): slot k
    // serves every other database, so one is loaded while the other's
    // pass runs; `mask` receives a build's residue-code mask
    ish::DbBuffers once[2];
    DevBuf mask;
    // one-shot batches are uploaded and scored in slices, each slice's
    // upload overlapping the scoring of the slice before: 0 sizes them
    // automatically, > 0 is a fixed size in residues, < 0 never slices
    int64_t slice_bytes = 0;
    // timing
    std::vector<hipEvent_t> ev;
    int launches = 0;
    float last_ms = 0.f;
    // a batch submitted with ish_hip_align_batch_submit, running on its own
    // host thread until ish_hip_align_batch_wait
    std::thread bg;
    int32_t bg_rc = 0;
};

// The context's host worker pool: up to 8 threads in all, with the caller.
static ish::HostPool &host_pool(ish_hip_ctx *ctx) {
    if (!ctx->pool) {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        ctx->pool = std::make_unique<ish::HostPool>((int)std::min(8u, hw) - 1);
    }
    return *ctx->pool;
}

static int32_t fail(ish_hip_ctx *ctx, int32_t code, const char *fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

#define HIP_TRY(ctx, expr)                                                                \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            return fail(ctx, _e == hipErrorOutOfMemory ? ISH_E_NOMEM : ISH_E_HIP,         \
                        "%s failed: %s", #expr, hipGetErrorString(_e));                   \
    } while (0)

// ---------------------------------------------------------------------------
// tables
// ---------------------------------------------------------------------------
extern "C" int32_t ish_hip_version(void) { return ISH_HIP_VERSION; }

extern "C" int32_t ish_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    int ok = 0;
    for (int d = 0; d < n; ++d) {
        hipDeviceProp_t pr;
        if (hipGetDeviceProperties(&pr, d) == hipSuccess && strncmp(pr.gcnArchName, "gfx950", 6) == 0)
            ++ok;
    }
    return ok;
}

int32_t ish::alphabet_of(int32_t kind) {
    switch (kind) {
    case ISH_MATRIX_ASCII: return 256;
    case ISH_MATRIX_ACTGN:
    case ISH_MATRIX_ACTGN0: return 5;
    case ISH_MATRIX_BLOSUM62: return 24;
    default: return -1;
    }
}

extern "C" int32_t ish_hip_matrix(int32_t kind, int8_t *out, int32_t cap) {
    const int32_t A = alphabet_of(kind);
    if (A < 0 || !out || cap < A * A) return ISH_E_INVALID;
    for (int a = 0; a < A; ++a)
        for (int b = 0; b < A; ++b) {
            int8_t v;
            if (kind == ISH_MATRIX_ASCII) v = a == b ? 2 : -2;
            else if (kind == ISH_MATRIX_BLOSUM62) v = kBlosum62[a * 24 + b];
            else v = nt_score(a, b, kind == ISH_MATRIX_ACTGN);
            out[a * A + b] = v;
        }
    return A;
}

extern "C" int32_t ish_hip_encoding(int32_t kind, uint8_t *out) {
    if (!out) return ISH_E_INVALID;
    switch (kind) {
    case ISH_MATRIX_ASCII:
        for (int c = 0; c < 256; ++c) out[c] = (uint8_t)c;
        return ISH_OK;
    case ISH_MATRIX_ACTGN:
    case ISH_MATRIX_ACTGN0:
        for (int c = 0; c < 256; ++c) {
            const int u = (c >= 'a' && c <= 'z') ? c - 32 : c;
            out[c] = u == 'A' ? 0 : u == 'C' ? 1 : u == 'G' ? 2 : (u == 'T' || u == 'U') ? 3 : 4;
        }
        return ISH_OK;
    case ISH_MATRIX_BLOSUM62:
        for (int c = 0; c < 256; ++c) {
            const int u = (c >= 'a' && c <= 'z') ? c - 32 : c;
            uint8_t code = 23;
            if (u >= 'A' && u <= 'Z')
                for (int k = 0; k < 23; ++k)
                    if (kAminoLetters[k] == u) code = (uint8_t)k;
            out[c] = code;
        }
        return ISH_OK;
    default: return ISH_E_INVALID;
    }
}

extern "C" int32_t ish_hip_decoding(int32_t kind, uint8_t *out, int32_t cap) {
    const int32_t A = alphabet_of(kind);
    if (A < 0 || !out || cap < A) return ISH_E_INVALID;
    for (int c = 0; c < A; ++c)
        out[c] = kind == ISH_MATRIX_ASCII ? (uint8_t)c
                 : kind == ISH_MATRIX_BLOSUM62 ? (uint8_t)kAminoLetters[c]
                                               : (uint8_t)kNucLetters[c];
    return A;
}

extern "C" int64_t ish_hip_encode_and_score(int32_t kind, const uint8_t *ascii, int64_t n,
                                            uint8_t *out) {
    uint8_t enc[256];
    const int32_t A = alphabet_of(kind);
    if (A < 0 || ish_hip_encoding(kind, enc) != ISH_OK || (n > 0 && (!ascii || !out)))
        return INT64_MIN;
    std::vector<int8_t> m((size_t)A * A);
    ish_hip_matrix(kind, m.data(), A * A);
    int64_t s = 0;
    for (int64_t i = 0; i < n; ++i) {
        const uint8_t c = enc[ascii[i]];
        out[i] = c;
        s += m[(size_t)c * A + c];
    }
    return s;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int32_t ish_hip_create(int32_t device, ish_hip_ctx **out) {
    if (!out) return ISH_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return ISH_E_NO_DEVICE;
    if (device < 0 || device >= n) return ISH_E_INVALID;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, device) != hipSuccess) return ISH_E_HIP;
    if (strncmp(pr.gcnArchName, "gfx950", 6) != 0) return ISH_E_NO_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return ISH_E_HIP;
    // the device address of the no-allocation pass's static scratch, once
    // per device
    if (ish::fixed_scratch_init(device) != hipSuccess) return ISH_E_HIP;
    ish_hip_ctx *c = new ish_hip_ctx();
    c->device = device;
    c->cus = pr.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return ISH_E_HIP;
    }
    bool ok = true;
    for (int k = 0; k < 2 && ok; ++k)
        ok = hipEventCreateWithFlags(&c->pass_done[k], hipEventDisableTiming) == hipSuccess &&
             hipEventCreate(&c->span[k]) == hipSuccess;
    if (!ok) {
        for (int k = 0; k < 2; ++k) {
            if (c->pass_done[k]) (void)hipEventDestroy(c->pass_done[k]);
            if (c->span[k]) (void)hipEventDestroy(c->span[k]);
        }
        (void)hipStreamDestroy(c->stream);
        delete c;
        return ISH_E_HIP;
    }
    *out = c;
    return ISH_OK;
}

extern "C" void ish_hip_destroy(ish_hip_ctx *ctx) {
    if (!ctx) return;
    if (ctx->bg.joinable()) ctx->bg.join();
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ish::one_shot_end(ctx);
    for (auto e : ctx->ev) (void)hipEventDestroy(e);
    for (int k = 0; k < 2; ++k) {
        if (ctx->stage[k]) (void)hipHostFree(ctx->stage[k]);
        if (ctx->stage_done[k]) (void)hipEventDestroy(ctx->stage_done[k]);
    }
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->once[0].release();
    ctx->once[1].release();
    for (DevBuf *b : {&ctx->prof_t, &ctx->prof_rt, &ctx->prof_p, &ctx->lut, &ctx->pprof, &ctx->plut,
                      &ctx->scratch, &ctx->counter, &ctx->mask,
                      &ctx->in_res, &ctx->in_off, &ctx->out_score, &ctx->out_er, &ctx->out_eq,
                      &ctx->out_st})
        b->release();
    if (ctx->res_host) (void)hipHostFree(ctx->res_host);
    if (ctx->mask_host) (void)hipHostFree(ctx->mask_host);
    if (ctx->mask_done) (void)hipEventDestroy(ctx->mask_done);
    if (ctx->tab_done) (void)hipEventSynchronize(ctx->tab_done);
    if (ctx->tab_host) (void)hipHostFree(ctx->tab_host);
    if (ctx->tab_done) (void)hipEventDestroy(ctx->tab_done);
    for (int k = 0; k < 2; ++k) {
        (void)hipEventDestroy(ctx->pass_done[k]);
        (void)hipEventDestroy(ctx->span[k]);
    }
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" const char *ish_hip_last_error(const ish_hip_ctx *ctx) {
    return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int32_t ish_hip_set_tile(ish_hip_ctx *ctx, int32_t query_rows_per_lane) {
    if (!ctx) return ISH_E_INVALID;
    if (query_rows_per_lane != 16 && query_rows_per_lane != 24 && query_rows_per_lane != 32 &&
        query_rows_per_lane != 40)
        return fail(ctx, ISH_E_INVALID, "tile height must be 16, 24, 32 or 40 (got %d)", query_rows_per_lane);
    ctx->chunk = query_rows_per_lane;
    ctx->tuned = false; // an explicit tile overrides the tuning table
    return ISH_OK;
}

extern "C" int32_t ish_hip_set_schedule(ish_hip_ctx *ctx, int32_t schedule) {
    if (!ctx) return ISH_E_INVALID;
    if (schedule != ISH_SCHEDULE_AUTO && schedule != ISH_SCHEDULE_WAVE && schedule != ISH_SCHEDULE_WORKGROUP)
        return fail(ctx, ISH_E_INVALID, "unknown schedule %d", schedule);
    ctx->schedule = schedule;
    return ISH_OK;
}

extern "C" int32_t ish_hip_set_f16(ish_hip_ctx *ctx, int32_t enable) {
    if (!ctx) return ISH_E_INVALID;
    ctx->f16 = enable != 0;
    return ISH_OK;
}

extern "C" int32_t ish_hip_set_engine(ish_hip_ctx *ctx, int32_t engine) {
    if (!ctx) return ISH_E_INVALID;
    if (engine != ISH_ENGINE_BATCHED && engine != ISH_ENGINE_PLAIN)
        return fail(ctx, ISH_E_INVALID, "unknown engine %d", engine);
    ctx->engine = engine;
    return ISH_OK;
}

// Longest target the packed int16 kernel may score for this query: every
// value the offset-form recurrence produces (H^, E^, F^, their sums with the
// profile and gap terms, the tracking keys) is then proven to lie strictly
// inside int16, so int16 arithmetic equals int32 arithmetic. The bounds:
//   H(j,i) <= (j+1)*M                       (each row adds at most M=max S)
//   H(j,i) >= -2go - ge*(i+j)               (gap from the boundary)
//   H^ = H + (i+j)*ge in [-2go, Q*M + (R+Q)*ge]
// plus one profile term (<= M + 2ge) and one gap term (>= -go - m).
int64_t ish::int16_max_len(int32_t Q, int32_t go, int32_t ge, int32_t max_s, int32_t min_s) {
    if (go < ge) return -1; // no top padding rows (engine.h, top_pad)
    const int64_t M = std::max(max_s, 0), m = std::max(-min_s, 0);
    const int64_t lim = 32767 - 16;
    if (3LL * go + 2LL * ge + m > lim) return -1;
    int64_t best = 65534;
    // upper: Q*M + (R+Q)*ge + M + 2ge + go <= lim
    const int64_t up = lim - (int64_t)Q * M - M - 2LL * ge - go;
    if (up < (int64_t)Q * ge) return -1;
    if (ge > 0) best = std::min(best, up / ge - Q);
    // lower (plain values, the tracking keys and the int16 floor):
    // 2go + ge*(R+Q) + max(go, m) + 2 <= lim
    const int64_t lo = lim - 2LL * go - std::max<int64_t>(go, m) - 2;
    if (lo < (int64_t)Q * ge) return -1;
    if (ge > 0) best = std::min(best, lo / ge - Q);
    if ((int64_t)Q > 65534) return -1;
    return best < 0 ? -1 : best;
}

// Columns a group may have for the semi-global offset form to run in
// binary16 lanes (-1: none), and the shift C of the values they hold.
// Every value such a pass produces is an integer in [L, U]:
//   H(j,i) <= P_j = sum over rows k <= j of max(0, max_r S(r, k)), since an
//     alignment scores at most its best match per query row and boundary
//     and gap terms are <= 0; E, F and Hdiag + S are bounded the same way;
//     with the offset (i+j)*ge and i <= cols-1, j <= Q-1:
//     U = P_{Q-1} + (cols + Q - 2)*ge (the top padding rows hold (i-1)*ge);
//   E^ starts at the left boundary plus ge-go >= -2go and only rises, H^ is
//     at least E^, F^ starts at >= -2go, G = H^ + ge - go >= -3go + ge, and
//     Hdiag^ + S' >= -2go + min S': L = -3go - 2ge - m - 2, m = max(0, -min S).
// Held minus C = L + 2040, they lie in [-2040, 2040] when U <= L + 4080,
// where binary16 arithmetic on integers is exact.
// rowmax: max over the profile rows of each query row's score S.
static int64_t f16_max_cols(int64_t P, int64_t Q, int32_t go, int32_t ge, int32_t min_s, int32_t *shift) {
    if (go < ge || Q == 0) return -1; // no top padding rows (engine.h, top_pad)
    const int64_t m = std::max(-min_s, 0);
    const int64_t L = -3LL * go - 2LL * ge - m - 2;
    if (L < -2040) return -1;
    *shift = (int32_t)(L + 2040);
    const int64_t room = L + 4080 - P; // for (cols + Q - 2) * ge
    if (room < 0) return -1;
    if (ge == 0) return INT64_MAX;
    const int64_t cols = room / ge - Q + 2;
    return cols < 4 ? -1 : cols;
}

// P = sum over query rows of max(0, the row's best score over the profile
// rows a pass can read): the classes of the residue codes set in
// present[8] and of code 0, which lanes past their target's end read (every
// row when present is null). Profile entries are S + bias.
static int64_t best_possible(const Profile &pf, int32_t Q, int32_t bias, const uint32_t *present) {
    const int32_t rw = pf.stride / 2;
    std::vector<char> use((size_t)pf.nrows, present ? 0 : 1);
    if (present)
        for (int c = 0; c < 256; ++c)
            if (c == 0 || ((present[c >> 5] >> (c & 31)) & 1u)) use[(size_t)(pf.lut[c] / rw)] = 1;
    std::vector<int32_t> rowmax((size_t)Q, 0); // max(0, best) per row
    for (int32_t r = 0; r < pf.nrows; ++r) {
        if (!use[(size_t)r]) continue;
        const int16_t *row = pf.fwd.data() + (size_t)r * pf.stride + pf.front;
        for (int32_t j = 0; j < Q; ++j) rowmax[(size_t)j] = std::max<int32_t>(rowmax[(size_t)j], row[j] - bias);
    }
    int64_t P = 0;
    for (int32_t v : rowmax) P += v;
    return P;
}

// f16_max_cols for a semi-global query profile over the rows a pass can
// read (best_possible).
int64_t ish::f16_bound(const Profile &pf, int32_t Q, int32_t go, int32_t ge, int32_t min_s,
                       const uint32_t *present, int32_t *shift) {
    if (!pf.top || Q <= 0) return -1;
    return f16_max_cols(best_possible(pf, Q, 2 * ge, present), Q, go, ge, min_s, shift);
}

// Local alignment (column-offset form, state shifted down every 256
// columns) in binary16 lanes: whether every value stays exact, and the
// shift C of the state. With P the query's best possible score over the
// rows a pass can read (best_possible: an alignment scores at most its best
// match per row):
//   H~ = H + i'*ge, i' < 256, lies in [0, P + 255ge] (H >= 0 and the
//   floor is i'*ge); E~ >= -go, F~ >= -go and Hdiag~ + S' >= min S (after a
//   shift the carried state is at least -ge below the floor) give
//   L = -(go + ge + m) - 2 (m = max(0, -min S)), U = P + 264ge + 2.
// The state is held minus C = L + 2040: exact when U <= L + 4080. The
// chunk and column maxima and the target's best are plain values in
// [-1, P], exact when P <= 2040; the shift of 256ge is exact when
// 256ge <= 2040.
bool ish::sw_f16_bound(const Profile &pf, int32_t Q, int32_t go, int32_t ge, int32_t min_s,
                       const uint32_t *present, int32_t *shift) {
    if (Q <= 0 || 256LL * ge > 2040) return false;
    const int64_t P = best_possible(pf, Q, ge, present);
    const int64_t m = std::max(-min_s, 0);
    const int64_t L = -(int64_t)go - ge - m - 2;
    const int64_t U = P + 264LL * ge + 2;
    if (P > 2040 || L < -2040 || U > L + 4080) return false;
    *shift = (int32_t)(L + 2040);
    return true;
}

// Local alignment in packed int16: with the state shifted down every 256
// columns, every value lies in [-(go+ge+|min S|), Q*M + (256+Q+8)*ge + M],
// whatever the target length.
static bool local_fits_int16(int32_t qpad, int32_t go, int32_t ge, int32_t max_s, int32_t min_s) {
    const int64_t M = std::max(max_s, 0), m = std::max(-min_s, 0);
    const int64_t lim = 32767 - 64;
    const int64_t up = (int64_t)qpad * M + (int64_t)(256 + qpad + 8) * ge + M + go;
    const int64_t lo = 3LL * go + 3LL * ge + m;
    return up <= lim && lo <= lim && qpad <= 65534;
}

void ish::build_profile(const uint8_t *q, int32_t Q, const std::vector<int8_t> &mat, int32_t A,
                        int32_t bias, bool want_rev_plain, bool top_rows, int32_t top_fwd,
                        int32_t top_rev, Profile &pf) {
    std::map<std::vector<int8_t>, int> row_of;
    std::vector<std::vector<int8_t>> rows;
    for (int c = 0; c < 256; ++c) {
        const int cc = c < A ? c : A - 1; // out-of-alphabet codes score as the last code
        std::vector<int8_t> col((size_t)Q);
        for (int32_t j = 0; j < Q; ++j) col[j] = mat[(size_t)cc * A + q[j]];
        auto it = row_of.find(col);
        if (it == row_of.end()) {
            it = row_of.emplace(col, (int)rows.size()).first;
            rows.push_back(col);
        }
        pf.lut[c] = (uint16_t)it->second;
    }
    pf.qpad = std::max((Q + kChunkAlign - 1) / kChunkAlign * kChunkAlign,
                       (Q + kChunkAlign40 - 1) / kChunkAlign40 * kChunkAlign40);
    pf.top = top_rows;
    pf.front = top_rows ? top_front(Q) : 0;
    // rows start 16 bytes apart modulo 128: 16-byte slices of different rows
    // spread over the LDS banks and every slice read is 16-byte aligned
    pf.stride = (pf.width() + 7) / 8 * 8 + 8;
    pf.nrows = (int32_t)rows.size();
    for (int c = 0; c < 256; ++c) pf.lut[c] = (uint16_t)(pf.lut[c] * (pf.stride / 2));
    const size_t elems = (size_t)pf.nrows * pf.stride;
    pf.fwd.assign(elems, 0);
    if (want_rev_plain) {
        pf.rev.assign(elems, 0);
        pf.plain.assign(elems, 0);
    }
    for (int32_t r = 0; r < pf.nrows; ++r) {
        int16_t *fw = pf.fwd.data() + (size_t)r * pf.stride;
        for (int32_t e = 0; e < pf.front; ++e) fw[e] = (int16_t)top_fwd;
        if (want_rev_plain)
            for (int32_t e = 0; e < pf.front; ++e) pf.rev[(size_t)r * pf.stride + e] = (int16_t)top_rev;
        for (int32_t j = 0; j < Q; ++j) {
            const int16_t v = (int16_t)(rows[r][j] + bias);
            fw[pf.front + j] = v;
            if (want_rev_plain) {
                pf.plain[(size_t)r * pf.stride + j] = rows[r][j];
                // reversed query: the profile of the reverse passes
                // (Profile of rev_pattern, striped_semi_global_matcher.mojo:68-75)
                pf.rev[(size_t)r * pf.stride + pf.front + (Q - 1 - j)] = v;
            }
        }
    }
}

bool ish::build_pairs(const Profile &pf, const uint32_t present[8], int32_t max_classes,
                      PairProfile &out) {
    const int32_t rw = pf.stride / 2;
    std::vector<int32_t> remap((size_t)pf.nrows, -1), cls_of;
    for (int c = 0; c < 256; ++c) {
        if (c != 0 && !((present[c >> 5] >> (c & 31)) & 1u)) continue;
        const int32_t cls = pf.lut[c] / rw;
        if (remap[(size_t)cls] < 0) {
            remap[(size_t)cls] = (int32_t)cls_of.size();
            cls_of.push_back(cls);
            if ((int32_t)cls_of.size() > max_classes) return false;
        }
    }
    const int32_t C = (int32_t)cls_of.size();
    // the table's uint16 offsets must reach the last pair row (tile 32)
    if ((int64_t)(C * C - 1) * pair_row_bytes(32) > 65535) return false;
    out.classes = C;
    out.qpad = (pf.width() + 7) / 8 * 8; // <= stride: every entry lies in the row
    out.words.assign((size_t)C * C * out.qpad, 0);
    for (int32_t a = 0; a < C; ++a)
        for (int32_t b = 0; b < C; ++b) {
            const int16_t *ra = pf.fwd.data() + (size_t)cls_of[(size_t)a] * pf.stride;
            const int16_t *rb = pf.fwd.data() + (size_t)cls_of[(size_t)b] * pf.stride;
            int32_t *w = out.words.data() + ((size_t)a * C + b) * out.qpad;
            for (int32_t e = 0; e < out.qpad; ++e)
                w[e] = (int32_t)((uint32_t)(uint16_t)ra[e] | ((uint32_t)(uint16_t)rb[e] << 16));
        }
    // codes the database does not hold take class 0's rows (never read);
    // a tile height whose last pair row lies beyond the uint16 offsets
    // keeps a zero table (its paired launches take at most kPairMaxRows
    // classes, whose rows fit)
    out.table.assign((size_t)kPairTables * 512, 0);
    for (int ti = 0; ti < kPairTables; ++ti) {
        const int32_t rowb = pair_row_bytes(pair_table_rows(ti));
        if ((int64_t)(C * C - 1) * rowb > 65535) continue;
        for (int c = 0; c < 256; ++c) {
            const int32_t r = std::max(remap[(size_t)(pf.lut[c] / rw)], 0);
            out.table[(size_t)ti * 512 + c] = (uint16_t)(r * C * rowb);
            out.table[(size_t)ti * 512 + 256 + c] = (uint16_t)(r * rowb);
        }
    }
    return true;
}

// matrix_len is the number of int8 entries the caller's matrix holds: the
// length of ScoringMatrix.values that GpuMatcher.matrix_len() returns
// (striped_semi_global_matcher.mojo:143-144), A*A for the kind's alphabet A.
int32_t ish::matrix_of(const int8_t *matrix, int64_t matrix_len, int32_t kind, int32_t &A,
                       std::vector<int8_t> &mat, std::string &err) {
    A = alphabet_of(kind);
    if (A < 0) {
        err = "unknown matrix kind " + std::to_string(kind);
        return ISH_E_INVALID;
    }
    if (!matrix) {
        mat.resize((size_t)A * A);
        ish_hip_matrix(kind, mat.data(), A * A);
        return ISH_OK;
    }
    if (matrix_len != (int64_t)A * A) {
        err = "matrix_len " + std::to_string(matrix_len) + " is not " + std::to_string(A) + "*" +
              std::to_string(A) + " = " + std::to_string(A * A) + ", the entries of matrix kind " +
              std::to_string(kind);
        return ISH_E_INVALID;
    }
    mat.assign(matrix, matrix + (size_t)A * A);
    return ISH_OK;
}

static int32_t resolve_matrix(ish_hip_ctx *ctx, const int8_t *matrix, int32_t matrix_len, int32_t kind,
                              int32_t &A, std::vector<int8_t> &mat) {
    std::string err;
    const int32_t rc = matrix_of(matrix, matrix_len, kind, A, mat, err);
    if (rc) return fail(ctx, rc, "%s", err.c_str());
    return ISH_OK;
}

// int16 floor seeding the last-row maximum (striped.mojo:329-334), from the
// matrix minimum seeded at 0 (:62-71).
int32_t ish::neg_limit_of(int32_t go, int32_t mn) {
    const int32_t nl = (-go < mn) ? -32768 + go + 1 : -32768 - mn + 1;
    return (int32_t)(int16_t)nl;
}

extern "C" int32_t ish_hip_set_query(ish_hip_ctx *ctx, const uint8_t *q, int32_t Q,
                                     const int8_t *matrix, int32_t matrix_len, int32_t kind, int32_t go,
                                     int32_t ge, uint32_t ends_free, int32_t algo) {
    if (!ctx) return ISH_E_INVALID;
    if (Q < 0 || (Q > 0 && !q)) return fail(ctx, ISH_E_INVALID, "bad query");
    if (algo != ISH_ALGO_SEMI_GLOBAL && algo != ISH_ALGO_LOCAL)
        return fail(ctx, ISH_E_INVALID, "unknown algo %d", algo);
    if (go < 0 || ge < 0) return fail(ctx, ISH_E_INVALID, "gap penalties must be >= 0");
    std::vector<int8_t> mat;
    int32_t A = 0;
    int32_t rc = resolve_matrix(ctx, matrix, matrix_len, kind, A, mat);
    if (rc) return rc;
    for (int32_t j = 0; j < Q; ++j)
        if (q[j] >= A) return fail(ctx, ISH_E_INVALID, "query code %d at %d >= alphabet %d", q[j], j, A);
    HIP_TRY(ctx, hipSetDevice(ctx->device));

    int32_t mn = 0, mx = 0; // both seeded at 0 (striped.mojo:62-71)
    for (int8_t v : mat) {
        mn = std::min<int32_t>(mn, v);
        mx = std::max<int32_t>(mx, v);
    }
    Profile pf;
    // offset-form scores: S + 2ge for the semi-global kernels, S + ge for
    // the local ones (their offsets are explained in the kernel sources)
    // semi-global with go >= ge: top padding rows (engine.h, top_pad); the
    // reverse passes swap the free ends, so their target start is te
    const bool top = algo != ISH_ALGO_LOCAL && go >= ge;
    build_profile(q, Q, mat, A, algo == ISH_ALGO_LOCAL ? ge : 2 * ge, true, top,
                  top_row_score(ends_free, ge), top_row_score((ends_free >> 1) & 4u, ge), pf);
    const size_t lds = pf.lds();
    if (lds > kLdsLimit)
        return fail(ctx, ISH_E_RANGE, "query profile needs %zu bytes of LDS (limit %zu)", lds,
                    kLdsLimit);
    const size_t bytes = pf.fwd.size() * 2;
    HIP_TRY(ctx, ctx->prof_t.ensure(bytes));
    HIP_TRY(ctx, ctx->prof_p.ensure(bytes));
    HIP_TRY(ctx, ctx->prof_rt.ensure(bytes));
    HIP_TRY(ctx, ctx->lut.ensure(512));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->prof_t.p, pf.fwd.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->prof_rt.p, pf.rev.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->prof_p.p, pf.plain.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->lut.p, pf.lut, 512, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));

    ctx->Q = Q;
    ctx->algo = algo;
    ctx->go = go;
    ctx->ge = ge;
    ctx->kind = kind;
    ctx->flags = ends_free & 15u;
    ctx->nchunks = 0; // set per launch from the tile height
    ctx->stride = pf.stride;
    ctx->front = pf.front;
    ctx->top = pf.top;
    ctx->rows = pf.nrows;
    ctx->neg_limit = neg_limit_of(go, mn);
    ctx->min_score = mn;
    ctx->max_score = mx;
    ctx->prof_words = pf.words();
    ctx->lds = lds;
    ctx->r16max = algo == ISH_ALGO_LOCAL ? (local_fits_int16(pf.qpad, go, ge, mx, mn) ? INT64_MAX : -1)
                                         : int16_max_len(Q, go, ge, mx, mn);
    ctx->w16 = -1; // set with the paired profile, per database code set
    ctx->fshift = 0;
    ctx->hprof = std::move(pf);
    ctx->query_gen++;
    ctx->have_query = true;
    return ISH_OK;
}

// ---------------------------------------------------------------------------
// timing helpers
// ---------------------------------------------------------------------------
static hipError_t timed_begin(ish_hip_ctx *ctx) {
    const size_t need = (size_t)(ctx->launches + 1) * 2;
    while (ctx->ev.size() < need) {
        hipEvent_t e;
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) return r;
        ctx->ev.push_back(e);
    }
    return hipEventRecord(ctx->ev[(size_t)ctx->launches * 2], ctx->stream);
}
static hipError_t timed_end(ish_hip_ctx *ctx) {
    hipError_t r = hipEventRecord(ctx->ev[(size_t)ctx->launches * 2 + 1], ctx->stream);
    ctx->launches++;
    return r;
}
static void collect_timing(ish_hip_ctx *ctx) {
    float total = 0.f;
    for (int k = 0; k < ctx->launches; ++k) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev[(size_t)k * 2], ctx->ev[(size_t)k * 2 + 1]) == hipSuccess)
            total += ms;
    }
    ctx->last_ms = total;
}

extern "C" int32_t ish_hip_last_timing(const ish_hip_ctx *ctx, float *ms, int32_t *launches) {
    if (!ctx) return ISH_E_INVALID;
    if (ms) *ms = ctx->last_ms;
    if (launches) *launches = ctx->launches;
    return ISH_OK;
}

extern "C" int32_t ish_hip_sync(ish_hip_ctx *ctx) {
    if (!ctx) return ISH_E_INVALID;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    collect_timing(ctx);
    return ISH_OK;
}

// Scratch for the boundary rows: one slot of `cols` columns x 64 lanes x 8
// bytes per resident wave. The grid is capped so that scratch stays within
// half of the free device memory.
hipError_t ish::persistent_grid(DevBuf &scratch, int occ_blocks, int cus, int64_t cols,
                                int *grid_out, int64_t *slot_stride, int waves_per_block) {
    int64_t grid = (int64_t)occ_blocks * cus;
    const int64_t per_block = (int64_t)waves_per_block * std::max<int64_t>(cols, 4) * 64 * 8;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)8 << 30;
    const int64_t budget = std::min<int64_t>((int64_t)(free_b / 2) + (int64_t)scratch.cap,
                                             (int64_t)64 << 30);
    if (grid * per_block > budget) grid = std::max<int64_t>(1, budget / per_block);
    hipError_t e = scratch.ensure((size_t)(grid * per_block));
    if (e != hipSuccess) return e;
    *grid_out = (int)grid;
    *slot_stride = std::max<int64_t>(cols, 4) * 64;
    return hipSuccess;
}

static int32_t size_grid(ish_hip_ctx *ctx, int occ_blocks, int64_t cols, int *grid_out,
                         int64_t *slot_stride, int waves_per_block = kBlock / 64) {
    HIP_TRY(ctx, persistent_grid(ctx->scratch, occ_blocks, ctx->cus, cols, grid_out, slot_stride,
                                 waves_per_block));
    return ISH_OK;
}

// ---------------------------------------------------------------------------
// device-resident database
// ---------------------------------------------------------------------------
// The host half of a database build: targets sorted by length (counting
// sort, descending, stable in input order) into groups of kGroup slots.
// It needs only the offsets, so it can run while the residues upload.
struct SortPlan {
    std::vector<int32_t> sorted, lens; // per slot: target index (-1: empty), length
    std::vector<GroupDesc> groups;
    std::vector<int32_t> group_maxlen;
    int32_t maxcols = 0;
    uint64_t words = 0; // uint2 words of all groups' blocks
};

static void plan_sort(const std::vector<uint64_t> &h_off, SortPlan &sp) {
    const int64_t n = (int64_t)h_off.size() - 1;
    int64_t maxlen = 0;
    for (int64_t t = 0; t < n; ++t)
        maxlen = std::max<int64_t>(maxlen, (int64_t)(h_off[(size_t)t + 1] - h_off[(size_t)t]));
    std::vector<int64_t> hist((size_t)maxlen + 2, 0);
    for (int64_t t = 0; t < n; ++t) hist[(size_t)(maxlen - (int64_t)(h_off[(size_t)t + 1] - h_off[(size_t)t]))]++;
    int64_t acc = 0;
    for (auto &h : hist) {
        const int64_t v = h;
        h = acc;
        acc += v;
    }
    const int32_t ng = (int32_t)((n + kGroup - 1) / kGroup);
    const int64_t nslots = (int64_t)ng * kGroup;
    sp.sorted.assign((size_t)nslots, -1);
    sp.lens.assign((size_t)nslots, 0);
    for (int64_t t = 0; t < n; ++t) {
        const int64_t L = (int64_t)(h_off[(size_t)t + 1] - h_off[(size_t)t]);
        const int64_t pos = hist[(size_t)(maxlen - L)]++;
        sp.sorted[(size_t)pos] = (int32_t)t;
        sp.lens[(size_t)pos] = (int32_t)L;
    }
    sp.groups.resize((size_t)ng);
    sp.group_maxlen.resize((size_t)ng);
    uint64_t off = 0;
    for (int32_t g = 0; g < ng; ++g) {
        const int32_t L = sp.lens[(size_t)g * kGroup];
        sp.groups[g].res_off = off;
        sp.groups[g].ncols = (L + 3) & ~3;
        sp.groups[g].first = g * kGroup;
        sp.group_maxlen[g] = L;
        sp.maxcols = std::max(sp.maxcols, sp.groups[g].ncols);
        off += (uint64_t)(sp.groups[g].ncols / 4) * 64;
    }
    sp.words = off;
}

// Host-to-device copies of a build's small host arrays (offsets, group
// tables) through the context's table staging buffer, each at its own
// offset: a pageable copy blocks the host for longer than the copy takes.
struct TableCopy {
    void *dst;
    const void *src;
    size_t bytes;
};
static hipError_t staged_tables(ish_hip_ctx *ctx, hipStream_t st, const TableCopy *c, size_t nc) {
    size_t total = 0;
    for (size_t k = 0; k < nc; ++k) total += (c[k].bytes + 255) & ~(size_t)255;
    if (total == 0) return hipSuccess;
    hipError_t e = hipSuccess;
    if (!ctx->tab_done) {
        e = hipEventCreateWithFlags(&ctx->tab_done, hipEventDisableTiming);
        if (e != hipSuccess) return e;
        e = hipEventRecord(ctx->tab_done, st);
        if (e != hipSuccess) return e;
    }
    e = hipEventSynchronize(ctx->tab_done); // the last table copies have been read
    if (e != hipSuccess) return e;
    if (ctx->tab_bytes < total) {
        if (ctx->tab_host) (void)hipHostFree(ctx->tab_host);
        ctx->tab_host = nullptr;
        ctx->tab_bytes = 0;
        e = hipHostMalloc((void **)&ctx->tab_host, total + total / 4, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        ctx->tab_bytes = total + total / 4;
    }
    // host copies (one thread per array when there is much to copy), then
    // the copies to the device
    std::vector<size_t> at(nc, 0);
    for (size_t k = 1; k < nc; ++k) at[k] = at[k - 1] + ((c[k - 1].bytes + 255) & ~(size_t)255);
    if (total >= ((size_t)4 << 20)) {
        host_pool(ctx).run((int)nc, [&](int k) {
            if (c[k].bytes) memcpy(ctx->tab_host + at[k], c[k].src, c[k].bytes);
        });
    } else {
        for (size_t k = 0; k < nc; ++k)
            if (c[k].bytes) memcpy(ctx->tab_host + at[k], c[k].src, c[k].bytes);
    }
    for (size_t k = 0; k < nc && e == hipSuccess; ++k)
        if (c[k].bytes)
            e = hipMemcpyAsync(c[k].dst, ctx->tab_host + at[k], c[k].bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipEventRecord(ctx->tab_done, st);
    return e;
}

// A device buffer of `bytes`: from the reused slot b when given, else
// allocated.
template <typename T>
static hipError_t db_alloc(ish::DevBuf *b, T **p, size_t bytes) {
    if (!b) return hipMalloc((void **)p, bytes);
    const hipError_t e = b->ensure(bytes);
    *p = e == hipSuccess ? (T *)b->p : nullptr;
    return e;
}

// Builds the packed, length-sorted layout from residues already on the
// device (d_raw, offsets d_off rebased to 0) and the sort plan of h_off,
// on stream st, and returns when it is done. Takes ownership of both
// buffers (or borrows them from reuse), which the database keeps for
// start-coordinate passes. A one-shot database built on the context's
// stream (a batch of one slice, whose pass is queued on that stream next)
// returns once its code mask is on the host, while the table copies and
// the packing still run.
static int32_t db_build(ish_hip_ctx *ctx, hipStream_t st, uint8_t *d_raw, uint64_t *d_off,
                        std::vector<uint64_t> h_off, SortPlan &sp, ish::DbBuffers *reuse,
                        ish_hip_db **out, bool copy_off = false) {
    uint64_t *const off_dst = copy_off ? d_off : nullptr; // offsets still on the host
    const int64_t n = (int64_t)h_off.size() - 1;
    ish_hip_db *db = new ish_hip_db();
    db->borrowed = reuse;
    db->device = ctx->device;
    db->n = n;
    db->nres = n > 0 ? (int64_t)h_off[(size_t)n] : 0;
    db->d_raw = d_raw;
    db->d_off = d_off;
    const int32_t ng = (int32_t)sp.groups.size();
    const int64_t nslots = (int64_t)ng * kGroup;
    db->ngroups = ng;
    db->group_maxlen = std::move(sp.group_maxlen);
    db->maxcols = sp.maxcols;
    const std::vector<GroupDesc> &groups = sp.groups;
    const std::vector<int32_t> &sorted = sp.sorted, &lens = sp.lens;
    const uint64_t off = sp.words;
    db->h_off = std::move(h_off);
    const uint64_t res_words = off + 64; // one spare block: empty groups read block 0
    uint32_t *d_mask = nullptr;
    auto cleanup = [&](int32_t code) {
        (void)hipStreamSynchronize(st);
        ish_hip_db_free(db);
        return code;
    };
    hipError_t e;
#define DB_TRY(expr)                                                                          \
    if ((e = (expr)) != hipSuccess)                                                           \
        return cleanup(fail(ctx, e == hipErrorOutOfMemory ? ISH_E_NOMEM : ISH_E_HIP, "%s: %s", \
                            #expr, hipGetErrorString(e)));
    DB_TRY(db_alloc(reuse ? &reuse->groups : nullptr, &db->d_groups, sizeof(GroupDesc) * std::max(ng, 1)));
    DB_TRY(db_alloc(reuse ? &reuse->res : nullptr, &db->d_res, sizeof(uint2) * res_words));
    DB_TRY(db_alloc(reuse ? &reuse->lens : nullptr, &db->d_lens, sizeof(int32_t) * std::max<int64_t>(nslots, 1)));
    DB_TRY(db_alloc(reuse ? &reuse->orig : nullptr, &db->d_orig, sizeof(int32_t) * std::max<int64_t>(nslots, 1)));
    DB_TRY(db_alloc(reuse ? &reuse->score : nullptr, &db->d_score, sizeof(int32_t) * std::max<int64_t>(n, 1)));
    DB_TRY(db_alloc(reuse ? &reuse->er : nullptr, &db->d_end_ref, sizeof(int32_t) * std::max<int64_t>(n, 1)));
    DB_TRY(db_alloc(reuse ? &reuse->eq : nullptr, &db->d_end_q, sizeof(int32_t) * std::max<int64_t>(n, 1)));
    DB_TRY(db_alloc(reuse ? &reuse->st : nullptr, &db->d_status, std::max<int64_t>(n, 1)));
    // the packer writes every block of every group; the spare block is zero
    DB_TRY(hipMemsetAsync(db->d_res + off, 0, sizeof(uint2) * 64, st));
    // residue codes present (paired profiles are built over their classes)
    const bool lazy = reuse && !ctx->copy && st == ctx->stream;
    DB_TRY(db_alloc(&ctx->mask, &d_mask, 8 * sizeof(uint32_t)));
    DB_TRY(launch_code_mask(db->d_raw, (uint64_t)db->nres, d_mask, st));
    if (!ctx->mask_host) {
        DB_TRY(hipHostMalloc((void **)&ctx->mask_host, 8 * sizeof(uint32_t), hipHostMallocDefault));
    }
    DB_TRY(hipMemcpyAsync(ctx->mask_host, d_mask, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    if (lazy) {
        if (!ctx->mask_done) {
            DB_TRY(hipEventCreateWithFlags(&ctx->mask_done, hipEventDisableTiming));
        }
        DB_TRY(hipEventRecord(ctx->mask_done, st));
    }
    {
        const TableCopy tc[4] = {
            {db->d_groups, groups.data(), sizeof(GroupDesc) * (size_t)ng},
            {db->d_lens, lens.data(), sizeof(int32_t) * (size_t)nslots},
            {db->d_orig, sorted.data(), sizeof(int32_t) * (size_t)nslots},
            {off_dst, db->h_off.data(), off_dst ? sizeof(uint64_t) * (size_t)(n + 1) : 0}};
        DB_TRY(staged_tables(ctx, st, tc, 4));
    }
    if (ng > 0)
        DB_TRY(launch_pack(db->d_raw, db->d_off, db->d_orig, db->d_groups, ng, n, off / 64, db->d_res,
                           st));
    if (lazy) {
        DB_TRY(hipEventSynchronize(ctx->mask_done));
    } else {
        DB_TRY(hipStreamSynchronize(st));
    }
    memcpy(db->code_mask, ctx->mask_host, sizeof db->code_mask);
#undef DB_TRY
    *out = db;
    return ISH_OK;
}

// Host-to-device copy of the concatenation of pieces into dst, through two
// pinned staging buffers: host threads copy one part into a buffer while the
// other buffer's part is on its way to the device, so the upload runs at
// the pinned DMA rate instead of the pageable path's (which stages small
// parts one after the other). The copies run on stream st; it returns once
// every part is enqueued there (the last copies may still run, but the
// pieces are no longer read).
constexpr size_t kStageBytes = (size_t)64 << 20;
static hipError_t staged_h2d(ish_hip_ctx *ctx, hipStream_t st, uint8_t *dst, const ish::HostPiece *pieces,
                             size_t np) {
    for (int k = 0; k < 2; ++k) {
        if (!ctx->stage[k]) {
            hipError_t e = hipHostMalloc((void **)&ctx->stage[k], kStageBytes, hipHostMallocDefault);
            if (e != hipSuccess) return e;
        }
        if (!ctx->stage_done[k]) {
            hipError_t e = hipEventCreateWithFlags(&ctx->stage_done[k], hipEventDisableTiming);
            if (e != hipSuccess) return e;
            e = hipEventRecord(ctx->stage_done[k], st);
            if (e != hipSuccess) return e;
        }
    }
    const int threads = host_pool(ctx).size();
    int buf = 0;
    uint64_t at = 0;
    size_t pk = 0;
    uint64_t in_piece = 0; // bytes of pieces[pk] already staged
    // parts of 4, 8, 16, 32, then 64 MB: the first copy to the device
    // starts after a short host copy
    size_t part = kStageBytes / 16;
    while (pk < np) {
        // fill one staging buffer with up to kStageBytes of the pieces
        hipError_t e = hipEventSynchronize(ctx->stage_done[buf]);
        if (e != hipSuccess) return e;
        struct Span {
            const uint8_t *src;
            size_t off, n;
        };
        std::vector<Span> spans;
        size_t fill = 0;
        while (pk < np && fill < part) {
            const uint64_t left = pieces[pk].n - in_piece;
            const size_t take = (size_t)std::min<uint64_t>(left, part - fill);
            if (take) spans.push_back({pieces[pk].p + in_piece, fill, take});
            fill += take;
            in_piece += take;
            if (in_piece == pieces[pk].n) {
                ++pk;
                in_piece = 0;
            }
        }
        if (fill == 0) continue;
        // the host copy, split over threads by byte range
        uint8_t *stage = ctx->stage[buf];
        auto copy_range = [&](size_t lo, size_t hi) {
            for (const Span &sp : spans) {
                const size_t a = std::max(lo, sp.off), b = std::min(hi, sp.off + sp.n);
                if (a < b) memcpy(stage + a, sp.src + (a - sp.off), b - a);
            }
        };
        const int tn = fill >= ((size_t)1 << 22) ? threads : 1;
        const size_t per = (fill + tn - 1) / tn;
        host_pool(ctx).run(tn, [&](int t) {
            copy_range(std::min(fill, (size_t)t * per), std::min(fill, (size_t)(t + 1) * per));
        });
        e = hipMemcpyAsync(dst + at, stage, fill, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipEventRecord(ctx->stage_done[buf], st);
        if (e != hipSuccess) return e;
        at += fill;
        buf ^= 1;
        part = std::min(kStageBytes, part * 2);
    }
    return hipSuccess;
}

// Uploads the residues (the concatenation of pieces) and builds the db;
// with `encode` the bytes are translated to residue codes on the device
// first. offsets: n + 1 entries, already checked; the pieces hold
// offsets[n] - offsets[0] bytes. During a one-shot call (the copy stream
// exists) a one-shot database (reuse given) is uploaded and built on the
// copy stream, beside the pass queued on the context's stream; any other
// load runs on the context's stream. The call returns when the database
// is complete.
static int32_t db_upload(ish_hip_ctx *ctx, const ish::HostPiece *pieces, size_t np,
                         const uint64_t *offsets, int64_t nt, const EncodeTable *encode,
                         ish::DbBuffers *reuse, ish_hip_db **out) {
    const size_t n = (size_t)nt;
    const uint64_t nres = n > 0 ? offsets[n] - offsets[0] : 0;
    hipStream_t st = reuse && ctx->copy ? ctx->copy : ctx->stream;
    uint8_t *d_raw = nullptr;
    uint64_t *d_off = nullptr;
    // the offsets rebased to 0 and the length sort, beside the upload
    std::vector<uint64_t> h_off;
    SortPlan sp;
    bool plan_ok = true; // an exception must not leave the thread
    std::thread planner([&] {
        try {
            h_off.resize(n + 1);
            const uint64_t base = n > 0 ? offsets[0] : 0;
            h_off[0] = 0;
            for (size_t t = 1; t <= n; ++t) h_off[t] = offsets[t] - base;
            plan_sort(h_off, sp);
        } catch (...) {
            plan_ok = false;
        }
    });
    hipError_t e = db_alloc(reuse ? &reuse->raw : nullptr, &d_raw, nres + kRawPad);
    if (e == hipSuccess) e = db_alloc(reuse ? &reuse->off : nullptr, &d_off, sizeof(uint64_t) * (n + 1));
    if (e == hipSuccess) e = staged_h2d(ctx, st, d_raw, pieces, np);
    planner.join();
    if (e == hipSuccess && !plan_ok) e = hipErrorOutOfMemory; // the host plan could not be allocated
    if (e == hipSuccess && encode) e = launch_encode_bytes(d_raw, nres, *encode, st);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st);
        if (!reuse && d_raw) (void)hipFree(d_raw);
        if (!reuse && d_off) (void)hipFree(d_off);
        return fail(ctx, e == hipErrorOutOfMemory ? ISH_E_NOMEM : ISH_E_HIP, "db upload: %s", hipGetErrorString(e));
    }
    // the offsets go up with the group tables
    return db_build(ctx, st, d_raw, d_off, std::move(h_off), sp, reuse, out, true);
}

// Checks the offsets (non-decreasing, targets within the int32 limit) and
// returns the longest target in *maxlen; large arrays are split over up to
// 8 host threads.
static int32_t db_offsets(ish_hip_ctx *ctx, const uint64_t *offsets, int64_t n,
                          uint64_t *maxlen = nullptr) {
    if (n > (int64_t)INT32_MAX - kGroup) return fail(ctx, ISH_E_RANGE, "too many targets");
    struct Part {
        uint64_t maxlen = 0;
        int bad = 0; // 1: decreasing, 2: too long
    };
    auto scan = [offsets](int64_t lo, int64_t hi, Part &r) {
        uint64_t m = 0;
        bool dec = false;
        for (int64_t t = lo; t < hi; ++t) {
            dec |= offsets[t + 1] < offsets[t];
            m = std::max<uint64_t>(m, offsets[t + 1] - offsets[t]);
        }
        r.bad = dec ? 1 : m > (uint64_t)INT32_MAX / 2 ? 2 : 0;
        r.maxlen = m;
    };
    const int nt = n >= ((int64_t)1 << 20) ? host_pool(ctx).size() : 1;
    std::vector<Part> parts((size_t)nt);
    const int64_t per = (n + nt - 1) / nt;
    host_pool(ctx).run(nt, [&](int k) {
        scan(std::min(n, k * per), std::min(n, (k + 1) * per), parts[(size_t)k]);
    });
    uint64_t m = 0;
    for (const Part &r : parts) {
        if (r.bad == 1) return fail(ctx, ISH_E_INVALID, "offsets must be non-decreasing");
        if (r.bad == 2) return fail(ctx, ISH_E_RANGE, "target too long");
        m = std::max(m, r.maxlen);
    }
    if (maxlen) *maxlen = m;
    return ISH_OK;
}

// slot >= 0: the database borrows the context's one-shot buffers
// once[slot]; checked: the offsets were checked by the caller
static int32_t db_load_impl(ish_hip_ctx *ctx, const uint8_t *residues, const uint64_t *offsets,
                            int64_t n, const EncodeTable *encode, ish_hip_db **out, int slot = -1,
                            bool checked = false) {
    if (!ctx || !out || n < 0 || (n > 0 && (!offsets || !residues)))
        return fail(ctx, ISH_E_INVALID, "bad db arguments");
    *out = nullptr;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int32_t rc = checked ? ISH_OK : db_offsets(ctx, offsets, n);
    if (rc) return rc;
    const ish::HostPiece one{n > 0 ? residues + offsets[0] : residues, n > 0 ? offsets[n] - offsets[0] : 0};
    return db_upload(ctx, &one, 1, offsets, n, encode, slot >= 0 ? &ctx->once[slot & 1] : nullptr, out);
}

int32_t ish::db_load_ascii_pieces(ish_hip_ctx *ctx, const HostPiece *pieces, size_t np,
                                  const uint64_t *offsets, int64_t n, int32_t matrix_kind,
                                  ish_hip_db **out, int slot) {
    if (!ctx || !out || n < 0 || !offsets || (np > 0 && !pieces) || offsets[0] != 0)
        return fail(ctx, ISH_E_INVALID, "bad db arguments");
    *out = nullptr;
    EncodeTable t;
    if (ish_hip_encoding(matrix_kind, t.code) != ISH_OK)
        return fail(ctx, ISH_E_INVALID, "unknown matrix kind %d", matrix_kind);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int32_t rc = db_offsets(ctx, offsets, n);
    if (rc) return rc;
    uint64_t total = 0;
    for (size_t k = 0; k < np; ++k) total += pieces[k].n;
    if (total != offsets[n]) return fail(ctx, ISH_E_INVALID, "pieces do not match the offsets");
    return db_upload(ctx, pieces, np, offsets, n, &t, slot >= 0 ? &ctx->once[slot & 1] : nullptr, out);
}

extern "C" int32_t ish_hip_db_load(ish_hip_ctx *ctx, const uint8_t *residues,
                                   const uint64_t *offsets, int64_t n, ish_hip_db **out) {
    return db_load_impl(ctx, residues, offsets, n, nullptr, out);
}

extern "C" int32_t ish_hip_db_load_ascii(ish_hip_ctx *ctx, const uint8_t *ascii,
                                         const uint64_t *offsets, int64_t n, int32_t matrix_kind,
                                         ish_hip_db **out) {
    EncodeTable t;
    if (ish_hip_encoding(matrix_kind, t.code) != ISH_OK)
        return fail(ctx, ISH_E_INVALID, "unknown matrix kind %d", matrix_kind);
    return db_load_impl(ctx, ascii, offsets, n, &t, out);
}

extern "C" void ish_hip_db_free(ish_hip_db *db) {
    if (!db) return;
    if (db->borrowed) { // the context keeps the buffers
        delete db;
        return;
    }
    (void)hipSetDevice(db->device);
    for (void *p : {(void *)db->d_groups, (void *)db->d_res, (void *)db->d_lens, (void *)db->d_orig,
                    (void *)db->d_score, (void *)db->d_end_ref, (void *)db->d_end_q,
                    (void *)db->d_status, (void *)db->d_raw, (void *)db->d_off})
        if (p) (void)hipFree(p);
    delete db;
}

extern "C" int32_t ish_hip_db_size(const ish_hip_db *db, int64_t *n, int64_t *nres) {
    if (!db) return ISH_E_INVALID;
    if (n) *n = db->n;
    if (nres) *nres = db->nres;
    return ISH_OK;
}

// Enqueues the scoring kernels over every group of db with the given
// profile (forward or reversed query) and free-end flags.
// One persistent launch of a db pass.
struct DbLaunch {
    int32_t kind; // kSg16, kSg32, kSw16, kSw32, kSwPipe, kSgHybrid, kSgLock
    int32_t chunk;
    int32_t g_begin, g_end;
    int32_t grid;
    int32_t n_long; // kSgHybrid / kSgLock: leading groups scored by whole workgroups
    int32_t paired; // kSwPipe / kSgHybrid: the paired-profile variant (kSgLock: always)
    int32_t waves;  // waves per workgroup (kSgHybrid: 2 or 4; the others: 4)
    int32_t f16;    // kSgLock: the tiles run in binary16 lanes; paired kSgHybrid: every group
    int64_t slot_stride;
    uint64_t lds;
    ScoreParams p;
};
enum { kSg16 = 0, kSg32 = 1, kSw16 = 2, kSw32 = 3, kSwPipe = 4, kSgHybrid = 5, kSgLock = 6 };

// The launches of one db pass: at most one over the int32 groups and one
// over the int16 groups. A fixed array, so that planning a pass allocates
// nothing.
struct DbPlan {
    DbLaunch l[2];
    int n = 0;
    // plans are compared bytewise (graph replay, the autotuner's
    // duplicates): the padding between fields must be zero too
    DbPlan() { memset((void *)l, 0, sizeof l); }
    void push(const DbLaunch &L) { memcpy((void *)&l[n++], (const void *)&L, sizeof L); }
    DbLaunch *begin() { return l; }
    DbLaunch *end() { return l + n; }
    const DbLaunch *begin() const { return l; }
    const DbLaunch *end() const { return l + n; }
    bool same(const DbPlan &o) const { return n == o.n && memcmp(l, o.l, sizeof(DbLaunch) * (size_t)n) == 0; }
};

// Caller-owned device scratch of a pass that must not allocate: the work
// counter in its first 256 bytes, the boundary rows after it. The grid of
// each launch shrinks to the wave slots the rest holds.
struct FixedScratch {
    uint8_t *base;
    size_t bytes;
};
constexpr size_t kFixedCounterBytes = 256;

static int32_t len_bucket(int32_t maxlen) {
    for (int32_t b = 0; b < kLenBuckets - 1; ++b)
        if (maxlen >= kLenBucketLo[b]) return b;
    return kLenBuckets - 1;
}

static TuneKey tune_key(const ish_hip_ctx *ctx, const ish_hip_db *db) {
    int32_t qb = 0;
    while ((1 << qb) < ctx->Q) ++qb;
    const int32_t longest = db->ngroups > 0 ? db->group_maxlen[0] : 0;
    return TuneKey{ctx->kind, ctx->algo, qb, len_bucket(longest)};
}

// Heuristic schedule of the local int16 groups [g0, g1) at one tile height:
// the workgroup pipeline when the longest group is far longer than the
// average work per wave (long reads), else one wave per group.
static bool pipe_heuristic(const ish_hip_ctx *ctx, const ish_hip_db *db, int32_t g0, int32_t g1,
                           int32_t chunk) {
    const int nch = (ctx->Q + chunk - 1) / chunk;
    const int rounds = (nch + 3) / 4;
    if (nch < 4 || 4 * rounds - nch > rounds) return false;
    double cols = 0;
    for (int32_t g = g0; g < g1; ++g) cols += db->group_maxlen[g] + 3;
    const int occ = occupancy_sw_batched_i16(chunk, kBlock, ctx->lds);
    const double per_wave = cols / ((double)occ * ctx->cus * (kBlock / 64));
    return db->group_maxlen[g0] > 1.5 * per_wave;
}

// Semi-global groups [g0, g1) whose longest target exceeds half the
// average work per resident wave: the leading ones that a hybrid launch
// scores with whole workgroups (so that no single wave holds the launch
// open); 0 when none is that long.
static int32_t long_groups(const ish_hip_ctx *ctx, const ish_hip_db *db, int32_t g0, int32_t g1,
                           int32_t chunk, uint32_t flags) {
    double cols = 0;
    for (int32_t g = g0; g < g1; ++g) cols += db->group_maxlen[g] + 3;
    const int occ = occupancy_sg_hybrid_i16(chunk, false, false, 4, ctx->lds + sg_pipe_extra_lds(), flags);
    const double per_wave = cols / ((double)occ * ctx->cus * (kBlock / 64));
    int32_t n = 0;
    while (g0 + n < g1 && db->group_maxlen[g0 + n] > 0.5 * per_wave) ++n;
    return n;
}

// Paired profile of the current query for db's residue codes, built and
// uploaded when the query or the code set changed. Returns its class count
// (0: those codes need more than kPairMaxShared classes) or an ISH_E_*
// code. Per-wave paired launches take at most kPairMaxRows classes, the
// workgroup-shared one (kSgLock) up to kPairMaxShared.
static int32_t ensure_pairs(ish_hip_ctx *ctx, const ish_hip_db *db) {
    if (ctx->pair_gen == ctx->query_gen && memcmp(ctx->pair_mask, db->code_mask, sizeof db->code_mask) == 0)
        return ctx->pair_rows;
    // a queued launch may still read the buffers
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->pair_gen = ctx->query_gen;
    memcpy(ctx->pair_mask, db->code_mask, sizeof db->code_mask);
    ctx->pair_rows = 0;
    if (!build_pairs(ctx->hprof, db->code_mask, kPairMaxShared, ctx->pairs)) return 0;
    const size_t wb = ctx->pairs.words.size() * 4, tb = ctx->pairs.table.size() * 2;
    HIP_TRY(ctx, ctx->pprof.ensure(wb));
    HIP_TRY(ctx, ctx->plut.ensure(tb));
    if (wb) HIP_TRY(ctx, hipMemcpyAsync(ctx->pprof.p, ctx->pairs.words.data(), wb, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->plut.p, ctx->pairs.table.data(), tb, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->pair_rows = ctx->pairs.classes;
    ctx->pqpad = ctx->pairs.qpad;
    // binary16 lanes: semi-global, the widest group admitted; local, any
    // group (INT64_MAX) or none
    if (ctx->algo == ISH_ALGO_LOCAL)
        ctx->w16 = sw_f16_bound(ctx->hprof, ctx->Q, ctx->go, ctx->ge, ctx->min_score, db->code_mask, &ctx->fshift)
                       ? INT64_MAX
                       : -1;
    else
        ctx->w16 = f16_bound(ctx->hprof, ctx->Q, ctx->go, ctx->ge, ctx->min_score, db->code_mask, &ctx->fshift);
    return ctx->pair_rows;
}

// Plans the launches that score groups [g_lo, g_hi) of db (int32 groups
// first), sizes the boundary scratch for all of them, then fills in every
// launch's parameters. `forced` >= 0 fixes (chunk, schedule) for the int16
// groups (autotuning), and forced_waves > 0 the width of pipelined
// semi-global groups; otherwise the tuning table or the defaults decide.
// With `fixed`, the pass allocates nothing: no paired profile is built (the
// launches read the int16 profile of set_query), and the counter and the
// boundary rows live in the fixed scratch.
static int32_t plan_db(ish_hip_ctx *ctx, ish_hip_db *db, const int16_t *prof, uint32_t flags,
                       int32_t g_lo, int32_t g_hi, int forced_chunk, int forced_sched,
                       DbPlan &out, int forced_waves = 0, const FixedScratch *fixed = nullptr) {
    out.n = 0;
    if (g_hi <= g_lo) return ISH_OK;
    if (!fixed) HIP_TRY(ctx, ctx->counter.ensure(256));
    const bool local = ctx->algo == ISH_ALGO_LOCAL;
    // groups whose longest target exceeds the int16 bound come first
    // (descending lengths); they run in the int32 kernel.
    int32_t g16 = 0;
    {
        int32_t lo = 0, hi = db->ngroups;
        while (lo < hi) {
            const int32_t mid = (lo + hi) / 2;
            if ((int64_t)db->group_maxlen[mid] > ctx->r16max) lo = mid + 1;
            else hi = mid;
        }
        g16 = lo;
    }
    const int32_t a = g_lo, m16 = std::max(g_lo, std::min(g16, g_hi));
    if (m16 > a) {
        DbLaunch L;
        memset((void *)&L, 0, sizeof L);
        L.kind = local ? kSw32 : kSg32;
        L.chunk = ctx->chunk == 40 || ctx->chunk == 24 ? 32 : ctx->chunk; // the int32 kernels have 16 and 32
        L.g_begin = a;
        L.g_end = m16;
        out.push(L);
    }
    const size_t pipe_lds = ctx->lds + sw_pipe_extra_lds();
    const size_t hybrid_lds = ctx->lds + sg_pipe_extra_lds();
    // forward passes over a database whose residue codes fall into few
    // profile classes read the paired profile (pipelined local launches and
    // hybrid semi-global launches, which then also serve the one-wave
    // schedule with no pipelined groups)
    int32_t pc = 0;
    if (prof == ctx->prof_t.p && !fixed) {
        pc = ensure_pairs(ctx, db);
        if (pc < 0) return pc;
    }
    auto add16 = [&](int32_t g0, int32_t g1, int chunk, int sched, int waves) {
        if (g1 <= g0) return;
        // the 24-row tile exists for the workgroup-shared lockstep launch
        // only; the others take 32
        if (chunk == 24 && (local || pc <= kPairMaxRows || sg_hybrid_lock_lds(24, pc, ctx->lds) > kLdsLimit))
            chunk = 32;
        // the 40-row tile exists for the hybrid semi-global kernel only
        // (per-wave paired or over the int16 profile); the others take 32
        if (chunk == 40 &&
            (local || pc > kPairMaxRows ||
             !(hybrid_lds <= kLdsLimit || (pc > 0 && sg_hybrid_paired_lds(40, pc, 4) <= kLdsLimit))))
            chunk = 32;
        DbLaunch L;
        memset((void *)&L, 0, sizeof L);
        bool pipe = false;
        int32_t n_long = 0;
        const bool sg_paired = !local && pc > 0 && pc <= kPairMaxRows &&
                               sg_hybrid_paired_lds(chunk, pc, 4) <= kLdsLimit;
        const bool sg_lock = !local && pc > kPairMaxRows &&
                             sg_hybrid_lock_lds(chunk, pc, ctx->lds) <= kLdsLimit;
        if (local && pipe_lds <= kLdsLimit) {
            if (sched == ISH_SCHEDULE_WORKGROUP) pipe = true;
            else if (sched == ISH_SCHEDULE_AUTO) pipe = pipe_heuristic(ctx, db, g0, g1, chunk);
        } else if (!local && (hybrid_lds <= kLdsLimit || sg_paired || sg_lock)) {
            if (sched == ISH_SCHEDULE_WORKGROUP) n_long = g1 - g0;
            else if (sched == ISH_SCHEDULE_AUTO) n_long = long_groups(ctx, db, g0, g1, chunk, flags);
        }
        L.kind = pipe ? kSwPipe : local ? kSw16 : sg_lock ? kSgLock : (n_long > 0 || sg_paired) ? kSgHybrid : kSg16;
        if (chunk == 40 && L.kind != kSgHybrid) {
            // no 40-row variant of these launches: the hybrid kernel with no
            // pipelined group is the one-wave schedule
            L.kind = kSgHybrid;
        }
        L.waves = 4;
        if (L.kind == kSgHybrid && n_long > 0) {
            // a pipelined group runs its chunks round by round over the
            // workgroup's waves. Two waves leave fewer idle chunk slots in
            // the last round when the chunk count is 1 or 2 modulo 4 (five
            // chunks: 5 of 6 slots busy against 5 of 8). From 8 chunks on,
            // four waves win anyway: a group then takes fewer rounds, so the
            // longest groups, which end the launch, finish sooner (a 150
            // row query over 200 k targets of 100-300 residues, 1,563
            // groups: ten 16-row chunks took 0.90 ms over four waves and
            // 0.99 ms over two; five 32-row chunks 1.02 ms over two and
            // 1.18 ms over four).
            const int nch = (ctx->Q + chunk - 1) / chunk;
            if (nch < 8 && 2 * ((nch + 1) / 2) < 4 * ((nch + 3) / 4)) L.waves = 2;
            if (waves == 2 || waves == 4) L.waves = waves;
            // the 40-row tile runs on four waves only (its two-wave
            // variants would spill registers)
            if (chunk == 40) L.waves = 4;
        }
        if (L.kind == kSgLock && ctx->f16 && ctx->w16 >= 4) {
            // binary16 tiles: the groups wider than the range bound join the
            // pipelined long groups (a prefix, groups being sorted by
            // length), if they are few: at most 4, 1/16 of the groups, or
            // the long groups the schedule pipelines anyway
            int32_t n = 0;
            while (g0 + n < g1 && (((int64_t)db->group_maxlen[g0 + n] + 3) & ~3LL) > ctx->w16) ++n;
            if (n <= std::max<int32_t>(std::max<int32_t>(n_long, (g1 - g0) / 16), 4)) {
                L.f16 = 1;
                n_long = std::max(n_long, n);
            }
        }
        L.chunk = chunk;
        L.g_begin = g0;
        L.g_end = g1;
        L.n_long = n_long;
        out.push(L);
    };
    if (forced_chunk > 0) {
        add16(m16, g_hi, forced_chunk, forced_sched, forced_waves);
    } else {
        auto it = ctx->tuned ? ctx->tune.find(tune_key(ctx, db)) : ctx->tune.end();
        if (it == ctx->tune.end()) {
            add16(m16, g_hi, ctx->chunk, ctx->schedule, 0);
        } else {
            // the tuned configuration for the whole pass (one launch: a
            // launch per length range would add a tail per launch)
            const TuneChoice &c = it->second;
            add16(m16, g_hi, c.chunk, ctx->schedule != ISH_SCHEDULE_AUTO ? ctx->schedule : c.schedule,
                  c.waves);
        }
    }
    // grids and scratch for every launch before any is enqueued (growing the
    // scratch between launches would free memory a queued launch uses)
    for (DbLaunch &L : out) {
        L.lds = L.kind == kSwPipe ? pipe_lds : L.kind == kSgHybrid ? hybrid_lds : ctx->lds;
        if (L.kind == kSgLock) {
            L.paired = 1;
            L.lds = sg_hybrid_lock_lds(L.chunk, pc, ctx->lds);
        } else if (L.kind == kSwPipe && pc > 0 && pc <= kPairMaxRows &&
                   sw_pipe_paired_lds(L.chunk, pc) <= kLdsLimit) {
            L.paired = 1;
            L.lds = sw_pipe_paired_lds(L.chunk, pc);
            // binary16 lanes: the local bound holds for any target length
            if (ctx->f16 && ctx->w16 == INT64_MAX) L.f16 = 1;
        } else if (L.kind == kSgHybrid && pc > 0 && pc <= kPairMaxRows &&
                   sg_hybrid_paired_lds(L.chunk, pc, 4) <= kLdsLimit) {
            L.paired = 1;
            L.lds = sg_hybrid_paired_lds(L.chunk, pc, L.waves);
            // binary16 lanes when every group of the launch lies within
            // the query's range bound
            if (ctx->f16 && ctx->w16 >= (((int64_t)db->group_maxlen[L.g_begin] + 3) & ~3LL)) L.f16 = 1;
        }
        int occ = 1;
        switch (L.kind) {
        case kSg32: occ = occupancy_sg_batched_i32(L.chunk, kBlock, L.lds, flags); break;
        case kSg16: occ = occupancy_sg_batched_i16(L.chunk, kBlock, L.lds, flags); break;
        case kSw32: occ = occupancy_sw_batched_i32(L.chunk, kBlock, L.lds); break;
        case kSw16: occ = occupancy_sw_batched_i16(L.chunk, kBlock, L.lds); break;
        case kSgHybrid:
            occ = occupancy_sg_hybrid_i16(L.chunk, L.paired != 0, L.f16 != 0, L.waves, L.lds, flags);
            break;
        case kSgLock: occ = occupancy_sg_hybrid_lock(L.chunk, L.f16 != 0, L.lds, flags); break;
        default: occ = occupancy_sw_pipe_i16(L.chunk, L.paired != 0, L.f16 != 0, L.lds); break;
        }
        int grid;
        int64_t ss;
        const int wpb = L.kind == kSgHybrid ? L.waves : kBlock / 64;
        const int64_t cols = db->group_maxlen[L.g_begin] + 3;
        if (fixed) {
            // as many workgroups as the fixed scratch holds boundary rows for
            const int64_t per_block = (int64_t)wpb * std::max<int64_t>(cols, 4) * 64 * 8;
            const int64_t fit = (int64_t)(fixed->bytes - kFixedCounterBytes) / per_block;
            if (fit < 1)
                return fail(ctx, ISH_E_RANGE,
                            "a target of %lld residues needs %lld bytes of boundary rows per "
                            "workgroup; the fixed scratch holds %zu",
                            (long long)db->group_maxlen[L.g_begin], (long long)per_block,
                            fixed->bytes - kFixedCounterBytes);
            grid = (int)std::min<int64_t>((int64_t)occ * ctx->cus, fit);
            ss = std::max<int64_t>(cols, 4) * 64;
        } else {
            int32_t rc = size_grid(ctx, occ, cols, &grid, &ss, wpb);
            if (rc) return rc;
        }
        L.grid = grid;
        L.slot_stride = ss;
    }
    for (DbLaunch &L : out) {
        ScoreParams &p = L.p;
        p.prof = prof;
        p.lut = (const uint16_t *)ctx->lut.p;
        p.prof_words = ctx->prof_words;
        p.stride = ctx->stride;
        p.Q = ctx->Q;
        p.nchunks = (ctx->Q + L.chunk - 1) / L.chunk;
        p.pad = ctx->top ? top_pad(ctx->Q, L.chunk) : 0;
        p.front = ctx->front;
        p.go = ctx->go;
        p.ge = ctx->ge;
        p.flags = flags;
        p.neg_limit = ctx->neg_limit;
        p.fshift = ctx->fshift;
        p.groups = db->d_groups;
        p.res = db->d_res;
        p.lens = db->d_lens;
        p.orig = db->d_orig;
        p.counter = fixed ? (uint32_t *)fixed->base : (uint32_t *)ctx->counter.p;
        p.score = db->d_score;
        p.end_ref = db->d_end_ref;
        p.end_q = db->d_end_q;
        p.status = db->d_status;
        p.bnd = fixed ? (uint2 *)(fixed->base + kFixedCounterBytes) : (uint2 *)ctx->scratch.p;
        p.slot_stride = L.slot_stride;
        p.g_begin = L.g_begin;
        p.g_end = L.g_end;
        p.n_long = L.n_long;
        p.pprof = L.paired ? (const int32_t *)ctx->pprof.p : nullptr;
        p.plut = L.paired ? (const uint16_t *)ctx->plut.p : nullptr;
        p.pnrows = L.paired ? ctx->pair_rows : 0;
        p.pqpad = L.paired ? ctx->pqpad : 0;
    }
    return ISH_OK;
}

static hipError_t enqueue_launch(ish_hip_ctx *ctx, const DbLaunch &L, bool timed) {
    hipError_t e = launch_reset(L.p.counter, ctx->stream);
    if (e == hipSuccess && timed) e = timed_begin(ctx);
    if (e != hipSuccess) return e;
    switch (L.kind) {
    case kSg32: e = launch_sg_batched_i32(L.p, L.chunk, L.grid, kBlock, L.lds, ctx->stream); break;
    case kSg16: e = launch_sg_batched_i16(L.p, L.chunk, L.grid, kBlock, L.lds, ctx->stream); break;
    case kSw32: e = launch_sw_batched_i32(L.p, L.chunk, L.grid, kBlock, L.lds, ctx->stream); break;
    case kSw16: e = launch_sw_batched_i16(L.p, L.chunk, L.grid, kBlock, L.lds, ctx->stream); break;
    case kSgHybrid:
        e = launch_sg_hybrid_i16(L.p, L.chunk, L.paired != 0, L.f16 != 0, L.waves, L.grid, L.lds, ctx->stream);
        break;
    case kSgLock: e = launch_sg_hybrid_lock(L.p, L.chunk, L.f16 != 0, L.grid, L.lds, ctx->stream); break;
    default: e = launch_sw_pipe_i16(L.p, L.chunk, L.paired != 0, L.f16 != 0, L.grid, L.lds, ctx->stream); break;
    }
    if (e == hipSuccess && timed) e = timed_end(ctx);
    return e;
}

static void drop_graph(ish_hip_ctx *ctx) {
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->graph_exec = nullptr;
    ctx->graph = nullptr;
    ctx->graph_sig.clear();
}

// Scores db against the context's query: eager launches, or a replay of the
// HIP graph captured for exactly these launches (same kernels, grids and
// parameters, pointers included) when graphs are on.
static int32_t launch_db(ish_hip_ctx *ctx, ish_hip_db *db, const int16_t *prof, uint32_t flags,
                         bool allow_graph = true) {
    ctx->launches = 0;
    if (db->ngroups == 0) return ISH_OK;
    DbPlan plan;
    int32_t rc = plan_db(ctx, db, prof, flags, 0, db->ngroups, 0, 0, plan);
    if (rc) return rc;
    if (!(ctx->graphs && allow_graph)) {
        for (const DbLaunch &L : plan) HIP_TRY(ctx, enqueue_launch(ctx, L, true));
        return ISH_OK;
    }
    std::vector<uint8_t> sig((size_t)plan.n * sizeof(DbLaunch));
    if (plan.n) memcpy(sig.data(), plan.l, sig.size());
    if (!ctx->graph_exec || sig != ctx->graph_sig) {
        drop_graph(ctx);
        HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (const DbLaunch &L : plan)
            if (e == hipSuccess) e = enqueue_launch(ctx, L, false);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
        if (e == hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            ctx->graph_exec = nullptr;
            return fail(ctx, ISH_E_HIP, "graph capture: %s", hipGetErrorString(e));
        }
        ctx->graph = g;
        ctx->graph_sig = sig;
    }
    HIP_TRY(ctx, timed_begin(ctx));
    HIP_TRY(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
    HIP_TRY(ctx, timed_end(ctx));
    return ISH_OK;
}

extern "C" int32_t ish_hip_set_graphs(ish_hip_ctx *ctx, int32_t on) {
    if (!ctx) return ISH_E_INVALID;
    ctx->graphs = on != 0;
    if (!ctx->graphs) {
        (void)hipStreamSynchronize(ctx->stream);
        drop_graph(ctx);
    }
    return ISH_OK;
}

// Times whole passes over db for each candidate plan: tile height 32 or
// 16; schedule (one wave per group, automatic, one workgroup per group);
// and, for pipelined semi-global groups, two or four waves per workgroup.
// Plans that come out identical are timed once. The fastest is kept under
// (matrix kind, algorithm, query-length bucket, length bucket of db's
// longest group).
extern "C" int32_t ish_hip_autotune(ish_hip_ctx *ctx, ish_hip_db *db, int32_t reps) {
    if (!ctx || !db) return ISH_E_INVALID;
    if (!ctx->have_query) return fail(ctx, ISH_E_NO_QUERY, "no query set");
    if (db->device != ctx->device) return fail(ctx, ISH_E_INVALID, "db belongs to another device");
    if (ctx->engine != ISH_ENGINE_BATCHED) return fail(ctx, ISH_E_INVALID, "autotuning needs the batched engine");
    if (reps <= 0) reps = 2;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int32_t chunks[4] = {32, 16, 40, 24};
    const int scheds[3] = {ISH_SCHEDULE_WAVE, ISH_SCHEDULE_AUTO, ISH_SCHEDULE_WORKGROUP};
    TuneChoice best{(int8_t)kChunkDefault, (int8_t)ISH_SCHEDULE_AUTO, 0};
    float best_ms = 1e30f;
    const int saved_launches = ctx->launches;
    // what a plan launches, without the buffers it points at (which a
    // grown scratch changes): candidates that plan the same launches are
    // timed once
    auto launch_key = [](const DbPlan &plan) {
        std::vector<int32_t> k;
        for (const DbLaunch &L : plan) {
            for (int32_t v : {L.kind, L.chunk, L.g_begin, L.g_end, L.grid, L.n_long, L.paired, L.waves, L.f16})
                k.push_back(v);
        }
        return k;
    };
    std::vector<std::vector<int32_t>> seen;
    // pipelined semi-global groups also try both workgroup widths (a plan
    // that has none comes out the same for each and is timed once)
    const int widths[3] = {0, 2, 4};
    struct Candidate {
        int32_t chunk;
        int sched, waves;
    };
    std::vector<Candidate> cands;
    for (int32_t ch : chunks)
        for (int sched : scheds)
            for (int wv : widths) cands.push_back({ch, sched, wv});
    for (const Candidate &c : cands) {
        // whole passes: the launch tail is part of what is being tuned
        DbPlan plan;
        int32_t rc = plan_db(ctx, db, (const int16_t *)ctx->prof_t.p, ctx->flags, 0, db->ngroups,
                             c.chunk, c.sched, plan, c.waves);
        if (rc) return rc;
        std::vector<int32_t> key = launch_key(plan);
        bool dup = false; // a candidate that plans the same launches as one already timed
        for (const auto &q : seen) dup = dup || q == key;
        if (dup) continue;
        seen.push_back(std::move(key));
        // the tile height the int16 launch runs with (a candidate height
        // that launch lacks falls back to another)
        int32_t eff = c.chunk;
        for (const DbLaunch &L : plan)
            if (L.kind != kSg32 && L.kind != kSw32) eff = L.chunk;
        float total = 0.f;
        for (int rep = 0; rep <= reps; ++rep) {
            ctx->launches = 0;
            for (const DbLaunch &L : plan) HIP_TRY(ctx, enqueue_launch(ctx, L, true));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            collect_timing(ctx);
            if (rep > 0) total += ctx->last_ms; // the first run warms up
        }
        if (total < best_ms) {
            best_ms = total;
            best = TuneChoice{(int8_t)eff, (int8_t)c.sched, (int8_t)c.waves};
        }
    }
    ctx->launches = saved_launches;
    ctx->tune[tune_key(ctx, db)] = best;
    ctx->tuned = true;
    return ISH_OK;
}

extern "C" int32_t ish_hip_tuning(const ish_hip_ctx *ctx, const ish_hip_db *db, int32_t *tile,
                                  int32_t *schedule) {
    if (!ctx || !db) return ISH_E_INVALID;
    if (!ctx->have_query) return ISH_E_NO_QUERY;
    auto it = ctx->tune.find(tune_key(ctx, db));
    if (it == ctx->tune.end()) return 0;
    if (tile) *tile = it->second.chunk;
    if (schedule) *schedule = it->second.schedule;
    return 1;
}

extern "C" int32_t ish_hip_align_db_async(ish_hip_ctx *ctx, ish_hip_db *db) {
    if (!ctx || !db) return ISH_E_INVALID;
    if (!ctx->have_query) return fail(ctx, ISH_E_NO_QUERY, "no query set");
    if (db->device != ctx->device) return fail(ctx, ISH_E_INVALID, "db belongs to another device");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return launch_db(ctx, db, (const int16_t *)ctx->prof_t.p, ctx->flags);
}

int32_t ish::ctx_cus(const ish_hip_ctx *ctx) { return ctx ? ctx->cus : 0; }
int32_t ish::ctx_device(const ish_hip_ctx *ctx) { return ctx ? ctx->device : -1; }

int32_t ish::one_shot_begin(ish_hip_ctx *ctx) {
    if (ctx->copy) return ISH_OK;
    HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->copy, hipStreamNonBlocking));
    return ISH_OK;
}

void ish::one_shot_end(ish_hip_ctx *ctx) {
    if (!ctx->copy) return;
    (void)hipStreamSynchronize(ctx->copy);
    // Events must not outlive the stream they were last recorded on: HIP
    // looks at that stream when an event is waited for. The staging events
    // may have been recorded on the copy stream; they are complete now, and
    // are created again (and recorded) at their next use.
    for (hipEvent_t *e : {&ctx->stage_done[0], &ctx->stage_done[1], &ctx->tab_done}) {
        if (*e) (void)hipEventDestroy(*e);
        *e = nullptr;
    }
    (void)hipStreamDestroy(ctx->copy);
    ctx->copy = nullptr;
}

int32_t ish::db_score_once_async(ish_hip_ctx *ctx, ish_hip_db *db, int slot) {
    if (!ctx || !db) return ISH_E_INVALID;
    if (!ctx->have_query) return fail(ctx, ISH_E_NO_QUERY, "no query set");
    if (db->device != ctx->device) return fail(ctx, ISH_E_INVALID, "db belongs to another device");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int32_t rc = launch_db(ctx, db, (const int16_t *)ctx->prof_t.p, ctx->flags, false);
    if (rc) return rc;
    HIP_TRY(ctx, hipEventRecord(ctx->pass_done[slot & 1], ctx->stream));
    return ISH_OK;
}

// Copies db's results to the host arrays on stream st (pinned landing
// buffer, then one host thread per array) and returns when they are there.
static int32_t fetch_results(ish_hip_ctx *ctx, ish_hip_db *db, hipStream_t st, int32_t *score,
                             int32_t *end_ref, int32_t *end_q, uint8_t *status) {
    const int64_t n = db->n;
    if (n <= 0) {
        HIP_TRY(ctx, hipStreamSynchronize(st));
        return ISH_OK;
    }
    const size_t need = (size_t)n * 13;
    if (ctx->res_host_bytes < need) {
        if (ctx->res_host) (void)hipHostFree(ctx->res_host);
        ctx->res_host = nullptr;
        ctx->res_host_bytes = 0;
        HIP_TRY(ctx, hipHostMalloc((void **)&ctx->res_host, need, hipHostMallocDefault));
        ctx->res_host_bytes = need;
    }
    uint8_t *h = ctx->res_host;
    const size_t w = sizeof(int32_t) * (size_t)n;
    HIP_TRY(ctx, hipMemcpyAsync(h, db->d_score, w, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipMemcpyAsync(h + w, db->d_end_ref, w, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipMemcpyAsync(h + 2 * w, db->d_end_q, w, hipMemcpyDeviceToHost, st));
    if (status) HIP_TRY(ctx, hipMemcpyAsync(h + 3 * w, db->d_status, (size_t)n, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    struct Out {
        void *dst;
        const uint8_t *src;
        size_t bytes;
    };
    std::vector<Out> outs = {{score, h, w}, {end_ref, h + w, w}, {end_q, h + 2 * w, w}};
    if (status) outs.push_back({status, h + 3 * w, (size_t)n});
    if (n < ((int64_t)1 << 16)) {
        for (const Out &o : outs) memcpy(o.dst, o.src, o.bytes);
    } else {
        host_pool(ctx).run((int)outs.size(), [&](int k) { memcpy(outs[k].dst, outs[k].src, outs[k].bytes); });
    }
    return ISH_OK;
}

int32_t ish::db_fetch_after(ish_hip_ctx *ctx, ish_hip_db *db, int slot, int32_t *score,
                            int32_t *end_ref, int32_t *end_q, uint8_t *status) {
    if (!ctx || !db || (db->n > 0 && (!score || !end_ref || !end_q))) return ISH_E_INVALID;
    if (!ctx->copy) return fail(ctx, ISH_E_INVALID, "one-shot fetch outside a one-shot call");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamWaitEvent(ctx->copy, ctx->pass_done[slot & 1], 0));
    return fetch_results(ctx, db, ctx->copy, score, end_ref, end_q, status);
}

extern "C" int32_t ish_hip_db_fetch(ish_hip_ctx *ctx, ish_hip_db *db, int32_t *score,
                                    int32_t *end_ref, int32_t *end_q, uint8_t *status) {
    if (!ctx || !db || (db->n > 0 && (!score || !end_ref || !end_q)))
        return ISH_E_INVALID;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int32_t rc = fetch_results(ctx, db, ctx->stream, score, end_ref, end_q, status);
    collect_timing(ctx);
    return rc;
}

extern "C" int32_t ish_hip_align_db(ish_hip_ctx *ctx, ish_hip_db *db, int32_t *score,
                                    int32_t *end_ref, int32_t *end_q, uint8_t *status) {
    int32_t rc = ish_hip_align_db_async(ctx, db);
    if (rc) return rc;
    return ish_hip_db_fetch(ctx, db, score, end_ref, end_q, status);
}

// ish_hip_align_db with caller-owned device outputs and the boundary rows
// and work counter in fixed device scratch: no device memory is allocated,
// no paired profile is built, no event is created and nothing is copied
// between host and device. Waits for the pass.
int32_t ish::align_db_fixed(ish_hip_ctx *ctx, ish_hip_db *db, void *scratch, size_t scratch_bytes,
                            int32_t *d_score, int32_t *d_end_ref, int32_t *d_end_q,
                            uint8_t *d_status) {
    if (!ctx || !db) return ISH_E_INVALID;
    if (!ctx->have_query) return fail(ctx, ISH_E_NO_QUERY, "no query set");
    if (db->device != ctx->device) return fail(ctx, ISH_E_INVALID, "db belongs to another device");
    if (db->n > 0 && (!d_score || !d_end_ref || !d_end_q || !d_status))
        return fail(ctx, ISH_E_INVALID, "null device outputs");
    if (ctx->engine != ISH_ENGINE_BATCHED)
        return fail(ctx, ISH_E_INVALID, "the fixed-scratch pass runs the batched engine only");
    if (!scratch || scratch_bytes <= kFixedCounterBytes)
        return fail(ctx, ISH_E_INVALID, "no fixed scratch on device %d", ctx->device);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    ctx->launches = 0;
    if (db->ngroups == 0) return ISH_OK;
    const FixedScratch fs{(uint8_t *)scratch, scratch_bytes};
    DbPlan plan;
    int32_t rc = plan_db(ctx, db, (const int16_t *)ctx->prof_t.p, ctx->flags, 0, db->ngroups, 0, 0,
                         plan, 0, &fs);
    if (rc) return rc;
    for (DbLaunch &L : plan) {
        L.p.score = d_score;
        L.p.end_ref = d_end_ref;
        L.p.end_q = d_end_q;
        L.p.status = d_status;
        HIP_TRY(ctx, enqueue_launch(ctx, L, false));
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return ISH_OK;
}

// ---------------------------------------------------------------------------
// one synchronous batch
// ---------------------------------------------------------------------------
static int32_t align_plain(ish_hip_ctx *ctx, const uint8_t *residues, const uint64_t *offsets,
                           int64_t n, int32_t *score, int32_t *end_ref, int32_t *end_q,
                           uint8_t *status) {
    const uint64_t base = offsets[0];
    const uint64_t nres = offsets[n] - base;
    int64_t maxlen = 0;
    std::vector<uint64_t> rebased((size_t)n + 1);
    for (int64_t t = 0; t <= n; ++t) {
        if (t < n && offsets[t + 1] < offsets[t]) return fail(ctx, ISH_E_INVALID, "offsets must be non-decreasing");
        rebased[(size_t)t] = offsets[t] - base;
        if (t < n) maxlen = std::max<int64_t>(maxlen, (int64_t)(offsets[t + 1] - offsets[t]));
    }
    HIP_TRY(ctx, ctx->in_res.ensure(std::max<uint64_t>(nres, 1)));
    HIP_TRY(ctx, ctx->in_off.ensure(sizeof(uint64_t) * (n + 1)));
    HIP_TRY(ctx, ctx->out_score.ensure(sizeof(int32_t) * n));
    HIP_TRY(ctx, ctx->out_er.ensure(sizeof(int32_t) * n));
    HIP_TRY(ctx, ctx->out_eq.ensure(sizeof(int32_t) * n));
    HIP_TRY(ctx, ctx->out_st.ensure((size_t)n));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->in_res.p, residues + base, nres, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->in_off.p, rebased.data(), sizeof(uint64_t) * (n + 1), hipMemcpyHostToDevice, ctx->stream));
    PlainParams p{};
    p.prof = (const int16_t *)ctx->prof_p.p;
    p.lut = (const uint16_t *)ctx->lut.p;
    p.prof_words = ctx->prof_words;
    p.stride = ctx->stride;
    p.Q = ctx->Q;
    p.nchunks = (ctx->Q + kChunkPlain - 1) / kChunkPlain;
    p.go = ctx->go;
    p.ge = ctx->ge;
    p.flags = ctx->flags;
    p.neg_limit = ctx->neg_limit;
    p.res = (const uint8_t *)ctx->in_res.p;
    p.off = (const uint64_t *)ctx->in_off.p;
    p.n = n;
    p.score = (int32_t *)ctx->out_score.p;
    p.end_ref = (int32_t *)ctx->out_er.p;
    p.end_q = (int32_t *)ctx->out_eq.p;
    p.status = (uint8_t *)ctx->out_st.p;
    int grid;
    int64_t ss;
    const int occ = occupancy_sg_plain(kBlock, ctx->lds, ctx->flags);
    int32_t rc = size_grid(ctx, occ, maxlen, &grid, &ss);
    if (rc) return rc;
    const int64_t waves_needed = (n + 63) / 64;
    grid = (int)std::min<int64_t>(grid, (waves_needed + kBlock / 64 - 1) / (kBlock / 64));
    p.bnd = (int2 *)ctx->scratch.p;
    p.slot_stride = ss;
    ctx->launches = 0;
    HIP_TRY(ctx, timed_begin(ctx));
    HIP_TRY(ctx, launch_sg_plain(p, grid, kBlock, ctx->lds, ctx->stream));
    HIP_TRY(ctx, timed_end(ctx));
    HIP_TRY(ctx, hipMemcpyAsync(score, p.score, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(end_ref, p.end_ref, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(end_q, p.end_q, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (status)
        HIP_TRY(ctx, hipMemcpyAsync(status, p.status, (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    collect_timing(ctx);
    return ISH_OK;
}

// Emission and start coordinates collected slice by slice during a one-shot
// batch (ish_hip_align_batch_starts): once a slice's results are on the
// host, its emitted targets' reversed prefixes rev(r[0..end_ref]) are
// gathered on the device from the slice's residues, into one buffer that
// the reverse pass scores after the last slice.
struct BatchStarts {
    int64_t max_score = 0;
    float threshold = 0.f;
    uint8_t *emit = nullptr;
    int32_t *tstart = nullptr, *qstart = nullptr;
    std::vector<int64_t> hit;       // emitted targets, batch order
    std::vector<uint64_t> off{0};   // their prefixes' offsets in acc
    uint8_t *acc = nullptr;         // gathered prefixes (device)
    size_t acc_cap = 0;
    std::vector<void *> to_free;    // device buffers freed at the end
    std::vector<std::vector<uint64_t>> host_u64; // kept until the copies are done
    std::vector<std::vector<int64_t>> host_i64;
    void release() {
        for (void *p : to_free) (void)hipFree(p);
        to_free.clear();
        if (acc) (void)hipFree(acc);
        acc = nullptr;
    }
};

// Emission of targets [t0, t1) of the batch (db holds them, rebased), and
// the device gather of the emitted targets' reversed prefixes on stream st:
// the stream that later reuses db's buffers, so that the gather reads them
// first.
static int32_t collect_starts(ish_hip_ctx *ctx, BatchStarts &bs, const ish_hip_db *db, int64_t t0,
                              int64_t t1, hipStream_t st, const int32_t *score, const int32_t *end_ref,
                              const int32_t *end_q) {
    const int64_t m = t1 - t0;
    if (m <= 0) return ISH_OK;
    if (ish_hip_emit(ctx->algo, score + t0, end_ref + t0, end_q + t0, m, bs.max_score, bs.threshold,
                     bs.emit + t0) != ISH_OK)
        return fail(ctx, ISH_E_INVALID, "emission failed");
    std::vector<int64_t> sel;
    std::vector<uint64_t> dst{0};
    for (int64_t t = t0; t < t1; ++t) {
        if (!bs.emit[t]) continue;
        const int64_t R = (int64_t)(db->h_off[(size_t)(t - t0) + 1] - db->h_off[(size_t)(t - t0)]);
        const int64_t L = std::min<int64_t>(std::max<int64_t>((int64_t)end_ref[t] + 1, 0), R);
        sel.push_back(t - t0);
        dst.push_back(dst.back() + (uint64_t)L);
        bs.hit.push_back(t);
    }
    if (sel.empty()) return ISH_OK;
    const size_t base = (size_t)bs.off.back(), add = (size_t)dst.back();
    for (size_t k = 1; k < dst.size(); ++k) bs.off.push_back(base + dst[k]);
    hipError_t e = hipSuccess;
    if (base + add + kRawPad > bs.acc_cap) {
        // grow: a new buffer, the prefixes so far copied over in stream order
        const size_t cap = std::max(base + add + kRawPad, 2 * bs.acc_cap);
        uint8_t *nb = nullptr;
        e = hipMalloc((void **)&nb, cap);
        if (e == hipSuccess && base > 0) e = hipMemcpyAsync(nb, bs.acc, base, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) {
            if (nb) (void)hipFree(nb);
            return fail(ctx, e == hipErrorOutOfMemory ? ISH_E_NOMEM : ISH_E_HIP, "start prefixes: %s",
                        hipGetErrorString(e));
        }
        if (bs.acc) bs.to_free.push_back(bs.acc);
        bs.acc = nb;
        bs.acc_cap = cap;
    }
    if (add == 0) return ISH_OK;
    int64_t *d_sel = nullptr;
    uint64_t *d_dst = nullptr;
    e = hipMalloc((void **)&d_sel, sizeof(int64_t) * sel.size());
    if (e == hipSuccess) bs.to_free.push_back(d_sel);
    if (e == hipSuccess) e = hipMalloc((void **)&d_dst, sizeof(uint64_t) * dst.size());
    if (e == hipSuccess) bs.to_free.push_back(d_dst);
    bs.host_i64.push_back(std::move(sel));
    bs.host_u64.push_back(std::move(dst));
    const auto &hs = bs.host_i64.back();
    const auto &hd = bs.host_u64.back();
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_sel, hs.data(), sizeof(int64_t) * hs.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_dst, hd.data(), sizeof(uint64_t) * hd.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = launch_gather_reversed(db->d_raw, db->d_off, d_sel, d_dst, (int64_t)hs.size(), bs.acc + base, st);
    if (e != hipSuccess)
        return fail(ctx, e == hipErrorOutOfMemory ? ISH_E_NOMEM : ISH_E_HIP, "start prefixes: %s",
                    hipGetErrorString(e));
    return ISH_OK;
}

static int32_t reverse_starts(ish_hip_ctx *ctx, uint8_t *d_raw, uint64_t *d_off, std::vector<uint64_t> off,
                              const int32_t *end_ref, const int32_t *end_query, int64_t n_sel,
                              int32_t *target_start, int32_t *query_start);

// After the last slice (every gather done): the reverse pass over the
// gathered prefixes, and the start coordinates of the emitted targets
// (-1 for the others).
static int32_t finish_starts(ish_hip_ctx *ctx, BatchStarts &bs, int64_t n, const int32_t *end_ref,
                             const int32_t *end_q) {
    for (int64_t t = 0; t < n; ++t) bs.tstart[t] = bs.qstart[t] = -1;
    const int64_t nh = (int64_t)bs.hit.size();
    if (nh == 0) return ISH_OK;
    std::vector<int32_t> er((size_t)nh), eq((size_t)nh), ts((size_t)nh), qs((size_t)nh);
    for (int64_t k = 0; k < nh; ++k) {
        er[(size_t)k] = end_ref[bs.hit[(size_t)k]];
        eq[(size_t)k] = end_q[bs.hit[(size_t)k]];
    }
    uint64_t *d_off = nullptr;
    HIP_TRY(ctx, hipMalloc((void **)&d_off, sizeof(uint64_t) * (size_t)(nh + 1)));
    hipError_t e = hipMemcpy(d_off, bs.off.data(), sizeof(uint64_t) * (size_t)(nh + 1), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d_off);
        return fail(ctx, ISH_E_HIP, "start offsets: %s", hipGetErrorString(e));
    }
    uint8_t *acc = bs.acc;
    bs.acc = nullptr; // reverse_starts takes it over
    // the forward pass's timing stays the call's timing
    const float ms = ctx->last_ms;
    const int launches = ctx->launches;
    const int32_t rc = reverse_starts(ctx, acc, d_off, bs.off, er.data(), eq.data(), nh, ts.data(), qs.data());
    ctx->last_ms = ms;
    ctx->launches = launches;
    if (rc) return rc;
    for (int64_t k = 0; k < nh; ++k) {
        bs.tstart[bs.hit[(size_t)k]] = ts[(size_t)k];
        bs.qstart[bs.hit[(size_t)k]] = qs[(size_t)k];
    }
    return ISH_OK;
}

static int32_t align_batch_impl(ish_hip_ctx *ctx, const uint8_t *residues, const uint64_t *offsets,
                                int64_t n, int32_t *score, int32_t *end_ref, int32_t *end_q,
                                uint8_t *status, BatchStarts *bs) {
    if (!ctx) return ISH_E_INVALID;
    if (n < 0 || (n > 0 && (!offsets || !residues || !score || !end_ref || !end_q)))
        return fail(ctx, ISH_E_INVALID, "bad batch arguments");
    if (!ctx->have_query) return fail(ctx, ISH_E_NO_QUERY, "no query set");
    if (n == 0) return ISH_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (ctx->engine == ISH_ENGINE_PLAIN) {
        if (ctx->algo != ISH_ALGO_SEMI_GLOBAL)
            return fail(ctx, ISH_E_INVALID, "the plain engine scores semi-global alignments only");
        if (bs) return fail(ctx, ISH_E_INVALID, "start coordinates need the batched engine");
        return align_plain(ctx, residues, offsets, n, score, end_ref, end_q, status);
    }
    // slices of about equal residues, cut at target boundaries
    uint64_t maxlen = 0;
    const int32_t checked = db_offsets(ctx, offsets, n, &maxlen);
    if (checked) return checked;
    const uint64_t base = offsets[0];
    const uint64_t total = offsets[n] - base;
    int64_t ns = 1;
    if (ctx->slice_bytes > 0) {
        ns = (int64_t)((total + (uint64_t)ctx->slice_bytes - 1) / (uint64_t)ctx->slice_bytes);
    } else if (ctx->slice_bytes == 0 && total >= 2 * (uint64_t)kSliceMin) {
        // Each launch ends with its longest group, pipelined over a
        // workgroup's four waves; a slice pays for that tail only if the
        // device's other waves have as much work: about cus x 256 residues
        // per residue of the longest target (8 resident waves per CU, 128
        // targets per group).
        const uint64_t need = std::max<uint64_t>((uint64_t)kSliceMin, (uint64_t)ctx->cus * 256 * maxlen);
        // More slices hide more of the upload (the first slice's upload
        // and the last one's results are exposed, ~1/ns of the total) but
        // each adds a tail; T(ns) ~ a/ns + b*ns is lowest near
        // sqrt(a/b) ~ sqrt(3 x total/need). Measured on a 4.19 GB protein
        // batch (10 M targets, longest 5,000): 13 slices 177.5 ms, 5
        // slices 168 ms, 4 172 ms, 3 182 ms, 2 198 ms.
        const double r = (double)total / (double)need;
        ns = std::min<int64_t>((int64_t)r, (int64_t)(std::sqrt(3.0 * r) + 0.5));
    }
    std::vector<int64_t> cut{0};
    if (ns > 1) {
        for (int64_t s = 1; s < ns; ++s) {
            const uint64_t want = base + total / (uint64_t)ns * (uint64_t)s;
            const int64_t t = (int64_t)(std::lower_bound(offsets, offsets + n + 1, want) - offsets);
            if (t > cut.back() && t < n) cut.push_back(t);
        }
    }
    cut.push_back(n);
    // Slice s is uploaded and built on the copy stream while slice s-1 is
    // scored, then queued behind it; slice s-1's results are fetched on the
    // copy stream once its pass ends, while slice s scores. The slices use
    // the context's two one-shot buffer slots in turn, so nothing is
    // allocated or freed between passes. No graphs: every slice is a new
    // database.
    if (cut.size() == 2) {
        // one slice: nothing to overlap, so no copy stream
        ish_hip_db *db = nullptr;
        int32_t rc = db_load_impl(ctx, residues, offsets, n, nullptr, &db, 0, true);
        if (rc == ISH_OK) rc = ish::db_score_once_async(ctx, db, 0);
        if (rc == ISH_OK) {
            rc = fetch_results(ctx, db, ctx->stream, score, end_ref, end_q, status);
            collect_timing(ctx);
        } else {
            (void)hipStreamSynchronize(ctx->stream);
        }
        if (rc == ISH_OK && bs) {
            rc = collect_starts(ctx, *bs, db, 0, n, ctx->stream, score, end_ref, end_q);
            const hipError_t e = hipStreamSynchronize(ctx->stream);
            if (rc == ISH_OK && e != hipSuccess) rc = fail(ctx, ISH_E_HIP, "start prefixes: %s", hipGetErrorString(e));
        }
        if (db) ish_hip_db_free(db);
        if (rc == ISH_OK && bs) rc = finish_starts(ctx, *bs, n, end_ref, end_q);
        return rc;
    }
    int32_t rc = ish::one_shot_begin(ctx);
    int32_t launches = 0;
    ish_hip_db *prev = nullptr;
    int64_t prev_t0 = 0, prev_t1 = 0;
    for (size_t s = 0; s + 1 < cut.size() && rc == ISH_OK; ++s) {
        const int64_t t0 = cut[s], t1 = cut[s + 1];
        const int slot = (int)(s & 1);
        ish_hip_db *db = nullptr;
        rc = db_load_impl(ctx, residues, offsets + t0, t1 - t0, nullptr, &db, slot, true);
        if (rc == ISH_OK && s == 0) {
            const hipError_t e = hipEventRecord(ctx->span[0], ctx->stream);
            if (e != hipSuccess) rc = fail(ctx, ISH_E_HIP, "span event: %s", hipGetErrorString(e));
        }
        if (rc == ISH_OK) {
            rc = ish::db_score_once_async(ctx, db, slot);
            launches += ctx->launches;
        }
        if (prev) {
            int32_t r = ish::db_fetch_after(ctx, prev, slot ^ 1, score + prev_t0, end_ref + prev_t0,
                                            end_q + prev_t0, status ? status + prev_t0 : nullptr);
            // the gather runs on the copy stream, before the next slice's
            // upload overwrites prev's buffers
            if (r == ISH_OK && bs) r = collect_starts(ctx, *bs, prev, prev_t0, prev_t1, ctx->copy, score, end_ref, end_q);
            if (rc == ISH_OK) rc = r;
            ish_hip_db_free(prev);
        }
        prev = db;
        prev_t0 = t0;
        prev_t1 = t1;
    }
    if (prev && rc == ISH_OK) {
        const int slot = (int)((cut.size() - 2) & 1);
        hipError_t e = hipEventRecord(ctx->span[1], ctx->stream);
        if (e != hipSuccess) rc = fail(ctx, ISH_E_HIP, "span event: %s", hipGetErrorString(e));
        if (rc == ISH_OK)
            rc = ish::db_fetch_after(ctx, prev, slot, score + prev_t0, end_ref + prev_t0, end_q + prev_t0,
                                     status ? status + prev_t0 : nullptr);
        if (rc == ISH_OK && bs)
            rc = collect_starts(ctx, *bs, prev, prev_t0, prev_t1, ctx->copy, score, end_ref, end_q);
        float ms = 0.f;
        if (rc == ISH_OK && hipEventSynchronize(ctx->span[1]) == hipSuccess &&
            hipEventElapsedTime(&ms, ctx->span[0], ctx->span[1]) == hipSuccess)
            ctx->last_ms = ms; // first launch to the end of the last one
    }
    if (rc != ISH_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        if (ctx->copy) (void)hipStreamSynchronize(ctx->copy);
    }
    if (prev) ish_hip_db_free(prev);
    ish::one_shot_end(ctx); // waits for the copy stream: every gather is done
    ctx->launches = launches;
    if (rc == ISH_OK && bs) rc = finish_starts(ctx, *bs, n, end_ref, end_q);
    return rc;
}

extern "C" int32_t ish_hip_align_batch(ish_hip_ctx *ctx, const uint8_t *residues,
                                       const uint64_t *offsets, int64_t n, int32_t *score,
                                       int32_t *end_ref, int32_t *end_q, uint8_t *status) {
    return align_batch_impl(ctx, residues, offsets, n, score, end_ref, end_q, status, nullptr);
}

extern "C" int32_t ish_hip_align_batch_starts(ish_hip_ctx *ctx, const uint8_t *residues,
                                              const uint64_t *offsets, int64_t n, int64_t max_score,
                                              float threshold, int32_t *score, int32_t *end_ref,
                                              int32_t *end_q, uint8_t *status, uint8_t *emit,
                                              int32_t *target_start, int32_t *query_start) {
    if (!ctx) return ISH_E_INVALID;
    if (n > 0 && (!emit || !target_start || !query_start))
        return fail(ctx, ISH_E_INVALID, "bad batch arguments");
    BatchStarts bs;
    bs.max_score = max_score;
    bs.threshold = threshold;
    bs.emit = emit;
    bs.tstart = target_start;
    bs.qstart = query_start;
    const int32_t rc = align_batch_impl(ctx, residues, offsets, n, score, end_ref, end_q, status, &bs);
    if (rc) {
        (void)hipStreamSynchronize(ctx->stream);
        if (ctx->copy) (void)hipStreamSynchronize(ctx->copy);
    }
    bs.release();
    return rc;
}

extern "C" int32_t ish_hip_align_batch_submit(ish_hip_ctx *ctx, const uint8_t *residues,
                                              const uint64_t *offsets, int64_t n, int64_t max_score,
                                              float threshold, int32_t *score, int32_t *end_ref,
                                              int32_t *end_q, uint8_t *status, uint8_t *emit,
                                              int32_t *target_start, int32_t *query_start) {
    if (!ctx) return ISH_E_INVALID;
    if (ctx->bg.joinable()) return fail(ctx, ISH_E_INVALID, "a submitted batch is still in flight");
    const bool starts = emit || target_start || query_start;
    if (starts && n > 0 && (!emit || !target_start || !query_start))
        return fail(ctx, ISH_E_INVALID, "emit, target_start and query_start go together");
    ctx->bg_rc = ISH_OK;
    try {
        ctx->bg = std::thread([=] {
            ctx->bg_rc = starts ? ish_hip_align_batch_starts(ctx, residues, offsets, n, max_score, threshold,
                                                             score, end_ref, end_q, status, emit,
                                                             target_start, query_start)
                                : ish_hip_align_batch(ctx, residues, offsets, n, score, end_ref, end_q,
                                                      status);
        });
    } catch (...) {
        return fail(ctx, ISH_E_NOMEM, "could not start the batch thread");
    }
    return ISH_OK;
}

extern "C" int32_t ish_hip_align_batch_wait(ish_hip_ctx *ctx) {
    if (!ctx) return ISH_E_INVALID;
    if (!ctx->bg.joinable()) return fail(ctx, ISH_E_INVALID, "no batch was submitted");
    ctx->bg.join();
    return ctx->bg_rc;
}

extern "C" int32_t ish_hip_set_slice_bytes(ish_hip_ctx *ctx, int64_t bytes) {
    if (!ctx) return ISH_E_INVALID;
    ctx->slice_bytes = bytes;
    return ISH_OK;
}

// ---------------------------------------------------------------------------
// many queries against one database, one persistent launch per cell width
// ---------------------------------------------------------------------------
int32_t ish::enqueue_multi(const ish_hip_db *db, const std::vector<MultiItem> &items, int32_t go,
                           int32_t ge, int32_t mn, int32_t mx, int chunk, int cus, int32_t *score,
                           int32_t *end_ref, int32_t *end_q, uint8_t *status, MultiBuffers &mb,
                           hipStream_t st, std::string &err,
                           const std::function<hipError_t(bool)> &on_launch, bool f16) {
    const size_t nq = items.size();
    if (nq == 0 || db->n == 0) return ISH_OK;
    auto hip_fail = [&](hipError_t e, const char *what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? ISH_E_NOMEM : ISH_E_HIP;
    };
    // profiles and luts of every query, back to back
    std::vector<size_t> pofs(nq + 1, 0);
    size_t lds = 0;
    for (size_t k = 0; k < nq; ++k) {
        const Profile &pf = *items[k].pf;
        lds = std::max(lds, pf.lds());
        pofs[k + 1] = pofs[k] + ((pf.fwd.size() * 2 + 511) & ~(size_t)255);
    }
    if (lds > kLdsLimit) {
        err = "a query profile needs " + std::to_string(lds) + " bytes of LDS";
        return ISH_E_RANGE;
    }
    const size_t luts_off = pofs[nq];
    // paired profiles when every query's classes over the database's codes
    // are few: the packed launches then read those instead, one copy of a
    // chunk's pair rows per wave (at most kPairMaxRows classes) or one per
    // workgroup with the tile's groups in lockstep (at most kPairMaxShared)
    std::vector<PairProfile> pairs(nq);
    bool paired = true;
    int32_t max_classes = 0;
    for (size_t k = 0; k < nq && paired; ++k) {
        paired = build_pairs(*items[k].pf, db->code_mask, kPairMaxShared, pairs[k]);
        max_classes = std::max(max_classes, pairs[k].classes);
    }
    const size_t wave_pair_lds = kPairTableLds + (size_t)(kBlock / 64) * max_classes * max_classes * pair_row_bytes(chunk);
    const size_t lock_lds = kPairTableLds + (size_t)max_classes * max_classes * pair_row_bytes(chunk);
    enum { kUnpaired = 0, kPerWave = 1, kShared = 2 };
    int pmode = kUnpaired;
    if (paired && max_classes <= kPairMaxRows && wave_pair_lds <= kLdsLimit) pmode = kPerWave;
    else if (paired && lock_lds <= kLdsLimit) pmode = kShared;
    paired = pmode != kUnpaired;
    const size_t paired_lds = pmode == kPerWave ? wave_pair_lds : lock_lds;
    std::vector<size_t> pp_ofs(nq + 1, luts_off + nq * 512);
    if (paired)
        for (size_t k = 0; k < nq; ++k)
            pp_ofs[k + 1] = pp_ofs[k] + ((pairs[k].words.size() * 4 + pairs[k].table.size() * 2 + 255) & ~(size_t)255);
    std::vector<uint8_t> &blob = *mb.host_blob;
    blob.assign(pp_ofs[nq], 0);
    for (size_t k = 0; k < nq; ++k) {
        const Profile &pf = *items[k].pf;
        memcpy(blob.data() + pofs[k], pf.fwd.data(), pf.fwd.size() * 2);
        memcpy(blob.data() + luts_off + k * 512, pf.lut, 512);
        if (paired) {
            memcpy(blob.data() + pp_ofs[k], pairs[k].words.data(), pairs[k].words.size() * 4);
            memcpy(blob.data() + pp_ofs[k] + pairs[k].words.size() * 4, pairs[k].table.data(),
                   pairs[k].table.size() * 2);
        }
    }
    hipError_t e = mb.blob->ensure(blob.size());
    if (e == hipSuccess) e = mb.desc->ensure(sizeof(QueryDesc) * (nq + 1) * 4);
    if (e == hipSuccess)
        e = hipMemcpyAsync(mb.blob->p, blob.data(), blob.size(), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "multi-query setup");
    // queries whose end rules need the last-row / last-column tracking, and
    // queries that only need the corner cell, run in different variants
    std::vector<QueryDesc> &hd = *mb.host_desc;
    hd.assign((nq + 1) * 4, QueryDesc{});
    const int waves = kBlock / 64;
    struct Launch {
        ScoreParams p;
        bool packed, paired; // paired: pmode's variant
        int32_t maxcols, ntab;
    };
    std::vector<Launch> launches;
    for (int cls = 0; cls < 2; ++cls) {
        std::vector<size_t> members;
        uint32_t vflags = 0;
        for (size_t k = 0; k < nq; ++k) {
            const bool track = items[k].flags & 10u;
            if (track == (cls == 1)) {
                members.push_back(k);
                vflags |= items[k].flags & 10u;
            }
        }
        if (members.empty()) continue;
        for (int pass = 0; pass < 2; ++pass) {
            const bool packed = pass == 1;
            const int table = (int)launches.size();
            QueryDesc *qd = hd.data() + (size_t)table * (nq + 1);
            int64_t tiles = 0;
            int32_t maxcols = 0;
            for (size_t m = 0; m < members.size(); ++m) {
                const MultiItem &it = items[members[m]];
                const int64_t r16 = int16_max_len(it.Q, go, ge, mx, mn);
                int32_t lo = 0, hi = db->ngroups;
                while (lo < hi) {
                    const int32_t mid = (lo + hi) / 2;
                    if ((int64_t)db->group_maxlen[mid] > r16) lo = mid + 1;
                    else hi = mid;
                }
                QueryDesc &d = qd[m];
                d.prof = (const int16_t *)((uint8_t *)mb.blob->p + pofs[members[m]]);
                d.lut = (const uint16_t *)((uint8_t *)mb.blob->p + luts_off + members[m] * 512);
                d.prof_words = it.pf->words();
                d.Q = it.Q;
                d.nchunks = (it.Q + chunk - 1) / chunk;
                d.pad = it.pf->top ? top_pad(it.Q, chunk) : 0;
                d.front = it.pf->front;
                d.neg_limit = neg_limit_of(go, mn);
                d.g_begin = packed ? lo : 0;
                d.g_end = packed ? db->ngroups : lo;
                d.flags = it.flags;
                d.tile_base = tiles;
                d.out_off = it.out_off;
                if (paired && packed) {
                    const PairProfile &pq = pairs[members[m]];
                    uint8_t *base = (uint8_t *)mb.blob->p + pp_ofs[members[m]];
                    d.pprof = (const int32_t *)base;
                    d.plut = (const uint16_t *)(base + pq.words.size() * 4);
                    d.pnrows = pq.classes;
                    d.pqpad = pq.qpad;
                    // binary16 lanes when every group of the query lies
                    // within its range bound
                    if (f16 && d.g_end > d.g_begin) {
                        int32_t sh = 0;
                        const int64_t w16 = f16_bound(*it.pf, it.Q, go, ge, mn, db->code_mask, &sh);
                        if (w16 >= (((int64_t)db->group_maxlen[d.g_begin] + 3) & ~3LL)) {
                            d.f16 = 1;
                            d.fshift = sh;
                        }
                    }
                }
                const int64_t witems = (int64_t)(d.g_end - d.g_begin) * (packed ? 1 : 2);
                tiles += (witems + waves - 1) / waves;
                if (d.g_end > d.g_begin) maxcols = std::max(maxcols, db->group_maxlen[d.g_begin] + 3);
            }
            qd[members.size()].tile_base = tiles;
            if (tiles == 0) continue;
            Launch L{};
            ScoreParams &p = L.p;
            p.go = go;
            p.ge = ge;
            p.flags = vflags;
            p.groups = db->d_groups;
            p.res = db->d_res;
            p.lens = db->d_lens;
            p.orig = db->d_orig;
            p.counter = mb.counter;
            p.score = score;
            p.end_ref = end_ref;
            p.end_q = end_q;
            p.status = status;
            p.queries = (QueryDesc *)mb.desc->p + (size_t)table * (nq + 1);
            p.nqueries = (int32_t)members.size();
            p.pnrows = paired && packed ? max_classes : 0;
            L.packed = packed;
            L.paired = paired && packed;
            L.maxcols = maxcols;
            L.ntab = (int32_t)members.size() + 1;
            launches.push_back(L);
        }
    }
    // size the boundary scratch for every launch before the first one is
    // enqueued (growing it later would free memory a queued launch uses)
    std::vector<int> grids(launches.size());
    for (size_t l = 0; l < launches.size(); ++l) {
        int64_t ss;
        const size_t llds = launches[l].paired ? paired_lds : lds;
        const int occ = launches[l].paired && pmode == kShared
                            ? occupancy_sg_multi_lock(chunk, llds, launches[l].p.flags)
                            : occupancy_sg_multi(launches[l].packed, launches[l].paired, chunk, kBlock,
                                                 llds, launches[l].p.flags);
        e = persistent_grid(*mb.scratch, occ, cus, launches[l].maxcols, &grids[l], &ss);
        if (e != hipSuccess) return hip_fail(e, "boundary scratch");
        launches[l].p.slot_stride = ss;
    }
    for (size_t l = 0; l < launches.size(); ++l) {
        ScoreParams &p = launches[l].p;
        p.bnd = (uint2 *)mb.scratch->p;
        e = hipMemcpyAsync((void *)p.queries, hd.data() + l * (nq + 1),
                           sizeof(QueryDesc) * (size_t)launches[l].ntab, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = launch_reset(p.counter, st);
        if (e == hipSuccess) e = on_launch(true);
        if (e == hipSuccess) {
            if (launches[l].paired && pmode == kShared)
                e = launch_sg_multi_lock(p, chunk, grids[l], paired_lds, st);
            else
                e = launch_sg_multi(p, launches[l].packed, launches[l].paired, chunk, grids[l], kBlock,
                                    launches[l].paired ? paired_lds : lds, st);
        }
        if (e == hipSuccess) e = on_launch(false);
        if (e != hipSuccess) return hip_fail(e, "multi-query launch");
    }
    return ISH_OK;
}

extern "C" int32_t ish_hip_align_db_multi(ish_hip_ctx *ctx, ish_hip_db *db,
                                          const uint8_t *queries, const uint64_t *qoffsets,
                                          int32_t nq, const int8_t *matrix, int32_t matrix_len,
                                          int32_t kind, int32_t go, int32_t ge,
                                          uint32_t ends_free, int32_t *score, int32_t *end_ref,
                                          int32_t *end_q, uint8_t *status) {
    if (!ctx || !db || nq < 0 || (nq > 0 && (!queries || !qoffsets)))
        return fail(ctx, ISH_E_INVALID, "bad multi-query arguments");
    if (go < 0 || ge < 0) return fail(ctx, ISH_E_INVALID, "gap penalties must be >= 0");
    if (db->device != ctx->device) return fail(ctx, ISH_E_INVALID, "db belongs to another device");
    const int64_t n = db->n;
    if (nq == 0 || n == 0) return ISH_OK;
    if (!score || !end_ref || !end_q) return fail(ctx, ISH_E_INVALID, "null outputs");
    std::vector<int8_t> mat;
    int32_t A = 0;
    int32_t rc = resolve_matrix(ctx, matrix, matrix_len, kind, A, mat);
    if (rc) return rc;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int32_t mn = 0, mx = 0;
    for (int8_t v : mat) {
        mn = std::min<int32_t>(mn, v);
        mx = std::max<int32_t>(mx, v);
    }
    std::vector<Profile> pfs((size_t)nq);
    std::vector<MultiItem> items((size_t)nq);
    for (int32_t k = 0; k < nq; ++k) {
        if (qoffsets[k + 1] < qoffsets[k]) return fail(ctx, ISH_E_INVALID, "query offsets must be non-decreasing");
        const int32_t Q = (int32_t)(qoffsets[k + 1] - qoffsets[k]);
        const uint8_t *q = queries + qoffsets[k];
        for (int32_t j = 0; j < Q; ++j)
            if (q[j] >= A) return fail(ctx, ISH_E_INVALID, "query %d code %d >= alphabet %d", k, q[j], A);
        build_profile(q, Q, mat, A, 2 * ge, false, go >= ge, top_row_score(ends_free, ge), 0,
                      pfs[(size_t)k]);
        items[(size_t)k] = MultiItem{&pfs[(size_t)k], Q, ends_free & 15u, (int64_t)k * n};
    }
    const size_t tot = (size_t)nq * (size_t)n;
    DevBuf d_blob, d_desc, d_out;
    std::vector<uint8_t> h_blob;
    std::vector<QueryDesc> h_desc;
    hipError_t e = d_out.ensure(tot * 13);
    if (e == hipSuccess) e = ctx->counter.ensure(256);
    if (e != hipSuccess) {
        d_out.release();
        return fail(ctx, e == hipErrorOutOfMemory ? ISH_E_NOMEM : ISH_E_HIP, "multi-query outputs: %s",
                    hipGetErrorString(e));
    }
    int32_t *o_score = (int32_t *)d_out.p;
    int32_t *o_er = o_score + tot;
    int32_t *o_eq = o_er + tot;
    uint8_t *o_st = (uint8_t *)(o_eq + tot);
    MultiBuffers mb;
    mb.blob = &d_blob;
    mb.desc = &d_desc;
    mb.scratch = &ctx->scratch;
    mb.counter = (uint32_t *)ctx->counter.p;
    mb.host_blob = &h_blob;
    mb.host_desc = &h_desc;
    ctx->launches = 0;
    std::string err;
    rc = enqueue_multi(db, items, go, ge, mn, mx, ctx->chunk == 40 || ctx->chunk == 24 ? 32 : ctx->chunk, ctx->cus, o_score,
                       o_er, o_eq, o_st,
                       mb, ctx->stream, err,
                       [ctx](bool begin) { return begin ? timed_begin(ctx) : timed_end(ctx); }, ctx->f16);
    if (rc) {
        fail(ctx, rc, "%s", err.c_str());
    } else {
        e = hipMemcpyAsync(score, o_score, tot * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(end_ref, o_er, tot * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(end_q, o_eq, tot * 4, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess && status) e = hipMemcpyAsync(status, o_st, tot, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(ctx, ISH_E_HIP, "multi-query results: %s", hipGetErrorString(e));
        collect_timing(ctx);
    }
    (void)hipStreamSynchronize(ctx->stream);
    d_blob.release();
    d_desc.release();
    d_out.release();
    return rc;
}

// ---------------------------------------------------------------------------
// start coordinates
// ---------------------------------------------------------------------------
// The reverse passes of the start coordinates: d_raw holds rev(r[0..L_k))
// of n_sel targets at offsets off (n_sel + 1 entries, device copy d_off);
// both device buffers are taken over (and freed). Target k's forward end
// coordinates are end_ref[k], end_query[k].
static int32_t reverse_starts(ish_hip_ctx *ctx, uint8_t *d_raw, uint64_t *d_off, std::vector<uint64_t> off,
                              const int32_t *end_ref, const int32_t *end_query, int64_t n_sel,
                              int32_t *target_start, int32_t *query_start) {
    const bool local = ctx->algo == ISH_ALGO_LOCAL;
    ish_hip_db *rdb = nullptr;
    SortPlan sp;
    plan_sort(off, sp);
    int32_t rc = db_build(ctx, ctx->stream, d_raw, d_off, std::move(off), sp, nullptr, &rdb);
    if (rc) return rc;
    // free ends swap for the reverse pass (striped.mojo:165-183)
    const uint32_t f = ctx->flags;
    const uint32_t rflags = ((f & 1u) << 1) | ((f >> 1) & 1u) | ((f & 4u) << 1) | ((f >> 1) & 4u);
    rc = launch_db(ctx, rdb, (const int16_t *)ctx->prof_rt.p, local ? 0u : rflags, false);
    std::vector<int32_t> rs((size_t)n_sel), rer((size_t)n_sel), req((size_t)n_sel);
    if (!rc) rc = ish_hip_db_fetch(ctx, rdb, rs.data(), rer.data(), req.data(), nullptr);
    ish_hip_db_free(rdb);
    if (rc) return rc;
    for (int64_t k = 0; k < n_sel; ++k) {
        if (local) {
            if (end_ref[k] <= 0 || end_query[k] <= 0) {
                target_start[k] = -1;
                query_start[k] = -1;
            } else {
                // begin = position where the reverse running maximum last
                // rose, in original coordinates (local_aln/striped.mojo:232-262)
                target_start[k] = rer[(size_t)k] < 0 ? -1 : end_ref[k] - rer[(size_t)k];
                query_start[k] = end_query[k] - req[(size_t)k];
            }
        } else {
            // semi_global_aln_start_end (striped.mojo:185-191)
            target_start[k] = end_ref[k] - rer[(size_t)k];
            query_start[k] = ctx->Q - req[(size_t)k] - 1;
        }
    }
    return ISH_OK;
}

extern "C" int32_t ish_hip_find_starts(ish_hip_ctx *ctx, ish_hip_db *db, const int64_t *sel,
                                       const int32_t *end_ref, const int32_t *end_query,
                                       int64_t n_sel, int32_t *target_start,
                                       int32_t *query_start) {
    if (!ctx || !db || n_sel < 0 ||
        (n_sel > 0 && (!sel || !end_ref || !end_query || !target_start || !query_start)))
        return fail(ctx, ISH_E_INVALID, "bad find_starts arguments");
    if (!ctx->have_query) return fail(ctx, ISH_E_NO_QUERY, "no query set");
    if (db->device != ctx->device) return fail(ctx, ISH_E_INVALID, "db belongs to another device");
    if (n_sel == 0) return ISH_OK;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const bool local = ctx->algo == ISH_ALGO_LOCAL;
    // reversed prefixes rev(r[0..end_ref]) of the selected targets
    std::vector<uint64_t> off((size_t)n_sel + 1, 0);
    std::vector<int64_t> src((size_t)n_sel);
    for (int64_t k = 0; k < n_sel; ++k) {
        const int64_t t = sel[k];
        if (t < 0 || t >= db->n) return fail(ctx, ISH_E_INVALID, "selected target %lld out of range", (long long)t);
        const int64_t R = (int64_t)(db->h_off[(size_t)t + 1] - db->h_off[(size_t)t]);
        int64_t L = (int64_t)end_ref[k] + 1;
        if (L < 0 || L > R) return fail(ctx, ISH_E_INVALID, "end_ref %d out of range for target %lld", end_ref[k], (long long)t);
        if (local && (end_ref[k] <= 0 || end_query[k] <= 0)) L = 0; // no begin pass (ssw_align)
        src[(size_t)k] = t;
        off[(size_t)k + 1] = off[(size_t)k] + (uint64_t)L;
    }
    const uint64_t nres = off[(size_t)n_sel];
    uint8_t *d_raw = nullptr;
    uint64_t *d_off = nullptr;
    int64_t *d_sel = nullptr;
    hipError_t e = hipMalloc((void **)&d_raw, nres + kRawPad);
    if (e == hipSuccess) e = hipMalloc((void **)&d_off, sizeof(uint64_t) * (size_t)(n_sel + 1));
    if (e == hipSuccess) e = hipMalloc((void **)&d_sel, sizeof(int64_t) * (size_t)n_sel);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_off, off.data(), sizeof(uint64_t) * (size_t)(n_sel + 1), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_sel, src.data(), sizeof(int64_t) * (size_t)n_sel, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && nres > 0)
        e = launch_gather_reversed(db->d_raw, db->d_off, d_sel, d_off, n_sel, d_raw, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (d_sel) (void)hipFree(d_sel);
    if (e != hipSuccess) {
        if (d_raw) (void)hipFree(d_raw);
        if (d_off) (void)hipFree(d_off);
        return fail(ctx, e == hipErrorOutOfMemory ? ISH_E_NOMEM : ISH_E_HIP, "reverse gather: %s", hipGetErrorString(e));
    }
    return reverse_starts(ctx, d_raw, d_off, std::move(off), end_ref, end_query, n_sel, target_start,
                          query_start);
}

// ---------------------------------------------------------------------------
// emission
// ---------------------------------------------------------------------------
extern "C" int32_t ish_hip_emit(int32_t algo, const int32_t *score, const int32_t *end_ref,
                                const int32_t *end_q, int64_t n, int64_t max_score,
                                float threshold, uint8_t *emit) {
    if (n < 0 || (n > 0 && (!score || !emit))) return ISH_E_INVALID;
    if (algo == ISH_ALGO_LOCAL && n > 0 && (!end_ref || !end_q)) return ISH_E_INVALID;
    const volatile float ms = (float)max_score;
    for (int64_t t = 0; t < n; ++t) {
        const volatile float s = (float)score[t];
        bool ok;
        if (algo == ISH_ALGO_SEMI_GLOBAL) {
            const volatile float cutoff = ms * threshold;
            ok = !(s < cutoff);
            if (ok) {
                const volatile float ratio = s / ms;
                ok = ratio >= threshold;
            }
        } else {
            const volatile float cutoff = threshold * ms;
            ok = s > cutoff && end_ref[t] > 0 && end_q[t] > 0;
            if (ok) {
                const volatile float ratio = s / ms;
                ok = ratio >= threshold;
            }
        }
        emit[t] = ok ? 1 : 0;
    }
    return ISH_OK;
}
