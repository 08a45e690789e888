// Host-side internals shared by the C ABI (ish_hip.cpp) and the serving
// loop (server.cpp). Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ish_hip.h"
#include "engine.h"

namespace ish {

constexpr size_t kLdsLimit = 160 * 1024; // gfx950 LDS per workgroup
constexpr int kBlock = 256;              // 4 waves per workgroup

// Grow-only device buffer. Growing frees the old buffer (hipFree waits for
// the whole device) and loses its contents.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap && p) return hipSuccess;
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            cap = 0;
        }
        // powers of two up to 64 MiB, then 1/8 headroom in 2 MiB steps
        size_t c = 256;
        if (bytes <= ((size_t)64 << 20)) {
            while (c < bytes) c <<= 1;
        } else {
            const size_t step = (size_t)2 << 20;
            c = (bytes + bytes / 8 + step - 1) / step * step;
        }
        hipError_t e = hipMalloc(&p, c);
        if (e == hipSuccess) cap = c;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Query profile in the LDS layout the kernels read: one row of `stride`
// int16 scores per group of residue codes whose score column over the query
// is identical (for the ascii matrix this collapses 256 codes to the
// distinct query bytes plus one "absent" row); lut maps a code to the dword
// offset of its row.
// With top rows (semi-global, go >= ge), fwd and rev rows hold front =
// top_front(Q) padding-row entries before query row 0 (plain rows start at
// row 0 regardless).
struct Profile {
    int32_t stride = 0, nrows = 0, qpad = 0, front = 0;
    bool top = false; // top padding rows (front > 0)
    std::vector<int16_t> fwd, rev, plain; // offset form, reversed query, S
    uint16_t lut[256];
    size_t lds() const { return (size_t)nrows * stride * 2 + kLutLds; }
    int32_t words() const { return nrows * stride / 2; }
    // entries of a row the kernels may read: the front and the padded query
    int32_t width() const { return front + qpad; }
};

// Profile entries are S + bias (the kernels' offset form); entries past the
// query are 0. top_rows: fwd rows start with top_front(Q) entries of value
// top_fwd and rev rows with entries of value top_rev (the padding rows'
// score: ge where that pass's target start is free, else 0).
void build_profile(const uint8_t *q, int32_t Q, const std::vector<int8_t> &mat, int32_t A,
                   int32_t bias, bool want_rev_plain, bool top_rows, int32_t top_fwd,
                   int32_t top_rev, Profile &pf);
// Padding-row score of a semi-global pass whose free ends are `flags`.
inline int32_t top_row_score(uint32_t flags, int32_t ge) { return (flags & 4u) ? ge : 0; }
// Paired profile of a query over the profile classes of the residue codes
// a database holds (the layout is described with ScoreParams.pprof):
// `classes` classes, words[(a * classes + b) * qpad + e] for the profile's
// row entries e (front included; qpad = width rounded up to 8), and the pair
// tables of tile heights 16, 32 and 40, each [A: 256][B: 256].
struct PairProfile {
    int32_t classes = 0, qpad = 0;
    std::vector<int32_t> words;
    std::vector<uint16_t> table;
};
// Builds it for the codes set in present[8] (code 0, the padding code past
// targets' ends, always counts). False when those codes fall into more
// than max_classes classes.
bool build_pairs(const Profile &pf, const uint32_t present[8], int32_t max_classes,
                 PairProfile &out);
// Alphabet size of a built-in matrix kind (-1 when unknown).
int32_t alphabet_of(int32_t kind);
// The A*A matrix to use (A the alphabet of `kind`): the caller's, whose
// matrix_len entries must be A*A, or the built-in one of `kind` when matrix
// is null. Returns 0 or an ISH_E_* code with a message in err.
int32_t matrix_of(const int8_t *matrix, int64_t matrix_len, int32_t kind, int32_t &A,
                  std::vector<int8_t> &mat, std::string &err);
// Longest target the packed int16 semi-global kernels may score for a
// query (-1: none; always so when go < ge, since the packed kernels read
// row Q-1 from the last register row, which takes top padding rows).
int64_t int16_max_len(int32_t Q, int32_t go, int32_t ge, int32_t max_s, int32_t min_s);
// Widest group (columns) whose semi-global pass may run in binary16 lanes
// for this profile over the residue codes in present[8] (null: any code),
// and the shift of the values those lanes hold; -1 when none.
int64_t f16_bound(const Profile &pf, int32_t Q, int32_t go, int32_t ge, int32_t min_s,
                  const uint32_t *present, int32_t *shift);
// Whether a local pass of this profile over those codes may run in
// binary16 lanes, and the shift of the state those lanes hold.
bool sw_f16_bound(const Profile &pf, int32_t Q, int32_t go, int32_t ge, int32_t min_s,
                  const uint32_t *present, int32_t *shift);
// int16 floor seeding the last-row maximum.
int32_t neg_limit_of(int32_t go, int32_t mn);
// Grid of a persistent launch (occ_blocks per CU on `cus` CUs) and the
// per-wave boundary scratch for groups of up to `cols` columns; grows
// `scratch` and may shrink the grid to stay within the memory budget.
hipError_t persistent_grid(DevBuf &scratch, int occ_blocks, int cus, int64_t cols, int *grid,
                           int64_t *slot_stride, int waves_per_block = 4);

// Host worker threads that run one fork-join job at a time (a context is
// used from one thread): run(n, fn) calls fn(0) .. fn(n - 1) on the workers
// and the calling thread and returns when all have returned. The workers
// live as long as the pool, so the host copies of a call do not pay for
// creating threads.
class HostPool {
  public:
    explicit HostPool(int workers) {
        for (int k = 0; k < workers; ++k) th_.emplace_back([this] { loop(); });
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    int size() const { return (int)th_.size() + 1; }
    void run(int n, const std::function<void(int)> &fn) {
        if (n <= 0) return;
        if (n == 1 || th_.empty()) {
            for (int k = 0; k < n; ++k) fn(k);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &fn;
            njobs_ = n;
            next_.store(0);
            busy_ = (int)th_.size();
            ++gen_;
        }
        cv_.notify_all();
        work(fn, n);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [this] { return busy_ == 0; });
        job_ = nullptr;
    }

  private:
    void work(const std::function<void(int)> &fn, int n) {
        for (int k; (k = next_.fetch_add(1)) < n;) fn(k);
    }
    void loop() {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            const std::function<void(int)> *fn = job_;
            const int n = njobs_;
            lk.unlock();
            work(*fn, n);
            lk.lock();
            if (--busy_ == 0) done_.notify_all();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)> *job_ = nullptr;
    int njobs_ = 0, busy_ = 0;
    std::atomic<int> next_{0};
    uint64_t gen_ = 0;
    bool stop_ = false;
};

// Device buffers of a database, kept by a context and reused (grow-only)
// by the databases of one-shot passes, so that those allocate nothing once
// warm: hipFree waits for every stream of the device, which would stall an
// upload overlapping a pass.
struct DbBuffers {
    DevBuf groups, res, lens, orig, score, er, eq, st, raw, off;
    void release() {
        for (DevBuf *b : {&groups, &res, &lens, &orig, &score, &er, &eq, &st, &raw, &off}) b->release();
    }
};

} // namespace ish

// Device-resident target database (ish_hip_db of the C ABI).
struct ish_hip_db {
    int device = 0;
    // buffers borrowed from a context (one-shot passes): not freed with
    // the database
    ish::DbBuffers *borrowed = nullptr;
    int64_t n = 0, nres = 0;
    int32_t ngroups = 0, maxcols = 0;
    std::vector<int32_t> group_maxlen; // per group, descending
    ish::GroupDesc *d_groups = nullptr;
    uint2 *d_res = nullptr;
    int32_t *d_lens = nullptr, *d_orig = nullptr;
    int32_t *d_score = nullptr, *d_end_ref = nullptr, *d_end_q = nullptr;
    uint8_t *d_status = nullptr;
    // raw residues (offsets rebased to 0) kept for start-coordinate passes
    uint8_t *d_raw = nullptr;
    uint64_t *d_off = nullptr;
    std::vector<uint64_t> h_off;
    // residue codes present: bit c of code_mask[c >> 5]
    uint32_t code_mask[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

namespace ish {

// The no-allocation pass (ish_hip_align_db_noalloc): the device address of
// its static scratch array is looked up once per device, at context
// creation with that device current (fixed_scratch_init); align_db_fixed
// scores db into caller-owned device arrays with the counter and boundary
// rows in that scratch, and waits for the pass.
hipError_t fixed_scratch_init(int device);
int32_t align_db_fixed(ish_hip_ctx *ctx, ish_hip_db *db, void *scratch, size_t scratch_bytes,
                       int32_t *d_score, int32_t *d_end_ref, int32_t *d_end_q, uint8_t *d_status);

// One query of a multi-query pass over a database.
struct MultiItem {
    const Profile *pf;
    int32_t Q;
    uint32_t flags;  // free ends
    int64_t out_off; // offset of its results in the output arrays
};

// Buffers a multi-query pass uses; the host vectors must stay alive until
// the stream has consumed the copies.
struct MultiBuffers {
    DevBuf *blob = nullptr;    // profiles + luts
    DevBuf *desc = nullptr;    // QueryDesc tables
    DevBuf *scratch = nullptr; // boundary rows
    uint32_t *counter = nullptr;
    std::vector<uint8_t> *host_blob = nullptr;
    std::vector<QueryDesc> *host_desc = nullptr;
};

// Enqueues on `st` every launch that scores the items against db
// (semi-global): int32 groups first, then packed int16, one persistent
// launch per (cell width, end-rule variant) that has work. `on_launch` is
// called with true before and false after each kernel launch (timing).
// f16: the lockstep tiles of a query whose groups all lie within its
// binary16 range bound run in binary16 lanes (f16_bound).
// Returns 0 or an ISH_E_* code with a message in err.
int32_t enqueue_multi(const ish_hip_db *db, const std::vector<MultiItem> &items, int32_t go,
                      int32_t ge, int32_t mn, int32_t mx, int chunk, int cus, int32_t *score,
                      int32_t *end_ref, int32_t *end_q, uint8_t *status, MultiBuffers &mb,
                      hipStream_t st, std::string &err,
                      const std::function<hipError_t(bool)> &on_launch, bool f16 = true);

} // namespace ish
