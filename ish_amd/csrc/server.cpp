// Serving loop of the ish_hip engine: a device-resident target database
// answers a stream of semi-global queries with continuous batching.
//
// Requests are admitted into launch slots. A slot owns a stream, its own
// work counter, boundary scratch, result buffers for up to max_batch
// queries and a device hit list. Whenever a slot is free and requests are
// waiting, the worker thread orders the waiting requests shortest first
// with aging (estimated scoring time minus `aging` times the time waited,
// so that a long request is not passed over for ever), takes them in that
// order while they fit the slot, and enqueues one multi-query pass for
// them plus the device-side --score emission. It does not wait for earlier
// slots, so a new batch is queued on the GPU behind the one running and its
// workgroups start as soon as the running launch's workgroups retire. Only
// the emitted hits are copied back.
//
// What fits a batch: at most max_batch queries; and, beyond its first
// query, either a fixed cell budget (batch_cells) or, by default, a budget
// from the database and the queue. A batch always grows until its work
// items (groups of 128 targets per query) give every resident wave four of
// them, since a smaller launch leaves the device idle; past that, it takes
// its share of the waiting cells (the waiting cells over the slots), up to
// batch_seconds of scoring at the measured rate. On a database large
// enough that one query fills the device, batches therefore hold one query
// while the queue is short, and grow when a queue builds up.
//
// Per-request results are independent of which other requests share the
// batch: every (query, target) cell is computed by the same kernels with
// that query's own profile and end rules, and nothing is shared across
// queries except read-only target data.
//
// The reference scores one pattern per process (ishlib/searcher/
// do_search.mojo:24-274) and has no serving mode; the emission rule is the
// matchers' (striped_semi_global_matcher.mojo:77-107).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ish_hip.h"
#include "engine.h"
#include "host.h"

using namespace ish;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
    int64_t id = 0;
    int32_t Q = 0;
    uint32_t flags = 0;
    int64_t max_score = 0;
    int64_t cells = 0;
    Profile pf;
    double t_submit = 0, t_launch = 0, t_done = 0;
    int32_t status = 1; // 1 pending, 0 done, < 0 error
    std::string err;
    int32_t batch_id = -1, batch_size = 0;
    std::vector<int4> hits; // (target, score, end_ref, end_query), by target
    std::vector<int32_t> score, end_ref, end_q;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    std::vector<hipEvent_t> kev; // begin/end pairs of the kernel launches
    int nlaunch = 0;
    DevBuf blob, desc, scratch, out, hits, counts, maxsc;
    uint32_t *h_counts = nullptr; // pinned
    std::vector<uint8_t> h_blob;
    std::vector<QueryDesc> h_desc;
    std::vector<float> h_maxsc;
    std::vector<std::shared_ptr<Request>> reqs;
    bool busy = false;
    int32_t batch_id = -1;
};

} // namespace

struct ish_hip_server {
    ish_hip_db *db = nullptr;
    int device = 0, cus = 0;
    // configuration
    std::vector<int8_t> mat;
    int32_t A = 0, go = 0, ge = 0, mn = 0, mx = 0;
    float threshold = 0.8f;
    int32_t max_batch = 16, nslots = 2;
    int64_t batch_cells = 0, cap = 0;
    double aging = 1.0, batch_seconds = 0.5;
    bool fifo = false;
    // scoring rate in cells per second, measured over finished batches
    double rate = 1e13;
    bool keep = false;
    std::vector<int8_t> diag;
    // state
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<std::shared_ptr<Request>> pending;
    std::unordered_map<int64_t, std::shared_ptr<Request>> reqs;
    std::vector<Slot> slots;
    int64_t next_id = 1;
    int32_t next_batch = 0;
    bool stop = false;
    std::thread worker;
    std::string err;
    // statistics
    int64_t st_batches = 0, st_queries = 0, st_cells = 0;
    double st_kernel_ms = 0;
};

static int32_t sfail(ish_hip_server *s, int32_t code, const std::string &msg) {
    if (s) {
        std::lock_guard<std::mutex> g(s->mu);
        s->err = msg;
    }
    return code;
}

// Host form of the emission rule, used when a query's hits overflow the
// device hit list (same float32 arithmetic as ish_hip_emit).
static bool emit_sg(int32_t score, float ms, float thr) {
    const volatile float s = (float)score;
    const volatile float cutoff = ms * thr;
    if (s < cutoff) return false;
    const volatile float ratio = s / ms;
    return ratio >= thr;
}

// Enqueues the batch held in slot `sl` (requests already attached).
static int32_t launch_slot(ish_hip_server *s, Slot &sl, std::string &err) {
    const int64_t n = s->db->n;
    const size_t nb = sl.reqs.size();
    std::vector<MultiItem> items(nb);
    sl.h_maxsc.resize(nb);
    for (size_t k = 0; k < nb; ++k) {
        const Request &r = *sl.reqs[k];
        items[k] = MultiItem{&r.pf, r.Q, r.flags, (int64_t)k * n};
        sl.h_maxsc[k] = (float)r.max_score;
    }
    int32_t *o_score = (int32_t *)sl.out.p;
    int32_t *o_er = o_score + (size_t)s->max_batch * n;
    int32_t *o_eq = o_er + (size_t)s->max_batch * n;
    uint8_t *o_st = (uint8_t *)(o_eq + (size_t)s->max_batch * n);
    uint32_t *counter = (uint32_t *)sl.counts.p;         // work queue counter
    uint32_t *hit_count = (uint32_t *)sl.counts.p + 64;  // per query
    hipError_t e = launch_fill_u32(hit_count, (int64_t)nb, 0u, sl.st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(sl.maxsc.p, sl.h_maxsc.data(), nb * sizeof(float), hipMemcpyHostToDevice, sl.st);
    if (e != hipSuccess) {
        err = std::string("batch setup: ") + hipGetErrorString(e);
        return ISH_E_HIP;
    }
    MultiBuffers mb;
    mb.blob = &sl.blob;
    mb.desc = &sl.desc;
    mb.scratch = &sl.scratch;
    mb.counter = counter;
    mb.host_blob = &sl.h_blob;
    mb.host_desc = &sl.h_desc;
    sl.nlaunch = 0;
    auto timer = [&sl](bool begin) -> hipError_t {
        const size_t need = (size_t)(sl.nlaunch + 1) * 2;
        while (sl.kev.size() < need) {
            hipEvent_t ev;
            hipError_t r = hipEventCreate(&ev);
            if (r != hipSuccess) return r;
            sl.kev.push_back(ev);
        }
        if (begin) return hipEventRecord(sl.kev[(size_t)sl.nlaunch * 2], sl.st);
        hipError_t r = hipEventRecord(sl.kev[(size_t)sl.nlaunch * 2 + 1], sl.st);
        sl.nlaunch++;
        return r;
    };
    int32_t rc = enqueue_multi(s->db, items, s->go, s->ge, s->mn, s->mx, kChunkDefault, s->cus,
                               o_score, o_er, o_eq, o_st, mb, sl.st, err, timer);
    if (rc) return rc;
    EmitParams ep{};
    ep.score = o_score;
    ep.end_ref = o_er;
    ep.end_q = o_eq;
    ep.n = n;
    ep.nq = (int32_t)nb;
    ep.algo = ISH_ALGO_SEMI_GLOBAL;
    ep.max_score = (const float *)sl.maxsc.p;
    ep.threshold = s->threshold;
    ep.hits = (int4 *)sl.hits.p;
    ep.cap = s->cap;
    ep.count = hit_count;
    e = launch_emit_hits(ep, sl.st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(sl.h_counts, hit_count, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, sl.st);
    if (e == hipSuccess) e = hipEventRecord(sl.done, sl.st);
    if (e != hipSuccess) {
        err = std::string("batch launch: ") + hipGetErrorString(e);
        return ISH_E_HIP;
    }
    return ISH_OK;
}

// Copies a finished slot's results into its requests (worker thread, no
// lock held: the requests of a busy slot are only touched here).
static int32_t collect_slot(ish_hip_server *s, Slot &sl, std::string &err, double &kernel_ms) {
    const int64_t n = s->db->n;
    const size_t nb = sl.reqs.size();
    kernel_ms = 0;
    for (int k = 0; k < sl.nlaunch; ++k) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, sl.kev[(size_t)k * 2], sl.kev[(size_t)k * 2 + 1]) == hipSuccess)
            kernel_ms += ms;
    }
    const int32_t *o_score = (const int32_t *)sl.out.p;
    const int32_t *o_er = o_score + (size_t)s->max_batch * n;
    const int32_t *o_eq = o_er + (size_t)s->max_batch * n;
    hipError_t e = hipSuccess;
    for (size_t k = 0; k < nb && e == hipSuccess; ++k) {
        Request &r = *sl.reqs[k];
        const int64_t cnt = sl.h_counts[k];
        const size_t off = k * (size_t)n;
        const bool overflow = cnt > s->cap;
        if (s->keep || overflow) {
            r.score.resize((size_t)n);
            r.end_ref.resize((size_t)n);
            r.end_q.resize((size_t)n);
            e = hipMemcpyAsync(r.score.data(), o_score + off, (size_t)n * 4, hipMemcpyDeviceToHost, sl.st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(r.end_ref.data(), o_er + off, (size_t)n * 4, hipMemcpyDeviceToHost, sl.st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(r.end_q.data(), o_eq + off, (size_t)n * 4, hipMemcpyDeviceToHost, sl.st);
        }
        if (!overflow) {
            r.hits.resize((size_t)cnt);
            if (cnt > 0 && e == hipSuccess)
                e = hipMemcpyAsync(r.hits.data(), (const int4 *)sl.hits.p + k * (size_t)s->cap,
                                   (size_t)cnt * sizeof(int4), hipMemcpyDeviceToHost, sl.st);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(sl.st);
    if (e != hipSuccess) {
        err = std::string("batch results: ") + hipGetErrorString(e);
        return ISH_E_HIP;
    }
    for (size_t k = 0; k < nb; ++k) {
        Request &r = *sl.reqs[k];
        const int64_t cnt = sl.h_counts[k];
        if (cnt > s->cap) {
            const float ms = (float)r.max_score;
            r.hits.clear();
            for (int64_t t = 0; t < n; ++t)
                if (emit_sg(r.score[(size_t)t], ms, s->threshold))
                    r.hits.push_back(make_int4((int)t, r.score[(size_t)t], r.end_ref[(size_t)t], r.end_q[(size_t)t]));
            if (!s->keep) {
                std::vector<int32_t>().swap(r.score);
                std::vector<int32_t>().swap(r.end_ref);
                std::vector<int32_t>().swap(r.end_q);
            }
        } else {
            std::sort(r.hits.begin(), r.hits.end(), [](const int4 &a, const int4 &b) { return a.x < b.x; });
        }
    }
    return ISH_OK;
}

// Moves the waiting requests that form the next batch into the free slot
// sl (lock held): shortest estimated scoring time first, less `aging`
// times the time waited (FIFO when aging < 0), while they fit.
static void admit(ish_hip_server *s, Slot &sl) {
    const double now = now_s();
    auto &P = s->pending;
    std::vector<std::pair<double, size_t>> order(P.size());
    int64_t backlog = 0;
    for (size_t k = 0; k < P.size(); ++k) {
        const Request &r = *P[k];
        const double key = s->fifo ? r.t_submit : (double)r.cells / s->rate - s->aging * (now - r.t_submit);
        order[k] = {key, k};
        backlog += r.cells;
    }
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<double, size_t> &a, const std::pair<double, size_t> &b) {
                         return a.first < b.first;
                     });
    // work items that keep every resident wave (8 per CU) busy with four
    const int64_t fill_items = (int64_t)s->cus * 8 * 4;
    const int64_t per_query_items = std::max<int64_t>(s->db->ngroups, 1);
    const int64_t cap = s->batch_cells > 0
                            ? s->batch_cells
                            : (int64_t)std::min((double)backlog / s->nslots, s->batch_seconds * s->rate);
    int64_t cells = 0; // beyond the first query
    std::vector<bool> taken(P.size(), false);
    for (const auto &o : order) {
        if ((int32_t)sl.reqs.size() >= s->max_batch) break;
        const auto &r = P[o.second];
        if (!sl.reqs.empty()) {
            const bool fills = (int64_t)sl.reqs.size() * per_query_items >= fill_items;
            if ((fills || s->batch_cells > 0) && cells + r->cells > cap) break;
            cells += r->cells;
        }
        sl.reqs.push_back(r);
        taken[o.second] = true;
    }
    size_t w = 0;
    for (size_t k = 0; k < P.size(); ++k)
        if (!taken[k]) P[w++] = P[k];
    P.resize(w);
}

static void worker_main(ish_hip_server *s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // 1. finished slots
        for (Slot &sl : s->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            lk.unlock();
            std::string err;
            double kms = 0;
            int32_t rc = q == hipSuccess ? collect_slot(s, sl, err, kms) : ISH_E_HIP;
            if (q != hipSuccess && q != hipErrorNotReady) err = std::string("batch: ") + hipGetErrorString(q);
            const double t = now_s();
            lk.lock();
            for (auto &r : sl.reqs) {
                r->t_done = t;
                r->status = rc;
                if (rc) r->err = err;
                s->st_queries++;
                s->st_cells += r->cells;
            }
            s->st_kernel_ms += kms;
            if (rc == ISH_OK && kms > 0) {
                int64_t bc = 0;
                for (auto &r : sl.reqs) bc += r->cells;
                s->rate = 0.7 * s->rate + 0.3 * ((double)bc / (kms * 1e-3));
            }
            sl.reqs.clear();
            sl.busy = false;
            s->cv_done.notify_all();
        }
        // 2. admission into free slots
        for (Slot &sl : s->slots) {
            if (sl.busy || s->pending.empty()) continue;
            admit(s, sl);
            sl.batch_id = s->next_batch++;
            const double t = now_s();
            for (auto &r : sl.reqs) {
                r->t_launch = t;
                r->batch_id = sl.batch_id;
                r->batch_size = (int32_t)sl.reqs.size();
            }
            sl.busy = true;
            s->st_batches++;
            lk.unlock();
            std::string err;
            const int32_t rc = launch_slot(s, sl, err);
            lk.lock();
            if (rc) {
                (void)hipStreamSynchronize(sl.st);
                for (auto &r : sl.reqs) {
                    r->status = rc;
                    r->err = err;
                    r->t_done = now_s();
                }
                sl.reqs.clear();
                sl.busy = false;
                s->cv_done.notify_all();
            }
        }
        bool any_busy = false, any_free = false;
        for (Slot &sl : s->slots) (sl.busy ? any_busy : any_free) = true;
        if (s->stop && !any_busy) break;
        if (!any_busy && s->pending.empty()) {
            s->cv_work.wait(lk, [s] { return s->stop || !s->pending.empty(); });
        } else if (any_busy) {
            // poll the in-flight slots; a submission into a free slot wakes
            // the worker at once
            s->cv_work.wait_for(lk, std::chrono::microseconds(100));
        }
        (void)any_free;
    }
    // shutdown: fail whatever is still waiting
    for (auto &r : s->pending) {
        r->status = ISH_E_INVALID;
        r->err = "server shut down";
        r->t_done = now_s();
    }
    s->pending.clear();
    s->cv_done.notify_all();
}

extern "C" int32_t ish_hip_server_create(ish_hip_db *db, const ish_hip_server_config *cfg,
                                         ish_hip_server **out) {
    if (!db || !cfg || !out) return ISH_E_INVALID;
    *out = nullptr;
    auto s = std::make_unique<ish_hip_server>();
    s->db = db;
    s->device = db->device;
    int32_t A = 0;
    std::string err;
    int32_t rc = matrix_of(cfg->matrix, cfg->matrix_len, cfg->matrix_kind, A, s->mat, err);
    if (rc) return rc;
    s->A = A;
    if (cfg->gap_open < 0 || cfg->gap_extend < 0) return ISH_E_INVALID;
    s->go = cfg->gap_open;
    s->ge = cfg->gap_extend;
    for (int8_t v : s->mat) {
        s->mn = std::min<int32_t>(s->mn, v);
        s->mx = std::max<int32_t>(s->mx, v);
    }
    s->diag.resize((size_t)A);
    for (int32_t a = 0; a < A; ++a) s->diag[(size_t)a] = s->mat[(size_t)a * A + a];
    s->threshold = cfg->threshold;
    s->max_batch = cfg->max_batch > 0 ? cfg->max_batch : 16;
    s->nslots = cfg->slots > 0 ? cfg->slots : 2;
    if (s->max_batch > 256 || s->nslots > 4) return ISH_E_RANGE;
    s->batch_cells = cfg->batch_cells > 0 ? cfg->batch_cells : 0; // 0: from the database and the queue
    s->aging = cfg->aging > 0 ? cfg->aging : 1.0;
    s->fifo = cfg->aging < 0;
    s->batch_seconds = cfg->batch_seconds > 0 ? cfg->batch_seconds : 0.5;
    s->cap = cfg->hits_per_query > 0 ? cfg->hits_per_query : 65536;
    s->keep = cfg->keep_scores != 0;
    if (hipSetDevice(s->device) != hipSuccess) return ISH_E_HIP;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, s->device) != hipSuccess) return ISH_E_HIP;
    s->cus = pr.multiProcessorCount;
    const int64_t n = db->n;
    s->slots.resize((size_t)s->nslots);
    hipError_t e = hipSuccess;
    for (Slot &sl : s->slots) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e == hipSuccess) e = sl.out.ensure((size_t)s->max_batch * (size_t)std::max<int64_t>(n, 1) * 13);
        if (e == hipSuccess) e = sl.hits.ensure((size_t)s->max_batch * (size_t)s->cap * sizeof(int4));
        if (e == hipSuccess) e = sl.counts.ensure((64 + 256) * sizeof(uint32_t));
        if (e == hipSuccess) e = sl.maxsc.ensure(256 * sizeof(float));
        if (e == hipSuccess) e = hipHostMalloc((void **)&sl.h_counts, 256 * sizeof(uint32_t), hipHostMallocDefault);
    }
    if (e != hipSuccess) {
        for (Slot &sl : s->slots) {
            if (sl.st) (void)hipStreamDestroy(sl.st);
            if (sl.done) (void)hipEventDestroy(sl.done);
            for (DevBuf *b : {&sl.out, &sl.hits, &sl.counts, &sl.maxsc}) b->release();
            if (sl.h_counts) (void)hipHostFree(sl.h_counts);
        }
        return e == hipErrorOutOfMemory ? ISH_E_NOMEM : ISH_E_HIP;
    }
    ish_hip_server *raw = s.release();
    raw->worker = std::thread(worker_main, raw);
    *out = raw;
    return ISH_OK;
}

extern "C" void ish_hip_server_destroy(ish_hip_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->device);
    for (Slot &sl : s->slots) {
        (void)hipStreamSynchronize(sl.st);
        for (auto ev : sl.kev) (void)hipEventDestroy(ev);
        (void)hipEventDestroy(sl.done);
        (void)hipStreamDestroy(sl.st);
        for (DevBuf *b : {&sl.blob, &sl.desc, &sl.scratch, &sl.out, &sl.hits, &sl.counts, &sl.maxsc})
            b->release();
        if (sl.h_counts) (void)hipHostFree(sl.h_counts);
    }
    delete s;
}

// The message is copied under the lock into a string of the calling thread:
// other threads may set a new error at any time, and the pointer stays
// valid until this thread's next call.
extern "C" const char *ish_hip_server_last_error(ish_hip_server *s) {
    if (!s) return "null server";
    thread_local std::string msg;
    {
        std::lock_guard<std::mutex> g(s->mu);
        msg = s->err;
    }
    return msg.c_str();
}

extern "C" double ish_hip_server_clock(void) { return now_s(); }

extern "C" int32_t ish_hip_server_submit(ish_hip_server *s, const uint8_t *enc_query, int32_t qlen,
                                         uint32_t ends_free, int64_t *ticket) {
    if (!s || !ticket || qlen < 0 || (qlen > 0 && !enc_query)) return sfail(s, ISH_E_INVALID, "bad request");
    auto r = std::make_shared<Request>();
    r->t_submit = now_s();
    r->Q = qlen;
    r->flags = ends_free & 15u;
    int64_t ms = 0;
    for (int32_t j = 0; j < qlen; ++j) {
        if (enc_query[j] >= s->A)
            return sfail(s, ISH_E_INVALID, "query code " + std::to_string(enc_query[j]) + " >= alphabet");
        ms += s->diag[enc_query[j]];
    }
    r->max_score = ms;
    r->cells = (int64_t)qlen * s->db->nres;
    build_profile(enc_query, qlen, s->mat, s->A, 2 * s->ge, false, s->go >= s->ge,
                  top_row_score(r->flags, s->ge), 0, r->pf);
    if (r->pf.lds() > kLdsLimit)
        return sfail(s, ISH_E_RANGE, "query profile needs " + std::to_string(r->pf.lds()) + " bytes of LDS");
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return ISH_E_INVALID;
        r->id = s->next_id++;
        s->reqs[r->id] = r;
        s->pending.push_back(r);
        *ticket = r->id;
    }
    s->cv_work.notify_all();
    return ISH_OK;
}

static std::shared_ptr<Request> find_req(ish_hip_server *s, int64_t ticket) {
    auto it = s->reqs.find(ticket);
    return it == s->reqs.end() ? nullptr : it->second;
}

extern "C" int32_t ish_hip_server_wait(ish_hip_server *s, int64_t ticket, int32_t timeout_ms,
                                       ish_hip_request_info *info) {
    if (!s) return ISH_E_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto r = find_req(s, ticket);
    if (!r) {
        s->err = "unknown ticket";
        return ISH_E_INVALID;
    }
    auto ready = [&] { return r->status != 1; };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return ISH_E_TIMEOUT;
    if (info) {
        info->n_hits = (int64_t)r->hits.size();
        info->cells = r->cells;
        info->t_submit = r->t_submit;
        info->t_launch = r->t_launch;
        info->t_done = r->t_done;
        info->batch_id = r->batch_id;
        info->batch_size = r->batch_size;
        info->status = r->status;
    }
    if (r->status < 0) s->err = r->err;
    return r->status;
}

extern "C" int32_t ish_hip_server_hits(ish_hip_server *s, int64_t ticket, int64_t *target,
                                       int32_t *score, int32_t *end_ref, int32_t *end_query,
                                       int64_t cap) {
    if (!s || cap < 0) return ISH_E_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    auto r = find_req(s, ticket);
    if (!r || r->status != 0) {
        s->err = "ticket has no results";
        return ISH_E_INVALID;
    }
    const int64_t m = std::min<int64_t>(cap, (int64_t)r->hits.size());
    for (int64_t k = 0; k < m; ++k) {
        const int4 &h = r->hits[(size_t)k];
        if (target) target[k] = h.x;
        if (score) score[k] = h.y;
        if (end_ref) end_ref[k] = h.z;
        if (end_query) end_query[k] = h.w;
    }
    return ISH_OK;
}

extern "C" int32_t ish_hip_server_scores(ish_hip_server *s, int64_t ticket, int32_t *score,
                                         int32_t *end_ref, int32_t *end_query) {
    if (!s) return ISH_E_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    auto r = find_req(s, ticket);
    if (!r || r->status != 0 || r->score.size() != (size_t)s->db->n) {
        s->err = "full results are kept only with keep_scores";
        return ISH_E_INVALID;
    }
    const size_t n = (size_t)s->db->n;
    if (score) memcpy(score, r->score.data(), n * 4);
    if (end_ref) memcpy(end_ref, r->end_ref.data(), n * 4);
    if (end_query) memcpy(end_query, r->end_q.data(), n * 4);
    return ISH_OK;
}

extern "C" int32_t ish_hip_server_release(ish_hip_server *s, int64_t ticket) {
    if (!s) return ISH_E_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    auto r = find_req(s, ticket);
    if (!r) return ISH_E_INVALID;
    if (r->status == 1) {
        s->err = "request still in flight";
        return ISH_E_INVALID;
    }
    s->reqs.erase(ticket);
    return ISH_OK;
}

extern "C" int32_t ish_hip_server_get_stats(ish_hip_server *s, ish_hip_server_stats *out) {
    if (!s || !out) return ISH_E_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    out->batches = s->st_batches;
    out->queries = s->st_queries;
    out->cells = s->st_cells;
    out->kernel_ms = s->st_kernel_ms;
    out->pending = (int64_t)s->pending.size();
    int64_t busy = 0;
    for (const Slot &sl : s->slots) busy += sl.busy ? 1 : 0;
    out->busy_slots = busy;
    return ISH_OK;
}
