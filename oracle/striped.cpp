/*
 * Farrar-striped restatement of the reference's CPU SIMD kernels.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): differential check on oracle.c and
 * the timed CPU baseline. Never linked into the product.
 *
 *  - sg_striped<W>: semi_global_aln (ishlib/matcher/alignment/semi_global_aln/
 *    striped.mojo:252-498) with int16 saturating lanes, the profile layout of
 *    Profile.generate_query_profile (:92-118), the SWPS3 lazy-F loop
 *    (:396-434), last-row tracking (:436-441) and the free-end passes in the
 *    reference's strided visiting order (:446-477).
 *  - sw_striped<T,W>: sw (local_aln/striped.mojo:272-434) in biased unsigned
 *    lanes; ssw_align's u8 pass with a u16 rerun when best+bias >= 255
 *    (:155-190).
 * Lanes are plain fixed-size arrays so that the compiler vectorises each
 * lane loop at the host's native width; W is the number of lanes, exactly
 * the reference's simd_width_of (matcher/__init__.mojo:11-30).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#if defined(__AVX2__)
#include <immintrin.h>
#endif

#include "oracle.h"

namespace {

template <typename T> struct Lim;
template <> struct Lim<int16_t> {
    static constexpr int32_t MIN = -32768, MAX = 32767;
};
template <> struct Lim<uint8_t> {
    static constexpr int32_t MIN = 0, MAX = 255;
};
template <> struct Lim<uint16_t> {
    static constexpr int32_t MIN = 0, MAX = 65535;
};

template <typename T> static inline T sat(int32_t v) {
    return (T)(v < Lim<T>::MIN ? Lim<T>::MIN : (v > Lim<T>::MAX ? Lim<T>::MAX : v));
}

template <typename T, int W> struct Vec {
    T v[W];
};

/* Lane access (the algorithms touch single lanes only outside the
 * per-segment loops). */
template <typename T, int W> static inline T lget(const Vec<T, W> &a, int k) { return a.v[k]; }
template <typename T, int W> static inline void lset(Vec<T, W> &a, int k, T x) { a.v[k] = x; }

template <typename T, int W> static inline bool vall_eq(const Vec<T, W> &a, const Vec<T, W> &b) {
    bool eq = true;
    for (int k = 0; k < W; ++k) eq &= a.v[k] == b.v[k];
    return eq;
}
template <typename T, int W> static inline int32_t vhmax(const Vec<T, W> &a) {
    int32_t m = 0;
    for (int k = 0; k < W; ++k) m = std::max<int32_t>(m, a.v[k]);
    return m;
}

template <typename T, int W> static inline Vec<T, W> vset(int32_t x) {
    Vec<T, W> r;
    for (int k = 0; k < W; ++k) r.v[k] = (T)x;
    return r;
}
template <typename T, int W>
static inline Vec<T, W> vadds(const Vec<T, W> &a, const Vec<T, W> &b) {
    Vec<T, W> r;
    for (int k = 0; k < W; ++k) r.v[k] = sat<T>((int32_t)a.v[k] + (int32_t)b.v[k]);
    return r;
}
template <typename T, int W>
static inline Vec<T, W> vsubs(const Vec<T, W> &a, const Vec<T, W> &b) {
    Vec<T, W> r;
    for (int k = 0; k < W; ++k) r.v[k] = sat<T>((int32_t)a.v[k] - (int32_t)b.v[k]);
    return r;
}
template <typename T, int W>
static inline Vec<T, W> vmax(const Vec<T, W> &a, const Vec<T, W> &b) {
    Vec<T, W> r;
    for (int k = 0; k < W; ++k) r.v[k] = a.v[k] > b.v[k] ? a.v[k] : b.v[k];
    return r;
}
/* shift_right[1]: lane k takes lane k-1, lane 0 becomes zero. */
template <typename T, int W> static inline Vec<T, W> vshr1(const Vec<T, W> &a) {
    Vec<T, W> r;
    r.v[0] = 0;
    for (int k = 1; k < W; ++k) r.v[k] = a.v[k - 1];
    return r;
}
template <typename T, int W>
static inline bool vany_gt(const Vec<T, W> &a, const Vec<T, W> &b) {
    bool any = false;
    for (int k = 0; k < W; ++k) any |= a.v[k] > b.v[k];
    return any;
}

/* Native-width lanes. The generic helpers above are the definition; these
 * overloads compute the same thing with AVX2 / AVX-512BW instructions
 * (saturating add/sub, max, whole-vector shift by one lane), so that the
 * timed CPU baseline runs at the speed of the reference's SIMD code. */
#if defined(__AVX2__)
#define ORACLE_SIMD_DEFS(TYPE, LANES, REG, PFX, ADDS, SUBS, MAX, GT_ANY, SHR1)       \
    template <> struct Vec<TYPE, LANES> {                                             \
        REG m;                                                                        \
    };                                                                                \
    template <> inline TYPE lget<TYPE, LANES>(const Vec<TYPE, LANES> &a, int k) {     \
        alignas(64) TYPE t[LANES];                                                    \
        memcpy(t, &a.m, sizeof t);                                                    \
        return t[k];                                                                  \
    }                                                                                 \
    template <> inline void lset<TYPE, LANES>(Vec<TYPE, LANES> & a, int k, TYPE x) {  \
        alignas(64) TYPE t[LANES];                                                    \
        memcpy(t, &a.m, sizeof t);                                                    \
        t[k] = x;                                                                     \
        memcpy(&a.m, t, sizeof t);                                                    \
    }                                                                                 \
    template <> inline bool vall_eq<TYPE, LANES>(const Vec<TYPE, LANES> &a, const Vec<TYPE, LANES> &b) { \
        return memcmp(&a.m, &b.m, sizeof(REG)) == 0;                                  \
    }                                                                                 \
    template <> inline int32_t vhmax<TYPE, LANES>(const Vec<TYPE, LANES> &a) {        \
        alignas(64) TYPE t[LANES];                                                    \
        memcpy(t, &a.m, sizeof t);                                                    \
        int32_t m = 0;                                                                \
        for (int k = 0; k < LANES; ++k) m = std::max<int32_t>(m, t[k]);               \
        return m;                                                                     \
    }                                                                                 \
    template <> inline Vec<TYPE, LANES> vset<TYPE, LANES>(int32_t x) {                \
        alignas(64) TYPE t[LANES];                                                    \
        for (int k = 0; k < LANES; ++k) t[k] = (TYPE)x;                               \
        Vec<TYPE, LANES> r;                                                           \
        memcpy(&r.m, t, sizeof t);                                                    \
        return r;                                                                     \
    }                                                                                 \
    static inline Vec<TYPE, LANES> vadds(const Vec<TYPE, LANES> &a, const Vec<TYPE, LANES> &b) { \
        Vec<TYPE, LANES> r;                                                           \
        r.m = ADDS(a.m, b.m);                                                         \
        return r;                                                                     \
    }                                                                                 \
    static inline Vec<TYPE, LANES> vsubs(const Vec<TYPE, LANES> &a, const Vec<TYPE, LANES> &b) { \
        Vec<TYPE, LANES> r;                                                           \
        r.m = SUBS(a.m, b.m);                                                         \
        return r;                                                                     \
    }                                                                                 \
    static inline Vec<TYPE, LANES> vmax(const Vec<TYPE, LANES> &a, const Vec<TYPE, LANES> &b) { \
        Vec<TYPE, LANES> r;                                                           \
        r.m = MAX(a.m, b.m);                                                          \
        return r;                                                                     \
    }                                                                                 \
    static inline bool vany_gt(const Vec<TYPE, LANES> &a, const Vec<TYPE, LANES> &b) { \
        return GT_ANY(a.m, b.m);                                                      \
    }                                                                                 \
    static inline Vec<TYPE, LANES> vshr1(const Vec<TYPE, LANES> &a) {                 \
        Vec<TYPE, LANES> r;                                                           \
        r.m = SHR1(a.m);                                                              \
        return r;                                                                     \
    }

/* 256 bits: lane shift toward higher indices by `bytes`, zero fill. */
template <int BYTES> static inline __m256i shr_lanes256(__m256i a) {
    const __m256i t = _mm256_permute2x128_si256(a, a, 0x08); /* [0, a.lo] */
    return _mm256_alignr_epi8(a, t, 16 - BYTES);
}
static inline bool gt_any_i16_256(__m256i a, __m256i b) {
    return _mm256_movemask_epi8(_mm256_cmpgt_epi16(a, b)) != 0;
}
static inline bool gt_any_u8_256(__m256i a, __m256i b) {
    return _mm256_movemask_epi8(_mm256_cmpeq_epi8(_mm256_max_epu8(a, b), b)) != -1;
}
static inline bool gt_any_u16_256(__m256i a, __m256i b) {
    return _mm256_movemask_epi8(_mm256_cmpeq_epi16(_mm256_max_epu16(a, b), b)) != -1;
}
ORACLE_SIMD_DEFS(int16_t, 16, __m256i, 256, _mm256_adds_epi16, _mm256_subs_epi16, _mm256_max_epi16,
                 gt_any_i16_256, shr_lanes256<2>)
ORACLE_SIMD_DEFS(uint8_t, 32, __m256i, 256, _mm256_adds_epu8, _mm256_subs_epu8, _mm256_max_epu8,
                 gt_any_u8_256, shr_lanes256<1>)
ORACLE_SIMD_DEFS(uint16_t, 16, __m256i, 256, _mm256_adds_epu16, _mm256_subs_epu16, _mm256_max_epu16,
                 gt_any_u16_256, shr_lanes256<2>)
#endif

#if defined(__AVX512BW__)
/* 512 bits: [0, a.l0, a.l1, a.l2] by 128-bit lanes, then a byte align. */
template <int BYTES> static inline __m512i shr_lanes512(__m512i a) {
    const __m512i t = _mm512_alignr_epi64(a, _mm512_setzero_si512(), 6);
    return _mm512_alignr_epi8(a, t, 16 - BYTES);
}
static inline bool gt_any_i16_512(__m512i a, __m512i b) { return _mm512_cmpgt_epi16_mask(a, b) != 0; }
static inline bool gt_any_u8_512(__m512i a, __m512i b) { return _mm512_cmpgt_epu8_mask(a, b) != 0; }
static inline bool gt_any_u16_512(__m512i a, __m512i b) { return _mm512_cmpgt_epu16_mask(a, b) != 0; }
ORACLE_SIMD_DEFS(int16_t, 32, __m512i, 512, _mm512_adds_epi16, _mm512_subs_epi16, _mm512_max_epi16,
                 gt_any_i16_512, shr_lanes512<2>)
ORACLE_SIMD_DEFS(uint8_t, 64, __m512i, 512, _mm512_adds_epu8, _mm512_subs_epu8, _mm512_max_epu8,
                 gt_any_u8_512, shr_lanes512<1>)
ORACLE_SIMD_DEFS(uint16_t, 32, __m512i, 512, _mm512_adds_epu16, _mm512_subs_epu16, _mm512_max_epu16,
                 gt_any_u16_512, shr_lanes512<2>)
#endif

/* The widest int16 lane count of this build (the reference's
 * simd_width_of for int16 on this host). */
#if defined(__AVX512BW__)
constexpr int kNativeW = 32;
#else
constexpr int kNativeW = 16;
#endif

/* Profile.generate_query_profile: P[c*seg + s].lane[k] = S[c][q[s + k*seg]],
 * padding rows score `pad`. */
template <typename T, int W>
static void make_profile(const uint8_t *q, int32_t Q, const int8_t *S,
                         int32_t A, int32_t bias, std::vector<Vec<T, W>> &P,
                         int32_t &seg) {
    seg = (Q + W - 1) / W;
    P.assign((size_t)A * seg, Vec<T, W>());
    for (int32_t c = 0; c < A; ++c)
        for (int32_t s = 0; s < seg; ++s)
            for (int k = 0; k < W; ++k) {
                int32_t j = s + k * seg;
                int32_t val = j >= Q ? bias : (int32_t)S[(int64_t)c * A + q[j]] + bias;
                lset(P[(size_t)c * seg + s], k, (T)val);
            }
}

/* Query-side state of semi_global_aln (the matcher builds it once per
 * query, striped_semi_global_matcher.mojo:41-75). */
template <int W> struct SgQuery {
    typedef Vec<int16_t, W> V;
    std::vector<V> P;
    int32_t seg = 0, Q = 0, go = 0, ge = 0, min_score = 0, max_score = 0;
    uint32_t flags = 0;
    void prepare(const uint8_t *q, int32_t Qn, const int8_t *S, int32_t A, int32_t gop,
                 int32_t gex, uint32_t fl) {
        Q = Qn;
        go = gop;
        ge = gex;
        flags = fl;
        min_score = max_score = 0;
        for (int64_t k = 0; k < (int64_t)A * A; ++k) {
            min_score = std::min<int32_t>(min_score, S[k]);
            max_score = std::max<int32_t>(max_score, S[k]);
        }
        if (Q > 0) make_profile<int16_t, W>(q, Q, S, A, 0, P, seg);
    }
};

/* Per-thread buffers of semi_global_aln (the reference allocates them per
 * call, striped.mojo:288-293; reusing them changes no result). */
template <int W> struct SgWork {
    std::vector<Vec<int16_t, W>> h_store, h_load, pv_e;
    std::vector<int16_t> boundary;
};

template <int W>
oracle_end_t sg_striped_run(const SgQuery<W> &qy, const uint8_t *r, int32_t R, SgWork<W> &wk) {
    typedef int16_t T;
    typedef Vec<T, W> V;
    const int32_t MIN = Lim<T>::MIN, MAX = Lim<T>::MAX;
    oracle_end_t res = {0, 0, 0};
    const int32_t Q = qy.Q, go = qy.go, ge = qy.ge, seg = qy.seg;
    if (Q == 0 || R == 0) return res;
    const uint32_t flags = qy.flags;
    const bool fqs = flags & ORACLE_QS, fqe = flags & ORACLE_QE;
    const bool fts = flags & ORACLE_TS, fte = flags & ORACLE_TE;
    const int32_t min_score = qy.min_score;
    const std::vector<V> &P = qy.P;

    int32_t end_query = Q - 1, end_reference = -1;
    const int32_t offset = (Q - 1) % seg;
    const int32_t position = (W - 1) - (Q - 1) / seg;
    const int pos_lane = W - 1 - position; /* lane holding row Q-1 */

    std::vector<V> &h_store = wk.h_store, &h_load = wk.h_load, &pv_e = wk.pv_e;
    std::vector<T> &boundary = wk.boundary;
    h_store.resize(seg);
    h_load.resize(seg);
    pv_e.resize(seg);
    boundary.resize((size_t)R + 1);
    for (int32_t s = 0; s < seg; ++s) {
        V h = vset<T, W>(0), e = vset<T, W>(0);
        for (int k = 0; k < W; ++k) {
            int32_t tmp = fqs ? 0 : (int32_t)(int16_t)(-go - (int16_t)(ge * (k * seg + s)));
            lset(h, k, (T)(tmp < MIN ? MIN : tmp));
            tmp = (int32_t)(int16_t)(tmp - go);
            lset(e, k, (T)(tmp < MIN ? MIN : tmp));
        }
        h_store[s] = h;
        pv_e[s] = e;
    }
    boundary[0] = 0;
    for (int32_t i = 1; i <= R; ++i) {
        int32_t tmp = fts ? 0 : -go - ge * (i - 1);
        boundary[i] = (T)(tmp < MIN ? MIN : (int16_t)tmp);
    }
    const V v_go = vset<T, W>(go), v_ge = vset<T, W>(ge);
    int32_t nl = (-go < min_score) ? MIN + go + 1 : MIN - min_score + 1;
    const V v_neg_limit = vset<T, W>((int16_t)nl);
    T score = (T)(int16_t)nl;
    V v_max_h = v_neg_limit;

    for (int32_t i = 0; i < R; ++i) {
        V v_e;
        V v_f = v_neg_limit;
        V v_h = vshr1(h_store[seg - 1]);
        const V *prof = &P[(size_t)r[i] * seg];
        std::swap(h_load, h_store);
        lset(v_h, 0, boundary[i]);
        for (int32_t j = 0; j < seg; ++j) {
            v_h = vadds(v_h, prof[j]);
            v_e = pv_e[j];
            v_h = vmax(v_h, v_e);
            v_h = vmax(v_h, v_f);
            h_store[j] = v_h;
            v_h = vsubs(v_h, v_go);
            v_e = vsubs(v_e, v_ge);
            v_e = vmax(v_e, v_h);
            pv_e[j] = v_e;
            v_f = vsubs(v_f, v_ge);
            v_f = vmax(v_f, v_h);
            v_h = h_load[j];
        }
        /* lazy-F (SWPS3): H is corrected, E is not */
        bool done = false;
        for (int kk = 0; kk < W && !done; ++kk) {
            int32_t tmp = fts ? -go : (int32_t)boundary[i + 1] - go;
            T tmp2 = (T)(tmp < MIN ? MIN : tmp);
            v_f = vshr1(v_f);
            lset(v_f, 0, tmp2);
            for (int32_t j = 0; j < seg; ++j) {
                v_h = h_store[j];
                v_h = vmax(v_h, v_f);
                h_store[j] = v_h;
                v_h = vsubs(v_h, v_go);
                v_f = vsubs(v_f, v_ge);
                if (!vany_gt(v_f, v_h)) {
                    done = true;
                    break;
                }
            }
        }
        v_h = h_store[offset];
        bool cmp = lget(v_h, pos_lane) > lget(v_max_h, pos_lane);
        v_max_h = vmax(v_h, v_max_h);
        if (cmp) end_reference = i;
    }
    if (fte) {
        score = lget(v_max_h, pos_lane);
        end_query = Q - 1;
    }
    if (fqe) {
        for (int32_t s = 0; s < seg; ++s)
            for (int k = 0; k < W; ++k) {
                int32_t temp = s + k * seg;
                if (temp >= Q) continue;
                T hv = lget(h_store[s], k);
                if (hv > score) {
                    score = hv;
                    end_query = temp;
                    end_reference = R - 1;
                } else if (hv == score && end_reference == R - 1 && temp < end_query) {
                    end_query = temp;
                }
            }
    }
    if (!fte && !fqe) {
        score = lget(h_store[offset], pos_lane);
        end_reference = R - 1;
        end_query = Q - 1;
    }
    (void)MAX;
    res.score = score;
    res.end_ref = end_reference;
    res.end_query = end_query;
    return res;
}


template <int W>
oracle_end_t sg_striped(const uint8_t *q, int32_t Q, const uint8_t *r, int32_t R, const int8_t *S,
                        int32_t A, int32_t go, int32_t ge, uint32_t flags) {
    SgQuery<W> qy;
    qy.prepare(q, Q, S, A, go, ge, flags);
    SgWork<W> wk;
    return sg_striped_run<W>(qy, r, R, wk);
}

/* sw[dt, width] (local_aln/striped.mojo:272-434), forward direction,
 * terminate disabled. Returns the best with overflow flagged by
 * best+bias >= MAX (the caller reruns wider). */
template <typename T, int W> struct SwWork {
    std::vector<Vec<T, W>> h_store, h_load, pv_e, h_max;
};

template <typename T, int W>
oracle_end_t sw_striped_t(const std::vector<Vec<T, W>> &P, int32_t seg, int32_t Q,
                          const uint8_t *r, int32_t R, int32_t go, int32_t ge, int32_t bias,
                          bool *overflow, SwWork<T, W> &wk) {
    typedef Vec<T, W> V;
    const int32_t MAXV = Lim<T>::MAX;
    std::vector<V> &h_store = wk.h_store, &h_load = wk.h_load, &pv_e = wk.pv_e, &h_max = wk.h_max;
    h_store.assign(seg, vset<T, W>(0));
    h_load.assign(seg, vset<T, W>(0));
    pv_e.assign(seg, vset<T, W>(0));
    h_max.assign(seg, vset<T, W>(0));
    const V zero = vset<T, W>(0), v_go = vset<T, W>(go), v_ge = vset<T, W>(ge),
            v_bias = vset<T, W>(bias);
    V v_max_score = zero, v_max_mark = zero;
    int32_t max_score = 0, end_query = Q - 1, end_reference = -1;
    *overflow = false;
    for (int32_t i = 0; i < R; ++i) {
        V e, v_f = zero, v_max_col = zero;
        V v_h = vshr1(h_store[seg - 1]);
        const V *prof = &P[(size_t)r[i] * seg];
        std::swap(h_load, h_store);
        for (int32_t j = 0; j < seg; ++j) {
            v_h = vadds(v_h, prof[j]);
            v_h = vsubs(v_h, v_bias);
            e = pv_e[j];
            v_h = vmax(v_h, e);
            v_h = vmax(v_h, v_f);
            v_max_col = vmax(v_max_col, v_h);
            h_store[j] = v_h;
            v_h = vsubs(v_h, v_go);
            e = vsubs(e, v_ge);
            e = vmax(e, v_h);
            pv_e[j] = e;
            v_f = vsubs(v_f, v_ge);
            v_f = vmax(v_f, v_h);
            v_h = h_load[j];
        }
        bool brk = false;
        for (int kk = 0; kk < W && !brk; ++kk) {
            v_f = vshr1(v_f);
            for (int32_t j = 0; j < seg; ++j) {
                v_h = h_store[j];
                v_h = vmax(v_h, v_f);
                v_max_col = vmax(v_max_col, v_h);
                h_store[j] = v_h;
                v_h = vsubs(v_h, v_go);
                v_f = vsubs(v_f, v_ge);
                if (!vany_gt(v_f, v_h)) {
                    brk = true;
                    break;
                }
            }
        }
        v_max_score = vmax(v_max_score, v_max_col);
        if (!vall_eq(v_max_mark, v_max_score)) {
            const int32_t temp = vhmax(v_max_score);
            v_max_mark = v_max_score;
            if (temp > max_score) {
                max_score = temp;
                if (max_score + bias >= MAXV) {
                    *overflow = true;
                    break;
                }
                end_reference = i;
                for (int32_t j = 0; j < seg; ++j) h_max[j] = h_store[j];
            }
        }
    }
    for (int32_t s = 0; s < seg; ++s)
        for (int k = 0; k < W; ++k)
            if ((int32_t)lget(h_max[s], k) == max_score) {
                int32_t temp = s + k * seg;
                if (temp < end_query) end_query = temp;
            }
    oracle_end_t res = {max_score, end_reference, end_query};
    return res;
}

/* Query-side state of ssw_align: the biased u8 and u16 profiles
 * (local_aln/striped.mojo:110-139), built once per query. */
template <int W> struct SwQuery {
    std::vector<Vec<uint8_t, 2 * W>> P8;
    std::vector<Vec<uint16_t, W>> P16;
    int32_t seg8 = 0, seg16 = 0, Q = 0, go = 0, ge = 0, bias = 0;
    void prepare(const uint8_t *q, int32_t Qn, const int8_t *S, int32_t A, int32_t gop, int32_t gex) {
        Q = Qn;
        go = gop;
        ge = gex;
        bias = 0;
        for (int64_t k = 0; k < (int64_t)A * A; ++k) bias = std::min<int32_t>(bias, S[k]);
        bias = -bias;
        make_profile<uint8_t, 2 * W>(q, Q, S, A, bias, P8, seg8);
        make_profile<uint16_t, W>(q, Q, S, A, bias, P16, seg16);
    }
};
template <int W> struct SwWorkAll {
    SwWork<uint8_t, 2 * W> w8;
    SwWork<uint16_t, W> w16;
};

template <int W>
oracle_end_t sw_striped_run(const SwQuery<W> &qy, const uint8_t *r, int32_t R, SwWorkAll<W> &wk) {
    bool ov = false;
    oracle_end_t e = sw_striped_t<uint8_t, 2 * W>(qy.P8, qy.seg8, qy.Q, r, R, qy.go, qy.ge, qy.bias,
                                                  &ov, wk.w8);
    if (e.score + qy.bias >= 255)
        e = sw_striped_t<uint16_t, W>(qy.P16, qy.seg16, qy.Q, r, R, qy.go, qy.ge, qy.bias, &ov,
                                      wk.w16);
    return e;
}

template <int W>
oracle_end_t sw_striped(const uint8_t *q, int32_t Q, const uint8_t *r, int32_t R, const int8_t *S,
                        int32_t A, int32_t go, int32_t ge) {
    SwQuery<W> qy;
    qy.prepare(q, Q, S, A, go, ge);
    SwWorkAll<W> wk;
    return sw_striped_run<W>(qy, r, R, wk);
}

/* Threaded batch over targets with the query state built once, at lane
 * count W. */
template <int W>
int32_t striped_batch(int32_t algo, const uint8_t *q, int32_t Q, const uint8_t *res,
                      const uint64_t *off, int64_t n, const int8_t *S, int32_t A, int32_t go,
                      int32_t ge, uint32_t flags, int32_t *score, int32_t *end_ref,
                      int32_t *end_query, int32_t nthreads) {
    SgQuery<W> sgq;
    SwQuery<W> swq;
    if (algo == 0) sgq.prepare(q, Q, S, A, go, ge, flags);
    else swq.prepare(q, Q, S, A, go, ge);
    std::atomic<int64_t> next(0);
    auto worker = [&]() {
        SgWork<W> sgw;
        SwWorkAll<W> sww;
        for (;;) {
            const int64_t lo = next.fetch_add(64);
            if (lo >= n) break;
            const int64_t hi = std::min<int64_t>(n, lo + 64);
            for (int64_t t = lo; t < hi; ++t) {
                const uint8_t *r = res + off[t];
                const int32_t R = (int32_t)(off[t + 1] - off[t]);
                const oracle_end_t e = algo == 0 ? sg_striped_run<W>(sgq, r, R, sgw)
                                                 : sw_striped_run<W>(swq, r, R, sww);
                score[t] = e.score;
                end_ref[t] = e.end_ref;
                end_query[t] = e.end_query;
            }
        }
    };
    std::vector<std::thread> th;
    for (int32_t k = 0; k < nthreads; ++k) th.emplace_back(worker);
    for (auto &x : th) x.join();
    return nthreads;
}

} // namespace

extern "C" int32_t oracle_batch_striped(int32_t algo, const uint8_t *q, int32_t Q,
                                        const uint8_t *residues, const uint64_t *offsets,
                                        int64_t n, const int8_t *S, int32_t A, int32_t go,
                                        int32_t ge, uint32_t flags, int32_t *score,
                                        int32_t *end_ref, int32_t *end_query, int32_t nthreads) {
    if (nthreads <= 0) nthreads = (int32_t)std::thread::hardware_concurrency();
    nthreads = std::max(1, std::min(nthreads, 256));
    return striped_batch<kNativeW>(algo, q, Q, residues, offsets, n, S, A, go, ge, flags, score,
                                   end_ref, end_query, nthreads);
}

extern "C" oracle_end_t oracle_sg_striped(const uint8_t *q, int32_t Q,
                                          const uint8_t *r, int32_t R,
                                          const int8_t *S, int32_t A,
                                          int32_t go, int32_t ge,
                                          uint32_t flags, int32_t W) {
    switch (W) {
    case 8: return sg_striped<8>(q, Q, r, R, S, A, go, ge, flags);
    case 32: return sg_striped<32>(q, Q, r, R, S, A, go, ge, flags);
    case 16: return sg_striped<16>(q, Q, r, R, S, A, go, ge, flags);
    default: return sg_striped<kNativeW>(q, Q, r, R, S, A, go, ge, flags);
    }
}

extern "C" oracle_end_t oracle_sw_striped(const uint8_t *q, int32_t Q,
                                          const uint8_t *r, int32_t R,
                                          const int8_t *S, int32_t A,
                                          int32_t go, int32_t ge, int32_t W) {
    switch (W) {
    case 8: return sw_striped<8>(q, Q, r, R, S, A, go, ge);
    case 32: return sw_striped<32>(q, Q, r, R, S, A, go, ge);
    case 16: return sw_striped<16>(q, Q, r, R, S, A, go, ge);
    default: return sw_striped<kNativeW>(q, Q, r, R, S, A, go, ge);
    }
}

/* Test hooks: the lane operations and the profile layout above, on their
 * own, so that tests can check them against the reference's vectors for
 * saturating_add / saturating_sub (tests/matcher/alignment/
 * test_striped_utils.mojo:9-66) and for the ssw profile
 * (tests/test_local_and_profile.mojo:19-152). The overloads called are the
 * ones the kernels above resolve to at each lane count (AVX2 / AVX-512 at
 * their native widths, the generic loops elsewhere). */
namespace {
template <typename T, int W>
int32_t sat_lanes(int32_t op, const int32_t *a, const int32_t *b, int32_t *r) {
    Vec<T, W> x = vset<T, W>(0), y = vset<T, W>(0);
    for (int k = 0; k < W; ++k) {
        lset(x, k, (T)a[k]);
        lset(y, k, (T)b[k]);
    }
    const Vec<T, W> z = op == 0 ? vadds(x, y) : vsubs(x, y);
    for (int k = 0; k < W; ++k) r[k] = (int32_t)lget(z, k);
    return 0;
}
template <typename T>
int32_t sat_lanes_w(int32_t width, int32_t op, const int32_t *a, const int32_t *b, int32_t *r) {
    switch (width) {
    case 8: return sat_lanes<T, 8>(op, a, b, r);
    case 16: return sat_lanes<T, 16>(op, a, b, r);
    case 32: return sat_lanes<T, 32>(op, a, b, r);
    case 64: return sat_lanes<T, 64>(op, a, b, r);
    default: return -1;
    }
}
template <typename T, int W>
int32_t profile_lanes(const uint8_t *q, int32_t Q, const int8_t *S, int32_t A, int32_t bias, int32_t *out) {
    std::vector<Vec<T, W>> P;
    int32_t seg = 0;
    make_profile<T, W>(q, Q, S, A, bias, P, seg);
    for (size_t v = 0; v < P.size(); ++v)
        for (int k = 0; k < W; ++k) out[v * W + (size_t)k] = (int32_t)lget(P[v], k);
    return seg;
}
} // namespace

extern "C" int32_t oracle_sat_lanes(int32_t type, int32_t width, int32_t op, const int32_t *a,
                                    const int32_t *b, int32_t *r) {
    switch (type) {
    case 0: return sat_lanes_w<uint8_t>(width, op, a, b, r);
    case 1: return sat_lanes_w<int16_t>(width, op, a, b, r);
    case 2: return sat_lanes_w<uint16_t>(width, op, a, b, r);
    default: return -1;
    }
}

extern "C" int32_t oracle_local_profile(const uint8_t *q, int32_t Q, const int8_t *S, int32_t A,
                                        int32_t bias, int32_t type, int32_t width, int32_t *out) {
    if (type == 0 && width == 16) return profile_lanes<uint8_t, 16>(q, Q, S, A, bias, out);
    if (type == 0 && width == 64) return profile_lanes<uint8_t, 64>(q, Q, S, A, bias, out);
    if (type == 0 && width == 32) return profile_lanes<uint8_t, 32>(q, Q, S, A, bias, out);
    if (type == 2 && width == 8) return profile_lanes<uint16_t, 8>(q, Q, S, A, bias, out);
    if (type == 2 && width == 16) return profile_lanes<uint16_t, 16>(q, Q, S, A, bias, out);
    if (type == 2 && width == 32) return profile_lanes<uint16_t, 32>(q, Q, S, A, bias, out);
    return -1;
}
