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
#include <vector>

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
                P[(size_t)c * seg + s].v[k] = (T)val;
            }
}

template <int W>
oracle_end_t sg_striped(const uint8_t *q, int32_t Q, const uint8_t *r,
                        int32_t R, const int8_t *S, int32_t A, int32_t go,
                        int32_t ge, uint32_t flags) {
    typedef int16_t T;
    typedef Vec<T, W> V;
    const int32_t MIN = Lim<T>::MIN, MAX = Lim<T>::MAX;
    oracle_end_t res = {0, 0, 0};
    if (Q == 0 || R == 0) return res;
    const bool fqs = flags & ORACLE_QS, fqe = flags & ORACLE_QE;
    const bool fts = flags & ORACLE_TS, fte = flags & ORACLE_TE;

    int32_t min_score = 0, max_score = 0;
    for (int64_t k = 0; k < (int64_t)A * A; ++k) {
        min_score = std::min<int32_t>(min_score, S[k]);
        max_score = std::max<int32_t>(max_score, S[k]);
    }
    std::vector<V> P;
    int32_t seg;
    make_profile<T, W>(q, Q, S, A, 0, P, seg);

    int32_t end_query = Q - 1, end_reference = -1;
    const int32_t offset = (Q - 1) % seg;
    const int32_t position = (W - 1) - (Q - 1) / seg;
    const int pos_lane = W - 1 - position; /* lane holding row Q-1 */

    std::vector<V> h_store(seg), h_load(seg), pv_e(seg);
    std::vector<T> boundary((size_t)R + 1);
    for (int32_t s = 0; s < seg; ++s) {
        V h, e;
        for (int k = 0; k < W; ++k) {
            int32_t tmp = fqs ? 0 : (int32_t)(int16_t)(-go - (int16_t)(ge * (k * seg + s)));
            h.v[k] = (T)(tmp < MIN ? MIN : tmp);
            tmp = (int32_t)(int16_t)(tmp - go);
            e.v[k] = (T)(tmp < MIN ? MIN : tmp);
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
        v_h.v[0] = boundary[i];
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
            v_f.v[0] = tmp2;
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
        bool cmp = v_h.v[pos_lane] > v_max_h.v[pos_lane];
        v_max_h = vmax(v_h, v_max_h);
        if (cmp) end_reference = i;
    }
    if (fte) {
        score = v_max_h.v[pos_lane];
        end_query = Q - 1;
    }
    if (fqe) {
        for (int32_t s = 0; s < seg; ++s)
            for (int k = 0; k < W; ++k) {
                int32_t temp = s + k * seg;
                if (temp >= Q) continue;
                T hv = h_store[s].v[k];
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
        score = h_store[offset].v[pos_lane];
        end_reference = R - 1;
        end_query = Q - 1;
    }
    (void)MAX;
    res.score = score;
    res.end_ref = end_reference;
    res.end_query = end_query;
    return res;
}

/* sw[dt, width] (local_aln/striped.mojo:272-434), forward direction,
 * terminate disabled. Returns the best with overflow flagged by
 * best+bias >= MAX (the caller reruns wider). */
template <typename T, int W>
oracle_end_t sw_striped_t(const uint8_t *q, int32_t Q, const uint8_t *r,
                          int32_t R, const int8_t *S, int32_t A, int32_t go,
                          int32_t ge, int32_t bias, bool *overflow) {
    typedef Vec<T, W> V;
    const int32_t MAXV = Lim<T>::MAX;
    std::vector<V> P;
    int32_t seg;
    make_profile<T, W>(q, Q, S, A, bias, P, seg);
    std::vector<V> h_store(seg, vset<T, W>(0)), h_load(seg, vset<T, W>(0)),
        pv_e(seg, vset<T, W>(0)), h_max(seg, vset<T, W>(0));
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
        bool all_eq = true;
        for (int k = 0; k < W; ++k) all_eq &= v_max_mark.v[k] == v_max_score.v[k];
        if (!all_eq) {
            int32_t temp = 0;
            for (int k = 0; k < W; ++k) temp = std::max<int32_t>(temp, v_max_score.v[k]);
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
            if ((int32_t)h_max[s].v[k] == max_score) {
                int32_t temp = s + k * seg;
                if (temp < end_query) end_query = temp;
            }
    oracle_end_t res = {max_score, end_reference, end_query};
    return res;
}

template <int W>
oracle_end_t sw_striped(const uint8_t *q, int32_t Q, const uint8_t *r,
                        int32_t R, const int8_t *S, int32_t A, int32_t go,
                        int32_t ge) {
    int32_t bias = 0;
    for (int64_t k = 0; k < (int64_t)A * A; ++k) bias = std::min<int32_t>(bias, S[k]);
    bias = -bias;
    bool ov = false;
    oracle_end_t e = sw_striped_t<uint8_t, 2 * W>(q, Q, r, R, S, A, go, ge, bias, &ov);
    if (e.score + bias >= 255)
        e = sw_striped_t<uint16_t, W>(q, Q, r, R, S, A, go, ge, bias, &ov);
    return e;
}

} // namespace

extern "C" oracle_end_t oracle_sg_striped(const uint8_t *q, int32_t Q,
                                          const uint8_t *r, int32_t R,
                                          const int8_t *S, int32_t A,
                                          int32_t go, int32_t ge,
                                          uint32_t flags, int32_t W) {
    switch (W) {
    case 8: return sg_striped<8>(q, Q, r, R, S, A, go, ge, flags);
    case 32: return sg_striped<32>(q, Q, r, R, S, A, go, ge, flags);
    default: return sg_striped<16>(q, Q, r, R, S, A, go, ge, flags);
    }
}

extern "C" oracle_end_t oracle_sw_striped(const uint8_t *q, int32_t Q,
                                          const uint8_t *r, int32_t R,
                                          const int8_t *S, int32_t A,
                                          int32_t go, int32_t ge, int32_t W) {
    switch (W) {
    case 8: return sw_striped<8>(q, Q, r, R, S, A, go, ge);
    case 32: return sw_striped<32>(q, Q, r, R, S, A, go, ge);
    default: return sw_striped<16>(q, Q, r, R, S, A, go, ge);
    }
}
