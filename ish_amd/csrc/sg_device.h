// Semi-global (Gotoh, configurable free ends) scoring for gfx950: the
// device code shared by the kernels of sg_batched.hip, sg_hybrid.hip,
// sg_lock.hip, sg_multi.hip and sg_misc.hip.
//
// Semantics: semi_global_aln of the reference's CPU striped path
// (ishlib/matcher/alignment/semi_global_aln/striped.mojo:252-498):
//   H(-1,-1)=0, H(j,-1)=qs?0:-go-ge*j, H(-1,i)=ts?0:-go-ge*i,
//   E(j,i)=max(E(j,i-1)-ge, H(j,i-1)-go), F(j,i)=max(F(j-1,i)-ge, H(j-1,i)-go),
//   H(j,i)=max(H(j-1,i-1)+S[r_i][q_j], E, F);
//   end rules: first column of the strict last-row maximum seeded at the
//   int16 floor (:436-441, :446-450), smallest row of the strict last-column
//   maximum with the end_ref == R-1 tie rule (:453-468), corner otherwise
//   (:471-477).
//
// Layout of the work: one lane owns whole targets and a chunk of K query
// rows in registers; it streams the target's residues column by column.
// The bottom row of a chunk (H and F of every column) goes to a per-wave
// boundary buffer and is read back by the next chunk. There is no
// cross-lane traffic and no data-dependent control flow in the cell loop.
//
// Batched kernels (persistent, one launch drains a device work queue over
// length-sorted groups of 128 targets):
//   - sg_batched<short2>: two targets per lane in packed int16 halves.
//   - sg_batched<int>   : one target per lane in int32, for groups whose
//                         proven score range does not fit int16.
// Both use the offset form of the recurrence, H^ = H + (i+j)*ge (likewise
// E^, F^), which removes the two gap-extension subtractions:
//   G  = H^ + (ge-go)
//   E^ = max(E^, G)          (next column)
//   F^ = max(F^, G)          (next row)
//   H^ = max(H^diag + (S+2ge), E^, F^)
// i.e. 6 integer ops per cell instead of 8.
//
// Plain kernel (one launch per batch, the comparison baseline): one target
// per lane in input order, int32, textbook 8-op recurrence, residues read
// straight from the caller's contiguous layout.
#pragma once

#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include <type_traits>

#include "engine.h"
#include "kernel_common.h"

namespace ish {

// Running (value, first index) maximum as one ordered key per target:
// value in the high 16 bits (signed), 0xFFFF - index in the low 16 bits, so
// the integer maximum picks the largest value and, among equals, the
// smallest index. Valid while values fit int16 and indices < 65536, which
// the int16 range bound guarantees for every group routed to this kernel.
__device__ __forceinline__ int key_lo(uint32_t packed, int idx) {
    return (int)((packed << 16) | (uint32_t)(0xFFFF - idx));
}
__device__ __forceinline__ int key_hi(uint32_t packed, int idx) {
    return (int)((packed & 0xFFFF0000u) | (uint32_t)(0xFFFF - idx));
}

struct TrackI16 {
    int te_a, te_b, qe_a, qe_b, corner_a, corner_b;
};

// Int32 tracking keeps (value, index) pairs explicitly.
struct TrackI32 {
    int te_v, te_i, qe_v, qe_j, corner;
};

// Final end-coordinate rules shared by every kernel (values already
// converted back from the offset form).
__device__ __forceinline__ void finalize(int R, int Q, bool te, bool qe, bool has_track,
                                         int track_val, int track_idx, bool has_qe, int qe_val,
                                         int qe_j, int corner, int neg_limit, int &score,
                                         int &end_ref, int &end_q) {
    if (R == 0 || Q == 0) {
        score = 0;
        end_ref = 0;
        end_q = 0;
        return;
    }
    // The last-row running maximum (seeded at the int16 floor) always runs;
    // it gives end_ref, and the score when the target end is free.
    score = neg_limit;
    end_ref = -1;
    end_q = Q - 1;
    if (has_track && track_val > neg_limit) {
        end_ref = track_idx;
        if (te) score = track_val;
    }
    if (qe && has_qe) {
        if (qe_val > score) {
            score = qe_val;
            end_q = qe_j;
            end_ref = R - 1;
        } else if (qe_val == score && end_ref == R - 1 && qe_j < end_q) {
            end_q = qe_j;
        }
    }
    if (!te && !qe) {
        score = corner;
        end_ref = R - 1;
        end_q = Q - 1;
    }
}

// ---------------------------------------------------------------------------
// Batched persistent kernel.
//   V = s16x2 : lane owns targets (first + lane) and (first + 64 + lane)
//   V = int   : work item = (group, half); lane owns one target
// ---------------------------------------------------------------------------
// Per-query values a group pass needs (one query, or one of many).
struct QueryView {
    int Q, nchunks, neg_limit;
    int pad, front;  // top padding rows, profile entries before row 0
    uint32_t flags;  // free ends of this query (qs|qe<<1|ts<<2|te<<3)
    int64_t out_off; // offset of this query's results in the output arrays
    int fshift;      // binary16 lanes: the shift of every value (read by those only)
};

// State of one lane while it scores one query chunk (K rows) of its
// target(s).
template <typename V, int K> struct SgChunk {
    V H[K], E[K];
    V diag_top; // H^ of the row above the chunk, previous column
};

// H^(j,-1), the left boundary of row j. Rows j < 0 are the top padding
// rows of the first chunk: with go >= ge and a profile entry of ge (target
// start free) or 0 per column, row j < 0 holds H^(j,i) = (i-1)*ge, resp.
// -go-ge, in every column i >= -1 (E^ and F^ never exceed it), which is the
// top boundary H^(-1,i) the query rows need; row -1's own left entry is the
// corner H^(-1,-1) = -2ge.
__device__ __forceinline__ int left_boundary(int j, bool qs, bool ts, int go, int ge) {
    if (j >= 0) return qs ? (j - 1) * ge : -go - ge;
    if (j == -1) return -2 * ge;
    return ts ? -2 * ge : -go - ge;
}

// shift: binary16 lanes hold every value minus shift (0 for the integer
// lanes).
template <typename V, int K>
__device__ __forceinline__ void sg_chunk_init(SgChunk<V, K> &st, int j0, bool qs, bool ts, int go,
                                              int ge, int shift = 0) {
    const int cst = ge - go;
#pragma unroll
    for (int m = 0; m < K; ++m) {
        const int hl = left_boundary(j0 + m, qs, ts, go, ge) - shift; // H^(j,-1)
        st.H[m] = splat_uniform<V>(hl);
        st.E[m] = splat_uniform<V>(hl + cst); // E^(j,0)
    }
    st.diag_top = splat_uniform<V>(left_boundary(j0 - 1, qs, ts, go, ge) - shift);
}

__device__ __forceinline__ void track_init(TrackI16 &t16, TrackI32 &t32) {
    t16.te_a = t16.te_b = t16.qe_a = t16.qe_b = INT_MIN;
    t16.corner_a = t16.corner_b = 0;
    t32.te_v = INT_MIN;
    t32.te_i = -1;
    t32.qe_v = INT_MIN;
    t32.qe_j = 0;
    t32.corner = 0;
}

// Scores columns 4b .. 4b+3 of one chunk (rows j0 .. j0+K-1; rows below 0
// are top padding rows).
//   rr / rr_next: residue words of this block and the next (the profile
//   slices of the next column are fetched while a column is scored);
//   bn: the chunk-top rows (H^, F^ of the row above) unless `first`;
//   out: where the chunk-bottom rows go (t*64 + lane), null for the last
//   chunk, which tracks row Q-1 instead: register mstar, which the packed
//   variants take to be K-1 (their queries are top padded). DEFER (out in
//   global memory): the four columns' rows are stored together after the
//   block, so that the wait for the next block's residues (in the last
//   column) finds the previous block's stores long done.
// TE / QE say which end rules the variant can apply; qe_rt is the query's.
// rmin: the shortest target length of the wave (wave-uniform); columns
// below it lie inside every lane's targets, so their end tracking needs
// no per-lane test.
// PAIRED (packed lanes): chunk_base is the wave's paired chunk profile and
// slut the pair table (see PairView).
template <typename V, int K, bool TE, bool QE, bool PAIRED = false, bool DEFER = false>
__device__ __forceinline__ void sg_block(SgChunk<V, K> &st, int b, int j0, bool first, bool last,
                                         int mstar, uint2 rr, uint2 rr_next, const uint2 (&bn)[4],
                                         short8 (&sa)[K / 8], short8 (&sb)[K / 8],
                                         const uint8_t *chunk_base, const uint32_t *slut,
                                         int half, int RA, int RB, int rmin, bool ts, bool qe_rt, int go,
                                         int ge, TrackI16 &t16, TrackI32 &t32, uint2 *out,
                                         int lane, int shift = 0) {
    constexpr bool PACKED = sizeof(V) == 4 && !__is_same(V, int);
    constexpr bool F16 = __is_same(V, h16x2);
    static_assert(!F16 || PAIRED, "binary16 lanes read the paired profile");
    constexpr bool FIXED_LAST = PACKED;
    const int cst = ge - go;
    const V vcst = splat<V>(cst);
    auto row_ptr = [&](uint32_t code) -> const short8 * {
        return (const short8 *)(chunk_base + 4u * (uint32_t)slut[code]);
    };
    uint2 bottom[4]; // chunk-bottom rows (H^, F^) of the block's columns
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int i = b * 4 + t;
        // residue codes of the next column (past the end: any code)
        const uint32_t wa = t < 3 ? (half ? rr.y : rr.x) : (half ? rr_next.y : rr_next.x);
        const uint32_t wb = t < 3 ? rr.y : rr_next.y;
        const int sh = 8 * ((t + 1) & 3);
        const short8 *na, *nb;
        if constexpr (PAIRED) {
            na = pair_row(chunk_base, slut, (wa >> sh) & 0xffu, (wb >> sh) & 0xffu);
            nb = na;
        } else {
            na = row_ptr((wa >> sh) & 0xffu);
            nb = row_ptr((wb >> sh) & 0xffu);
        }
        V htop, f;
        if (first) {
            const int ht = (ts ? (i - 1) * ge : -go - ge) - shift; // H^(-1,i)
            htop = splat<V>(ht);
            f = splat<V>(ht + cst); // F^(0,i)
        } else {
            htop = from_u32<V>(bn[t].x);
            f = from_u32<V>(bn[t].y);
        }
        // Rows run software-pipelined: row m's F chain (h = max(t2, f),
        // gg = h + c, f = max(f, gg)) alternates with work that does not
        // depend on it (t1 and t2 of row m+2, the E update of row m), so no
        // packed operation reads the result of the one just before it,
        // which would cost a wait state. The scheduling barriers keep this
        // order.
        if constexpr (F16) {
            // Binary16 lanes: H = max3(H^diag + S', E^, F^) is one
            // operation, so a row is five: t1 = H^diag + S' (two rows
            // ahead), h = max3(t1, E, f), gg = h + c, f = max(f, gg),
            // E = max(E, gg). The F chain (max3, add, max) alternates with
            // t1 of row m+2 and the E update of row m.
            V t1[K];
            t1[0] = st.diag_top + slice_val<V, PAIRED, PACKED>(sa[0], sb[0], 0);
            t1[1] = st.H[0] + slice_val<V, PAIRED, PACKED>(sa[0], sb[0], 1);
#pragma unroll
            for (int m = 0; m < K; ++m) {
                const V h = vmax3(t1[m], st.E[m], f);
                __builtin_amdgcn_sched_barrier(0);
                // row m-1's E update stays in its slot (not sunk into the
                // next column, where its operands would stay live)
                if (m > 0) pin(st.E[m - 1]);
                if (m + 2 < K)
                    t1[m + 2] = st.H[m + 1] + slice_val<V, PAIRED, PACKED>(sa[(m + 2) / 8], sb[(m + 2) / 8],
                                                                           (m + 2) % 8);
                __builtin_amdgcn_sched_barrier(0);
                const V gg = h + vcst;
                __builtin_amdgcn_sched_barrier(0);
                f = vmax(f, gg);
                __builtin_amdgcn_sched_barrier(0);
                st.E[m] = vmax(st.E[m], gg);
                st.H[m] = h;
                __builtin_amdgcn_sched_barrier(0);
                if (m % 8 == 7) {
                    const int k = m / 8;
                    sa[k] = na[2 * k];
                    sb[k] = na[2 * k + 1];
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        } else {
        V t2[K]; // max(H^diag + S', E^) of the rows ahead
        t2[0] = vmax(st.diag_top + slice_val<V, PAIRED, PACKED>(sa[0], sb[0], 0), st.E[0]);
        t2[1] = vmax(st.H[0] + slice_val<V, PAIRED, PACKED>(sa[0], sb[0], 1), st.E[1]);
#pragma unroll
        for (int m = 0; m < K; ++m) {
            const V h = vmax(t2[m], f);
            __builtin_amdgcn_sched_barrier(0);
            // row m-1's E update stays in its slot (not sunk into the next
            // column); pinned one operation after it is written
            if (m > 0) pin(st.E[m - 1]);
            V t1;
            if (m + 2 < K) // H^diag of row m+2 is row m+1's value in the previous column
                t1 = st.H[m + 1] + slice_val<V, PAIRED, PACKED>(sa[(m + 2) / 8], sb[(m + 2) / 8], (m + 2) % 8);
            __builtin_amdgcn_sched_barrier(0);
            const V gg = h + vcst;
            __builtin_amdgcn_sched_barrier(0);
            if (m + 2 < K) t2[m + 2] = vmax(t1, st.E[m + 2]);
            __builtin_amdgcn_sched_barrier(0);
            f = vmax(f, gg);
            __builtin_amdgcn_sched_barrier(0);
            st.E[m] = vmax(st.E[m], gg);
            st.H[m] = h;
            __builtin_amdgcn_sched_barrier(0);
            if (m % 8 == 7) {
                // slice m/8 is consumed (its last row's t1 was taken two rows
                // ago): fetch it for the next column, after the rows that
                // read it, so only one extra slice is live at a time
                const int k = m / 8;
                if constexpr (PAIRED) {
                    sa[k] = na[2 * k];
                    sb[k] = na[2 * k + 1];
                } else {
                    sa[k] = na[k];
                    if (PACKED) sb[k] = nb[k];
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        }
        st.diag_top = htop;
        if (!last) {
            bottom[t].x = as_u32(st.H[K - 1]);
            bottom[t].y = as_u32(f);
            if constexpr (!DEFER) out[t * 64 + lane] = bottom[t];
        } else {
            // row Q-1 lives in register mstar of the last chunk
            V hl = st.H[K - 1];
            if constexpr (!FIXED_LAST) {
                hl = st.H[0];
#pragma unroll
                for (int m = 1; m < K; ++m)
                    if (m == mstar) hl = st.H[m];
            }
            // u = H^(Q-1,i) - i*ge = H(Q-1,i) + (Q-1)*ge; binary16 lanes go
            // back to int16 first (i*ge alone may leave the exact range)
            V u{};
            if constexpr (!F16) u = hl - splat<V>(i * ge);
            if constexpr (PACKED) {
                const uint32_t uu = F16 ? as_u32(to_s16x2(hl, shift) - splat<s16x2>(i * ge)) : as_u32(u);
                if (TE || QE) {
                    int ka = key_lo(uu, i), kb = key_hi(uu, i);
                    if (i >= rmin) { // past some lane's target end
                        // a real (scalar) branch: the empty asm keeps the
                        // compiler from selecting on every column instead
                        asm volatile("");
                        if (i >= RA) ka = INT_MIN;
                        if (i >= RB) kb = INT_MIN;
                    }
                    t16.te_a = max(t16.te_a, ka);
                    t16.te_b = max(t16.te_b, kb);
                } else if (i >= rmin - 1) {
                    if (i == RA - 1) t16.corner_a = (int)(short)(uu & 0xffffu);
                    if (i == RB - 1) t16.corner_b = (int)(short)(uu >> 16);
                }
            } else {
                const int uv = (int)as_u32(u);
                if (TE || QE) {
                    if (i < RA && uv > t32.te_v) {
                        t32.te_v = uv;
                        t32.te_i = i;
                    }
                } else if (i >= rmin - 1 && i == RA - 1) {
                    t32.corner = uv;
                }
            }
        }
        if (qe_rt) {
            const bool cap_a = i == RA - 1;
            const bool cap_b = PACKED && i == RB - 1;
            // wave-uniform test keeps this once-per-target path a real
            // branch instead of predicated code on every column
            if (__builtin_amdgcn_ballot_w64(cap_a || cap_b) != 0) {
#pragma unroll
                for (int m = 0; m < K; ++m) {
                    const int j = j0 + m;
                    if ((last && m > mstar) || j < 0) continue; // past Q-1, padding
                    // w = H^(j,R-1) - j*ge = H(j,R-1) + (R-1)*ge
                    V wv{};
                    if constexpr (!F16) wv = st.H[m] - splat<V>(j * ge);
                    if constexpr (PACKED) {
                        const uint32_t ww =
                            F16 ? as_u32(to_s16x2(st.H[m], shift) - splat<s16x2>(j * ge)) : as_u32(wv);
                        if (cap_a) t16.qe_a = max(t16.qe_a, key_lo(ww, j));
                        if (cap_b) t16.qe_b = max(t16.qe_b, key_hi(ww, j));
                    } else {
                        const int wi = (int)as_u32(wv);
                        if (cap_a && wi > t32.qe_v) {
                            t32.qe_v = wi;
                            t32.qe_j = j;
                        }
                    }
                }
            }
        }
    }
    if (DEFER && !last) {
#pragma unroll
        for (int t = 0; t < 4; ++t) out[t * 64 + lane] = bottom[t];
    }
}

// Final end rules and the result writes of one lane's target(s).
template <typename V>
__device__ __forceinline__ void sg_finish(const ScoreParams &p, const QueryView &qv, int tA, int tB,
                                          int RA, int RB, bool te_rt, bool qe_rt,
                                          const TrackI16 &t16, const TrackI32 &t32) {
    constexpr bool PACKED = sizeof(V) == 4 && !__is_same(V, int);
    const int Q = qv.Q, ge = p.ge;
    const int qoff = (Q - 1) * ge;
    if constexpr (PACKED) {
        // once per group: kept rolled, so that it adds few registers where
        // the chunk state is live
#pragma unroll 1
        for (int which = 0; which < 2; ++which) {
            const int t = which ? tB : tA;
            const int R = which ? RB : RA;
            const int o = p.orig[t];
            if (o < 0) continue;
            const int kte = which ? t16.te_b : t16.te_a;
            const int kqe = which ? t16.qe_b : t16.qe_a;
            const int cor = which ? t16.corner_b : t16.corner_a;
            const int te_val = (kte >> 16) - qoff;
            const int te_idx = 0xFFFF - (kte & 0xFFFF);
            const int qe_val = (kqe >> 16) - (R - 1) * ge;
            const int qe_j = 0xFFFF - (kqe & 0xFFFF);
            int sc, er, eq;
            finalize(R, Q, te_rt, qe_rt, kte != INT_MIN, te_val, te_idx, kqe != INT_MIN, qe_val,
                     qe_j, cor - qoff, qv.neg_limit, sc, er, eq);
            p.score[qv.out_off + o] = sc;
            p.end_ref[qv.out_off + o] = er;
            p.end_q[qv.out_off + o] = eq;
            p.status[qv.out_off + o] = __is_same(V, h16x2) ? 16 | 0x40 : 16;
        }
    } else {
        const int o = p.orig[tA];
        if (o >= 0) {
            int sc, er, eq;
            finalize(RA, Q, te_rt, qe_rt, t32.te_i >= 0, t32.te_v - qoff, t32.te_i,
                     t32.qe_v != INT_MIN, t32.qe_v - (RA - 1) * ge, t32.qe_j,
                     t32.corner - qoff, qv.neg_limit, sc, er, eq);
            p.score[qv.out_off + o] = sc;
            p.end_ref[qv.out_off + o] = er;
            p.end_q[qv.out_off + o] = eq;
            p.status[qv.out_off + o] = 32 | 0x80;
        }
    }
}

// One query chunk c (rows j0 = c*K - pad .. j0+K-1) of a group: every
// column of the lane's target(s). chunk_base / slut: the chunk's profile
// rows and their table (int16 profile: smem + 2*(front + j0) and the residue
// lut; PAIRED: the chunk's pair rows and the pair table).
// MID: c is neither the query's first nor its last chunk, known at
// compile time, so that the cell loop of these chunks (most of a long
// query's) carries no run-time tests of either.
template <typename V, int K, bool TE, bool QE, bool PAIRED, bool MID>
__device__ __forceinline__ void sg_chunk_v(const ScoreParams &p, const QueryView &qv,
                                           const uint8_t *chunk_base, const uint32_t *slut,
                                           uint2 *bnd, const uint2 *res, int nblk, int c, int half,
                                           int lane, int RA, int RB, TrackI16 &t16, TrackI32 &t32) {
    constexpr bool PACKED = sizeof(V) == 4 && !__is_same(V, int);
    constexpr int NS = K / 8; // profile slices of 8 rows (one 16-byte LDS read each)
    const int go = p.go, ge = p.ge, Q = qv.Q;
    const bool qs = qv.flags & 1u, ts = qv.flags & 4u;
    const bool qe_rt = QE && (qv.flags & 2u);
    const int j0 = c * K - qv.pad;
    const bool first = !MID && c == 0;
    const bool last = !MID && c == qv.nchunks - 1;
    const int mstar = Q - 1 - j0; // row Q-1 inside the last chunk
    const int rmin = wave_min(PACKED ? min(RA, RB) : RA);
    const int shift = __is_same(V, h16x2) ? qv.fshift : 0;
    SgChunk<V, K> st;
    sg_chunk_init(st, j0, qs, ts, go, ge, shift);
    uint2 rr_next = res[lane];
    uint2 bn_next[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) bn_next[t] = make_uint2(0u, 0u);
    if (!first) {
#pragma unroll
        for (int t = 0; t < 4; ++t) bn_next[t] = bnd[t * 64 + lane];
    }
    // profile slices of column 0
    short8 sa[NS], sb[NS];
    if constexpr (PAIRED) {
        const short8 *np = pair_row(chunk_base, slut, rr_next.x & 0xffu, rr_next.y & 0xffu);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            sa[k] = np[2 * k];
            sb[k] = np[2 * k + 1];
        }
    } else {
        const short8 *na = (const short8 *)(chunk_base + 4u * slut[(half ? rr_next.y : rr_next.x) & 0xffu]);
        const short8 *nb = (const short8 *)(chunk_base + 4u * slut[rr_next.y & 0xffu]);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            sa[k] = na[k];
            sb[k] = PACKED ? nb[k] : na[k];
        }
    }
    for (int b = 0; b < nblk; ++b) {
        const uint2 rr = rr_next;
        uint2 bn[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) bn[t] = bn_next[t];
        if (b + 1 < nblk) {
            const uint2 *rb = res + (size_t)(b + 1) * 64; // uniform base, lane offset
            rr_next = rb[lane];
            if (!first) {
                const uint2 *bb = bnd + (size_t)(b + 1) * 4 * 64;
#pragma unroll
                for (int t = 0; t < 4; ++t) bn_next[t] = bb[t * 64 + lane];
            }
        }
        sg_block<V, K, TE, QE, PAIRED, true>(st, b, j0, first, last, mstar, rr, rr_next, bn, sa,
                                             sb, chunk_base, slut, half, RA, RB, rmin, ts, qe_rt, go, ge,
                                             t16, t32, last ? nullptr : bnd + (size_t)b * 4 * 64,
                                             lane, shift);
    }
}

template <typename V, int K, bool TE, bool QE, bool PAIRED>
__device__ __forceinline__ void sg_chunk(const ScoreParams &p, const QueryView &qv,
                                         const uint8_t *chunk_base, const uint32_t *slut,
                                         uint2 *bnd, const uint2 *res, int nblk, int c, int half,
                                         int lane, int RA, int RB, TrackI16 &t16, TrackI32 &t32) {
    if (c > 0 && c < qv.nchunks - 1)
        sg_chunk_v<V, K, TE, QE, PAIRED, true>(p, qv, chunk_base, slut, bnd, res, nblk, c, half,
                                               lane, RA, RB, t16, t32);
    else
        sg_chunk_v<V, K, TE, QE, PAIRED, false>(p, qv, chunk_base, slut, bnd, res, nblk, c, half,
                                                lane, RA, RB, t16, t32);
}

// Scores one group (PACKED: 128 targets, two per lane) or one half group
// (int32: 64 targets) against the query whose profile is staged in smem,
// one wave through every query chunk in turn.
// TE / QE say which end rules the variant can apply; the query's own flags
// select among them at run time (a variant with TE or QE must only see
// queries that have te or qe free, since it tracks no corner cell).
// PAIRED: the profile is the paired one of pv (the wave copies each
// chunk's rows into its own buffer), slut the pair table.
template <typename V, int K, bool TE, bool QE, bool PAIRED = false>
__device__ __forceinline__ void sg_group(const ScoreParams &p, const QueryView &qv,
                                         const uint8_t *smem, const uint32_t *slut, uint2 *bnd,
                                         int g, int half, int lane, const PairView *pv = nullptr) {
    constexpr bool PACKED = sizeof(V) == 4 && !__is_same(V, int);
    const bool te_rt = TE && (qv.flags & 8u), qe_rt = QE && (qv.flags & 2u);
    const GroupDesc gd = p.groups[g];
    const int tA = gd.first + half * 64 + lane;
    const int tB = gd.first + 64 + lane;
    const int RA = p.lens[tA];
    const int RB = PACKED ? p.lens[tB] : 0;
    const uint2 *res = p.res + gd.res_off;
    const int nblk = gd.ncols >> 2;

    TrackI16 t16;
    TrackI32 t32;
    track_init(t16, t32);
    for (int c = 0; c < qv.nchunks; ++c) {
        const int e0 = qv.front + c * K - qv.pad; // profile entry of the chunk's first row
        // LDS address of this chunk's rows in profile row `code`
        const uint8_t *chunk_base = PAIRED ? pv->pbuf : smem + 2 * e0;
        if constexpr (PAIRED) stage_pairs<K, __is_same(V, h16x2)>(*pv, e0, lane);
        // one code path for every chunk: a group per wave runs beside
        // pipelined groups (sg_hybrid), whose code shares the instruction
        // cache
        sg_chunk_v<V, K, TE, QE, PAIRED, false>(p, qv, chunk_base, slut, bnd, res, nblk, c, half,
                                                lane, RA, RB, t16, t32);
    }
    sg_finish<V>(p, qv, tA, tB, RA, RB, te_rt, qe_rt, t16, t32);
}

// A tile of groups g0 + w (one per wave of the workgroup) over a paired
// profile shared by the workgroup: the waves step through the query chunks
// together, and before each chunk the workgroup copies that chunk's rows of
// every pair row into pbuf. Groups in a tile are neighbours in length
// order, so the waves reach each barrier at nearly the same time. A wave
// whose group is past the range only joins the barriers and the copies.
// pprof / npairs / pqpad: the query's paired profile; valid: the wave has
// a group (wave-uniform).
template <int K, bool TE, bool QE, bool F16 = false>
__device__ __forceinline__ void sg_tile_lockstep(const ScoreParams &p, const QueryView &qv,
                                                 const uint32_t *ptab, uint8_t *pbuf,
                                                 const int32_t *pprof, int npairs, int pqpad,
                                                 uint2 *bnd, int g, bool valid, int lane) {
    using V = typename std::conditional<F16, h16x2, s16x2>::type;
    const bool te_rt = TE && (qv.flags & 8u), qe_rt = QE && (qv.flags & 2u);
    GroupDesc gd;
    gd.res_off = 0;
    gd.ncols = 0;
    gd.first = 0;
    if (valid) gd = p.groups[g];
    const int tA = gd.first + lane;
    const int tB = gd.first + 64 + lane;
    const int RA = valid ? p.lens[tA] : 0;
    const int RB = valid ? p.lens[tB] : 0;
    const uint2 *res = p.res + gd.res_off;
    const int nblk = gd.ncols >> 2;
    constexpr int parts = K / 4; // 16-byte parts per pair row

    TrackI16 t16;
    TrackI32 t32;
    track_init(t16, t32);
    for (int c = 0; c < qv.nchunks; ++c) {
        const int e0 = qv.front + c * K - qv.pad; // profile entry of the chunk's first row
        __syncthreads(); // every wave is done with the previous chunk's rows
        for (int x = threadIdx.x; x < npairs * parts; x += blockDim.x) {
            const int pr = x / parts, part = x % parts;
            uint4 v = *(const uint4 *)(pprof + (size_t)pr * pqpad + e0 + part * 4);
            if constexpr (F16) {
                v.x = i16x2_to_h16x2(v.x);
                v.y = i16x2_to_h16x2(v.y);
                v.z = i16x2_to_h16x2(v.z);
                v.w = i16x2_to_h16x2(v.w);
            }
            *(uint4 *)(pbuf + pr * pair_row_bytes(K) + part * 16) = v;
        }
        __syncthreads();
        if (valid)
            sg_chunk<V, K, TE, QE, true>(p, qv, pbuf, ptab, bnd, res, nblk, c, 0, lane, RA, RB, t16,
                                         t32);
    }
    if (valid) sg_finish<V>(p, qv, tA, tB, RA, RB, te_rt, qe_rt, t16, t32);
}

// ---------------------------------------------------------------------------
// Workgroup pipeline (packed int16) for the longest groups: the four waves
// of a workgroup share one group of 128 targets and split its query
// chunks, wave w taking chunks w, w+4, ... Time advances in steps of one
// 4-column block closed by a workgroup barrier; at step s wave w works on
// block s - w - r*P of its round-r chunk (P = max(blocks, 5)), one block
// behind the wave owning the chunk above. Chunk-bottom rows flow to the
// next wave through a double-buffered LDS ring, and from wave 3 to wave 0
// of the next round through a global boundary slot (written P-4 or more
// steps before it is read, overwritten 3 steps after). The end-coordinate
// keys of the waves combine by max in LDS (the last-row keys and the corner
// come from the wave that owns the last chunk). Per-column code is
// sg_block, so results are those of sg_group.
// ---------------------------------------------------------------------------
constexpr int kSgPipeWaves = 4;
constexpr size_t kSgRingBytes = (size_t)(kSgPipeWaves - 1) * 2 * 4 * 64 * sizeof(uint2);
constexpr size_t kSgCombBytes = (size_t)kSgPipeWaves * 64 * 6 * sizeof(int);

// LDS after the ring: two end-key buffers (one for sg_pipe_group; the
// group parity's for sg_pipe_stream) and the stream's control words
constexpr size_t kSgCtlBytes = 32;
constexpr size_t kSgExtraBytes = kSgRingBytes + 2 * kSgCombBytes + kSgCtlBytes;

// W: the waves of the workgroup, which share the group's query chunks.
// F16 (PAIRED only): binary16 lanes, values held minus qv.fshift.
template <int K, bool TE, bool QE, bool PAIRED = false, int W = kSgPipeWaves, bool F16 = false>
__device__ __forceinline__ void sg_pipe_group(const ScoreParams &p, const QueryView &qv,
                                              const uint8_t *smem, const uint32_t *slut,
                                              uint2 *ring, int *comb, uint2 *gbnd, int g, int w,
                                              int lane, const PairView *pv = nullptr) {
    static_assert(!F16 || PAIRED, "binary16 lanes read the paired profile");
    using V = typename std::conditional<F16, h16x2, s16x2>::type;
    const int shift = F16 ? qv.fshift : 0;
    constexpr int NS = K / 8;
    const int go = p.go, ge = p.ge, Q = qv.Q;
    const bool qs = qv.flags & 1u, ts = qv.flags & 4u;
    const bool te_rt = TE && (qv.flags & 8u), qe_rt = QE && (qv.flags & 2u);
    const GroupDesc gd = p.groups[g];
    const int tA = gd.first + lane;
    const int tB = gd.first + 64 + lane;
    const int RA = p.lens[tA];
    const int RB = p.lens[tB];
    const uint2 *res = p.res + gd.res_off;
    const int nblk = gd.ncols >> 2;
    const int nch = qv.nchunks;
    const int rounds = (nch + W - 1) / W;
    const int P = nblk > 5 ? nblk : 5;
    const int steps = (rounds - 1) * P + nblk + (W - 1);
    const int last_w = (nch - 1) % W;
    const int rmin = wave_min(min(RA, RB));

    TrackI16 t16;
    TrackI32 t32;
    track_init(t16, t32);
    SgChunk<V, K> st;
    sg_chunk_init(st, 0, qs, ts, go, ge, shift);
    uint2 rr_next = make_uint2(0u, 0u);
    uint2 bn_next[4];
    short8 sa[NS], sb[NS];
#pragma unroll
    for (int t = 0; t < 4; ++t) bn_next[t] = make_uint2(0u, 0u);
#pragma unroll
    for (int k = 0; k < NS; ++k) sa[k] = sb[k] = (short8)0;

    int b = -w, r = 0; // block and round of this wave at the current step
    for (int s = 0; s < steps; ++s) {
        const int c = W * r + w;
        if (b >= 0 && b < nblk && r < rounds && c < nch) {
            const int j0 = c * K - qv.pad;
            const bool first = c == 0;
            const bool last = c == nch - 1;
            const int mstar = Q - 1 - j0;
            const int par = (r * P + b) & 1; // ring parity of this (round, block)
            const uint8_t *chunk_base = PAIRED ? pv->pbuf : smem + 2 * (qv.front + j0);
            if (b == 0) {
                sg_chunk_init(st, j0, qs, ts, go, ge, shift);
                rr_next = res[lane];
                if (!first && w == 0) {
#pragma unroll
                    for (int t = 0; t < 4; ++t) bn_next[t] = gbnd[t * 64 + lane];
                }
                if constexpr (PAIRED) {
                    stage_pairs<K, F16>(*pv, qv.front + j0, lane);
                    const short8 *np = pair_row(chunk_base, slut, rr_next.x & 0xffu, rr_next.y & 0xffu);
#pragma unroll
                    for (int k = 0; k < NS; ++k) {
                        sa[k] = np[2 * k];
                        sb[k] = np[2 * k + 1];
                    }
                } else {
                    const short8 *na = (const short8 *)(chunk_base + 4u * (uint32_t)slut[rr_next.x & 0xffu]);
                    const short8 *nb = (const short8 *)(chunk_base + 4u * (uint32_t)slut[rr_next.y & 0xffu]);
#pragma unroll
                    for (int k = 0; k < NS; ++k) {
                        sa[k] = na[k];
                        sb[k] = nb[k];
                    }
                }
            }
            const uint2 rr = rr_next;
            uint2 bn[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) bn[t] = make_uint2(0u, 0u);
            if (!first) {
                if (w == 0) {
#pragma unroll
                    for (int t = 0; t < 4; ++t) bn[t] = bn_next[t];
                    if (b + 1 < nblk) {
#pragma unroll
                        for (int t = 0; t < 4; ++t)
                            bn_next[t] = gbnd[(size_t)((b + 1) * 4 + t) * 64 + lane];
                    }
                } else {
                    const uint2 *src = ring + (size_t)((w - 1) * 2 + par) * 256;
#pragma unroll
                    for (int t = 0; t < 4; ++t) bn[t] = src[t * 64 + lane];
                }
            }
            if (b + 1 < nblk) rr_next = res[(size_t)(b + 1) * 64 + lane];
            uint2 *out = nullptr;
            if (!last)
                out = w < W - 1 ? ring + (size_t)(w * 2 + par) * 256
                                           : gbnd + (size_t)b * 4 * 64;
            sg_block<V, K, TE, QE, PAIRED>(st, b, j0, first, last, mstar, rr, rr_next, bn, sa, sb,
                                           chunk_base, slut, 0, RA, RB, rmin, ts, qe_rt, go, ge, t16,
                                           t32, out, lane, shift);
        }
        __syncthreads();
        if (++b == P) {
            b = 0;
            ++r;
        }
    }
    int *cw = comb + (size_t)(w * 64 + lane) * 6;
    cw[0] = t16.te_a;
    cw[1] = t16.te_b;
    cw[2] = t16.qe_a;
    cw[3] = t16.qe_b;
    cw[4] = t16.corner_a;
    cw[5] = t16.corner_b;
    __syncthreads();
    if (w == 0) {
        for (int v = 1; v < W; ++v) {
            const int *cv = comb + (size_t)(v * 64 + lane) * 6;
            t16.te_a = max(t16.te_a, cv[0]);
            t16.te_b = max(t16.te_b, cv[1]);
            t16.qe_a = max(t16.qe_a, cv[2]);
            t16.qe_b = max(t16.qe_b, cv[3]);
        }
        // the corner cell belongs to the wave that ran the last chunk
        const int *cl = comb + (size_t)(last_w * 64 + lane) * 6;
        t16.corner_a = cl[4];
        t16.corner_b = cl[5];
        sg_finish<V>(p, qv, tA, tB, RA, RB, te_rt, qe_rt, t16, t32);
    }
    __syncthreads(); // ring and comb are reused by the next item
}

// ---------------------------------------------------------------------------
// Streamed workgroup pipeline for queries of at most W chunks (one round):
// wave w always scores chunk w, and the workgroup runs through consecutive
// long groups of the queue without draining between them. Group k holds
// steps [start_k + w, start_k + w + nblk_k) of wave w, with
// start_{k+1} = start_k + max(nblk_k, W). Chunk-bottom rows pass to the
// next wave through the LDS ring at the parity of the step, so a wave that
// starts a group while the next wave still finishes the previous one does
// not overwrite what that wave reads. Wave w's end keys of group k go to
// comb buffer k & 1; the wave of the last chunk combines them and writes
// the group's results after its last block (the other waves wrote theirs
// at earlier steps, and the buffer is written again no sooner than 2W
// steps later). The paired chunk profile of a wave is staged once.
// Thread 0 fetches the next queue item when a group starts on wave 0; an
// item that is not a long group ends the stream after that group and is
// returned to the caller. The query needs at least one chunk: wave 0 must
// run, or no fetch ever ends the stream. ctl: [0..3] items of groups k (k & 3), [4] the
// stream's step count once its last group is known, [5] the returned
// item, [6] the number of groups in the stream.
// ---------------------------------------------------------------------------
// LAST: this wave scores the query's last chunk (its tracking and the
// other waves' ring stores are separate code, so that neither adds its
// registers to the other).
template <int K, bool TE, bool QE, bool PAIRED, int W, bool LAST, bool F16>
__device__ __forceinline__ int sg_pipe_stream_role(const ScoreParams &p, const QueryView &qv,
                                                   const uint8_t *smem, const uint32_t *slut,
                                                   uint2 *ring, int *comb, int *ctl, int item0,
                                                   int nlong, int w, int lane,
                                                   const PairView *pv) {
    static_assert(!F16 || PAIRED, "binary16 lanes read the paired profile");
    using V = typename std::conditional<F16, h16x2, s16x2>::type;
    const int shift = F16 ? qv.fshift : 0;
    constexpr int NS = K / 8;
    const int go = p.go, ge = p.ge, Q = qv.Q;
    const bool qs = qv.flags & 1u, ts = qv.flags & 4u;
    const bool te_rt = TE && (qv.flags & 8u), qe_rt = QE && (qv.flags & 2u);
    const int last_w = qv.nchunks - 1; // nchunks <= W
    const bool active = LAST || w < last_w; // wave-uniform
    const bool first = w == 0;
    constexpr bool last = LAST;
    const int j0 = w * K - qv.pad;
    const int mstar = Q - 1 - j0;
    const uint8_t *chunk_base = PAIRED ? pv->pbuf : smem + 2 * (qv.front + j0);
    if (threadIdx.x == 0) {
        ctl[0] = item0;
        ctl[4] = INT_MAX;
        ctl[6] = INT_MAX;
    }
    if constexpr (PAIRED) {
        if (active) stage_pairs<K, F16>(*pv, qv.front + j0, lane);
    }
    __syncthreads();

    TrackI16 t16;
    TrackI32 t32;
    track_init(t16, t32);
    SgChunk<V, K> st;
    sg_chunk_init(st, j0, qs, ts, go, ge, shift);
    short8 sa[NS], sb[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) sa[k] = sb[k] = (short8)0;
    uint2 rr_next = make_uint2(0u, 0u);
    // this wave's group k: first step of the group (on wave 0), blocks,
    // steps it holds, targets
    int k = 0, start = 0, nblk = 0, P = 0;
    int tA = 0, tB = 0, RA = 0, RB = 0, rmin = 0;
    const uint2 *res = p.res;
    int ngroups = INT_MAX; // groups of the stream, once known
    // group k is done on this wave: its end keys, and from the wave of the
    // last chunk the group's results
    auto group_done = [&]() {
        int *cb = comb + (size_t)(k & 1) * W * 64 * 6;
        if (!last) {
            int *cw = cb + (size_t)(w * 64 + lane) * 6;
            cw[0] = t16.te_a;
            cw[1] = t16.te_b;
            cw[2] = t16.qe_a;
            cw[3] = t16.qe_b;
        } else {
            for (int v = 0; v < last_w; ++v) {
                const int *cv = cb + (size_t)(v * 64 + lane) * 6;
                t16.te_a = max(t16.te_a, cv[0]);
                t16.te_b = max(t16.te_b, cv[1]);
                t16.qe_a = max(t16.qe_a, cv[2]);
                t16.qe_b = max(t16.qe_b, cv[3]);
            }
            sg_finish<V>(p, qv, tA, tB, RA, RB, te_rt, qe_rt, t16, t32);
        }
        start += P;
        ++k;
    };
    for (int s = 0;; ++s) {
        const int b = s - start - w; // block of group k on this wave
        if (active && b == 0 && k < ngroups) {
            // group k begins on this wave
            const int item = (int)wave_broadcast0((uint32_t)ctl[k & 3]);
            const GroupDesc gd = p.groups[p.g_begin + item];
            nblk = gd.ncols >> 2;
            P = nblk > W ? nblk : W;
            tA = gd.first + lane;
            tB = gd.first + 64 + lane;
            RA = p.lens[tA];
            RB = p.lens[tB];
            rmin = wave_min(min(RA, RB));
            res = p.res + gd.res_off;
            if (w == 0 && lane == 0) {
                // the next queue item; one that is not a long group ends
                // the stream after this group
                const int nx = (int)atomicAdd(p.counter, 1u);
                ctl[(k + 1) & 3] = nx;
                if (nx >= nlong) {
                    ctl[4] = start + last_w + (nblk > 0 ? nblk : 1);
                    ctl[5] = nx;
                    ctl[6] = k + 1;
                }
            }
            sg_chunk_init(st, j0, qs, ts, go, ge, shift);
            track_init(t16, t32);
            if (nblk == 0) {
                group_done(); // no column: results from the boundary alone
            } else {
                rr_next = res[lane];
                if constexpr (PAIRED) {
                    const short8 *np = pair_row(chunk_base, slut, rr_next.x & 0xffu, rr_next.y & 0xffu);
#pragma unroll
                    for (int q = 0; q < NS; ++q) {
                        sa[q] = np[2 * q];
                        sb[q] = np[2 * q + 1];
                    }
                } else {
                    const short8 *na = (const short8 *)(chunk_base + 4u * slut[rr_next.x & 0xffu]);
                    const short8 *nb = (const short8 *)(chunk_base + 4u * slut[rr_next.y & 0xffu]);
#pragma unroll
                    for (int q = 0; q < NS; ++q) {
                        sa[q] = na[q];
                        sb[q] = nb[q];
                    }
                }
            }
        }
        if (active && b >= 0 && b < nblk) {
            const uint2 rr = rr_next;
            uint2 bn[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) bn[t] = make_uint2(0u, 0u);
            if (!first) {
                const uint2 *src = ring + (size_t)((w - 1) * 2 + ((s - 1) & 1)) * 256;
#pragma unroll
                for (int t = 0; t < 4; ++t) bn[t] = src[t * 64 + lane];
            }
            if (b + 1 < nblk) rr_next = res[(size_t)(b + 1) * 64 + lane];
            uint2 *out = last ? nullptr : ring + (size_t)(w * 2 + (s & 1)) * 256;
            sg_block<V, K, TE, QE, PAIRED>(st, b, j0, first, last, mstar, rr, rr_next, bn, sa, sb,
                                           chunk_base, slut, 0, RA, RB, rmin, ts, qe_rt, go, ge, t16,
                                           t32, out, lane, shift);
            if (b == nblk - 1) group_done();
        }
        __syncthreads();
        if (s + 1 >= (int)wave_broadcast0((uint32_t)ctl[4])) break;
        ngroups = (int)wave_broadcast0((uint32_t)ctl[6]);
    }
    const int nx = (int)wave_broadcast0((uint32_t)ctl[5]);
    __syncthreads(); // ring, comb and ctl are reused by the next item
    return nx;
}

template <int K, bool TE, bool QE, bool PAIRED, int W, bool F16 = false>
__device__ __forceinline__ int sg_pipe_stream(const ScoreParams &p, const QueryView &qv,
                                              const uint8_t *smem, const uint32_t *slut,
                                              uint2 *ring, int *comb, int *ctl, int item0,
                                              int nlong, int w, int lane,
                                              const PairView *pv = nullptr) {
    // every wave runs the same steps and barriers in either role
    if (w == qv.nchunks - 1)
        return sg_pipe_stream_role<K, TE, QE, PAIRED, W, true, F16>(p, qv, smem, slut, ring, comb, ctl,
                                                                    item0, nlong, w, lane, pv);
    return sg_pipe_stream_role<K, TE, QE, PAIRED, W, false, F16>(p, qv, smem, slut, ring, comb, ctl,
                                                                 item0, nlong, w, lane, pv);
}

// Resident workgroups per CU of kernel k at `block` threads and `lds` bytes
// of dynamic LDS (at least 1).
template <typename KernelT> static int occ(KernelT k, int block, size_t lds) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, block, lds) != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

} // namespace ish
