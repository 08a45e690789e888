// Semi-global (Gotoh, configurable free ends) scoring kernels for gfx950.
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
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

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

template <typename V, int K>
__device__ __forceinline__ void sg_chunk_init(SgChunk<V, K> &st, int j0, bool qs, bool ts, int go,
                                              int ge) {
    const int cst = ge - go;
#pragma unroll
    for (int m = 0; m < K; ++m) {
        const int hl = left_boundary(j0 + m, qs, ts, go, ge); // H^(j,-1)
        st.H[m] = splat<V>(hl);
        st.E[m] = splat<V>(hl + cst); // E^(j,0)
    }
    st.diag_top = splat<V>(left_boundary(j0 - 1, qs, ts, go, ge));
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
// PAIRED (packed lanes): chunk_base is the wave's paired chunk profile and
// slut the pair table (see PairView).
template <typename V, int K, bool TE, bool QE, bool PAIRED = false, bool DEFER = false>
__device__ __forceinline__ void sg_block(SgChunk<V, K> &st, int b, int j0, bool first, bool last,
                                         int mstar, uint2 rr, uint2 rr_next, const uint2 (&bn)[4],
                                         short8 (&sa)[K / 8], short8 (&sb)[K / 8],
                                         const uint8_t *chunk_base, const uint16_t *slut,
                                         int half, int RA, int RB, bool ts, bool qe_rt, int go,
                                         int ge, TrackI16 &t16, TrackI32 &t32, uint2 *out,
                                         int lane) {
    constexpr bool PACKED = sizeof(V) == 4 && !__is_same(V, int);
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
            const int ht = ts ? (i - 1) * ge : -go - ge; // H^(-1,i)
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
            // u = H^(Q-1,i) - i*ge = H(Q-1,i) + (Q-1)*ge
            const V u = hl - splat<V>(i * ge);
            if constexpr (PACKED) {
                const uint32_t uu = as_u32(u);
                if (TE || QE) {
                    if (i < RA) t16.te_a = max(t16.te_a, key_lo(uu, i));
                    if (i < RB) t16.te_b = max(t16.te_b, key_hi(uu, i));
                } else {
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
                } else if (i == RA - 1) {
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
                    const V wv = st.H[m] - splat<V>(j * ge);
                    if constexpr (PACKED) {
                        const uint32_t ww = as_u32(wv);
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
            p.status[qv.out_off + o] = 16;
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
template <typename V, int K, bool TE, bool QE, bool PAIRED>
__device__ __forceinline__ void sg_chunk(const ScoreParams &p, const QueryView &qv,
                                         const uint8_t *chunk_base, const uint16_t *slut,
                                         uint2 *bnd, const uint2 *res, int nblk, int c, int half,
                                         int lane, int RA, int RB, TrackI16 &t16, TrackI32 &t32) {
    constexpr bool PACKED = sizeof(V) == 4 && !__is_same(V, int);
    constexpr int NS = K / 8; // profile slices of 8 rows (one 16-byte LDS read each)
    const int go = p.go, ge = p.ge, Q = qv.Q;
    const bool qs = qv.flags & 1u, ts = qv.flags & 4u;
    const bool qe_rt = QE && (qv.flags & 2u);
    const int j0 = c * K - qv.pad;
    const bool first = c == 0;
    const bool last = c == qv.nchunks - 1;
    const int mstar = Q - 1 - j0; // row Q-1 inside the last chunk
    SgChunk<V, K> st;
    sg_chunk_init(st, j0, qs, ts, go, ge);
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
        const short8 *na = (const short8 *)(chunk_base + 4u * (uint32_t)slut[(half ? rr_next.y : rr_next.x) & 0xffu]);
        const short8 *nb = (const short8 *)(chunk_base + 4u * (uint32_t)slut[rr_next.y & 0xffu]);
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
            rr_next = res[(size_t)(b + 1) * 64 + lane];
            if (!first) {
#pragma unroll
                for (int t = 0; t < 4; ++t) bn_next[t] = bnd[(size_t)((b + 1) * 4 + t) * 64 + lane];
            }
        }
        sg_block<V, K, TE, QE, PAIRED, true>(st, b, j0, first, last, mstar, rr, rr_next, bn, sa,
                                             sb, chunk_base, slut, half, RA, RB, ts, qe_rt, go, ge,
                                             t16, t32, last ? nullptr : bnd + (size_t)b * 4 * 64,
                                             lane);
    }
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
                                         const uint8_t *smem, const uint16_t *slut, uint2 *bnd,
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
        if constexpr (PAIRED) stage_pairs<K>(*pv, e0, lane);
        sg_chunk<V, K, TE, QE, PAIRED>(p, qv, chunk_base, slut, bnd, res, nblk, c, half, lane, RA,
                                       RB, t16, t32);
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
template <int K, bool TE, bool QE>
__device__ __forceinline__ void sg_tile_lockstep(const ScoreParams &p, const QueryView &qv,
                                                 const uint16_t *ptab, uint8_t *pbuf,
                                                 const int32_t *pprof, int npairs, int pqpad,
                                                 uint2 *bnd, int g, bool valid, int lane) {
    using V = s16x2;
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
            const uint4 v = *(const uint4 *)(pprof + (size_t)pr * pqpad + e0 + part * 4);
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

constexpr size_t kSgExtraBytes = kSgRingBytes + kSgCombBytes + 16;
size_t sg_pipe_extra_lds() { return kSgExtraBytes; }
// LDS of a paired hybrid launch of `waves` waves per workgroup: the pair
// table (1 KB), the pipeline's buffers, and one paired chunk profile per
// wave.
size_t sg_hybrid_paired_lds(int chunk, int nrows, int waves) {
    return 1024 + kSgExtraBytes + (size_t)waves * nrows * nrows * pair_row_bytes(chunk);
}

// W: the waves of the workgroup, which share the group's query chunks.
template <int K, bool TE, bool QE, bool PAIRED = false, int W = kSgPipeWaves>
__device__ __forceinline__ void sg_pipe_group(const ScoreParams &p, const QueryView &qv,
                                              const uint8_t *smem, const uint16_t *slut,
                                              uint2 *ring, int *comb, uint2 *gbnd, int g, int w,
                                              int lane, const PairView *pv = nullptr) {
    using V = s16x2;
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

    TrackI16 t16;
    TrackI32 t32;
    track_init(t16, t32);
    SgChunk<V, K> st;
    sg_chunk_init(st, 0, qs, ts, go, ge);
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
                sg_chunk_init(st, j0, qs, ts, go, ge);
                rr_next = res[lane];
                if (!first && w == 0) {
#pragma unroll
                    for (int t = 0; t < 4; ++t) bn_next[t] = gbnd[t * 64 + lane];
                }
                if constexpr (PAIRED) {
                    stage_pairs<K>(*pv, qv.front + j0, lane);
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
                                           chunk_base, slut, 0, RA, RB, ts, qe_rt, go, ge, t16,
                                           t32, out, lane);
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

template <typename V, int K, bool TE, bool QE>
__global__ void __launch_bounds__(256) sg_batched(ScoreParams p) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    stage_profile(p.prof, p.lut, p.prof_words, smem);
    const uint16_t *slut = (const uint16_t *)(smem + (size_t)p.prof_words * 4);

    constexpr bool PACKED = sizeof(V) == 4 && !__is_same(V, int);
    const int lane = threadIdx.x & 63;
    const int slot = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    uint2 *bnd = p.bnd + (size_t)slot * p.slot_stride;
    const int items = PACKED ? (p.g_end - p.g_begin) : 2 * (p.g_end - p.g_begin);
    QueryView qv;
    qv.Q = p.Q;
    qv.nchunks = p.nchunks;
    qv.neg_limit = p.neg_limit;
    qv.pad = p.pad;
    qv.front = p.front;
    qv.flags = p.flags;
    qv.out_off = 0;

    for (;;) {
        uint32_t w = 0;
        if (lane == 0) w = atomicAdd(p.counter, 1u);
        w = wave_broadcast0(__shfl(w, 0));
        if ((int)w >= items) break;
        const int g = p.g_begin + (PACKED ? (int)w : (int)(w >> 1));
        const int half = PACKED ? 0 : (int)(w & 1);
        sg_group<V, K, TE, QE>(p, qv, smem, slut, bnd, g, half, lane);
    }
}

// ---------------------------------------------------------------------------
// Hybrid persistent kernel (packed int16): a workgroup-level queue whose
// first p.n_long items are the longest groups, each scored by the four
// waves of a workgroup together (sg_pipe_group), and whose other items are
// tiles of four groups, one per wave (sg_group). Long groups then take a
// quarter of the time one wave needs, so they no longer set the length of
// the launch.
// ---------------------------------------------------------------------------
template <int K, bool TE, bool QE, bool PAIRED, int W>
__global__ void __launch_bounds__(256) sg_hybrid(ScoreParams p) {
    using V = s16x2;
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    __shared__ int s_item;
    size_t head;
    if constexpr (PAIRED) {
        stage_pair_table<K>(p.plut, smem);
        head = 1024;
    } else {
        stage_profile(p.prof, p.lut, p.prof_words, smem);
        head = (size_t)p.prof_words * 4 + 512;
    }
    const uint16_t *slut = (const uint16_t *)(PAIRED ? smem : smem + (size_t)p.prof_words * 4);
    uint8_t *tail = smem + ((head + 15) & ~(size_t)15);
    uint2 *ring = (uint2 *)tail;
    int *comb = (int *)(tail + kSgRingBytes);
    const int lane = threadIdx.x & 63;
    const int w = (int)wave_broadcast0(threadIdx.x >> 6); // wave-uniform
    const int waves = blockDim.x >> 6;
    PairView pv; // this wave's paired chunk profile (PAIRED only)
    pv.pprof = p.pprof;
    pv.npairs = PAIRED ? p.pnrows * p.pnrows : 0;
    pv.pqpad = p.pqpad;
    pv.pbuf = tail + kSgExtraBytes + (size_t)w * pv.npairs * pair_row_bytes(K);
    uint2 *bnd = p.bnd + (size_t)(blockIdx.x * waves + w) * p.slot_stride;
    uint2 *gbnd = p.bnd + (size_t)(blockIdx.x * waves) * p.slot_stride;
    QueryView qv;
    qv.Q = p.Q;
    qv.nchunks = p.nchunks;
    qv.neg_limit = p.neg_limit;
    qv.pad = p.pad;
    qv.front = p.front;
    qv.flags = p.flags;
    qv.out_off = 0;
    const int ng = p.g_end - p.g_begin;
    const int nlong = p.n_long < ng ? p.n_long : ng;
    const int items = nlong + (ng - nlong + waves - 1) / waves;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(p.counter, 1u);
        __syncthreads();
        const int item = s_item;
        __syncthreads(); // s_item is rewritten by the next fetch
        if (item >= items) break;
        if (item < nlong) {
            sg_pipe_group<K, TE, QE, PAIRED, W>(p, qv, smem, slut, ring, comb, gbnd, p.g_begin + item,
                                                w, lane, &pv);
        } else {
            const int g = p.g_begin + nlong + (item - nlong) * waves + w;
            if (g < p.g_end) sg_group<V, K, TE, QE, PAIRED>(p, qv, smem, slut, bnd, g, 0, lane, &pv);
        }
    }
}

// ---------------------------------------------------------------------------
// Hybrid launch over a workgroup-shared paired profile (databases whose
// codes fall into more profile classes than per-wave copies of a chunk's
// pair rows can hold in LDS, e.g. protein): the long groups run as in
// sg_hybrid over the int16 profile; the other items are tiles of four
// groups scored in lockstep (sg_tile_lockstep). After the pair table
// (1 KB), one LDS area holds either the int16 profile with the pipeline's
// buffers (long groups, which the queue hands out first) or the shared
// pair rows of one chunk (tiles).
// ---------------------------------------------------------------------------
size_t sg_hybrid_lock_lds(int chunk, int nrows, size_t profile_lds) {
    const size_t pairs = (size_t)nrows * nrows * pair_row_bytes(chunk);
    const size_t prof = ((profile_lds + 15) & ~(size_t)15) + kSgExtraBytes;
    return 1024 + (pairs > prof ? pairs : prof);
}

template <int K, bool TE, bool QE>
__global__ void __launch_bounds__(256) sg_hybrid_lock(ScoreParams p) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    __shared__ int s_item;
    stage_pair_table<K>(p.plut, smem);
    const uint16_t *ptab = (const uint16_t *)smem;
    uint8_t *area = smem + 1024;
    const uint16_t *slut = (const uint16_t *)(area + (size_t)p.prof_words * 4);
    uint8_t *tail = area + (((size_t)p.prof_words * 4 + 512 + 15) & ~(size_t)15);
    uint2 *ring = (uint2 *)tail;
    int *comb = (int *)(tail + kSgRingBytes);
    const int lane = threadIdx.x & 63;
    const int w = (int)wave_broadcast0(threadIdx.x >> 6); // wave-uniform
    const int waves = blockDim.x >> 6;
    uint2 *bnd = p.bnd + (size_t)(blockIdx.x * waves + w) * p.slot_stride;
    uint2 *gbnd = p.bnd + (size_t)(blockIdx.x * waves) * p.slot_stride;
    QueryView qv;
    qv.Q = p.Q;
    qv.nchunks = p.nchunks;
    qv.neg_limit = p.neg_limit;
    qv.pad = p.pad;
    qv.front = p.front;
    qv.flags = p.flags;
    qv.out_off = 0;
    const int ng = p.g_end - p.g_begin;
    const int nlong = p.n_long < ng ? p.n_long : ng;
    const int items = nlong + (ng - nlong + waves - 1) / waves;
    bool have_profile = false; // the int16 profile is in `area`
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(p.counter, 1u);
        __syncthreads();
        const int item = s_item;
        __syncthreads(); // s_item is rewritten by the next fetch
        if (item >= items) break;
        if (item < nlong) {
            if (!have_profile) {
                stage_profile(p.prof, p.lut, p.prof_words, area); // ends in a barrier
                have_profile = true;
            }
            sg_pipe_group<K, TE, QE, false>(p, qv, area, slut, ring, comb, gbnd, p.g_begin + item, w,
                                            lane);
        } else {
            have_profile = false; // the tile's pair rows overwrite it
            const int g = p.g_begin + nlong + (item - nlong) * waves + w;
            sg_tile_lockstep<K, TE, QE>(p, qv, ptab, area, p.pprof, p.pnrows * p.pnrows, p.pqpad, bnd,
                                        g, g < p.g_end, lane);
        }
    }
}

// ---------------------------------------------------------------------------
// Multi-query persistent kernel: one launch drains tiles of (query, 4 wave
// items) in query-major order. A workgroup restages the profile only when
// its next tile belongs to another query. Wave item = group (PACKED) or
// half group (int32).
// ---------------------------------------------------------------------------
template <typename V, int K, bool TE, bool QE, bool PAIRED>
__global__ void __launch_bounds__(256) sg_multi(ScoreParams p) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    __shared__ int s_tile[3]; // query, first wave item, items of that query
    constexpr bool PACKED = sizeof(V) == 4 && !__is_same(V, int);
    const int lane = threadIdx.x & 63;
    const int wave = (int)wave_broadcast0(threadIdx.x >> 6); // wave-uniform
    const int waves = blockDim.x >> 6;
    const int slot = blockIdx.x * waves + wave;
    uint2 *bnd = p.bnd + (size_t)slot * p.slot_stride;
    int cur_q = -1;
    const uint16_t *slut = nullptr;

    for (;;) {
        if (threadIdx.x == 0) {
            const int64_t tile = (int64_t)atomicAdd(p.counter, 1u);
            int q = -1;
            if (tile < p.queries[p.nqueries].tile_base) {
                int lo = 0, hi = p.nqueries - 1; // last query whose tile_base <= tile
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (p.queries[mid].tile_base <= tile) lo = mid;
                    else hi = mid - 1;
                }
                q = lo;
            }
            s_tile[0] = q;
            if (q >= 0) {
                const QueryDesc &qd = p.queries[q];
                const int nitems = PACKED ? (qd.g_end - qd.g_begin) : 2 * (qd.g_end - qd.g_begin);
                s_tile[1] = (int)((tile - qd.tile_base) * waves);
                s_tile[2] = nitems;
            }
        }
        __syncthreads();
        const int q = s_tile[0];
        if (q < 0) break;
        const QueryDesc qd = p.queries[q];
        if (q != cur_q) {
            // both end in a barrier
            if constexpr (PAIRED) {
                stage_pair_table<K>(qd.plut, smem);
                slut = (const uint16_t *)smem;
            } else {
                stage_profile(qd.prof, qd.lut, qd.prof_words, smem);
                slut = (const uint16_t *)(smem + (size_t)qd.prof_words * 4);
            }
            cur_q = q;
        }
        const int item = s_tile[1] + wave;
        if (item < s_tile[2]) {
            QueryView qv;
            qv.Q = qd.Q;
            qv.nchunks = qd.nchunks;
            qv.neg_limit = qd.neg_limit;
            qv.pad = qd.pad;
            qv.front = qd.front;
            qv.flags = qd.flags;
            qv.out_off = qd.out_off;
            const int g = qd.g_begin + (PACKED ? item : (item >> 1));
            const int half = PACKED ? 0 : (item & 1);
            // this wave's paired chunk profile (PAIRED only); p.pnrows is
            // the largest class count of the launch's queries
            PairView pv;
            pv.pprof = qd.pprof;
            pv.npairs = qd.pnrows * qd.pnrows;
            pv.pqpad = qd.pqpad;
            pv.pbuf = smem + 1024 + (size_t)wave * p.pnrows * p.pnrows * pair_row_bytes(K);
            sg_group<V, K, TE, QE, PAIRED>(p, qv, smem, slut, bnd, g, half, lane, &pv);
        }
        __syncthreads(); // everyone is done with s_tile and the profile
    }
}

// ---------------------------------------------------------------------------
// Multi-query kernel over workgroup-shared paired profiles (packed int16):
// as sg_multi, but a tile's four groups step through its query's chunks in
// lockstep over one LDS copy of the chunk's pair rows (sg_tile_lockstep).
// LDS: the query's pair table (1 KB), then the pair rows of one chunk
// (p.pnrows is the largest class count of the launch's queries).
// ---------------------------------------------------------------------------
template <int K, bool TE, bool QE>
__global__ void __launch_bounds__(256) sg_multi_lock(ScoreParams p) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    __shared__ int s_tile[3]; // query, first wave item, items of that query
    const int lane = threadIdx.x & 63;
    const int wave = (int)wave_broadcast0(threadIdx.x >> 6); // wave-uniform
    const int waves = blockDim.x >> 6;
    uint2 *bnd = p.bnd + (size_t)(blockIdx.x * waves + wave) * p.slot_stride;
    const uint16_t *ptab = (const uint16_t *)smem;
    uint8_t *area = smem + 1024;
    int cur_q = -1;
    for (;;) {
        if (threadIdx.x == 0) {
            const int64_t tile = (int64_t)atomicAdd(p.counter, 1u);
            int q = -1;
            if (tile < p.queries[p.nqueries].tile_base) {
                int lo = 0, hi = p.nqueries - 1; // last query whose tile_base <= tile
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (p.queries[mid].tile_base <= tile) lo = mid;
                    else hi = mid - 1;
                }
                q = lo;
            }
            s_tile[0] = q;
            if (q >= 0) {
                const QueryDesc &qd = p.queries[q];
                s_tile[1] = (int)((tile - qd.tile_base) * waves);
                s_tile[2] = qd.g_end - qd.g_begin;
            }
        }
        __syncthreads();
        const int q = s_tile[0];
        if (q < 0) break;
        const QueryDesc qd = p.queries[q];
        if (q != cur_q) {
            stage_pair_table<K>(qd.plut, smem); // ends in a barrier
            cur_q = q;
        }
        QueryView qv;
        qv.Q = qd.Q;
        qv.nchunks = qd.nchunks;
        qv.neg_limit = qd.neg_limit;
        qv.pad = qd.pad;
        qv.front = qd.front;
        qv.flags = qd.flags;
        qv.out_off = qd.out_off;
        const int item = s_tile[1] + wave;
        sg_tile_lockstep<K, TE, QE>(p, qv, ptab, area, qd.pprof, qd.pnrows * qd.pnrows, qd.pqpad,
                                    bnd, qd.g_begin + item, item < s_tile[2], lane);
        __syncthreads(); // everyone is done with s_tile and the pair rows
    }
}

// ---------------------------------------------------------------------------
// Plain kernel: textbook int32 recurrence, one target per lane, input order,
// grid-stride over targets (one launch per batch).
// ---------------------------------------------------------------------------
template <int K, bool TE, bool QE>
__global__ void __launch_bounds__(256) sg_plain(PlainParams p) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    stage_profile(p.prof, p.lut, p.prof_words, smem);
    const int16_t *sprof = (const int16_t *)smem;
    const uint16_t *slut = (const uint16_t *)(smem + (size_t)p.prof_words * 4);

    const int lane = threadIdx.x & 63;
    const int slot = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    int2 *bnd = p.bnd + (size_t)slot * p.slot_stride;
    const int go = p.go, ge = p.ge, Q = p.Q;
    const bool qs = p.flags & 1u, ts = p.flags & 4u;
    const int64_t nthreads = (int64_t)gridDim.x * blockDim.x;
    const int64_t base_stride = nthreads - (threadIdx.x & 63);
    (void)base_stride;

    // whole waves step together so that every lane of a wave uses the same
    // boundary slot rows
    for (int64_t t0 = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); t0 < p.n;
         t0 += nthreads) {
        const int64_t t = t0 + lane;
        const bool valid = t < p.n;
        const uint8_t *r = p.res + (valid ? p.off[t] : 0);
        const int R = valid ? (int)(p.off[t + 1] - p.off[t]) : 0;
        // the wave runs to its longest target
        int Rw = R;
        for (int d = 32; d > 0; d >>= 1) Rw = max(Rw, __shfl_xor(Rw, d));

        int te_v = p.neg_limit, te_i = -1, qe_v = INT_MIN, qe_j = 0, corner = 0;
        for (int c = 0; c < p.nchunks; ++c) {
            const int j0 = c * K;
            const bool first = c == 0, last = c == p.nchunks - 1;
            const int mstar = Q - 1 - j0;
            int H[K], E[K];
#pragma unroll
            for (int m = 0; m < K; ++m) {
                const int j = j0 + m;
                H[m] = qs ? 0 : -go - ge * j;
                E[m] = H[m] - go;
            }
            int diag_top = j0 == 0 ? 0 : (qs ? 0 : -go - ge * (j0 - 1));
            for (int i = 0; i < Rw; ++i) {
                const uint32_t code = i < R ? r[i] : 0u;
                const int16_t *pa = sprof + 2 * (int)slut[code] + j0;
                int htop, f;
                if (first) {
                    htop = ts ? 0 : -go - ge * i;
                    f = htop - go;
                } else {
                    const int2 bv = bnd[(size_t)i * 64 + lane];
                    htop = bv.x;
                    f = bv.y;
                }
                int diag = diag_top;
#pragma unroll
                for (int m = 0; m < K; ++m) {
                    int h = max(max(diag + (int)pa[m], E[m]), f);
                    diag = H[m];
                    H[m] = h;
                    const int hg = h - go;
                    E[m] = max(E[m] - ge, hg);
                    f = max(f - ge, hg);
                }
                diag_top = htop;
                if (!last) {
                    bnd[(size_t)i * 64 + lane] = make_int2(H[K - 1], f);
                } else {
                    int hl = H[0];
#pragma unroll
                    for (int m = 1; m < K; ++m)
                        if (m == mstar) hl = H[m];
                    if ((TE || QE) && i < R && hl > te_v) {
                        te_v = hl;
                        te_i = i;
                    }
                    if (!TE && !QE && i == R - 1) corner = hl;
                }
                if (QE && i == R - 1) {
#pragma unroll
                    for (int m = 0; m < K; ++m) {
                        if (last && m > mstar) continue;
                        if (H[m] > qe_v) {
                            qe_v = H[m];
                            qe_j = j0 + m;
                        }
                    }
                }
            }
        }
        if (valid) {
            int sc, er, eq;
            finalize(R, Q, TE, QE, te_i >= 0, te_v, te_i, qe_v != INT_MIN, qe_v, qe_j, corner,
                     p.neg_limit, sc, er, eq);
            p.score[t] = sc;
            p.end_ref[t] = er;
            p.end_q[t] = eq;
            p.status[t] = 32;
        }
    }
}

// ---------------------------------------------------------------------------
// Group packing: sorted targets -> coalesced uint2 blocks. Block b of group g
// (columns 4b .. 4b+3) is 64 uint2 words: lane l holds the four residues of
// target slot first+l in .x and of slot first+64+l in .y, zero past a
// target's end and in empty slots.
//
// The blocks of all groups form one sequence (group g starts at block
// res_off/64). Each wave packs kPackSpan consecutive blocks of it, four per
// step: a lane reads five aligned dwords of each of its two targets, shifts
// them into place (v_alignbyte) and writes four words, so the 512-byte rows
// are written whole and a lane walks its targets front to back. The last
// dword read can lie up to 3 bytes past the buffer's last residue; callers
// pad raw by kRawPad bytes.
// ---------------------------------------------------------------------------
constexpr int kPackSpan = 64;

__global__ void __launch_bounds__(256) pack_groups(const uint8_t *raw, const uint64_t *off,
                                                   const int32_t *sorted_idx,
                                                   const GroupDesc *groups, int32_t ngroups,
                                                   int64_t n, uint64_t total_blocks, uint2 *res) {
    const int lane = threadIdx.x & 63;
    const uint64_t wave = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    uint64_t gb = wave * kPackSpan;
    if (gb >= total_blocks) return;
    const uint64_t gend = gb + kPackSpan < total_blocks ? gb + kPackSpan : total_blocks;
    // the group holding block gb: the last one starting at or before it
    // (groups without blocks start where the next one does)
    int lo = 0, hi = ngroups - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (groups[mid].res_off / 64 <= gb) lo = mid;
        else hi = mid - 1;
    }
    for (int g = lo; g < ngroups && gb < gend; ++g) {
        const GroupDesc gd = groups[g];
        const uint64_t gstart = gd.res_off / 64;
        const uint64_t stop = min(gstart + (uint64_t)(gd.ncols >> 2), gend);
        if (gb >= stop) continue;
        uint64_t a[2], e[2]; // this lane's two targets: first byte, end
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int64_t si = (int64_t)gd.first + h * 64 + lane;
            a[h] = e[h] = 0;
            if (si < n) {
                const int32_t t = sorted_idx[si];
                a[h] = off[t];
                e[h] = off[t + 1];
            }
        }
        while (gb < stop) {
            const uint64_t lb = gb - gstart; // block inside the group
            const int u = stop - gb < 4 ? (int)(stop - gb) : 4;
            uint32_t w[2][4];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint64_t s0 = a[h] + lb * 4; // first byte of block lb
                const uint64_t base = s0 & ~(uint64_t)3;
                const uint32_t sh = (uint32_t)(s0 & 3);
                uint32_t d[5];
#pragma unroll
                for (int k = 0; k < 5; ++k) {
                    const uint64_t q = base + 4 * k;
                    d[k] = q < e[h] ? *(const uint32_t *)(raw + q) : 0u;
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t v = __builtin_amdgcn_alignbyte(d[j + 1], d[j], sh);
                    const int64_t left = (int64_t)e[h] - (int64_t)(s0 + 4 * j); // bytes of the target left
                    w[h][j] = left >= 4 ? v : left <= 0 ? 0u : v & ((1u << (8 * (uint32_t)left)) - 1u);
                }
            }
            uint2 *dst = res + gd.res_off + lb * 64 + lane;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (j < u) dst[(size_t)j * 64] = make_uint2(w[0][j], w[1][j]);
            gb += (uint64_t)u;
        }
    }
}

__global__ void reset_counter(uint32_t *c) { *c = 0u; }

// rev(raw[src_off[sel[k]] .. + L_k)) -> out[dst_off[k] .. dst_off[k+1]),
// L_k = dst_off[k+1] - dst_off[k] (callers check L_k <= target length).
__global__ void __launch_bounds__(256) gather_reversed(const uint8_t *raw, const uint64_t *src_off,
                                                       const int64_t *sel, const uint64_t *dst_off,
                                                       int64_t n_sel, uint8_t *out) {
    for (int64_t k = blockIdx.x; k < n_sel; k += gridDim.x) {
        const uint64_t s0 = src_off[sel[k]];
        const uint64_t d0 = dst_off[k];
        const uint64_t L = dst_off[k + 1] - d0;
        for (uint64_t x = threadIdx.x; x < L; x += blockDim.x) out[d0 + x] = raw[s0 + L - 1 - x];
    }
}

hipError_t launch_gather_reversed(const uint8_t *raw, const uint64_t *src_off, const int64_t *sel,
                                  const uint64_t *dst_off, int64_t n_sel, uint8_t *out,
                                  hipStream_t st) {
    const int64_t grid = n_sel < 65536 ? n_sel : 65536;
    hipLaunchKernelGGL(gather_reversed, dim3((unsigned)grid), dim3(256), 0, st, raw, src_off, sel,
                       dst_off, n_sel, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename V, int K>
static hipError_t launch_batched(const ScoreParams &p, int grid, int block, size_t lds,
                                 hipStream_t st) {
    const bool te = p.flags & 8u, qe = p.flags & 2u;
    if (te && qe)
        hipLaunchKernelGGL((sg_batched<V, K, true, true>), dim3(grid), dim3(block), lds, st, p);
    else if (te)
        hipLaunchKernelGGL((sg_batched<V, K, true, false>), dim3(grid), dim3(block), lds, st, p);
    else if (qe)
        hipLaunchKernelGGL((sg_batched<V, K, false, true>), dim3(grid), dim3(block), lds, st, p);
    else
        hipLaunchKernelGGL((sg_batched<V, K, false, false>), dim3(grid), dim3(block), lds, st, p);
    return hipGetLastError();
}

template <typename V, int K, bool PAIRED>
static hipError_t launch_multi(const ScoreParams &p, int grid, int block, size_t lds, hipStream_t st) {
    const bool te = p.flags & 8u, qe = p.flags & 2u;
    if (te && qe)
        hipLaunchKernelGGL((sg_multi<V, K, true, true, PAIRED>), dim3(grid), dim3(block), lds, st, p);
    else if (te)
        hipLaunchKernelGGL((sg_multi<V, K, true, false, PAIRED>), dim3(grid), dim3(block), lds, st, p);
    else if (qe)
        hipLaunchKernelGGL((sg_multi<V, K, false, true, PAIRED>), dim3(grid), dim3(block), lds, st, p);
    else
        hipLaunchKernelGGL((sg_multi<V, K, false, false, PAIRED>), dim3(grid), dim3(block), lds, st, p);
    return hipGetLastError();
}

template <int K>
static hipError_t launch_multi_lock(const ScoreParams &p, int grid, size_t lds, hipStream_t st) {
    const bool te = p.flags & 8u, qe = p.flags & 2u;
    if (te && qe)
        hipLaunchKernelGGL((sg_multi_lock<K, true, true>), dim3(grid), dim3(256), lds, st, p);
    else if (te)
        hipLaunchKernelGGL((sg_multi_lock<K, true, false>), dim3(grid), dim3(256), lds, st, p);
    else if (qe)
        hipLaunchKernelGGL((sg_multi_lock<K, false, true>), dim3(grid), dim3(256), lds, st, p);
    else
        hipLaunchKernelGGL((sg_multi_lock<K, false, false>), dim3(grid), dim3(256), lds, st, p);
    return hipGetLastError();
}

hipError_t launch_sg_multi_lock(const ScoreParams &p, int chunk, int grid, size_t lds, hipStream_t st) {
    return chunk == 16 ? launch_multi_lock<16>(p, grid, lds, st) : launch_multi_lock<32>(p, grid, lds, st);
}

hipError_t launch_sg_multi(const ScoreParams &p, bool packed, bool paired, int chunk, int grid,
                           int block, size_t lds, hipStream_t st) {
    if (packed && paired) return chunk == 16 ? launch_multi<s16x2, 16, true>(p, grid, block, lds, st)
                                             : launch_multi<s16x2, 32, true>(p, grid, block, lds, st);
    if (packed) return chunk == 16 ? launch_multi<s16x2, 16, false>(p, grid, block, lds, st)
                                   : launch_multi<s16x2, 32, false>(p, grid, block, lds, st);
    return chunk == 16 ? launch_multi<int, 16, false>(p, grid, block, lds, st)
                       : launch_multi<int, 32, false>(p, grid, block, lds, st);
}

template <int K, bool PAIRED, int W>
static hipError_t launch_hybrid(const ScoreParams &p, int grid, size_t lds, hipStream_t st) {
    const bool te = p.flags & 8u, qe = p.flags & 2u;
    const dim3 block(64 * W);
    if (te && qe)
        hipLaunchKernelGGL((sg_hybrid<K, true, true, PAIRED, W>), dim3(grid), block, lds, st, p);
    else if (te)
        hipLaunchKernelGGL((sg_hybrid<K, true, false, PAIRED, W>), dim3(grid), block, lds, st, p);
    else if (qe)
        hipLaunchKernelGGL((sg_hybrid<K, false, true, PAIRED, W>), dim3(grid), block, lds, st, p);
    else
        hipLaunchKernelGGL((sg_hybrid<K, false, false, PAIRED, W>), dim3(grid), block, lds, st, p);
    return hipGetLastError();
}

template <int K, bool PAIRED>
static hipError_t launch_hybrid_w(const ScoreParams &p, int waves, int grid, size_t lds, hipStream_t st) {
    return waves == 2 ? launch_hybrid<K, PAIRED, 2>(p, grid, lds, st) : launch_hybrid<K, PAIRED, 4>(p, grid, lds, st);
}

hipError_t launch_sg_hybrid_i16(const ScoreParams &p, int chunk, bool paired, int waves, int grid,
                                size_t lds, hipStream_t st) {
    if (paired)
        return chunk == 16   ? launch_hybrid_w<16, true>(p, waves, grid, lds, st)
               : chunk == 40 ? launch_hybrid_w<40, true>(p, waves, grid, lds, st)
                             : launch_hybrid_w<32, true>(p, waves, grid, lds, st);
    return chunk == 16   ? launch_hybrid_w<16, false>(p, waves, grid, lds, st)
           : chunk == 40 ? launch_hybrid_w<40, false>(p, waves, grid, lds, st)
                         : launch_hybrid_w<32, false>(p, waves, grid, lds, st);
}

template <int K>
static hipError_t launch_lock(const ScoreParams &p, int grid, size_t lds, hipStream_t st) {
    const bool te = p.flags & 8u, qe = p.flags & 2u;
    if (te && qe)
        hipLaunchKernelGGL((sg_hybrid_lock<K, true, true>), dim3(grid), dim3(256), lds, st, p);
    else if (te)
        hipLaunchKernelGGL((sg_hybrid_lock<K, true, false>), dim3(grid), dim3(256), lds, st, p);
    else if (qe)
        hipLaunchKernelGGL((sg_hybrid_lock<K, false, true>), dim3(grid), dim3(256), lds, st, p);
    else
        hipLaunchKernelGGL((sg_hybrid_lock<K, false, false>), dim3(grid), dim3(256), lds, st, p);
    return hipGetLastError();
}

hipError_t launch_sg_hybrid_lock(const ScoreParams &p, int chunk, int grid, size_t lds, hipStream_t st) {
    return chunk == 16 ? launch_lock<16>(p, grid, lds, st) : launch_lock<32>(p, grid, lds, st);
}

hipError_t launch_sg_batched_i16(const ScoreParams &p, int chunk, int grid, int block,
                                 size_t lds, hipStream_t st) {
    if (chunk == 16) return launch_batched<s16x2, 16>(p, grid, block, lds, st);
    return launch_batched<s16x2, 32>(p, grid, block, lds, st);
}
hipError_t launch_sg_batched_i32(const ScoreParams &p, int chunk, int grid, int block,
                                 size_t lds, hipStream_t st) {
    if (chunk == 16) return launch_batched<int, 16>(p, grid, block, lds, st);
    return launch_batched<int, 32>(p, grid, block, lds, st);
}

hipError_t launch_sg_plain(const PlainParams &p, int grid, int block, size_t lds,
                           hipStream_t st) {
    const bool te = p.flags & 8u, qe = p.flags & 2u;
    constexpr int K = kChunkPlain;
    if (te && qe)
        hipLaunchKernelGGL((sg_plain<K, true, true>), dim3(grid), dim3(block), lds, st, p);
    else if (te)
        hipLaunchKernelGGL((sg_plain<K, true, false>), dim3(grid), dim3(block), lds, st, p);
    else if (qe)
        hipLaunchKernelGGL((sg_plain<K, false, true>), dim3(grid), dim3(block), lds, st, p);
    else
        hipLaunchKernelGGL((sg_plain<K, false, false>), dim3(grid), dim3(block), lds, st, p);
    return hipGetLastError();
}

hipError_t launch_pack(const uint8_t *raw, const uint64_t *off, const int32_t *sorted_idx,
                       const GroupDesc *groups, int32_t ngroups, int64_t n, uint64_t total_blocks,
                       uint2 *res, hipStream_t st) {
    if (ngroups <= 0 || total_blocks == 0) return hipSuccess;
    const uint64_t waves = (total_blocks + kPackSpan - 1) / kPackSpan;
    const uint64_t grid = (waves + 3) / 4;
    if (grid > 0x7fffffffull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pack_groups, dim3((unsigned)grid), dim3(256), 0, st, raw, off, sorted_idx,
                       groups, ngroups, n, total_blocks, res);
    return hipGetLastError();
}

hipError_t launch_reset(uint32_t *counter, hipStream_t st) {
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, st, counter);
    return hipGetLastError();
}

template <typename KernelT> static int occ(KernelT k, int block, size_t lds) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, block, lds) != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

template <typename V, int K> static int occ_batched(int block, size_t lds, uint32_t flags) {
    const bool te = flags & 8u, qe = flags & 2u;
    if (te && qe) return occ(sg_batched<V, K, true, true>, block, lds);
    if (te) return occ(sg_batched<V, K, true, false>, block, lds);
    if (qe) return occ(sg_batched<V, K, false, true>, block, lds);
    return occ(sg_batched<V, K, false, false>, block, lds);
}
template <typename V, int K, bool PAIRED> static int occ_multi(int block, size_t lds, uint32_t flags) {
    const bool te = flags & 8u, qe = flags & 2u;
    if (te && qe) return occ(sg_multi<V, K, true, true, PAIRED>, block, lds);
    if (te) return occ(sg_multi<V, K, true, false, PAIRED>, block, lds);
    if (qe) return occ(sg_multi<V, K, false, true, PAIRED>, block, lds);
    return occ(sg_multi<V, K, false, false, PAIRED>, block, lds);
}
template <int K> static int occ_multi_lock(size_t lds, uint32_t flags) {
    const bool te = flags & 8u, qe = flags & 2u;
    if (te && qe) return occ(sg_multi_lock<K, true, true>, 256, lds);
    if (te) return occ(sg_multi_lock<K, true, false>, 256, lds);
    if (qe) return occ(sg_multi_lock<K, false, true>, 256, lds);
    return occ(sg_multi_lock<K, false, false>, 256, lds);
}
int occupancy_sg_multi_lock(int chunk, size_t lds, uint32_t flags) {
    return chunk == 16 ? occ_multi_lock<16>(lds, flags) : occ_multi_lock<32>(lds, flags);
}

int occupancy_sg_multi(bool packed, bool paired, int chunk, int block, size_t lds, uint32_t flags) {
    if (packed && paired) return chunk == 16 ? occ_multi<s16x2, 16, true>(block, lds, flags)
                                             : occ_multi<s16x2, 32, true>(block, lds, flags);
    if (packed) return chunk == 16 ? occ_multi<s16x2, 16, false>(block, lds, flags)
                                   : occ_multi<s16x2, 32, false>(block, lds, flags);
    return chunk == 16 ? occ_multi<int, 16, false>(block, lds, flags)
                       : occ_multi<int, 32, false>(block, lds, flags);
}

template <int K, bool PAIRED, int W> static int occ_hybrid(size_t lds, uint32_t flags) {
    const bool te = flags & 8u, qe = flags & 2u;
    if (te && qe) return occ(sg_hybrid<K, true, true, PAIRED, W>, 64 * W, lds);
    if (te) return occ(sg_hybrid<K, true, false, PAIRED, W>, 64 * W, lds);
    if (qe) return occ(sg_hybrid<K, false, true, PAIRED, W>, 64 * W, lds);
    return occ(sg_hybrid<K, false, false, PAIRED, W>, 64 * W, lds);
}

template <int K, bool PAIRED> static int occ_hybrid_w(int waves, size_t lds, uint32_t flags) {
    return waves == 2 ? occ_hybrid<K, PAIRED, 2>(lds, flags) : occ_hybrid<K, PAIRED, 4>(lds, flags);
}

int occupancy_sg_hybrid_i16(int chunk, bool paired, int waves, size_t lds, uint32_t flags) {
    if (paired)
        return chunk == 16   ? occ_hybrid_w<16, true>(waves, lds, flags)
               : chunk == 40 ? occ_hybrid_w<40, true>(waves, lds, flags)
                             : occ_hybrid_w<32, true>(waves, lds, flags);
    return chunk == 16   ? occ_hybrid_w<16, false>(waves, lds, flags)
           : chunk == 40 ? occ_hybrid_w<40, false>(waves, lds, flags)
                         : occ_hybrid_w<32, false>(waves, lds, flags);
}

template <int K> static int occ_lock(size_t lds, uint32_t flags) {
    const bool te = flags & 8u, qe = flags & 2u;
    if (te && qe) return occ(sg_hybrid_lock<K, true, true>, 256, lds);
    if (te) return occ(sg_hybrid_lock<K, true, false>, 256, lds);
    if (qe) return occ(sg_hybrid_lock<K, false, true>, 256, lds);
    return occ(sg_hybrid_lock<K, false, false>, 256, lds);
}

int occupancy_sg_hybrid_lock(int chunk, size_t lds, uint32_t flags) {
    return chunk == 16 ? occ_lock<16>(lds, flags) : occ_lock<32>(lds, flags);
}

int occupancy_sg_batched_i16(int chunk, int block, size_t lds, uint32_t flags) {
    return chunk == 16 ? occ_batched<s16x2, 16>(block, lds, flags)
                       : occ_batched<s16x2, 32>(block, lds, flags);
}
int occupancy_sg_batched_i32(int chunk, int block, size_t lds, uint32_t flags) {
    return chunk == 16 ? occ_batched<int, 16>(block, lds, flags)
                       : occ_batched<int, 32>(block, lds, flags);
}
int occupancy_sg_plain(int block, size_t lds, uint32_t flags) {
    const bool te = flags & 8u, qe = flags & 2u;
    if (te && qe) return occ(sg_plain<kChunkPlain, true, true>, block, lds);
    if (te) return occ(sg_plain<kChunkPlain, true, false>, block, lds);
    if (qe) return occ(sg_plain<kChunkPlain, false, true>, block, lds);
    return occ(sg_plain<kChunkPlain, false, false>, block, lds);
}

} // namespace ish
