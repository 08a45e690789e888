// Local (Smith-Waterman) scoring kernels for gfx950.
//
// Semantics: sw / ssw_align of the reference's CPU striped path
// (ishlib/matcher/alignment/local_aln/striped.mojo:142-434):
//   H = max(0, Hdiag + S[r_i][q_j], E, F), E and F as in Gotoh, every
//   boundary 0 (E and F never need their floor: H takes the max with 0);
//   score = max H; end_ref = first column whose column maximum reaches the
//   final maximum (-1 when it is 0); end_query = smallest row of that column
//   holding it (0 when the maximum is 0) (:395-428).
//
// Layout: the same as the semi-global batched kernel (lane = two targets in
// packed int16 halves, K query rows in registers, chunk-bottom rows handed
// to the next chunk).
//
// Column-offset form: with i' = i - i0 the column inside the current
// 256-column window, H~ = H + i'*ge (likewise E~, F~). Then
//   G   = H~ + (ge - go)
//   E~  = max(E~, G)            (next column)
//   F~  = max(F~, G) - ge       (next row)
//   H~  = max(H~diag + (S + ge), E~, F~, i'*ge)
// The floor i'*ge is one value for the whole column, and so is the offset
// of every row, so the column maximum is a plain max over the rows minus
// one scalar. The state is shifted down by 256*ge at every window start,
// so int16 stays exact on reads of any length.
//
// End coordinates are tracked per query chunk c: its maximum M_c, the first
// column i_c where the chunk reaches it and the smallest row r_c of that
// column holding it (the row scan runs only on columns where some lane's
// chunk maximum strictly rises). Chunks combine by value, then first
// column, then smallest row, which is exactly the whole-column rule.
// The padding rows below the query (profile entry 0, i.e. S = -ge <= 0)
// take part in the column maximum: their values derive from real rows of
// the same chunk through a non-positive diagonal score or a gap, so they
// never strictly exceed the real rows' maximum so far, and the row scan
// meets the real row first.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include <type_traits>

#include "engine.h"
#include "kernel_common.h"

namespace ish {

constexpr int kRenorm = 256; // columns per offset window (multiple of 4)

__device__ __forceinline__ void sw_combine(int mc, int ic, int rc, int &best, int &er, int &eq) {
    if (mc > best) {
        best = mc;
        er = ic;
        eq = rc;
    } else if (mc == best && mc > 0) {
        if (ic < er) {
            er = ic;
            eq = rc;
        } else if (ic == er && rc < eq) {
            eq = rc;
        }
    }
}

// State of one lane while it scores one query chunk of its target(s).
template <typename V, int K> struct SwChunk {
    V H[K], E[K];
    V diag_top; // H~ of the row above the chunk, previous column
    V run;      // chunk maximum so far (plain values)
    int icA, rcA, icB, rcB;
    int ibase; // first column of the current offset window
};

// shift: binary16 lanes hold the state (H~, E~, F~) minus shift; the
// chunk maximum `run` and the column maxima are plain values in either.
template <typename V, int K>
__device__ __forceinline__ void sw_chunk_init(SwChunk<V, K> &st, int go, int ge, int shift = 0) {
#pragma unroll
    for (int m = 0; m < K; ++m) {
        st.H[m] = splat_uniform<V>(-ge - shift); // H~(j, -1): H = 0 at column -1, offset -ge
        st.E[m] = splat_uniform<V>(-go - shift); // E~(j, 0) = H(j, -1) - go
    }
    st.diag_top = splat_uniform<V>(-ge - shift);
    st.run = splat<V>(0);
    st.icA = st.rcA = st.icB = st.rcB = 0;
    st.icA = st.icB = -1;
    st.ibase = 0;
}

// Scores the 4 columns of block b (columns 4b .. 4b+3) of one chunk.
//   rr / rr_next: residue words of this block and the next (the profile
//   slices of the next column are fetched while a column is scored);
//   bn: chunk-top rows (H~, F~ of the row above) when the chunk is not the
//   first; out: where the chunk-bottom rows go (4 x 64 uint2, t*64 + lane),
//   null for the last chunk. lb: a lower bound of the maximum over the
//   target's other chunks, per half (filters the row scans; it may also
//   include this chunk's own earlier maximum, which the test's run term
//   covers anyway). rmin: the shortest target length of the wave (wave-uniform);
//   columns below it need no end-of-target mask.
// PAIRED (packed lanes only): chunk_base is the wave's paired profile of
//   this chunk, one row per pair of profile classes (class of target A,
//   class of target B) holding K int32 words (S_A + ge, S_B + ge), and slut
//   maps a residue code to a byte offset (A's at [0, 256), B's at
//   [256, 512)), so the packed operand of a cell is loaded as it is used.
template <typename V, int K, bool PAIRED = false>
__device__ __forceinline__ void sw_block(SwChunk<V, K> &st, int b, int j0, bool first,
                                         uint2 rr, uint2 rr_next, const uint2 (&bn)[4],
                                         short8 (&sa)[K / 8], short8 (&sb)[K / 8],
                                         const uint8_t *chunk_base, const uint32_t *slut, int half,
                                         int RA, int RB, int rmin, V lb, int go, int ge,
                                         uint2 *out, int lane, int shift = 0) {
    constexpr bool PACKED = !__is_same(V, int);
    constexpr bool F16 = __is_same(V, h16x2);
    static_assert(!F16 || PAIRED, "binary16 lanes read the paired profile");
    constexpr int NS = K / 8;
    const V vcst = splat<V>(ge - go);
    const V vge = splat<V>(ge);
    auto row_ptr = [&](uint32_t code) -> const short8 * {
        return (const short8 *)(chunk_base + 4u * (uint32_t)slut[code]);
    };
    if (b > 0 && ((b * 4) % kRenorm) == 0) {
        // new offset window: shift the carried state down
        const V vshift = splat<V>(kRenorm * ge);
#pragma unroll
        for (int m = 0; m < K; ++m) {
            st.H[m] = st.H[m] - vshift;
            st.E[m] = st.E[m] - vshift;
        }
        st.diag_top = st.diag_top - vshift;
        st.ibase = b * 4;
    }
    V lbm1;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int i = b * 4 + t;
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
        const int zi = (i - st.ibase) * ge; // floor of column i (all rows)
        const V vz = splat<V>(zi - shift);
        V htop, f;
        if (first) {
            htop = vz;                       // H~(-1, i): H = 0 above the query
            f = splat<V>(zi - go - shift);   // F~(0, i) = H(-1, i) - go
        } else {
            htop = from_u32<V>(bn[t].x);
            f = from_u32<V>(bn[t].y);
        }
        // Rows run software-pipelined: row m's F chain (h = max(t3, f),
        // g = h + c, f = max(f, g) - ge) alternates with work that does not
        // depend on it (t1, t2, t3 of row m+2: the diagonal, E and floor
        // terms; the E update of row m), so no packed operation reads the
        // result of the one just before it, which would cost a wait state.
        // The scheduling barriers keep this order.
        V cms[NS]; // column maximum of each 8-row slice (every h includes the floor vz)
        if constexpr (F16) {
            // Binary16 lanes: h = max3(t3, E~, f) with t3 = max(H~diag + S',
            // floor) two rows ahead, and the slice maxima take two rows per
            // three-input maximum: 7.5 packed ops per row instead of 9. The
            // F chain (max3, add, max, sub) alternates with t1 and t3 of row
            // m+2, the E update of row m and the slice maximum.
            V t3[K];
            t3[0] = vmax(st.diag_top + slice_val<V, PAIRED, PACKED>(sa[0], sb[0], 0), vz);
            t3[1] = vmax(st.H[0] + slice_val<V, PAIRED, PACKED>(sa[0], sb[0], 1), vz);
#pragma unroll
            for (int m = 0; m < K; ++m) {
                const V h = vmax3(t3[m], st.E[m], f);
                __builtin_amdgcn_sched_barrier(0);
                if (m > 0) {
                    pin(st.E[m - 1]);
                    if ((m - 1) % 2 == 1) pin(cms[(m - 1) / 8]);
                }
                V t1;
                if (m + 2 < K)
                    t1 = st.H[m + 1] + slice_val<V, PAIRED, PACKED>(sa[(m + 2) / 8], sb[(m + 2) / 8], (m + 2) % 8);
                __builtin_amdgcn_sched_barrier(0);
                const V g = h + vcst;
                __builtin_amdgcn_sched_barrier(0);
                if (m + 2 < K) t3[m + 2] = vmax(t1, vz);
                __builtin_amdgcn_sched_barrier(0);
                const V f1 = vmax(f, g);
                __builtin_amdgcn_sched_barrier(0);
                st.E[m] = vmax(st.E[m], g);
                __builtin_amdgcn_sched_barrier(0);
                f = f1 - vge;
                __builtin_amdgcn_sched_barrier(0);
                // rows m-1 and m of the slice in one maximum
                if (m % 2 == 1) cms[m / 8] = m % 8 == 1 ? vmax(st.H[m - 1], h) : vmax3(cms[m / 8], st.H[m - 1], h);
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
        V t3[K];  // max(H~diag + S', E~, floor) of the rows ahead
        t3[0] = vmax(vmax(st.diag_top + slice_val<V, PAIRED, PACKED>(sa[0], sb[0], 0), st.E[0]), vz);
        t3[1] = vmax(vmax(st.H[0] + slice_val<V, PAIRED, PACKED>(sa[0], sb[0], 1), st.E[1]), vz);
#pragma unroll
        for (int m = 0; m < K; ++m) {
            const V h = vmax(t3[m], f);
            __builtin_amdgcn_sched_barrier(0);
            // row m-1's E update and slice-maximum step stay in their slot
            // (not sunk past the loop); pinned one operation after they are
            // written
            if (m > 0) {
                pin(st.E[m - 1]);
                pin(cms[(m - 1) / 8]);
            }
            V t1, t2;
            if (m + 2 < K) // H~diag of row m+2 is row m+1's value in the previous column
                t1 = st.H[m + 1] + slice_val<V, PAIRED, PACKED>(sa[(m + 2) / 8], sb[(m + 2) / 8], (m + 2) % 8);
            __builtin_amdgcn_sched_barrier(0);
            const V g = h + vcst;
            __builtin_amdgcn_sched_barrier(0);
            if (m + 2 < K) t2 = vmax(t1, st.E[m + 2]);
            __builtin_amdgcn_sched_barrier(0);
            const V f1 = vmax(f, g);
            __builtin_amdgcn_sched_barrier(0);
            if (m + 2 < K) t3[m + 2] = vmax(t2, vz);
            __builtin_amdgcn_sched_barrier(0);
            f = f1 - vge;
            __builtin_amdgcn_sched_barrier(0);
            st.E[m] = vmax(st.E[m], g);
            st.H[m] = h;
            // the slice maxima run as chains one row apart
            cms[m / 8] = m % 8 == 0 ? h : vmax(cms[m / 8], h);
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
        if (out) {
            uint2 o;
            o.x = as_u32(st.H[K - 1]);
            o.y = as_u32(f);
            out[t * 64 + lane] = o;
        }
        // column maximum over the slices' maxima
        V cm = cms[0];
#pragma unroll
        for (int k = 1; k < NS; ++k) cm = vmax(cm, cms[k]);
        cm = cm - vz; // plain column maximum (>= 0)
        // columns past a target's end do not count; only columns at or past
        // the wave's shortest target need the mask
        if (i >= rmin) {
            if constexpr (PACKED) {
                const uint32_t mask = (i < RA ? 0x0000FFFFu : 0u) | (i < RB ? 0xFFFF0000u : 0u);
                cm = from_u32<V>(as_u32(cm) & mask);
            } else {
                cm = i < RA ? cm : 0;
            }
        }
        // A rise matters only when it reaches the target's best over the
        // other chunks: a chunk that ends below it cannot win, and one that
        // ends at or above it records its last rise here, which is its
        // (M_c, i_c, r_c). Per half, "cm > max(run, lb - 1)" is exactly
        // "the chunk maximum strictly rises, to at least lb"; one packed max
        // and one 32-bit compare test both halves.
        // lb - 1 is formed at the first column's test, so a caller's loads
        // of lb are waited for behind that column's rows
        if (t == 0) lbm1 = lb - splat<V>(1);
        const V thr = vmax(st.run, lbm1);
        // wave-uniform tests keep these rare paths real branches
        if (__builtin_amdgcn_ballot_w64(as_u32(vmax(cm, thr)) != as_u32(thr)) != 0) {
            bool up_a, up_b;
            if constexpr (PACKED) {
                up_a = cm.x > thr.x;
                up_b = cm.y > thr.y;
            } else {
                up_a = cm > thr;
                up_b = false;
            }
            // smallest row of this column holding the new maximum
            const V key = cm + vz;
            int ra = K, rb = K;
            if constexpr (PACKED) {
                // Both halves at once, branch-free per row: x = (H ^ key) | nu
                // is zero in a half exactly where that half rises and H
                // equals the key, so min(x, 255) * 256 + m is the row m
                // there and at least 256 elsewhere; the running minimum is
                // the first such row. A slice whose maximum is below the
                // key (for every rising half of every lane) holds no such
                // row and is skipped.
                const uint32_t ku = as_u32(key);
                const uint32_t nu = (up_a ? 0u : 0x0000FFFFu) | (up_b ? 0u : 0xFFFF0000u);
                const u16x2 one = {1, 1}, cap = {255, 255}, scale = {256, 256};
                u16x2 r = {0xFFFF, 0xFFFF};
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    const u16x2 t = __builtin_elementwise_min(
                        __builtin_bit_cast(u16x2, (as_u32(cms[k]) ^ ku) | nu), one);
                    if (__builtin_amdgcn_ballot_w64(__builtin_bit_cast(uint32_t, t) != 0x00010001u) != 0) {
#pragma unroll
                        for (int mm = 0; mm < 8; ++mm) {
                            const int m = k * 8 + mm;
                            const u16x2 x = __builtin_elementwise_min(
                                __builtin_bit_cast(u16x2, (as_u32(st.H[m]) ^ ku) | nu), cap);
                            r = __builtin_elementwise_min(
                                r, x * scale + (u16x2){(unsigned short)m, (unsigned short)m});
                        }
                    }
                }
                ra = r.x < 256 ? (int)r.x : K;
                rb = r.y < 256 ? (int)r.y : K;
            } else {
                // the first 8-row slice whose maximum is the key holds it
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    const bool ha = up_a && ra == K && cms[k] == key;
                    if (__builtin_amdgcn_ballot_w64(ha) != 0) {
#pragma unroll
                        for (int mm = 0; mm < 8; ++mm) {
                            const int m = k * 8 + mm;
                            if (ha && ra == K && st.H[m] == key) ra = m;
                        }
                    }
                }
            }
            if (up_a) {
                st.icA = i;
                st.rcA = j0 + ra;
            }
            if (up_b) {
                st.icB = i;
                st.rcB = j0 + rb;
            }
        }
        st.run = vmax(st.run, cm);
    }
}

// One wave per group: the wave runs every query chunk of its group in turn.
template <typename V, int K>
__global__ void __launch_bounds__(256) sw_batched(ScoreParams p) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    stage_profile(p.prof, p.lut, p.prof_words, smem);
    const uint32_t *slut = (const uint32_t *)(smem + (size_t)p.prof_words * 4);

    constexpr bool PACKED = !__is_same(V, int);
    constexpr int NS = K / 8;
    const int lane = threadIdx.x & 63;
    const int slot = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    uint2 *bnd = p.bnd + (size_t)slot * p.slot_stride;
    const int go = p.go, ge = p.ge;
    const int items = PACKED ? (p.g_end - p.g_begin) : 2 * (p.g_end - p.g_begin);

    for (;;) {
        uint32_t w = 0;
        if (lane == 0) w = atomicAdd(p.counter, 1u);
        w = wave_broadcast0(__shfl(w, 0));
        if ((int)w >= items) break;
        const int g = p.g_begin + (PACKED ? (int)w : (int)(w >> 1));
        const int half = PACKED ? 0 : (int)(w & 1);
        const GroupDesc gd = p.groups[g];
        const int tA = gd.first + half * 64 + lane;
        const int tB = gd.first + 64 + lane;
        const int RA = p.lens[tA];
        const int RB = PACKED ? p.lens[tB] : 0;
        const uint2 *res = p.res + gd.res_off;
        const int nblk = gd.ncols >> 2;
        const int rmin = wave_min(PACKED ? min(RA, RB) : RA);

        int bestA = 0, erA = -1, eqA = 0, bestB = 0, erB = -1, eqB = 0;
        for (int c = 0; c < p.nchunks; ++c) {
            const int j0 = c * K;
            const bool first = c == 0;
            const bool last = c == p.nchunks - 1;
            const uint8_t *chunk_base = smem + 2 * j0;
            const V lb = pack2<V>(bestA, bestB);
            SwChunk<V, K> st;
            sw_chunk_init(st, go, ge);
            uint2 rr_next = nblk > 0 ? res[lane] : make_uint2(0u, 0u);
            uint2 bn_next[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) bn_next[t] = (!first && nblk > 0) ? bnd[t * 64 + lane] : make_uint2(0u, 0u);
            short8 sa[NS], sb[NS];
            {
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
                sw_block<V, K>(st, b, j0, first, rr, rr_next, bn, sa, sb, chunk_base, slut, half, RA,
                               RB, rmin, lb, go, ge, last ? nullptr : bnd + (size_t)b * 4 * 64, lane);
            }
            if constexpr (PACKED) {
                sw_combine((int)st.run.x, st.icA, st.rcA, bestA, erA, eqA);
                sw_combine((int)st.run.y, st.icB, st.rcB, bestB, erB, eqB);
            } else {
                sw_combine((int)st.run, st.icA, st.rcA, bestA, erA, eqA);
            }
        }

        for (int which = 0; which < (PACKED ? 2 : 1); ++which) {
            const int t = which ? tB : tA;
            const int o = p.orig[t];
            if (o < 0) continue;
            const int best = which ? bestB : bestA;
            p.score[o] = best;
            p.end_ref[o] = best > 0 ? (which ? erB : erA) : -1;
            p.end_q[o] = best > 0 ? (which ? eqB : eqA) : 0;
            p.status[o] = PACKED ? 16 : (32 | 0x80);
        }
    }
}

// ---------------------------------------------------------------------------
// Workgroup-pipelined local kernel (packed int16): the four waves of a
// workgroup share one group of 128 targets and split its query chunks,
// wave w taking chunks w, w+4, w+8, ...  Time advances in steps of one
// 4-column block, closed by a workgroup barrier; at step s wave w works on
// block s - w - r*P of its round-r chunk (P = max(blocks, 5) steps per
// round), one block behind the wave that owns the chunk above. The
// chunk-bottom rows (H~, F~) flow to the next wave through a double-
// buffered LDS ring, and from wave 3 to wave 0 of the next round through
// the workgroup's boundary slot in global memory (written P-4 or more
// steps before it is read, and overwritten 3 steps after). A group then
// takes a quarter of the time one wave needs, so reads of 50 kb no longer
// set the length of the launch. Results are the same as sw_batched's: each
// chunk runs the same per-column code, and the per-chunk (M_c, i_c, r_c)
// combine is order-free (value, then first column, then smallest row).
// ---------------------------------------------------------------------------
constexpr int kPipeWaves = 4;
constexpr size_t kPipeRingBytes = (size_t)(kPipeWaves - 1) * 2 * 4 * 64 * sizeof(uint2);
constexpr size_t kPipeCombBytes = (size_t)kPipeWaves * 64 * 6 * sizeof(int);
constexpr size_t kPipePubBytes = (size_t)kPipeWaves * 64 * sizeof(uint32_t);

constexpr size_t kPipeExtraBytes = kPipeRingBytes + kPipeCombBytes + kPipePubBytes + 16;
size_t sw_pipe_extra_lds() { return kPipeExtraBytes; }

// Paired variant: the LDS a launch needs is the pair table (2 KB), the
// pipeline's own buffers, and one paired chunk profile per wave.
size_t sw_pipe_paired_lds(int chunk, int nrows) {
    return kPairTableLds + sw_pipe_extra_lds() +
           (size_t)kPipeWaves * nrows * nrows * pair_row_bytes(chunk);
}

// F16 (PAIRED only): binary16 lanes for every group (the query's range
// bound holds for any target), the state held minus p.fshift.
template <int K, bool PAIRED, bool F16 = false>
__global__ void __launch_bounds__(256) sw_pipe(ScoreParams p) {
    static_assert(!F16 || PAIRED, "binary16 lanes read the paired profile");
    using V = typename std::conditional<F16, h16x2, s16x2>::type;
    const int shift = F16 ? p.fshift : 0;
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    __shared__ int s_item;
    size_t head;
    if constexpr (PAIRED) {
        stage_pair_table<K>(p.plut, smem);
        head = kPairTableLds;
    } else {
        stage_profile(p.prof, p.lut, p.prof_words, smem);
        head = (size_t)p.prof_words * 4 + kLutLds;
    }
    const uint32_t *slut = (const uint32_t *)(PAIRED ? smem : smem + (size_t)p.prof_words * 4);
    uint8_t *tail = smem + ((head + 15) & ~(size_t)15);
    uint2 *ring = (uint2 *)tail;                // [wave 0..2][parity][t][lane]
    int *comb = (int *)(tail + kPipeRingBytes); // [wave][lane][6]
    // each wave's best so far per lane (packed A|B): the other waves read it
    // as a lower bound of the maximum over the chunks they do not own
    uint32_t *pub = (uint32_t *)(tail + kPipeRingBytes + kPipeCombBytes); // [wave][lane]

    constexpr int NS = K / 8;
    const int lane = threadIdx.x & 63;
    const int w = (int)wave_broadcast0(threadIdx.x >> 6); // wave-uniform
    uint2 *gbnd = p.bnd + (size_t)blockIdx.x * p.slot_stride;
    const int go = p.go, ge = p.ge;
    const int items = p.g_end - p.g_begin;
    const int nch = p.nchunks;
    const int rounds = (nch + kPipeWaves - 1) / kPipeWaves;
    // this wave's paired chunk profile (PAIRED only)
    PairView pv;
    pv.pprof = p.pprof;
    pv.npairs = PAIRED ? p.pnrows * p.pnrows : 0;
    pv.pqpad = p.pqpad;
    pv.pbuf = tail + kPipeExtraBytes + (size_t)w * pv.npairs * pair_row_bytes(K);
    uint8_t *pbuf = pv.pbuf;

    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(p.counter, 1u);
        __syncthreads();
        const int item = (int)wave_broadcast0((uint32_t)s_item); // scalar from here on
        pub[w * 64 + lane] = 0u;
        __syncthreads(); // s_item is rewritten at the top of the next group;
                         // every pub entry is reset before anyone reads it
        if (item >= items) break;
        const GroupDesc gd = p.groups[p.g_begin + item];
        const int tA = gd.first + lane;
        const int tB = gd.first + 64 + lane;
        const int RA = p.lens[tA];
        const int RB = p.lens[tB];
        const uint2 *res = p.res + gd.res_off;
        const int nblk = gd.ncols >> 2;
        const int rmin = wave_min(min(RA, RB));
        const int P = nblk > 5 ? nblk : 5;
        const int steps = (rounds - 1) * P + nblk + (kPipeWaves - 1);

        int bestA = 0, erA = -1, eqA = 0, bestB = 0, erB = -1, eqB = 0;
        SwChunk<V, K> st;
        sw_chunk_init(st, go, ge, shift);
        uint2 rr_next = make_uint2(0u, 0u);
        uint2 bn_next[4];
        short8 sa[NS], sb[NS];
#pragma unroll
        for (int t = 0; t < 4; ++t) bn_next[t] = make_uint2(0u, 0u);
#pragma unroll
        for (int k = 0; k < NS; ++k) sa[k] = sb[k] = (short8)0;

        int b = -w, r = 0; // block and round of this wave at the current step
        for (int s = 0; s < steps; ++s) {
            const int c = kPipeWaves * r + w;
            if (b >= 0 && b < nblk && r < rounds && c < nch) {
                const int j0 = c * K;
                const bool first = c == 0;
                const bool last = c == nch - 1;
                const int par = (r * P + b) & 1; // ring parity of this (round, block)
                const uint8_t *chunk_base = PAIRED ? pbuf : smem + 2 * j0;
                if (b == 0) {
                    sw_chunk_init(st, go, ge, shift);
                    rr_next = res[lane];
                    if (!first && w == 0) {
#pragma unroll
                        for (int t = 0; t < 4; ++t) bn_next[t] = gbnd[t * 64 + lane];
                    }
                    if constexpr (PAIRED) {
                        stage_pairs<K, F16>(pv, j0, lane);
                        const short8 *np = pair_row(pbuf, slut, rr_next.x & 0xffu, rr_next.y & 0xffu);
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
                    out = w < kPipeWaves - 1 ? ring + (size_t)(w * 2 + par) * 256
                                             : gbnd + (size_t)b * 4 * 64;
                // Row-scan filter: a lower bound of the maximum over the
                // target's other chunks. This wave's finished chunks give
                // bestA / bestB; the other waves publish their best so far
                // (finished chunks and running chunk maxima, which only
                // grow, so a stale entry is still a lower bound). This
                // wave's own entry is read too, so that the four reads go
                // out together with no branch: it is at most max(best,
                // run), which the filter's own terms already cover.
                static_assert(kPipeWaves == 4, "one read per wave's entry");
                const V bestv = pack2<V>(bestA, bestB);
                const V lb = vmax3(vmax3(bestv, from_u32<V>(pub[lane]), from_u32<V>(pub[64 + lane])),
                                   from_u32<V>(pub[128 + lane]), from_u32<V>(pub[192 + lane]));
                sw_block<V, K, PAIRED>(st, b, j0, first, rr, rr_next, bn, sa, sb, chunk_base, slut, 0,
                                       RA, RB, rmin, lb, go, ge, out, lane, shift);
                pub[w * 64 + lane] = as_u32(vmax(bestv, st.run));
                if (b == nblk - 1) {
                    sw_combine((int)st.run.x, st.icA, st.rcA, bestA, erA, eqA);
                    sw_combine((int)st.run.y, st.icB, st.rcB, bestB, erB, eqB);
                }
            }
            __syncthreads();
            if (++b == P) {
                b = 0;
                ++r;
            }
        }
        int *cw = comb + (size_t)(w * 64 + lane) * 6;
        cw[0] = bestA;
        cw[1] = erA;
        cw[2] = eqA;
        cw[3] = bestB;
        cw[4] = erB;
        cw[5] = eqB;
        __syncthreads();
        if (w == 0) {
            for (int v = 1; v < kPipeWaves; ++v) {
                const int *cv = comb + (size_t)(v * 64 + lane) * 6;
                sw_combine(cv[0], cv[1], cv[2], bestA, erA, eqA);
                sw_combine(cv[3], cv[4], cv[5], bestB, erB, eqB);
            }
            for (int which = 0; which < 2; ++which) {
                const int t = which ? tB : tA;
                const int o = p.orig[t];
                if (o < 0) continue;
                const int best = which ? bestB : bestA;
                p.score[o] = best;
                p.end_ref[o] = best > 0 ? (which ? erB : erA) : -1;
                p.end_q[o] = best > 0 ? (which ? eqB : eqA) : 0;
                p.status[o] = F16 ? 16 | 0x40 : 16;
            }
        }
        __syncthreads(); // comb is rewritten by the next group
    }
}

hipError_t launch_sw_pipe_i16(const ScoreParams &p, int chunk, bool paired, bool f16, int grid, size_t lds,
                              hipStream_t st) {
    const dim3 g(grid), b(64 * kPipeWaves);
    if (paired && f16) {
        if (chunk == 16) hipLaunchKernelGGL((sw_pipe<16, true, true>), g, b, lds, st, p);
        else hipLaunchKernelGGL((sw_pipe<32, true, true>), g, b, lds, st, p);
    } else if (paired) {
        if (chunk == 16) hipLaunchKernelGGL((sw_pipe<16, true>), g, b, lds, st, p);
        else hipLaunchKernelGGL((sw_pipe<32, true>), g, b, lds, st, p);
    } else {
        if (chunk == 16) hipLaunchKernelGGL((sw_pipe<16, false>), g, b, lds, st, p);
        else hipLaunchKernelGGL((sw_pipe<32, false>), g, b, lds, st, p);
    }
    return hipGetLastError();
}

hipError_t launch_sw_batched_i16(const ScoreParams &p, int chunk, int grid, int block,
                                 size_t lds, hipStream_t st) {
    if (chunk == 16)
        hipLaunchKernelGGL((sw_batched<s16x2, 16>), dim3(grid), dim3(block), lds, st, p);
    else
        hipLaunchKernelGGL((sw_batched<s16x2, 32>), dim3(grid), dim3(block), lds, st, p);
    return hipGetLastError();
}

hipError_t launch_sw_batched_i32(const ScoreParams &p, int chunk, int grid, int block,
                                 size_t lds, hipStream_t st) {
    if (chunk == 16)
        hipLaunchKernelGGL((sw_batched<int, 16>), dim3(grid), dim3(block), lds, st, p);
    else
        hipLaunchKernelGGL((sw_batched<int, 32>), dim3(grid), dim3(block), lds, st, p);
    return hipGetLastError();
}

template <typename KernelT> static int occ_of(KernelT k, int block, size_t lds) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, block, lds) != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

int occupancy_sw_pipe_i16(int chunk, bool paired, bool f16, size_t lds) {
    if (paired && f16)
        return chunk == 16 ? occ_of(sw_pipe<16, true, true>, 64 * kPipeWaves, lds)
                           : occ_of(sw_pipe<32, true, true>, 64 * kPipeWaves, lds);
    if (paired)
        return chunk == 16 ? occ_of(sw_pipe<16, true>, 64 * kPipeWaves, lds)
                           : occ_of(sw_pipe<32, true>, 64 * kPipeWaves, lds);
    return chunk == 16 ? occ_of(sw_pipe<16, false>, 64 * kPipeWaves, lds)
                       : occ_of(sw_pipe<32, false>, 64 * kPipeWaves, lds);
}

int occupancy_sw_batched_i16(int chunk, int block, size_t lds) {
    return chunk == 16 ? occ_of(sw_batched<s16x2, 16>, block, lds)
                       : occ_of(sw_batched<s16x2, 32>, block, lds);
}
int occupancy_sw_batched_i32(int chunk, int block, size_t lds) {
    return chunk == 16 ? occ_of(sw_batched<int, 16>, block, lds)
                       : occ_of(sw_batched<int, 32>, block, lds);
}

} // namespace ish
