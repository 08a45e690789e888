// Hybrid semi-global kernel (packed int16): the longest groups pipelined
// across the waves of a workgroup, the others one group per wave;
// sg_hybrid, its LDS sizes, launchers and occupancy queries.
#include "sg_device.h"

namespace ish {

size_t sg_pipe_extra_lds() { return kSgExtraBytes; }
// LDS of a paired hybrid launch of `waves` waves per workgroup: the pair
// table (2 KB), the pipeline's buffers, and one paired chunk profile per
// wave.
size_t sg_hybrid_paired_lds(int chunk, int nrows, int waves) {
    return kPairTableLds + kSgExtraBytes + (size_t)waves * nrows * nrows * pair_row_bytes(chunk);
}

// ---------------------------------------------------------------------------
// Hybrid persistent kernel (packed int16): a workgroup-level queue whose
// first p.n_long items are the longest groups, each scored by the four
// waves of a workgroup together (sg_pipe_group), and whose other items are
// tiles of four groups, one per wave (sg_group). Long groups then take a
// quarter of the time one wave needs, so they no longer set the length of
// the launch.
// ---------------------------------------------------------------------------
// F16 (PAIRED only): every group runs in binary16 lanes (the launch's
// groups all lie within the query's range bound), values held minus
// p.fshift.
template <int K, bool TE, bool QE, bool PAIRED, int W, bool F16>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) sg_hybrid(ScoreParams p) {
    static_assert(!F16 || PAIRED, "binary16 lanes read the paired profile");
    using V = typename std::conditional<F16, h16x2, s16x2>::type;
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
    uint2 *ring = (uint2 *)tail;
    int *comb = (int *)(tail + kSgRingBytes);
    int *ctl = (int *)(tail + kSgRingBytes + 2 * kSgCombBytes);
    const int lane = threadIdx.x & 63;
    const int w = (int)wave_broadcast0(threadIdx.x >> 6); // wave-uniform
    constexpr int waves = W; // the launch has 64 * W threads
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
    qv.fshift = p.fshift;
    const int ng = p.g_end - p.g_begin;
    const int nlong = p.n_long < ng ? p.n_long : ng;
    const int items = nlong + (ng - nlong + waves - 1) / waves;
    auto fetch = [&]() {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(p.counter, 1u);
        __syncthreads();
        const int it = (int)wave_broadcast0((uint32_t)s_item); // scalar from here on
        __syncthreads(); // s_item is rewritten by the next fetch
        return it;
    };
    int item = fetch();
    while (item < items) {
        if (item < nlong) {
            // one round per group: stream the long groups through the
            // pipeline; the item that ends the stream is scored next. (Over
            // the int16 profile, or with the query-end rule's per-row keys,
            // the streamed code would push the kernel past 2 waves per SIMD.)
            if constexpr (PAIRED && !QE) {
                // (an empty query has no chunk and no wave to run the stream)
                if (qv.nchunks >= 1 && qv.nchunks <= W) {
                    item = sg_pipe_stream<K, TE, QE, PAIRED, W, F16>(p, qv, smem, slut, ring, comb, ctl,
                                                                    item, nlong, w, lane, &pv);
                    continue;
                }
            }
            sg_pipe_group<K, TE, QE, PAIRED, W, F16>(p, qv, smem, slut, ring, comb, gbnd, p.g_begin + item,
                                                     w, lane, &pv);
        } else {
            const int g = p.g_begin + nlong + (item - nlong) * waves + w;
            if (g < p.g_end) sg_group<V, K, TE, QE, PAIRED>(p, qv, smem, slut, bnd, g, 0, lane, &pv);
        }
        item = fetch();
    }
}

template <int K, bool PAIRED, int W, bool F16>
static hipError_t launch_hybrid(const ScoreParams &p, int grid, size_t lds, hipStream_t st) {
    const bool te = p.flags & 8u, qe = p.flags & 2u;
    const dim3 block(64 * W);
    if (te && qe)
        hipLaunchKernelGGL((sg_hybrid<K, true, true, PAIRED, W, F16>), dim3(grid), block, lds, st, p);
    else if (te)
        hipLaunchKernelGGL((sg_hybrid<K, true, false, PAIRED, W, F16>), dim3(grid), block, lds, st, p);
    else if (qe)
        hipLaunchKernelGGL((sg_hybrid<K, false, true, PAIRED, W, F16>), dim3(grid), block, lds, st, p);
    else
        hipLaunchKernelGGL((sg_hybrid<K, false, false, PAIRED, W, F16>), dim3(grid), block, lds, st, p);
    return hipGetLastError();
}

// The 40-row tile has four-wave variants only (two waves would spill
// registers); the planner never asks for two.
template <int K, bool PAIRED, bool F16>
static hipError_t launch_hybrid_w(const ScoreParams &p, int waves, int grid, size_t lds, hipStream_t st) {
    if constexpr (K == 40) {
        if (waves != 4) return hipErrorInvalidValue;
        return launch_hybrid<K, PAIRED, 4, F16>(p, grid, lds, st);
    } else {
        return waves == 2 ? launch_hybrid<K, PAIRED, 2, F16>(p, grid, lds, st)
                          : launch_hybrid<K, PAIRED, 4, F16>(p, grid, lds, st);
    }
}

template <bool PAIRED, bool F16>
static hipError_t launch_hybrid_k(const ScoreParams &p, int chunk, int waves, int grid, size_t lds,
                                  hipStream_t st) {
    return chunk == 16   ? launch_hybrid_w<16, PAIRED, F16>(p, waves, grid, lds, st)
           : chunk == 40 ? launch_hybrid_w<40, PAIRED, F16>(p, waves, grid, lds, st)
                         : launch_hybrid_w<32, PAIRED, F16>(p, waves, grid, lds, st);
}

hipError_t launch_sg_hybrid_i16(const ScoreParams &p, int chunk, bool paired, bool f16, int waves, int grid,
                                size_t lds, hipStream_t st) {
    if (paired && f16) return launch_hybrid_k<true, true>(p, chunk, waves, grid, lds, st);
    if (paired) return launch_hybrid_k<true, false>(p, chunk, waves, grid, lds, st);
    return launch_hybrid_k<false, false>(p, chunk, waves, grid, lds, st);
}

template <int K, bool PAIRED, int W, bool F16> static int occ_hybrid(size_t lds, uint32_t flags) {
    const bool te = flags & 8u, qe = flags & 2u;
    if (te && qe) return occ(sg_hybrid<K, true, true, PAIRED, W, F16>, 64 * W, lds);
    if (te) return occ(sg_hybrid<K, true, false, PAIRED, W, F16>, 64 * W, lds);
    if (qe) return occ(sg_hybrid<K, false, true, PAIRED, W, F16>, 64 * W, lds);
    return occ(sg_hybrid<K, false, false, PAIRED, W, F16>, 64 * W, lds);
}

template <int K, bool PAIRED, bool F16> static int occ_hybrid_w(int waves, size_t lds, uint32_t flags) {
    if constexpr (K == 40) {
        return occ_hybrid<K, PAIRED, 4, F16>(lds, flags);
    } else {
        return waves == 2 ? occ_hybrid<K, PAIRED, 2, F16>(lds, flags) : occ_hybrid<K, PAIRED, 4, F16>(lds, flags);
    }
}

template <bool PAIRED, bool F16> static int occ_hybrid_k(int chunk, int waves, size_t lds, uint32_t flags) {
    return chunk == 16   ? occ_hybrid_w<16, PAIRED, F16>(waves, lds, flags)
           : chunk == 40 ? occ_hybrid_w<40, PAIRED, F16>(waves, lds, flags)
                         : occ_hybrid_w<32, PAIRED, F16>(waves, lds, flags);
}

int occupancy_sg_hybrid_i16(int chunk, bool paired, bool f16, int waves, size_t lds, uint32_t flags) {
    if (paired && f16) return occ_hybrid_k<true, true>(chunk, waves, lds, flags);
    if (paired) return occ_hybrid_k<true, false>(chunk, waves, lds, flags);
    return occ_hybrid_k<false, false>(chunk, waves, lds, flags);
}

} // namespace ish
