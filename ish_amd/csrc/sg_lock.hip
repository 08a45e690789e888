// Hybrid semi-global kernel over a workgroup-shared paired profile
// (packed int16, tiles of four groups in lockstep): sg_hybrid_lock, its
// LDS size, launchers and occupancy queries.
#include "sg_device.h"

namespace ish {

// ---------------------------------------------------------------------------
// Hybrid launch over a workgroup-shared paired profile (databases whose
// codes fall into more profile classes than per-wave copies of a chunk's
// pair rows can hold in LDS, e.g. protein): the long groups run as in
// sg_hybrid over the int16 profile; the other items are tiles of four
// groups scored in lockstep (sg_tile_lockstep). After the pair table
// (2 KB), one LDS area holds either the int16 profile with the pipeline's
// buffers (long groups, which the queue hands out first) or the shared
// pair rows of one chunk (tiles).
// ---------------------------------------------------------------------------
size_t sg_hybrid_lock_lds(int chunk, int nrows, size_t profile_lds) {
    const size_t pairs = (size_t)nrows * nrows * pair_row_bytes(chunk);
    const size_t prof = ((profile_lds + 15) & ~(size_t)15) + kSgExtraBytes;
    return kPairTableLds + (pairs > prof ? pairs : prof);
}

// F16: the tiles run in binary16 lanes (every group past the first n_long
// is within the query's binary16 range bound; the long groups keep int16).
template <int K, bool TE, bool QE, bool F16>
__global__ void __launch_bounds__(256) sg_hybrid_lock(ScoreParams p) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    __shared__ int s_item;
    stage_pair_table<K>(p.plut, smem);
    const uint32_t *ptab = (const uint32_t *)smem;
    uint8_t *area = smem + kPairTableLds;
    const uint32_t *slut = (const uint32_t *)(area + (size_t)p.prof_words * 4);
    uint8_t *tail = area + (((size_t)p.prof_words * 4 + kLutLds + 15) & ~(size_t)15);
    uint2 *ring = (uint2 *)tail;
    int *comb = (int *)(tail + kSgRingBytes);
    const int lane = threadIdx.x & 63;
    const int w = (int)wave_broadcast0(threadIdx.x >> 6); // wave-uniform
    constexpr int waves = 4; // the launch has 256 threads
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
    bool have_profile = false; // the int16 profile is in `area`
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(p.counter, 1u);
        __syncthreads();
        const int item = (int)wave_broadcast0((uint32_t)s_item); // scalar from here on
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
            sg_tile_lockstep<K, TE, QE, F16>(p, qv, ptab, area, p.pprof, p.pnrows * p.pnrows, p.pqpad,
                                             bnd, g, g < p.g_end, lane);
        }
    }
}

template <int K, bool F16>
static hipError_t launch_lock(const ScoreParams &p, int grid, size_t lds, hipStream_t st) {
    const bool te = p.flags & 8u, qe = p.flags & 2u;
    if (te && qe)
        hipLaunchKernelGGL((sg_hybrid_lock<K, true, true, F16>), dim3(grid), dim3(256), lds, st, p);
    else if (te)
        hipLaunchKernelGGL((sg_hybrid_lock<K, true, false, F16>), dim3(grid), dim3(256), lds, st, p);
    else if (qe)
        hipLaunchKernelGGL((sg_hybrid_lock<K, false, true, F16>), dim3(grid), dim3(256), lds, st, p);
    else
        hipLaunchKernelGGL((sg_hybrid_lock<K, false, false, F16>), dim3(grid), dim3(256), lds, st, p);
    return hipGetLastError();
}

hipError_t launch_sg_hybrid_lock(const ScoreParams &p, int chunk, bool f16, int grid, size_t lds,
                                 hipStream_t st) {
    if (f16)
        return chunk == 16   ? launch_lock<16, true>(p, grid, lds, st)
               : chunk == 24 ? launch_lock<24, true>(p, grid, lds, st)
                             : launch_lock<32, true>(p, grid, lds, st);
    return chunk == 16   ? launch_lock<16, false>(p, grid, lds, st)
           : chunk == 24 ? launch_lock<24, false>(p, grid, lds, st)
                         : launch_lock<32, false>(p, grid, lds, st);
}

template <int K, bool F16> static int occ_lock(size_t lds, uint32_t flags) {
    const bool te = flags & 8u, qe = flags & 2u;
    if (te && qe) return occ(sg_hybrid_lock<K, true, true, F16>, 256, lds);
    if (te) return occ(sg_hybrid_lock<K, true, false, F16>, 256, lds);
    if (qe) return occ(sg_hybrid_lock<K, false, true, F16>, 256, lds);
    return occ(sg_hybrid_lock<K, false, false, F16>, 256, lds);
}

int occupancy_sg_hybrid_lock(int chunk, bool f16, size_t lds, uint32_t flags) {
    if (f16)
        return chunk == 16 ? occ_lock<16, true>(lds, flags)
               : chunk == 24 ? occ_lock<24, true>(lds, flags)
                             : occ_lock<32, true>(lds, flags);
    return chunk == 16 ? occ_lock<16, false>(lds, flags)
           : chunk == 24 ? occ_lock<24, false>(lds, flags)
                         : occ_lock<32, false>(lds, flags);
}

} // namespace ish
