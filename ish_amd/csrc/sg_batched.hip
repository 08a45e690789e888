// Semi-global batched persistent kernels, one wave per group of 128
// targets (packed int16, or int32 for groups whose proven score range
// needs it): sg_batched, its launchers and occupancy queries.
#include "sg_device.h"

namespace ish {

template <typename V, int K, bool TE, bool QE>
__global__ void __launch_bounds__(256) sg_batched(ScoreParams p) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    stage_profile(p.prof, p.lut, p.prof_words, smem);
    const uint32_t *slut = (const uint32_t *)(smem + (size_t)p.prof_words * 4);

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

template <typename V, int K> static int occ_batched(int block, size_t lds, uint32_t flags) {
    const bool te = flags & 8u, qe = flags & 2u;
    if (te && qe) return occ(sg_batched<V, K, true, true>, block, lds);
    if (te) return occ(sg_batched<V, K, true, false>, block, lds);
    if (qe) return occ(sg_batched<V, K, false, true>, block, lds);
    return occ(sg_batched<V, K, false, false>, block, lds);
}

int occupancy_sg_batched_i16(int chunk, int block, size_t lds, uint32_t flags) {
    return chunk == 16 ? occ_batched<s16x2, 16>(block, lds, flags)
                       : occ_batched<s16x2, 32>(block, lds, flags);
}
int occupancy_sg_batched_i32(int chunk, int block, size_t lds, uint32_t flags) {
    return chunk == 16 ? occ_batched<int, 16>(block, lds, flags)
                       : occ_batched<int, 32>(block, lds, flags);
}

} // namespace ish
