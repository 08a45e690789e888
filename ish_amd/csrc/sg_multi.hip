// Multi-query semi-global kernels (many queries in one persistent
// launch): sg_multi and sg_multi_lock, their launchers and occupancy
// queries.
#include "sg_device.h"

namespace ish {

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
    const uint32_t *slut = nullptr;

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
        const int q = (int)wave_broadcast0((uint32_t)s_tile[0]);
        if (q < 0) break;
        const QueryDesc qd = p.queries[q];
        if (q != cur_q) {
            // both end in a barrier
            if constexpr (PAIRED) {
                stage_pair_table<K>(qd.plut, smem);
                slut = (const uint32_t *)smem;
            } else {
                stage_profile(qd.prof, qd.lut, qd.prof_words, smem);
                slut = (const uint32_t *)(smem + (size_t)qd.prof_words * 4);
            }
            cur_q = q;
        }
        const int item = (int)wave_broadcast0((uint32_t)s_tile[1]) + wave;
        if (item < (int)wave_broadcast0((uint32_t)s_tile[2])) {
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
            pv.pbuf = smem + kPairTableLds + (size_t)wave * p.pnrows * p.pnrows * pair_row_bytes(K);
            if constexpr (PACKED && PAIRED) {
                // binary16 lanes when every group of the query lies within
                // its range bound (a wave-uniform choice per query)
                qv.fshift = qd.fshift;
                if (qd.f16)
                    sg_group<h16x2, K, TE, QE, true>(p, qv, smem, slut, bnd, g, half, lane, &pv);
                else
                    sg_group<V, K, TE, QE, true>(p, qv, smem, slut, bnd, g, half, lane, &pv);
            } else {
                sg_group<V, K, TE, QE, PAIRED>(p, qv, smem, slut, bnd, g, half, lane, &pv);
            }
        }
        __syncthreads(); // everyone is done with s_tile and the profile
    }
}

// ---------------------------------------------------------------------------
// Multi-query kernel over workgroup-shared paired profiles (packed int16):
// as sg_multi, but a tile's four groups step through its query's chunks in
// lockstep over one LDS copy of the chunk's pair rows (sg_tile_lockstep).
// LDS: the query's pair table (2 KB), then the pair rows of one chunk
// (p.pnrows is the largest class count of the launch's queries).
// ---------------------------------------------------------------------------
template <int K, bool TE, bool QE>
__global__ void __launch_bounds__(256) sg_multi_lock(ScoreParams p) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    __shared__ int s_tile[3]; // query, first wave item, items of that query
    const int lane = threadIdx.x & 63;
    const int wave = (int)wave_broadcast0(threadIdx.x >> 6); // wave-uniform
    constexpr int waves = 4; // the launch has 256 threads
    uint2 *bnd = p.bnd + (size_t)(blockIdx.x * waves + wave) * p.slot_stride;
    const uint32_t *ptab = (const uint32_t *)smem;
    uint8_t *area = smem + kPairTableLds;
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
        const int q = (int)wave_broadcast0((uint32_t)s_tile[0]);
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
        qv.fshift = qd.fshift;
        const int item = (int)wave_broadcast0((uint32_t)s_tile[1]) + wave;
        const bool valid = item < (int)wave_broadcast0((uint32_t)s_tile[2]);
        // the query's groups all lie within its binary16 range bound, or
        // all run in int16 (a wave-uniform choice per query)
        if (qd.f16)
            sg_tile_lockstep<K, TE, QE, true>(p, qv, ptab, area, qd.pprof, qd.pnrows * qd.pnrows, qd.pqpad,
                                              bnd, qd.g_begin + item, valid, lane);
        else
            sg_tile_lockstep<K, TE, QE, false>(p, qv, ptab, area, qd.pprof, qd.pnrows * qd.pnrows, qd.pqpad,
                                               bnd, qd.g_begin + item, valid, lane);
        __syncthreads(); // everyone is done with s_tile and the pair rows
    }
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

} // namespace ish
