// The plain one-launch-per-batch int32 semi-global kernel (the comparison
// baseline), database group packing, the reversed-target gather of the
// start pass and the queue-counter reset, with their launchers.
#include "sg_device.h"

namespace ish {

// ---------------------------------------------------------------------------
// Plain kernel: textbook int32 recurrence, one target per lane, input order,
// grid-stride over targets (one launch per batch).
// ---------------------------------------------------------------------------
template <int K, bool TE, bool QE>
__global__ void __launch_bounds__(256) sg_plain(PlainParams p) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    stage_profile(p.prof, p.lut, p.prof_words, smem);
    const int16_t *sprof = (const int16_t *)smem;
    const uint32_t *slut = (const uint32_t *)(smem + (size_t)p.prof_words * 4);

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

int occupancy_sg_plain(int block, size_t lds, uint32_t flags) {
    const bool te = flags & 8u, qe = flags & 2u;
    if (te && qe) return occ(sg_plain<kChunkPlain, true, true>, block, lds);
    if (te) return occ(sg_plain<kChunkPlain, true, false>, block, lds);
    if (qe) return occ(sg_plain<kChunkPlain, false, true>, block, lds);
    return occ(sg_plain<kChunkPlain, false, false>, block, lds);
}

} // namespace ish
