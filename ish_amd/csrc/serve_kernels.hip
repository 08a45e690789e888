// Device-side emission for the serving loop: applies the matchers' --score
// rule to the results of a multi-query pass and compacts the emitted
// targets into per-query hit lists, so that only hits cross PCIe.
//
// The rule is the float32 one of the reference matchers:
//   semi-global (striped_semi_global_matcher.mojo:77-107):
//     reject if f32(score) < f32(max_score) * thr, then emit iff
//     f32(score) / f32(max_score) >= thr;
//   local (striped_local_matcher.mojo:61-90, local_aln/striped.mojo:213-230):
//     reject unless f32(score) > thr * f32(max_score), end_ref > 0 and
//     end_query > 0, then emit iff f32(score) / f32(max_score) >= thr.
// __fmul_rn / __fdiv_rn give IEEE round-to-nearest products and quotients,
// the arithmetic of the host rule.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine.h"

namespace ish {

__global__ void __launch_bounds__(256) emit_hits(EmitParams p) {
    const int k = blockIdx.y;
    const float ms = p.max_score[k];
    const float thr = p.threshold;
    const int64_t base = (int64_t)k * p.n;
    uint32_t *cnt = p.count + k;
    int4 *hits = p.hits + (int64_t)k * p.cap;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < p.n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t sc = p.score[base + t];
        const float s = (float)sc;
        bool ok;
        if (p.algo == 0) {
            ok = !(s < __fmul_rn(ms, thr));
        } else {
            ok = s > __fmul_rn(thr, ms) && p.end_ref[base + t] > 0 && p.end_q[base + t] > 0;
        }
        if (ok) ok = __fdiv_rn(s, ms) >= thr;
        if (ok) {
            const uint32_t slot = atomicAdd(cnt, 1u);
            if ((int64_t)slot < p.cap)
                hits[slot] = make_int4((int)t, sc, p.end_ref[base + t], p.end_q[base + t]);
        }
    }
}

__global__ void __launch_bounds__(256) fill_u32(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// In-place byte translation (ASCII -> residue codes) of a device buffer:
// 16 bytes per thread, grid-stride.
__global__ void __launch_bounds__(256) encode_bytes(uint8_t *p, uint64_t n, EncodeTable tp) {
    __shared__ EncodeTable t;
    if (threadIdx.x < 64) ((uint32_t *)t.code)[threadIdx.x] = ((const uint32_t *)tp.code)[threadIdx.x];
    __syncthreads();
    const uint64_t nv = n / 16;
    uint4 *v = (uint4 *)p;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint4 x = v[i];
        uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t a = w[k];
            w[k] = (uint32_t)t.code[a & 0xffu] | ((uint32_t)t.code[(a >> 8) & 0xffu] << 8) |
                   ((uint32_t)t.code[(a >> 16) & 0xffu] << 16) | ((uint32_t)t.code[a >> 24] << 24);
        }
        v[i] = make_uint4(w[0], w[1], w[2], w[3]);
    }
    const uint64_t tail = nv * 16 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tail < n && tail < nv * 16 + 16) p[tail] = t.code[p[tail]];
}

// Which residue codes occur in p[0, n): bit c of mask[c >> 5] is set when
// code c does (mask zeroed by the caller). Block of 256 threads.
__global__ void __launch_bounds__(256) code_mask(const uint8_t *p, uint64_t n, uint32_t *mask) {
    __shared__ uint8_t seen[256];
    seen[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t nv = n / 16;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint4 *v = (const uint4 *)p;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
        const uint4 x = v[i];
        const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int b = 0; b < 4; ++b) seen[(w[k] >> (8 * b)) & 0xffu] = 1;
    }
    for (uint64_t t = nv * 16 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride)
        seen[p[t]] = 1;
    __syncthreads();
    // thread t reports code t: one 64-bit ballot per wave covers 64 codes
    const uint64_t bits = __builtin_amdgcn_ballot_w64(seen[threadIdx.x] != 0);
    if ((threadIdx.x & 63) == 0) {
        const int w = threadIdx.x >> 6;
        if ((uint32_t)bits) atomicOr(&mask[2 * w], (uint32_t)bits);
        if ((uint32_t)(bits >> 32)) atomicOr(&mask[2 * w + 1], (uint32_t)(bits >> 32));
    }
}

hipError_t launch_code_mask(const uint8_t *p, uint64_t n, uint32_t *mask, hipStream_t st) {
    hipLaunchKernelGGL(fill_u32, dim3(1), dim3(256), 0, st, mask, (int64_t)8, 0u);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || n == 0) return e;
    uint64_t b = (n / 16 + 255) / 256;
    if (b < 1) b = 1;
    if (b > 2048) b = 2048;
    hipLaunchKernelGGL(code_mask, dim3((unsigned)b), dim3(256), 0, st, p, n, mask);
    return hipGetLastError();
}

hipError_t launch_encode_bytes(uint8_t *p, uint64_t n, const EncodeTable &t, hipStream_t st) {
    if (n == 0) return hipSuccess;
    uint64_t b = (n / 16 + 255) / 256;
    if (b < 1) b = 1;
    if (b > 4096) b = 4096;
    hipLaunchKernelGGL(encode_bytes, dim3((unsigned)b), dim3(256), 0, st, p, n, t);
    return hipGetLastError();
}

hipError_t launch_emit_hits(const EmitParams &p, hipStream_t st) {
    if (p.n <= 0 || p.nq <= 0) return hipSuccess;
    int64_t bx = (p.n + 255) / 256;
    if (bx > 1024) bx = 1024;
    hipLaunchKernelGGL(emit_hits, dim3((unsigned)bx, (unsigned)p.nq), dim3(256), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    int64_t b = (n + 255) / 256;
    if (b > 1024) b = 1024;
    hipLaunchKernelGGL(fill_u32, dim3((unsigned)b), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

} // namespace ish
