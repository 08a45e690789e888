// Device helpers shared by the scoring kernels (semi-global and local).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ish {

typedef short s16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s16x2 vmax(s16x2 a, s16x2 b) {
    return __builtin_elementwise_max(a, b);
}
__device__ __forceinline__ int vmax(int a, int b) { return a > b ? a : b; }

template <typename V> __device__ __forceinline__ V splat(int v);
template <> __device__ __forceinline__ s16x2 splat<s16x2>(int v) {
    s16x2 r;
    r.x = (short)v;
    r.y = (short)v;
    return r;
}
template <> __device__ __forceinline__ int splat<int>(int v) { return v; }

__device__ __forceinline__ uint32_t as_u32(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t as_u32(int v) { return (uint32_t)v; }
template <typename V> __device__ __forceinline__ V from_u32(uint32_t v);
template <> __device__ __forceinline__ s16x2 from_u32<s16x2>(uint32_t v) {
    return __builtin_bit_cast(s16x2, v);
}
template <> __device__ __forceinline__ int from_u32<int>(uint32_t v) { return (int)v; }

__device__ __forceinline__ uint32_t wave_broadcast0(uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}

// Minimum of v over the 64 lanes of the wave, returned wave-uniform.
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = min(v, __shfl_xor(v, d));
    return __builtin_amdgcn_readfirstlane(v);
}

// (a, b) in the two int16 halves; the int32 form keeps a only.
template <typename V> __device__ __forceinline__ V pack2(int a, int b);
template <> __device__ __forceinline__ s16x2 pack2<s16x2>(int a, int b) {
    s16x2 r;
    r.x = (short)a;
    r.y = (short)b;
    return r;
}
template <> __device__ __forceinline__ int pack2<int>(int a, int) { return a; }

// Stage the query profile and the residue -> profile-row table into LDS.
// The table holds each row's offset in dwords (256 x uint16), so a residue
// code turns into an LDS address with one shift-add.
__device__ __forceinline__ void stage_profile(const int16_t *prof, const uint16_t *lut,
                                              int prof_words, uint8_t *smem) {
    const int *src = (const int *)prof;
    int *dst = (int *)smem;
    for (int x = threadIdx.x; x < prof_words; x += blockDim.x) dst[x] = src[x];
    const int *lsrc = (const int *)lut;
    int *ldst = (int *)(smem + (size_t)prof_words * 4);
    for (int x = threadIdx.x; x < 128; x += blockDim.x) ldst[x] = lsrc[x];
    __syncthreads();
}

typedef short short8 __attribute__((ext_vector_type(8)));

} // namespace ish
