// Device helpers shared by the scoring kernels (semi-global and local).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ish {

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
// Two cells in binary16 halves. Integers of magnitude <= 2048 are exact in
// binary16 and so are sums and maxima whose result stays in that range, so
// a recurrence whose values are proven to stay inside it computes exactly
// what the int16 form computes. gfx950's packed three-input maximum
// (v_pk_maximum3_f16) then takes H = max(Hdiag + S, E, F) in one operation.
typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ s16x2 vmax(s16x2 a, s16x2 b) {
    return __builtin_elementwise_max(a, b);
}
__device__ __forceinline__ int vmax(int a, int b) { return a > b ? a : b; }
// IEEE maximum (the values are never NaN); nested, it selects the
// three-input form.
__device__ __forceinline__ h16x2 vmax(h16x2 a, h16x2 b) {
    return __builtin_elementwise_maximum(a, b);
}
__device__ __forceinline__ h16x2 vmax3(h16x2 a, h16x2 b, h16x2 c) {
    return __builtin_elementwise_maximum(__builtin_elementwise_maximum(a, b), c);
}
__device__ __forceinline__ s16x2 vmax3(s16x2 a, s16x2 b, s16x2 c) { return vmax(vmax(a, b), c); }

template <typename V> __device__ __forceinline__ V splat(int v);
template <> __device__ __forceinline__ s16x2 splat<s16x2>(int v) {
    s16x2 r;
    r.x = (short)v;
    r.y = (short)v;
    return r;
}
template <> __device__ __forceinline__ int splat<int>(int v) { return v; }
template <> __device__ __forceinline__ h16x2 splat<h16x2>(int v) {
    const _Float16 h = (_Float16)v; // exact: |v| <= 2048 wherever it is used
    h16x2 r;
    r.x = h;
    r.y = h;
    return r;
}

__device__ __forceinline__ uint32_t as_u32(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t as_u32(int v) { return (uint32_t)v; }
__device__ __forceinline__ uint32_t as_u32(h16x2 v) { return __builtin_bit_cast(uint32_t, v); }
template <typename V> __device__ __forceinline__ V from_u32(uint32_t v);
template <> __device__ __forceinline__ s16x2 from_u32<s16x2>(uint32_t v) {
    return __builtin_bit_cast(s16x2, v);
}
template <> __device__ __forceinline__ int from_u32<int>(uint32_t v) { return (int)v; }
template <> __device__ __forceinline__ h16x2 from_u32<h16x2>(uint32_t v) {
    return __builtin_bit_cast(h16x2, v);
}

// A wave-uniform value as a cell pair. Binary16 lanes convert on the VALU;
// the empty asm keeps the conversion where the value is used, since the
// compiler would otherwise hoist a chunk's loop-invariant boundary values
// out of the pipeline's step loop and keep them all in vector registers.
template <typename V> __device__ __forceinline__ V splat_uniform(int v) {
    if constexpr (__is_same(V, h16x2)) asm volatile("" : "+s"(v));
    return splat<V>(v);
}

// The int16 pair of a cell pair: binary16 halves hold value - shift (exact
// integers), the int16 form the value itself.
__device__ __forceinline__ s16x2 to_s16x2(s16x2 v, int) { return v; }
__device__ __forceinline__ s16x2 to_s16x2(h16x2 v, int shift) {
    s16x2 r;
    r.x = (short)v.x; // exact integers of magnitude <= 2048
    r.y = (short)v.y;
    return r + splat<s16x2>(shift);
}

// A word of two int16 profile scores as the binary16 pair of the same
// integers (|score| <= 2048).
__device__ __forceinline__ uint32_t i16x2_to_h16x2(uint32_t w) {
    const s16x2 s = from_u32<s16x2>(w);
    h16x2 h;
    h.x = (_Float16)s.x;
    h.y = (_Float16)s.y;
    return as_u32(h);
}

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
template <> __device__ __forceinline__ h16x2 pack2<h16x2>(int a, int b) {
    h16x2 r;
    r.x = (_Float16)(short)a; // exact: |a|, |b| <= 2048 wherever it is used
    r.y = (_Float16)(short)b;
    return r;
}

// Stage the query profile and the residue -> profile-row table into LDS.
// The table holds each row's offset in dwords (256 x uint16 in memory,
// widened to uint32 in LDS), so a residue code turns into an LDS address
// with one load and one shift-add.
__device__ __forceinline__ void stage_profile(const int16_t *prof, const uint16_t *lut,
                                              int prof_words, uint8_t *smem) {
    const int *src = (const int *)prof;
    int *dst = (int *)smem;
    for (int x = threadIdx.x; x < prof_words; x += blockDim.x) dst[x] = src[x];
    uint32_t *ldst = (uint32_t *)(smem + (size_t)prof_words * 4);
    for (int x = threadIdx.x; x < 256; x += blockDim.x) ldst[x] = lut[x];
    __syncthreads();
}

typedef short short8 __attribute__((ext_vector_type(8)));

// Makes v be computed at this point: an empty asm that reads and writes
// it, past which the compiler cannot sink the computation.
template <typename V> __device__ __forceinline__ void pin(V &v) {
    uint32_t u = as_u32(v);
    asm volatile("" : "+v"(u));
    v = from_u32<V>(u);
}

// Profile value of row mm (0..7) of an 8-row slice: PAIRED, the packed
// word of rows mm (sa: rows 0..3, sb: rows 4..7); packed int16, the halves
// of the lane's two targets' slices; int32, the one slice's entry.
template <typename V, bool PAIRED, bool PACKED>
__device__ __forceinline__ V slice_val(const short8 &sa, const short8 &sb, int mm) {
    V s;
    if constexpr (__is_same(V, h16x2)) {
        // PAIRED only: the slice words hold binary16 pairs
        const short8 &src = mm < 4 ? sa : sb;
        s = from_u32<V>(__builtin_bit_cast(u32x4, src)[mm & 3]);
    } else if constexpr (PAIRED) {
        const short8 &src = mm < 4 ? sa : sb;
        s.x = src[2 * (mm & 3)];
        s.y = src[2 * (mm & 3) + 1];
    } else if constexpr (PACKED) {
        s.x = sa[mm];
        s.y = sb[mm];
    } else {
        s = (int)sa[mm];
    }
    return s;
}

// Paired profile (small alphabets): one row per pair of profile classes
// (class of a lane's target A, class of its target B) holding, per query
// row j, the int32 word (S_A[j] + bias) | (S_B[j] + bias) << 16, so that a
// packed cell's operand is one LDS word. A wave copies the K words of its
// current chunk of every pair row into its own LDS buffer, rows
// pair_row_bytes(K) apart (16 bytes over K words, so that rows start 16
// bytes apart modulo 128 and the 16-byte slices of different rows spread
// over the LDS banks). The pair table maps a residue code to the byte
// offset of its class's rows: for target A at [0, 256), for target B at
// [256, 512). pair_row_bytes is in engine.h.
struct PairView {
    const int32_t *pprof; // npairs rows of pqpad words
    int32_t npairs, pqpad;
    uint8_t *pbuf;        // this wave's LDS copy of one chunk
};

// Copies rows j0 .. j0+K-1 of every pair row into the wave's buffer. The
// previous chunk's reads of the buffer were issued earlier, and a wave's
// LDS accesses complete in order.
// F16: the copy holds each word's two scores as a binary16 pair.
template <int K, bool F16 = false>
__device__ __forceinline__ void stage_pairs(const PairView &pv, int j0, int lane) {
    constexpr int parts = K / 4; // 16-byte parts per row
    for (int x = lane; x < pv.npairs * parts; x += 64) {
        const int pr = x / parts, part = x % parts;
        uint4 v = *(const uint4 *)(pv.pprof + (size_t)pr * pv.pqpad + j0 + part * 4);
        if constexpr (F16) {
            v.x = i16x2_to_h16x2(v.x);
            v.y = i16x2_to_h16x2(v.y);
            v.z = i16x2_to_h16x2(v.z);
            v.w = i16x2_to_h16x2(v.w);
        }
        *(uint4 *)(pv.pbuf + pr * pair_row_bytes(K) + part * 16) = v;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// Pair row of the residue codes (a, b) in a wave's buffer.
__device__ __forceinline__ const short8 *pair_row(const uint8_t *pbuf, const uint32_t *plut,
                                                  uint32_t a, uint32_t b) {
    return (const short8 *)(pbuf + plut[a] + plut[256 + b]);
}

// Stages the pair table of tile height K (2 x 256 uint16 in memory) into
// smem as 512 uint32 entries (kPairTableLds bytes) and waits for the
// workgroup.
template <int K>
__device__ __forceinline__ void stage_pair_table(const uint16_t *plut, uint8_t *smem) {
    const uint16_t *src = plut + pair_table_index(K) * 512;
    for (int x = threadIdx.x; x < 512; x += blockDim.x) ((uint32_t *)smem)[x] = src[x];
    __syncthreads();
}

} // namespace ish
