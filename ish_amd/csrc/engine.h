// Internal declarations shared by the host side (ish_hip.cpp) and the HIP
// kernels of the ish_hip engine. Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ish {

// Targets per group of the device-resident layout: 64 lanes x 2 targets.
constexpr int kGroup = 128;
// Lanes per wavefront on CDNA.
constexpr int kWave = 64;

// One group of 128 length-sorted targets. Its residues are stored as
// ncols/4 blocks of 64 uint2: block b, lane l holds columns 4b..4b+3 of
// target `first + l` in .x and of target `first + 64 + l` in .y (little
// endian, column 4b in the low byte). A wave reading one block issues one
// coalesced 512-byte load.
struct GroupDesc {
    uint64_t res_off; // offset of the group's first block, in uint2 units
    int32_t ncols;    // longest target of the group rounded up to 4
    int32_t first;    // index of the group's first target in sorted arrays
};

// One query of a multi-query launch.
struct QueryDesc {
    const int16_t *prof;  // offset-form profile (staged into LDS)
    const uint16_t *lut;  // residue code -> profile row offset (dwords)
    int32_t prof_words;
    int32_t Q, nchunks, neg_limit;
    int32_t g_begin, g_end; // groups of this launch for this query
    uint32_t flags;         // free ends of this query: qs|qe<<1|ts<<2|te<<3
    int32_t pad, front;     // top padding rows and profile front (ScoreParams)
    int32_t f16, fshift;    // lockstep tiles: binary16 lanes, the shift of their values
    int64_t tile_base;      // first work tile of this query (prefix sum)
    int64_t out_off;        // offset of this query's results in the outputs
    // paired profile of this query (paired launches; see ScoreParams)
    const int32_t *pprof;
    const uint16_t *plut;
    int32_t pnrows, pqpad;
};

// Everything a semi-global / local scoring launch needs.
struct ScoreParams {
    // query profile (staged into LDS): nrows rows of `stride` int16 scores,
    // row r, query position j at prof[r * stride + j]; lut maps a residue
    // code to the dword offset of its profile row (r * stride / 2).
    const int16_t *prof;
    const uint16_t *lut;
    int32_t prof_words; // int32 words to stage for the profile
    int32_t stride;     // int16 elements per profile row (16-byte multiple + 16)
    int32_t Q;          // query length
    int32_t nchunks;    // ceil(Q / K)
    // semi-global top padding (see top_pad): chunk c holds query rows
    // c*K - pad .. c*K - pad + K - 1, rows below 0 being padding rows;
    // query row j of profile row r is at prof[r * stride + front + j]
    int32_t pad, front;
    int32_t go, ge;     // gap open / extend (non-negative)
    uint32_t flags;     // free ends: qs|qe<<1|ts<<2|te<<3
    int32_t neg_limit;  // int16 floor that seeds the last-row maximum
    int32_t fshift;     // binary16 lanes: every value is held minus fshift

    // device-resident targets
    const GroupDesc *groups;
    const uint2 *res;
    const int32_t *lens; // sorted order
    const int32_t *orig; // sorted order -> original index (-1 = padding)
    int32_t g_begin, g_end;

    // work queue counter (one uint32, zeroed before each launch)
    uint32_t *counter;

    // per-wave boundary rows: slot s owns bnd[s*slot_stride ...], column i,
    // lane l at [i*64 + l]
    uint2 *bnd;
    int64_t slot_stride;

    // outputs, original target order
    int32_t *score, *end_ref, *end_q;
    uint8_t *status;

    // multi-query launches: nqueries entries plus a sentinel whose
    // tile_base is the total number of tiles
    const QueryDesc *queries;
    int32_t nqueries;

    // hybrid launches: the first n_long groups of [g_begin, g_end) are
    // scored by whole workgroups
    int32_t n_long;

    // paired local profile (small alphabets), used by the paired pipeline:
    // pprof[pair * pqpad + j] = (S[A][j] + ge) | (S[B][j] + ge) << 16 for
    // the profile classes A, B of pair = A * pnrows + B (0 past the query);
    // plut = byte offsets of a class's pair rows in a wave's LDS copy, for
    // tiles 16, 32 and 40 (pair_table_index), each as [A: 256][B: 256]
    // uint16
    const int32_t *pprof;
    const uint16_t *plut;
    int32_t pnrows, pqpad;
};

// Plain (one launch per batch) parameters: targets in input order,
// contiguous residues with offsets.
struct PlainParams {
    const int16_t *prof;
    const uint16_t *lut;
    int32_t prof_words, stride, Q, nchunks, go, ge;
    uint32_t flags;
    int32_t neg_limit;
    const uint8_t *res;
    const uint64_t *off;
    int64_t n;
    int2 *bnd; // per wave slot: [i*64 + lane]
    int64_t slot_stride;
    int32_t *score, *end_ref, *end_q;
    uint8_t *status;
};

// Launchers (sg_*.hip, sw_kernels.hip). Return hipError_t of the launch.
hipError_t launch_reset(uint32_t *counter, hipStream_t st);
hipError_t launch_sg_batched_i16(const ScoreParams &p, int chunk, int grid,
                                 int block, size_t lds, hipStream_t st);
hipError_t launch_sg_batched_i32(const ScoreParams &p, int chunk, int grid,
                                 int block, size_t lds, hipStream_t st);
hipError_t launch_sw_batched_i16(const ScoreParams &p, int chunk, int grid,
                                 int block, size_t lds, hipStream_t st);
hipError_t launch_sw_batched_i32(const ScoreParams &p, int chunk, int grid,
                                 int block, size_t lds, hipStream_t st);
int occupancy_sw_batched_i16(int chunk, int block, size_t lds);
// Paired profiles (both algorithms): the bytes of one pair row of a wave's
// LDS copy of a chunk (K int32 words plus 16 bytes), and the largest class
// count that gets one.
constexpr int pair_row_bytes(int K) { return 4 * K + 16; }
// LDS bytes of the lookup tables, held as uint32 entries so that a lookup
// is one zero-extended load: the residue-code -> profile-row table (256
// dwords) after the int16 profile, and the pair table (2 x 256 dwords)
// at the start of a paired launch's LDS.
constexpr size_t kLutLds = 256 * 4;
constexpr size_t kPairTableLds = 512 * 4;
constexpr int kPairMaxRows = 8;
// Hybrid semi-global launch (packed int16, `waves` = 2 or 4 waves per
// workgroup): the first n_long groups are pipelined across a workgroup's
// waves, the others run as tiles of one group per wave. `paired` selects
// the variant over the paired profile (p.pprof, p.plut), whose whole LDS
// size is sg_hybrid_paired_lds(chunk, classes, waves).
size_t sg_pipe_extra_lds();
size_t sg_hybrid_paired_lds(int chunk, int nrows, int waves);
// f16 (paired only): binary16 lanes for every group of the launch.
hipError_t launch_sg_hybrid_i16(const ScoreParams &p, int chunk, bool paired, bool f16, int waves, int grid,
                                size_t lds, hipStream_t st);
int occupancy_sg_hybrid_i16(int chunk, bool paired, bool f16, int waves, size_t lds, uint32_t flags);
// Hybrid launch over a workgroup-shared paired profile (up to
// kPairMaxShared classes; tiles of four groups in lockstep, long groups over
// the int16 profile). Its LDS is sg_hybrid_lock_lds(chunk, classes, the int16
// profile's LDS). The pair table holds uint16 byte offsets: with 21 classes
// and K = 32 the last pair row starts at (21 * 21 - 1) * 144 = 63,360 bytes.
constexpr int kPairMaxShared = 21;
size_t sg_hybrid_lock_lds(int chunk, int nrows, size_t profile_lds);
hipError_t launch_sg_hybrid_lock(const ScoreParams &p, int chunk, bool f16, int grid, size_t lds,
                                 hipStream_t st);
int occupancy_sg_hybrid_lock(int chunk, bool f16, size_t lds, uint32_t flags);
// Workgroup-pipelined local kernel (packed int16, 4 waves share a group):
// extra LDS beyond the profile, launcher (block of 256) and occupancy.
// `paired` selects the variant that reads the paired profile (p.pprof,
// p.plut) instead of staging the int16 profile; its whole LDS size is
// sw_pipe_paired_lds(chunk, profile classes).
size_t sw_pipe_extra_lds();
size_t sw_pipe_paired_lds(int chunk, int nrows);
// f16 (paired only): binary16 lanes for every group of the launch.
hipError_t launch_sw_pipe_i16(const ScoreParams &p, int chunk, bool paired, bool f16, int grid, size_t lds,
                              hipStream_t st);
int occupancy_sw_pipe_i16(int chunk, bool paired, bool f16, size_t lds);
int occupancy_sw_batched_i32(int chunk, int block, size_t lds);
hipError_t launch_sg_plain(const PlainParams &p, int grid, int block,
                           size_t lds, hipStream_t st);
// Packs the length-sorted targets into the group layout (total_blocks =
// the blocks of all groups together). raw must hold kRawPad bytes past the
// last residue: the packer reads whole aligned dwords.
constexpr uint64_t kRawPad = 16;
hipError_t launch_pack(const uint8_t *raw, const uint64_t *off,
                       const int32_t *sorted_idx, const GroupDesc *groups,
                       int32_t ngroups, int64_t n, uint64_t total_blocks, uint2 *res,
                       hipStream_t st);
hipError_t launch_gather_reversed(const uint8_t *raw, const uint64_t *src_off,
                                  const int64_t *sel, const uint64_t *dst_off,
                                  int64_t n_sel, uint8_t *out, hipStream_t st);
// Multi-query launch. p.flags selects the kernel variant: its te / qe bits
// must be the union over the launch's queries, and either every query has
// te or qe free or none has (the per-query bits come from QueryDesc.flags).
// `paired` (packed only): every query reads its paired profile
// (QueryDesc.pprof / plut) and p.pnrows is the largest class count; the
// LDS is then 1 KB plus 4 waves x pnrows^2 pair rows.
hipError_t launch_sg_multi(const ScoreParams &p, bool packed, bool paired, int chunk, int grid,
                           int block, size_t lds, hipStream_t st);
int occupancy_sg_multi(bool packed, bool paired, int chunk, int block, size_t lds, uint32_t flags);
// Multi-query launch over workgroup-shared paired profiles (packed, queries
// of up to kPairMaxShared classes, p.pnrows the largest): LDS 1 KB plus
// pnrows^2 pair rows of the tile height; block of 256.
hipError_t launch_sg_multi_lock(const ScoreParams &p, int chunk, int grid, size_t lds, hipStream_t st);
int occupancy_sg_multi_lock(int chunk, size_t lds, uint32_t flags);
// Emission at the --score threshold for the results of nq queries (query
// k's results at k * n): hit (target, score, end_ref, end_query) records go
// to hits[k * cap ...] and count[k] counts every emitted target (it may
// exceed cap; the caller then falls back to the full arrays).
struct EmitParams {
    const int32_t *score, *end_ref, *end_q;
    int64_t n;
    int32_t nq;
    int32_t algo;
    const float *max_score; // per query, float32(max_score)
    float threshold;
    int4 *hits;
    int64_t cap;
    uint32_t *count;
};
hipError_t launch_emit_hits(const EmitParams &p, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t st);
// Byte -> residue code table, passed by value to the encoding kernel.
struct EncodeTable {
    uint8_t code[256];
};
// p[i] = t.code[p[i]] for i < n (p 16-byte aligned).
hipError_t launch_encode_bytes(uint8_t *p, uint64_t n, const EncodeTable &t, hipStream_t st);
// Sets bit c of mask[c >> 5] (8 words, cleared first) for every residue code
// c that occurs in p[0, n).
hipError_t launch_code_mask(const uint8_t *p, uint64_t n, uint32_t *mask, hipStream_t st);

// Occupancy helpers: blocks per CU of each kernel at the given LDS size.
int occupancy_sg_batched_i16(int chunk, int block, size_t lds, uint32_t flags);
int occupancy_sg_batched_i32(int chunk, int block, size_t lds, uint32_t flags);
int occupancy_sg_plain(int block, size_t lds, uint32_t flags);

// Rows of the query held in registers per lane (chunk height). The batched
// kernels are compiled for 16 and 32, the hybrid semi-global kernel also
// for 40 (a 150-row query is then four chunks, one round of a four-wave
// pipeline). Profiles are padded to a multiple of both kChunkAlign and
// kChunkAlign40, so that every tile height's last chunk lies inside them.
constexpr int kChunkAlign = 32;
constexpr int kChunkAlign40 = 40;
constexpr int kChunkDefault = 32;
constexpr int kChunkPlain = 32;
// Top padding of semi-global queries (used when go >= ge): pad rows above
// query row 0 make the query a whole number of K-row chunks, so that row Q-1
// is the last register row of the last chunk and the kernels read it
// without a select. The profile holds the padding rows' entries in front of
// row 0: front = 40 + (-Q mod 8) entries, at least the 39 rows a 40-row
// tile pads and such that every chunk start (front + c*K - pad) is a
// multiple of 8 entries, i.e. 16-byte aligned, for every tile height. Their
// value is ge when the target start is free and 0 otherwise: with go >= ge
// the padding rows then repeat the top boundary row (H^ = (i-1)*ge, resp.
// -go-ge), so the rows below see exactly the boundary they would without
// them.
constexpr int kTopFront = 40;
constexpr int top_front(int Q) { return kTopFront + (8 - Q % 8) % 8; }
constexpr int top_pad(int Q, int K) { return (K - Q % K) % K; }
// Pair tables are stored per tile height, 512 uint16 each, in this order
// (24 is a height of the workgroup-shared lockstep launch only).
constexpr int kPairTables = 4;
constexpr int pair_table_index(int K) { return K == 16 ? 0 : K == 32 ? 1 : K == 40 ? 2 : 3; }
constexpr int pair_table_rows(int index) { return index == 0 ? 16 : index == 1 ? 32 : index == 2 ? 40 : 24; }

// Smallest slice of a one-shot batch cut into input-order slices (each
// uploaded while the slice before is scored).
constexpr int64_t kSliceMin = (int64_t)64 << 20;

} // namespace ish
