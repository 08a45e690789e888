/*
 * CPU oracle for the ish alignment-scoring hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP engine
 * is compared against; it is never linked into, loaded by, or called from
 * the product path (ish_amd/). Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.
 *
 * Two independent restatements of the reference algorithm live here:
 *   - oracle.c : scalar int32 textbook Gotoh (semi-global) and Smith-Waterman
 *                with the reference's end-coordinate tie rules, reverse start
 *                pass and float32 emission rule.
 *   - striped.c: Farrar-striped int16 semi-global with the SWPS3 lazy-F loop
 *                and biased-unsigned u8->u16 striped SW, restating the
 *                reference's SIMD CPU kernels step by step. It is the timed
 *                CPU baseline and a differential check on oracle.c.
 *
 * Flag word for semi-global free ends (order of the --sg-ends-free letters,
 * ishlib/searcher_settings.mojo:404-407):
 *   bit0 query_start, bit1 query_end, bit2 target_start, bit3 target_end.
 *
 * Scoring lookup convention: S[r * A + q] with r the target residue code and
 * q the query residue code (ishlib/matcher/alignment/semi_global_aln/
 * striped.mojo:112-114 builds the profile from score_matrix.get(nt, query[j])).
 */
#ifndef ISH_ORACLE_H
#define ISH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_QS 1u
#define ORACLE_QE 2u
#define ORACLE_TS 4u
#define ORACLE_TE 8u

typedef struct {
    int32_t score;
    int32_t end_ref;   /* 0-based inclusive target end (-1 if none) */
    int32_t end_query; /* 0-based inclusive query end */
} oracle_end_t;

typedef struct {
    int32_t score;
    int32_t query_start, query_end;
    int32_t target_start, target_end;
} oracle_start_end_t;

typedef struct {
    int32_t score1;
    int32_t ref_begin1, ref_end1;
    int32_t read_begin1, read_end1;
} oracle_local_t;

/* ---- tables (oracle/tables.c) ------------------------------------------ */
/* matrix kinds: 0 ascii, 1 actgn, 2 actgn0, 3 blosum62
 * (ishlib/matcher/alignment/scoring_matrix.mojo:176-254) */
int32_t oracle_alphabet_size(int32_t kind);
/* Copies the A*A int8 matrix (65536 bytes for ascii). Returns A or -1. */
int32_t oracle_matrix(int32_t kind, int8_t *out);
/* Copies the 256-entry ascii->code table. Returns 0 or -1. */
int32_t oracle_encoding(int32_t kind, uint8_t *out);
/* Copies the A-entry code->ascii table. Returns A or -1. */
int32_t oracle_decoding(int32_t kind, uint8_t *out);
/* Encodes n bytes in place-free fashion and returns max_score = sum of the
 * matrix diagonal over the encoded bytes
 * (scoring_matrix.mojo:403-416 convert_ascii_to_encoding_and_score). */
int64_t oracle_encode_and_score(int32_t kind, const uint8_t *ascii, int64_t n,
                                uint8_t *out);

/* ---- scalar int32 oracle (oracle/oracle.c) ----------------------------- */
oracle_end_t oracle_sg(const uint8_t *q, int32_t Q, const uint8_t *r,
                       int32_t R, const int8_t *S, int32_t A, int32_t go,
                       int32_t ge, uint32_t flags);
/* Forward pass, cutoff test, reverse pass for the start coordinate.
 * Returns 1 and fills *out when f32(score) >= cutoff, else 0. */
int32_t oracle_sg_start_end(const uint8_t *q, int32_t Q, const uint8_t *r,
                            int32_t R, const int8_t *S, int32_t A, int32_t go,
                            int32_t ge, uint32_t flags, float cutoff,
                            oracle_start_end_t *out);
oracle_end_t oracle_sw(const uint8_t *q, int32_t Q, const uint8_t *r,
                       int32_t R, const int8_t *S, int32_t A, int32_t go,
                       int32_t ge);
/* ssw_align: forward, cutoff (strict <=), end>0 guard, reverse begin pass.
 * Returns 1 and fills *out when an Alignment is produced, else 0. */
int32_t oracle_sw_align(const uint8_t *q, int32_t Q, const uint8_t *r,
                        int32_t R, const int8_t *S, int32_t A, int32_t go,
                        int32_t ge, float cutoff, oracle_local_t *out);

/* Emission decisions (float32 arithmetic throughout). */
int32_t oracle_sg_emit(int32_t score, int64_t max_score, float threshold);
int32_t oracle_sw_emit(int32_t score, int32_t end_ref, int32_t end_query,
                       int64_t max_score, float threshold);

/* Batch drivers over targets given as residues + offsets[n+1].
 * algo 0 = semi-global, 1 = local. nthreads <= 0 means all cores. */
int32_t oracle_batch(int32_t algo, const uint8_t *q, int32_t Q,
                     const uint8_t *residues, const uint64_t *offsets,
                     int64_t n, const int8_t *S, int32_t A, int32_t go,
                     int32_t ge, uint32_t flags, int32_t *score,
                     int32_t *end_ref, int32_t *end_query, int32_t nthreads);

/* ---- Farrar-striped restatement (oracle/striped.c) --------------------- */
/* Striped int16 semi-global, width W lanes (8, 16 or 32; 0 = native). Follows
 * semi_global_aln (striped.mojo:252-498) including its lazy-F loop and its
 * strided visiting order in the free-query-end pass. */
oracle_end_t oracle_sg_striped(const uint8_t *q, int32_t Q, const uint8_t *r,
                               int32_t R, const int8_t *S, int32_t A,
                               int32_t go, int32_t ge, uint32_t flags,
                               int32_t W);
/* Striped biased-unsigned SW: u8 pass, u16 rerun when best+bias >= 255
 * (local_aln/striped.mojo:142-190, :272-434). */
oracle_end_t oracle_sw_striped(const uint8_t *q, int32_t Q, const uint8_t *r,
                               int32_t R, const int8_t *S, int32_t A,
                               int32_t go, int32_t ge, int32_t W);
/* Batch driver for the striped restatement at the host's native int16
 * width (32 lanes with AVX-512BW, else 16): the CPU baseline that is timed.
 * The query profile is built once per call, as the reference's matcher
 * does. Returns the number of threads used. */
int32_t oracle_batch_striped(int32_t algo, const uint8_t *q, int32_t Q,
                             const uint8_t *residues, const uint64_t *offsets,
                             int64_t n, const int8_t *S, int32_t A, int32_t go,
                             int32_t ge, uint32_t flags, int32_t *score,
                             int32_t *end_ref, int32_t *end_query,
                             int32_t nthreads);

/* Test hooks of the striped restatement (striped.cpp). oracle_sat_lanes:
 * r = a + b (op 0) or a - b (op 1) lane by lane with the saturating lanes of
 * `type` (0 uint8, 1 int16, 2 uint16) at `width` lanes (8, 16, 32 or 64);
 * returns 0, or -1 for an unknown pair. oracle_local_profile: the biased
 * profile ssw_align builds (local_aln/striped.mojo:110-139) for `type`
 * (0 uint8, 2 uint16) at `width` lanes, written vector by vector, lane by
 * lane (A * seg * width entries); returns seg, or -1. */
int32_t oracle_sat_lanes(int32_t type, int32_t width, int32_t op, const int32_t *a,
                         const int32_t *b, int32_t *r);
int32_t oracle_local_profile(const uint8_t *q, int32_t Q, const int8_t *S, int32_t A,
                             int32_t bias, int32_t type, int32_t width, int32_t *out);

#ifdef __cplusplus
}
#endif
#endif
