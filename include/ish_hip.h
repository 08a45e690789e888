/*
 * ish_hip: MI355X (gfx950) alignment-scoring engine for ish, C ABI.
 *
 * Plain C types only, so that ish's Mojo host can bind it with
 * external_call / DLHandle (the same mechanism ish uses for libz,
 * ishlib/vendor/zlib.mojo:61-100).
 *
 * What it replaces in the reference (BioRadOpenSource/ish):
 *   - the GPU kernel entry GpuMatcher.batch_match_coarse
 *     (ishlib/matcher/__init__.mojo:96-121, implemented in
 *     ishlib/matcher/striped_semi_global_matcher.mojo:165-298 and
 *     ishlib/gpu/kernels/semi_global.mojo:20-115);
 *   - the device manager SearcherDevice (ishlib/gpu/searcher_device.mojo:
 *     98-519: create_devices, set_host_inputs, copy_inputs_to_device,
 *     launch_kernel, copy_outputs_to_host, synchronize);
 *   - the per-target CPU kernels behind Matcher.first_match for the data-
 *     parallel batch executor (ishlib/cpu/process_batch.mojo:14-34):
 *     semi_global_aln (semi_global_aln/striped.mojo:252-498) and ssw_align
 *     (local_aln/striped.mojo:142-269).
 *
 * Conventions
 *   - Return value 0 = OK, < 0 = error code (ISH_E_*). No exceptions cross
 *     the ABI. ish_hip_last_error(ctx) describes the last failure on ctx.
 *   - The caller owns every host array; arrays are only read (or written,
 *     for outputs) during the call.
 *   - A context owns its device memory, stream and tables. A context is
 *     single-thread-affine; distinct contexts may be used concurrently.
 *   - End coordinates are 0-based inclusive, like the reference's CPU
 *     striped kernels (AlignmentEnd, striped_utils.mojo:101-105). A
 *     MatchResult(start, end+1) is formed by the caller.
 *   - Residues are encoded codes (ish_hip_encoding), one byte each.
 */
#ifndef ISH_HIP_H
#define ISH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ISH_HIP_VERSION 10000 /* 1.0.0 */

/* error codes */
#define ISH_OK 0
#define ISH_E_INVALID (-1)   /* bad argument */
#define ISH_E_HIP (-2)       /* HIP runtime error */
#define ISH_E_NOMEM (-3)     /* allocation failed */
#define ISH_E_NO_QUERY (-4)  /* ish_hip_set_query not called */
#define ISH_E_RANGE (-5)     /* size beyond what the engine supports */
#define ISH_E_NO_DEVICE (-6) /* no gfx950 device */
#define ISH_E_TIMEOUT (-7)   /* ish_hip_server_wait: not done in time */

/* matrix kinds, MatrixKind (ishlib/matcher/alignment/scoring_matrix.mojo:
 * 176-254) */
#define ISH_MATRIX_ASCII 0
#define ISH_MATRIX_ACTGN 1
#define ISH_MATRIX_ACTGN0 2
#define ISH_MATRIX_BLOSUM62 3

/* --match-algo (ishlib/searcher_settings.mojo:58-193) */
#define ISH_ALGO_SEMI_GLOBAL 0 /* striped-semi-global */
#define ISH_ALGO_LOCAL 1       /* striped-local */

/* --sg-ends-free letters in order (searcher_settings.mojo:404-407) */
#define ISH_FREE_QUERY_START 1u
#define ISH_FREE_QUERY_END 2u
#define ISH_FREE_TARGET_START 4u
#define ISH_FREE_TARGET_END 8u

/* engines */
#define ISH_ENGINE_BATCHED 0 /* default: length-bucketed persistent kernel,
                                packed int16 lanes, int32 for targets whose
                                range bound does not fit int16 */
#define ISH_ENGINE_PLAIN 1   /* one launch per batch, int32 cells, input
                                order, one target per lane */

/* per-target status byte: cell width used, plus a flag when the target was
 * routed to int32 because its proven score range exceeds int16, and a flag
 * when its 16-bit cells were binary16 halves (ish_hip_set_f16). */
#define ISH_STATUS_I16 16u
#define ISH_STATUS_I32 32u
#define ISH_STATUS_F16 0x40u
#define ISH_STATUS_WIDENED 0x80u

typedef struct ish_hip_ctx ish_hip_ctx;
typedef struct ish_hip_db ish_hip_db;

int32_t ish_hip_version(void);
/* Number of visible gfx950 devices (0 when none). */
int32_t ish_hip_device_count(void);

/* SearcherDevice.create_devices (searcher_device.mojo:152-174), one device
 * per context. */
int32_t ish_hip_create(int32_t device, ish_hip_ctx **out);
void ish_hip_destroy(ish_hip_ctx *ctx);
const char *ish_hip_last_error(const ish_hip_ctx *ctx);

/* Tables (host only, no device needed). ScoringMatrix values and
 * ascii_to_encoding / encoding_to_ascii (scoring_matrix.mojo:9-172,
 * :304-358). ish_hip_matrix writes A*A bytes (65536 for ascii) and returns
 * A; ish_hip_encoding writes 256 bytes; ish_hip_decoding writes A bytes and
 * returns A. */
int32_t ish_hip_matrix(int32_t matrix_kind, int8_t *out, int32_t cap);
int32_t ish_hip_encoding(int32_t matrix_kind, uint8_t *out256);
int32_t ish_hip_decoding(int32_t matrix_kind, uint8_t *out, int32_t cap);
/* convert_ascii_to_encoding_and_score (scoring_matrix.mojo:403-416):
 * encodes n bytes into out and returns the optimal score, the sum of the
 * matrix diagonal over the encoded bytes (the --score normaliser). */
int64_t ish_hip_encode_and_score(int32_t matrix_kind, const uint8_t *ascii,
                                 int64_t n, uint8_t *out);

/* Query setup: StripedSemiGlobalMatcher.__init__ /
 * StripedLocalMatcher.__init__ (striped_semi_global_matcher.mojo:41-75,
 * striped_local_matcher.mojo:33-59). `matrix` / `matrix_len` are what
 * GpuMatcher.matrix_bytes() / matrix_len() return (ishlib/matcher/
 * __init__.mojo:79-86; striped_semi_global_matcher.mojo:139-144): the
 * row-major int8 values of matrix_kind's ScoringMatrix and their count,
 * which must be A*A for the kind's alphabet A (65536 ascii, 576 blosum62,
 * 25 actgn and actgn0), else ISH_E_INVALID. NULL selects the built-in
 * matrix of matrix_kind (matrix_len is then ignored). gap_open / gap_extend
 * are the absolute values of the CLI flags (searcher_settings.mojo:218-219).
 * ends_free is only used by ISH_ALGO_SEMI_GLOBAL. */
int32_t ish_hip_set_query(ish_hip_ctx *ctx, const uint8_t *enc_query,
                          int32_t qlen, const int8_t *matrix, int32_t matrix_len,
                          int32_t matrix_kind, int32_t gap_open,
                          int32_t gap_extend, uint32_t ends_free,
                          int32_t algo);
int32_t ish_hip_set_engine(ish_hip_ctx *ctx, int32_t engine);
/* Tile height of the batched kernels: query rows each lane holds in
 * registers (16, 32; 40 for the hybrid semi-global kernel, 24 for the
 * workgroup-shared lockstep one; launches that have no variant of the
 * height use 32). A tuning knob only; results never depend on it.
 * No counterpart in the reference (its GPU kernel keeps whole rows in
 * private memory, ishlib/matcher/alignment/semi_global_aln/basic.mojo:
 * 305-306). */
int32_t ish_hip_set_tile(ish_hip_ctx *ctx, int32_t query_rows_per_lane);

/* How the batched kernels spread a group of 128 targets. WAVE: one wave
 * scores a group through every query chunk. WORKGROUP: the four waves of a
 * workgroup split the query chunks of one group and pipeline them through
 * LDS (every group). AUTO: local alignment pipelines every group when the
 * longest groups are far longer than the average work per wave (long
 * reads); semi-global pipelines just the groups longer than half the
 * average work per wave, in the same launch as the others. A tuning knob
 * only; results never depend on it. No counterpart in the reference. */
#define ISH_SCHEDULE_AUTO 0
#define ISH_SCHEDULE_WAVE 1
#define ISH_SCHEDULE_WORKGROUP 2
int32_t ish_hip_set_schedule(ish_hip_ctx *ctx, int32_t schedule);

/* Binary16 lanes (default on). A semi-global pass over the shared paired
 * profile runs the groups whose every value the query's range bound keeps
 * within [-2040, 2040] in binary16 halves, where integer arithmetic is
 * exact and one packed three-input maximum takes H = max(Hdiag+S, E, F):
 * five operations per cell pair instead of six. Other groups stay in
 * int16. 0 keeps every group in int16. A tuning knob only; results never
 * depend on it. No counterpart in the reference. */
int32_t ish_hip_set_f16(ish_hip_ctx *ctx, int32_t enable);

/* Autotuner: times whole passes over db for the context's current query
 * with every launch configuration (tile height 32, 16, 40 or 24; schedule WAVE,
 * AUTO or WORKGROUP), `reps` timed runs each after a warm-up (0 = 2), and
 * keeps the fastest under the key (matrix kind, algorithm,
 * ceil(log2(query length)), length bucket of the database's longest
 * target: >= 4096, >= 1024, >= 256, shorter). Later passes whose query and
 * database share that key use the tuned configuration, in one launch.
 * ish_hip_set_tile switches the table off until the next ish_hip_autotune.
 * Results never depend on it. No counterpart in the reference. */
int32_t ish_hip_autotune(ish_hip_ctx *ctx, ish_hip_db *db, int32_t reps);
/* The tuned (tile, schedule) for the current query and db's key: returns
 * 1, or 0 when that key has not been tuned. */
int32_t ish_hip_tuning(const ish_hip_ctx *ctx, const ish_hip_db *db, int32_t *tile,
                       int32_t *schedule);
/* HIP graphs for database passes: when on, ish_hip_align_db[_async]
 * captures its launches into a graph the first time and replays it while
 * the launches (kernels, grids, parameters and buffers) stay the same; a
 * change of query, database, tile or tuning re-captures. Counters are reset
 * by kernels, not memsets, inside the graph. */
int32_t ish_hip_set_graphs(ish_hip_ctx *ctx, int32_t on);

/* One synchronous batch (the batch_match_coarse + copy_outputs_to_host
 * step). Targets are residues[offsets[t] .. offsets[t+1]) for t < n.
 * Outputs (n entries each): score, 0-based end_ref / end_query, status.
 * status may be NULL. With the batched engine a batch of more residues
 * than the slice size is cut into input-order slices: each slice is
 * uploaded while the slice before it is scored. */
int32_t ish_hip_align_batch(ish_hip_ctx *ctx, const uint8_t *residues,
                            const uint64_t *offsets, int64_t n, int32_t *score,
                            int32_t *end_ref, int32_t *end_query,
                            uint8_t *status);
/* One batch with start coordinates: the per-batch step of the reference's
 * GPU path in one call, i.e. batch_match_coarse, then the threshold and
 * find_start for every target that passes (ishlib/gpu/process_batch.mojo:
 * 129-160 calling cpu/process_batch.mojo:36-68). Scores as
 * ish_hip_align_batch; then emit[t] = 1 for the targets that pass the
 * matchers' emission rule at `threshold` for `max_score` (ish_hip_emit, the
 * CPU matchers' float32 rule), and target_start / query_start as
 * ish_hip_find_starts for them, -1 for the others. The reverse passes run
 * on the GPU over the emitted targets' prefixes, gathered on the device
 * from the batch already uploaded (slice by slice), so no residue is
 * uploaded twice. Batched engine only. */
int32_t ish_hip_align_batch_starts(ish_hip_ctx *ctx, const uint8_t *residues,
                                   const uint64_t *offsets, int64_t n,
                                   int64_t max_score, float threshold,
                                   int32_t *score, int32_t *end_ref,
                                   int32_t *end_query, uint8_t *status,
                                   uint8_t *emit, int32_t *target_start,
                                   int32_t *query_start);
/* The same asynchronously, for overlapping host work with the GPU as the
 * reference's runners do while a launch is in flight (ishlib/gpu/
 * process_batch.mojo:124-128): submit starts the batch on a host thread of
 * the context and returns; with emit, target_start and query_start NULL it
 * is ish_hip_align_batch, else ish_hip_align_batch_starts. wait returns
 * when it is done, with its return code. One batch per context at a time;
 * until wait returns, the caller keeps every array alive and calls nothing
 * else on the context. */
int32_t ish_hip_align_batch_submit(ish_hip_ctx *ctx, const uint8_t *residues,
                                   const uint64_t *offsets, int64_t n,
                                   int64_t max_score, float threshold,
                                   int32_t *score, int32_t *end_ref,
                                   int32_t *end_query, uint8_t *status,
                                   uint8_t *emit, int32_t *target_start,
                                   int32_t *query_start);
int32_t ish_hip_align_batch_wait(ish_hip_ctx *ctx);
/* Slice size of ish_hip_align_batch in residue bytes: 0 (the default)
 * sizes slices automatically, a positive value is a fixed size, a
 * negative value turns slicing off. Automatic slices hold at least 64 MiB
 * and enough work that their longest target does not set the launch's
 * length (the minimum m); a batch of B bytes gets about sqrt(3 B / m) of
 * them, which balances the exposed first upload against each slice's
 * tail. No
 * counterpart in the reference, whose batches are sized by the host
 * (searcher_settings.mojo:44-50). */
int32_t ish_hip_set_slice_bytes(ish_hip_ctx *ctx, int64_t bytes);

/* Device-resident target database: uploaded once, sorted by length and
 * packed into coalesced groups on the device, then scored against any
 * number of queries. The packing replaces SearcherDevice.set_host_inputs
 * (searcher_device.mojo:419-450). */
int32_t ish_hip_db_load(ish_hip_ctx *ctx, const uint8_t *residues,
                        const uint64_t *offsets, int64_t n, ish_hip_db **out);
/* The same from raw record bytes (ASCII): they are translated to residue
 * codes with matrix_kind's encoding (ish_hip_encoding) on the device after
 * the upload, so the host only splits records. */
int32_t ish_hip_db_load_ascii(ish_hip_ctx *ctx, const uint8_t *ascii,
                              const uint64_t *offsets, int64_t n,
                              int32_t matrix_kind, ish_hip_db **out);
void ish_hip_db_free(ish_hip_db *db);
int32_t ish_hip_db_size(const ish_hip_db *db, int64_t *n_targets,
                        int64_t *n_residues);
/* Scores every target of db against the context's query; synchronous,
 * results copied to the host arrays (n_targets entries, original order). */
int32_t ish_hip_align_db(ish_hip_ctx *ctx, ish_hip_db *db, int32_t *score,
                         int32_t *end_ref, int32_t *end_query,
                         uint8_t *status);
/* Asynchronous form: enqueue on the context's stream; results stay on the
 * device until ish_hip_db_fetch. ish_hip_sync waits for the stream. */
int32_t ish_hip_align_db_async(ish_hip_ctx *ctx, ish_hip_db *db);
int32_t ish_hip_sync(ish_hip_ctx *ctx);
int32_t ish_hip_db_fetch(ish_hip_ctx *ctx, ish_hip_db *db, int32_t *score,
                         int32_t *end_ref, int32_t *end_query,
                         uint8_t *status);

/* No counterpart in the reference: this entry point is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). It is ish_hip_align_db with the
 * results written to caller-owned DEVICE arrays (n_targets entries each,
 * original order; d_status may not be NULL) and with the kernels' scratch,
 * the work counter and the boundary rows between query chunks, in a static
 * 512 MiB device array that the library holds on every device it runs on
 * until the process ends. It allocates nothing, builds no paired profile
 * (the pass reads the int16 profile of ish_hip_set_query), creates no event
 * and copies nothing between host and device; it returns when the pass is
 * done. Full occupancy holds for targets up to 512 residues, longer ones
 * run on fewer workgroups, and a database whose longest target needs more
 * boundary rows than the array holds for one workgroup (over 262,143
 * residues) is rejected with ISH_E_RANGE. All calls on one device share the
 * array and keep the last call's contents: calls that overlap on one device
 * (two contexts of one device used at once) race. ish_hip_last_timing is
 * not updated. */
int32_t ish_hip_align_db_noalloc(ish_hip_ctx *ctx, ish_hip_db *db, int32_t *d_score,
                                 int32_t *d_end_ref, int32_t *d_end_query,
                                 uint8_t *d_status);

/* Many queries against one database (semi-global): every target of db is
 * scored against each of nq queries, query k being
 * queries[qoffsets[k] .. qoffsets[k+1]), with one shared matrix, gap
 * penalties and free ends (matrix / matrix_len as in ish_hip_set_query).
 * One persistent launch per cell width drains a
 * work queue of (query, group) tiles. Outputs are [nq][n_targets]
 * row-major. This is the batched multi-query mode of the engine; the
 * reference scores one pattern per run (do_search.mojo:24-274). */
int32_t ish_hip_align_db_multi(ish_hip_ctx *ctx, ish_hip_db *db,
                               const uint8_t *queries, const uint64_t *qoffsets,
                               int32_t nq, const int8_t *matrix, int32_t matrix_len,
                               int32_t matrix_kind, int32_t gap_open,
                               int32_t gap_extend, uint32_t ends_free,
                               int32_t *score, int32_t *end_ref,
                               int32_t *end_query, uint8_t *status);

/* Start coordinates of selected targets of db, for the context's query:
 * the reverse pass of semi_global_aln_start_end (semi_global_aln/
 * striped.mojo:160-191; also StripedSemiGlobalMatcher.find_start,
 * striped_semi_global_matcher.mojo:109-126, called by the reference's GPU
 * path from cpu/process_batch.mojo:56-61) and of ssw_align (local_aln/
 * striped.mojo:232-262). For k < n_sel, target sel[k] with forward end
 * coordinates (end_ref[k], end_query[k]) gets:
 *   semi-global: target_start = end_ref - rev.end_ref,
 *                query_start  = qlen - rev.end_query - 1;
 *   local:       target_start (ref_begin1) = end_ref - rev.end_ref,
 *                query_start (read_begin1) = end_query - rev.end_query,
 *                or -1/-1 when end_ref <= 0 or end_query <= 0.
 * The reverse pass runs on the GPU over rev(target[0..end_ref]) with the
 * reversed query and swapped free ends. */
int32_t ish_hip_find_starts(ish_hip_ctx *ctx, ish_hip_db *db,
                            const int64_t *sel, const int32_t *end_ref,
                            const int32_t *end_query, int64_t n_sel,
                            int32_t *target_start, int32_t *query_start);

/* Timing of the last align call on ctx, measured with HIP events on the
 * context's stream around each kernel launch: total milliseconds and the
 * number of kernel launches it covered. */
int32_t ish_hip_last_timing(const ish_hip_ctx *ctx, float *kernel_ms,
                            int32_t *launches);

/* Emission rule of the matchers, float32 throughout:
 *  semi-global: f32(score) >= f32(max_score)*thr, then
 *               f32(score)/f32(max_score) >= thr
 *               (striped_semi_global_matcher.mojo:77-107);
 *  local:       f32(score) > thr*f32(max_score), end_ref > 0, end_query > 0,
 *               then f32(score)/f32(max_score) >= thr
 *               (striped_local_matcher.mojo:61-90, local_aln/striped.mojo:
 *               213-230).
 * Writes emit[t] = 0/1 for n targets. */
int32_t ish_hip_emit(int32_t algo, const int32_t *score,
                     const int32_t *end_ref, const int32_t *end_query,
                     int64_t n, int64_t max_score, float threshold,
                     uint8_t *emit);

/* ------------------------------------------------------------------------
 * File search runner: the search loop of ish's parallel runners over one
 * file (ishlib/parallel_line_search_runner.mojo:75-161,
 * ishlib/parallel_fastx_search_runner.mojo:79-182) with GPU scoring. Reads
 * line or FASTA/FASTQ records, plain or gzip (zlib, as the reference's
 * GZFile), with the reference's kseq record rules; encodes them with the
 * matrix's encoding; scores batches on the GPU while the next batch is
 * parsed; writes each emitted record to out_fd in input order, in the
 * runners' non-tty format:
 *   line:  decode(encode(line)) "\n"
 *   fastx: ">" name "\n" decode(encode(seq)) ["\n+\n" qual] "\n"
 * Binary files (a NUL byte in the first 8192 bytes) are skipped in line
 * mode. `pattern` is the ASCII pattern (encoded and scored as
 * convert_ascii_to_encoding_and_score does). Sets the context's query.
 * ------------------------------------------------------------------------ */
#define ISH_RECORD_LINE 0
#define ISH_RECORD_FASTX 1

typedef struct ish_hip_search_config {
    int32_t matrix_kind;  /* ISH_MATRIX_* */
    int32_t algo;         /* ISH_ALGO_* */
    int32_t gap_open;     /* absolute values */
    int32_t gap_extend;
    uint32_t ends_free;   /* ISH_FREE_* bits (semi-global) */
    float threshold;      /* --score */
    int32_t record_type;  /* ISH_RECORD_* */
    int64_t batch_bytes;  /* record bytes per GPU batch (0 = 256 MiB);
                             output never depends on it */
    int32_t reader_threads;       /* parser threads for plain files (0 = up
                                     to 16); gzip streams are read by one */
    int32_t reader_segment_bytes; /* bytes per parallel parse segment
                                     (0 = 8 MiB); output never depends on
                                     either */
    int32_t tty;          /* 1: the runners' terminal format, which the
                             reference uses when stdout is a tty
                             (parallel_line_search_runner.mojo:123-150,
                             parallel_fastx_search_runner.mojo:150-167):
                             line: PURPLE path RESET ":" GREEN line number
                             RESET ": " then the line with the match span
                             [target_start, end_ref+1) in RED..RESET;
                             fastx: the sequence with that span in
                             "\x1b[1;31m" RED..RESET. Starts come from the
                             GPU reverse pass (ish_hip_find_starts). */
} ish_hip_search_config;

typedef struct ish_hip_search_stats {
    int64_t records;
    int64_t residues;
    int64_t matches;
    int64_t batches;
    double cells;          /* residues * pattern length */
    double seconds_total;
    double seconds_parse;  /* parser thread, overlapped with the GPU */
    double seconds_gpu;    /* upload + scoring + results, per batch */
    double seconds_write;
    int32_t reader_status;  /* 0, or the reader code that ended the file
                               early (-2 truncated quality, -3 stream error;
                               the reference's runners stop there too) */
} ish_hip_search_stats;

/* The record reader of the runner, on its own (host only, no device):
 * reads every record of a file (raw bytes, not encoded) with the same rules.
 * flags: bit 0 = binary (a NUL in the first 8192 bytes), bit 1 = FASTQ
 * (decided from the first 8192 bytes). status: 0, or the reader's code that
 * ended the file (-2 truncated quality, -3 stream error); a FASTX record with
 * an empty sequence also ends the file (status 0). */
typedef struct ish_hip_records ish_hip_records;
int32_t ish_hip_records_read(const char *path, int32_t record_type,
                             ish_hip_records **out);
/* The same records, read with `threads` parser threads (0 = up to 16) over
 * segments of `segment_bytes` (0 = 8 MiB) when the file is not gzip: the
 * records never depend on either. No counterpart in the reference (its
 * kseq reader is serial). */
int32_t ish_hip_records_read_ex(const char *path, int32_t record_type,
                                int32_t threads, int64_t segment_bytes,
                                ish_hip_records **out);
int64_t ish_hip_records_count(const ish_hip_records *r);
int32_t ish_hip_records_flags(const ish_hip_records *r);
int32_t ish_hip_records_status(const ish_hip_records *r);
/* Pointers stay valid until ish_hip_records_free. name / qual are empty for
 * line records. */
int32_t ish_hip_records_get(const ish_hip_records *r, int64_t i,
                            const uint8_t **seq, int64_t *seq_len,
                            const uint8_t **name, int64_t *name_len,
                            const uint8_t **qual, int64_t *qual_len);
void ish_hip_records_free(ish_hip_records *r);

int32_t ish_hip_search_file(ish_hip_ctx *ctx, const char *path,
                            const uint8_t *pattern, int32_t pattern_len,
                            const ish_hip_search_config *cfg, int32_t out_fd,
                            ish_hip_search_stats *stats);

/* ------------------------------------------------------------------------
 * Serving loop: a device-resident database answers semi-global queries that
 * arrive over time, with continuous batching. No counterpart in the
 * reference, which scores one pattern per run (do_search.mojo:24-274);
 * per query it computes what StripedSemiGlobalMatcher.first_match computes
 * for every target (striped_semi_global_matcher.mojo:77-107) and returns
 * the emitted targets.
 *
 * A worker thread owns `slots` launch slots. Whenever a slot is free and
 * requests are waiting, it orders them shortest first with aging (the
 * request's estimated scoring time, from its cells and the measured rate,
 * minus `aging` times the seconds it has waited) and takes them in that
 * order while they fit: at most max_batch queries and, beyond the first,
 * at most batch_cells cells, or by default as many as make the launch
 * fill the device and then the slot's share of the waiting cells, up to
 * batch_seconds of scoring. It enqueues one multi-query pass plus the
 * --score emission on the GPU; a batch never waits for the batches already
 * in flight. A request's results do not depend on the other requests of
 * its batch.
 *
 * The database must stay loaded until ish_hip_server_destroy returns.
 * Submit / wait / hits / release may be called from any thread.
 * ------------------------------------------------------------------------ */
typedef struct ish_hip_server ish_hip_server;

typedef struct ish_hip_server_config {
    int32_t matrix_kind;      /* ISH_MATRIX_* */
    const int8_t *matrix;     /* as ish_hip_set_query: NULL = built-in */
    int32_t matrix_len;       /* entries of matrix, A*A */
    int32_t gap_open;         /* absolute values, as ish_hip_set_query */
    int32_t gap_extend;
    float threshold;          /* --score (float32 rule of ish_hip_emit) */
    int32_t max_batch;        /* queries per batch, 1..256 (0 = 16) */
    int32_t slots;            /* batches in flight, 1..4 (0 = 2) */
    int64_t batch_cells;      /* cells per batch beyond its first query
                                 (0 = from the database and the queue) */
    int64_t hits_per_query;   /* device hit-list capacity per query; a query
                                 with more hits is emitted on the host from
                                 its full results (0 = 65536) */
    int32_t keep_scores;      /* 1: keep every request's full per-target
                                 score / end_ref / end_query */
    double aging;             /* admission order: estimated seconds of
                                 scoring minus aging x seconds waited,
                                 smallest first (0 = 1.0; < 0 = FIFO) */
    double batch_seconds;     /* automatic budget: at most this much
                                 scoring per batch beyond its first query
                                 (0 = 0.5) */
} ish_hip_server_config;

typedef struct ish_hip_request_info {
    int64_t n_hits;           /* emitted targets */
    int64_t cells;            /* qlen * residues in the database */
    double t_submit;          /* ish_hip_server_clock() times (seconds) */
    double t_launch;
    double t_done;
    int32_t batch_id;
    int32_t batch_size;       /* queries in the request's batch */
    int32_t status;           /* 0 or ISH_E_* */
} ish_hip_request_info;

typedef struct ish_hip_server_stats {
    int64_t batches;
    int64_t queries;          /* completed */
    int64_t cells;            /* of completed queries */
    double kernel_ms;         /* scoring kernels, HIP events, summed */
    int64_t pending;          /* waiting for a slot */
    int64_t busy_slots;
} ish_hip_server_stats;

int32_t ish_hip_server_create(ish_hip_db *db, const ish_hip_server_config *cfg,
                              ish_hip_server **out);
/* Finishes every submitted request, then frees the server. */
void ish_hip_server_destroy(ish_hip_server *srv);
/* The server's last error message, copied for the calling thread: the
 * pointer stays valid until this thread calls it again. */
const char *ish_hip_server_last_error(ish_hip_server *srv);
/* Monotonic clock of the request timestamps, in seconds. */
double ish_hip_server_clock(void);
/* Queues one query (encoded codes, free ends as ISH_FREE_* bits). Rejects a
 * query whose profile does not fit the LDS with ISH_E_RANGE. */
int32_t ish_hip_server_submit(ish_hip_server *srv, const uint8_t *enc_query,
                              int32_t qlen, uint32_t ends_free,
                              int64_t *ticket);
/* Waits up to timeout_ms (< 0: no limit) for a request; returns 0, the
 * request's error code, or ISH_E_TIMEOUT. info may be NULL. */
int32_t ish_hip_server_wait(ish_hip_server *srv, int64_t ticket,
                            int32_t timeout_ms, ish_hip_request_info *info);
/* Copies up to cap emitted targets, ascending by target index (original DB
 * order), with score and 0-based end coordinates. NULL arrays are skipped. */
int32_t ish_hip_server_hits(ish_hip_server *srv, int64_t ticket,
                            int64_t *target, int32_t *score, int32_t *end_ref,
                            int32_t *end_query, int64_t cap);
/* Full per-target results (n_targets each); keep_scores servers only. */
int32_t ish_hip_server_scores(ish_hip_server *srv, int64_t ticket,
                              int32_t *score, int32_t *end_ref,
                              int32_t *end_query);
/* Forgets a finished request. */
int32_t ish_hip_server_release(ish_hip_server *srv, int64_t ticket);
int32_t ish_hip_server_get_stats(ish_hip_server *srv, ish_hip_server_stats *out);

#ifdef __cplusplus
}
#endif
#endif
