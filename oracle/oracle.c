/*
 * Scalar int32 oracle: textbook Gotoh semi-global and Smith-Waterman with the
 * reference CPU striped kernels' end-coordinate rules.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker, never the product.
 *
 * Why a scalar textbook DP restates the striped kernels exactly:
 *  - The striped kernel's lazy-F loop updates H but not E (SWPS3 rule,
 *    ishlib/matcher/alignment/semi_global_aln/striped.mojo:396-434). That only
 *    forbids a vertical gap directly followed by a horizontal gap; every such
 *    path has an equal-cost horizontal-then-vertical path ending in the same
 *    cell, so every H value equals the unrestricted Gotoh value.
 *  - Saturating int16 arithmetic never clamps an H value at the lengths and
 *    penalties the engine accepts for int16 (the HIP engine routes anything
 *    else to int32), so int32 arithmetic gives the same numbers.
 * Both claims are checked differentially against striped.c, which restates
 * the striped kernel step by step, and against brute-force enumeration.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "oracle.h"

#define I16_MIN (-32768)

/* min/max over the whole matrix, both starting at 0 exactly like the
 * reference's Profile constructor (striped.mojo:62-71). */
static void matrix_min_max(const int8_t *S, int32_t A, int32_t *mn,
                           int32_t *mx) {
    int32_t lo = 0, hi = 0;
    for (int64_t k = 0; k < (int64_t)A * A; ++k) {
        if (S[k] < lo) lo = S[k];
        if (S[k] > hi) hi = S[k];
    }
    *mn = lo;
    *mx = hi;
}

/* The int16 floor that seeds the last-row running maximum and the
 * free-query-end comparison (striped.mojo:329-334). */
static int32_t neg_limit_i16(int32_t go, int32_t min_score) {
    int32_t v = (-go < min_score) ? I16_MIN + go + 1 : I16_MIN - min_score + 1;
    return (int32_t)(int16_t)v;
}

/*
 * Semi-global Gotoh, int32 (striped.mojo:252-498 semantics):
 *   H(-1,-1)=0; H(j,-1)= qs?0:-go-ge*j; H(-1,i)= ts?0:-go-ge*i
 *   E(j,i)=max(E(j,i-1)-ge, H(j,i-1)-go), E(j,-1)=-inf
 *   F(j,i)=max(F(j-1,i)-ge, H(j-1,i)-go), F(0,i)=H(-1,i)-go
 *   H(j,i)=max(H(j-1,i-1)+S[r_i][q_j], E, F)
 * End rules: first column of the strict last-row maximum (striped.mojo:
 * 436-441, seeded at the int16 floor); te keeps it (:446-450); qe takes the
 * smallest row of the strict last-column maximum with the end_ref==R-1 tie
 * rule (:453-468); neither takes the corner (:471-477).
 */
oracle_end_t oracle_sg(const uint8_t *q, int32_t Q, const uint8_t *r,
                       int32_t R, const int8_t *S, int32_t A, int32_t go,
                       int32_t ge, uint32_t flags) {
    oracle_end_t res = {0, 0, 0};
    if (Q <= 0 || R <= 0) return res; /* striped.mojo:275-276 */
    const int qs = (flags & ORACLE_QS) != 0, qe = (flags & ORACLE_QE) != 0;
    const int ts = (flags & ORACLE_TS) != 0, te = (flags & ORACLE_TE) != 0;
    int32_t mn, mx;
    matrix_min_max(S, A, &mn, &mx);
    const int32_t neg_limit = neg_limit_i16(go, mn);

    int32_t *H = (int32_t *)malloc(sizeof(int32_t) * (size_t)Q);
    int32_t *E = (int32_t *)malloc(sizeof(int32_t) * (size_t)Q);
    for (int32_t j = 0; j < Q; ++j) {
        H[j] = qs ? 0 : -go - ge * j; /* H(j,-1) */
        E[j] = H[j] - go;             /* E(j,0) */
    }
    int32_t vmax = neg_limit, end_ref = -1;
    for (int32_t i = 0; i < R; ++i) {
        const int8_t *row = S + (int64_t)r[i] * A;
        int32_t diag = (i == 0) ? 0 : (ts ? 0 : -go - ge * (i - 1));
        int32_t top = ts ? 0 : -go - ge * i; /* H(-1,i) */
        int32_t F = top - go;                /* F(0,i) */
        for (int32_t j = 0; j < Q; ++j) {
            int32_t h = diag + row[q[j]];
            if (E[j] > h) h = E[j];
            if (F > h) h = F;
            diag = H[j];
            H[j] = h;
            int32_t hg = h - go;
            int32_t e = E[j] - ge;
            E[j] = e > hg ? e : hg;
            int32_t f = F - ge;
            F = f > hg ? f : hg;
        }
        if (H[Q - 1] > vmax) {
            vmax = H[Q - 1];
            end_ref = i;
        }
    }
    int32_t score = neg_limit, end_q = Q - 1;
    if (te) score = vmax;
    if (qe) {
        for (int32_t j = 0; j < Q; ++j) {
            if (H[j] > score) {
                score = H[j];
                end_q = j;
                end_ref = R - 1;
            } else if (H[j] == score && end_ref == R - 1 && j < end_q) {
                end_q = j;
            }
        }
    }
    if (!te && !qe) {
        score = H[Q - 1];
        end_ref = R - 1;
        end_q = Q - 1;
    }
    free(H);
    free(E);
    res.score = score;
    res.end_ref = end_ref;
    res.end_query = end_q;
    return res;
}

static void reverse_copy(const uint8_t *src, int32_t n, uint8_t *dst) {
    for (int32_t k = 0; k < n; ++k) dst[k] = src[n - 1 - k];
}

/* semi_global_aln_start_end (striped.mojo:121-191): reverse pass over
 * rev(ref[0..end_ref]) with rev(query) and the free-end flags swapped. */
int32_t oracle_sg_start_end(const uint8_t *q, int32_t Q, const uint8_t *r,
                            int32_t R, const int8_t *S, int32_t A, int32_t go,
                            int32_t ge, uint32_t flags, float cutoff,
                            oracle_start_end_t *out) {
    oracle_end_t fw = oracle_sg(q, Q, r, R, S, A, go, ge, flags);
    if ((float)fw.score < cutoff) return 0;
    int32_t n = fw.end_ref + 1;
    if (n < 0) n = 0;
    uint8_t *rr = (uint8_t *)malloc((size_t)(n > 0 ? n : 1));
    uint8_t *rq = (uint8_t *)malloc((size_t)(Q > 0 ? Q : 1));
    reverse_copy(r, n, rr);
    reverse_copy(q, Q, rq);
    uint32_t sw = ((flags & ORACLE_QS) ? ORACLE_QE : 0) |
                  ((flags & ORACLE_QE) ? ORACLE_QS : 0) |
                  ((flags & ORACLE_TS) ? ORACLE_TE : 0) |
                  ((flags & ORACLE_TE) ? ORACLE_TS : 0);
    oracle_end_t rv = oracle_sg(rq, Q, rr, n, S, A, go, ge, sw);
    free(rr);
    free(rq);
    out->score = fw.score;
    out->query_start = Q - rv.end_query - 1;
    out->query_end = fw.end_query;
    out->target_start = n - rv.end_ref - 1;
    out->target_end = fw.end_ref;
    return 1;
}

/*
 * Smith-Waterman, int32 (local_aln/striped.mojo:272-434 semantics):
 *   H = max(0, Hdiag+S, E, F) with E and F floored at 0 (unsigned
 *   saturation in the reference; the floor never changes H).
 * best = max H; end_ref = first column whose column maximum strictly raises
 * the running maximum to best (-1 if best == 0); end_query = smallest row of
 * that column holding best (0 when best == 0: the stored column is zeroed).
 * Columns are visited in increasing order when dir > 0 and decreasing order
 * when dir < 0 (ReferenceDirection.Reverse); end_ref is the visited index.
 */
static oracle_end_t sw_dir(const uint8_t *q, int32_t Q, const uint8_t *r,
                           int32_t R, const int8_t *S, int32_t A, int32_t go,
                           int32_t ge, int dir) {
    oracle_end_t res = {0, -1, 0};
    if (Q <= 0) return res;
    int32_t *H = (int32_t *)calloc((size_t)Q, sizeof(int32_t));
    int32_t *E = (int32_t *)calloc((size_t)Q, sizeof(int32_t));
    int32_t *Hmax = (int32_t *)calloc((size_t)Q, sizeof(int32_t));
    int32_t best = 0, end_ref = -1;
    for (int32_t step = 0; step < R; ++step) {
        int32_t i = dir > 0 ? step : R - 1 - step;
        const int8_t *row = S + (int64_t)r[i] * A;
        int32_t diag = 0, F = 0, colmax = 0;
        for (int32_t j = 0; j < Q; ++j) {
            int32_t h = diag + row[q[j]];
            if (h < 0) h = 0;
            if (E[j] > h) h = E[j];
            if (F > h) h = F;
            diag = H[j];
            H[j] = h;
            if (h > colmax) colmax = h;
            int32_t hg = h - go;
            int32_t e = E[j] - ge;
            e = e > hg ? e : hg;
            E[j] = e > 0 ? e : 0;
            int32_t f = F - ge;
            f = f > hg ? f : hg;
            F = f > 0 ? f : 0;
        }
        if (colmax > best) {
            best = colmax;
            end_ref = i;
            memcpy(Hmax, H, sizeof(int32_t) * (size_t)Q);
        }
    }
    int32_t end_q = Q - 1;
    for (int32_t j = 0; j < Q; ++j)
        if (Hmax[j] == best && j < end_q) end_q = j;
    free(H);
    free(E);
    free(Hmax);
    res.score = best;
    res.end_ref = end_ref;
    res.end_query = end_q;
    return res;
}

oracle_end_t oracle_sw(const uint8_t *q, int32_t Q, const uint8_t *r,
                       int32_t R, const int8_t *S, int32_t A, int32_t go,
                       int32_t ge) {
    return sw_dir(q, Q, r, R, S, A, go, ge, +1);
}

/* ssw_align (local_aln/striped.mojo:142-269). */
int32_t oracle_sw_align(const uint8_t *q, int32_t Q, const uint8_t *r,
                        int32_t R, const int8_t *S, int32_t A, int32_t go,
                        int32_t ge, float cutoff, oracle_local_t *out) {
    oracle_end_t fw = sw_dir(q, Q, r, R, S, A, go, ge, +1);
    if ((float)fw.score <= cutoff) return 0; /* :215 */
    out->score1 = fw.score;
    out->ref_end1 = fw.end_ref;
    out->read_end1 = fw.end_query;
    if (fw.end_ref <= 0 || fw.end_query <= 0) { /* :223-230 */
        out->ref_begin1 = -1;
        out->read_begin1 = -1;
        return 1;
    }
    /* reverse pass over ref[0..end_ref] with the full reversed query,
     * visiting columns from end_ref down to 0 (:232-258). */
    uint8_t *rq = (uint8_t *)malloc((size_t)Q);
    reverse_copy(q, Q, rq);
    oracle_end_t rv = sw_dir(rq, Q, r, fw.end_ref + 1, S, A, go, ge, -1);
    free(rq);
    out->ref_begin1 = rv.end_ref;
    out->read_begin1 = fw.end_query - rv.end_query;
    return 1;
}

/* Emission (striped_semi_global_matcher.mojo:77-107): the forward score is
 * rejected when f32(score) < f32(max_score)*thr, then emitted when
 * f32(score)/f32(max_score) >= thr. */
int32_t oracle_sg_emit(int32_t score, int64_t max_score, float threshold) {
    volatile float ms = (float)max_score;
    volatile float cutoff = ms * threshold;
    volatile float s = (float)score;
    if (s < cutoff) return 0;
    volatile float ratio = s / ms;
    return ratio >= threshold;
}

/* Local emission (striped_local_matcher.mojo:61-90 over ssw_align): strict
 * cutoff, end coordinates > 0 so that a begin exists, then the ratio test.
 * The begin pass always finds a positive maximum when the forward score is
 * positive, so ref_begin >= 0 follows from the other conditions. */
int32_t oracle_sw_emit(int32_t score, int32_t end_ref, int32_t end_query,
                       int64_t max_score, float threshold) {
    volatile float ms = (float)max_score;
    volatile float cutoff = threshold * ms;
    volatile float s = (float)score;
    if (s <= cutoff) return 0;
    if (end_ref <= 0 || end_query <= 0) return 0;
    volatile float ratio = s / ms;
    return ratio >= threshold;
}

/* ---- threaded batch driver -------------------------------------------- */
typedef struct {
    int32_t algo, striped;
    const uint8_t *q;
    int32_t Q;
    const uint8_t *res;
    const uint64_t *off;
    int64_t n;
    const int8_t *S;
    int32_t A, go, ge;
    uint32_t flags;
    int32_t *score, *end_ref, *end_q;
    int64_t next;
    pthread_mutex_t mu;
} batch_job_t;

static void *batch_worker(void *arg) {
    batch_job_t *jb = (batch_job_t *)arg;
    const int64_t chunk = 64;
    for (;;) {
        pthread_mutex_lock(&jb->mu);
        int64_t lo = jb->next;
        jb->next += chunk;
        pthread_mutex_unlock(&jb->mu);
        if (lo >= jb->n) break;
        int64_t hi = lo + chunk < jb->n ? lo + chunk : jb->n;
        for (int64_t t = lo; t < hi; ++t) {
            const uint8_t *r = jb->res + jb->off[t];
            int32_t R = (int32_t)(jb->off[t + 1] - jb->off[t]);
            oracle_end_t e;
            if (jb->striped)
                e = jb->algo == 0
                        ? oracle_sg_striped(jb->q, jb->Q, r, R, jb->S, jb->A,
                                            jb->go, jb->ge, jb->flags, 0)
                        : oracle_sw_striped(jb->q, jb->Q, r, R, jb->S, jb->A,
                                            jb->go, jb->ge, 0);
            else
                e = jb->algo == 0 ? oracle_sg(jb->q, jb->Q, r, R, jb->S, jb->A,
                                              jb->go, jb->ge, jb->flags)
                                  : oracle_sw(jb->q, jb->Q, r, R, jb->S, jb->A,
                                              jb->go, jb->ge);
            jb->score[t] = e.score;
            jb->end_ref[t] = e.end_ref;
            jb->end_q[t] = e.end_query;
        }
    }
    return NULL;
}

static int32_t run_batch(int32_t striped, int32_t algo, const uint8_t *q,
                         int32_t Q, const uint8_t *residues,
                         const uint64_t *offsets, int64_t n, const int8_t *S,
                         int32_t A, int32_t go, int32_t ge, uint32_t flags,
                         int32_t *score, int32_t *end_ref, int32_t *end_query,
                         int32_t nthreads) {
    if (nthreads <= 0) nthreads = (int32_t)sysconf(_SC_NPROCESSORS_ONLN);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    batch_job_t jb;
    jb.algo = algo;
    jb.striped = striped;
    jb.q = q;
    jb.Q = Q;
    jb.res = residues;
    jb.off = offsets;
    jb.n = n;
    jb.S = S;
    jb.A = A;
    jb.go = go;
    jb.ge = ge;
    jb.flags = flags;
    jb.score = score;
    jb.end_ref = end_ref;
    jb.end_q = end_query;
    jb.next = 0;
    pthread_mutex_init(&jb.mu, NULL);
    pthread_t th[256];
    for (int32_t k = 0; k < nthreads; ++k)
        pthread_create(&th[k], NULL, batch_worker, &jb);
    for (int32_t k = 0; k < nthreads; ++k) pthread_join(th[k], NULL);
    pthread_mutex_destroy(&jb.mu);
    return nthreads;
}

int32_t oracle_batch(int32_t algo, const uint8_t *q, int32_t Q,
                     const uint8_t *residues, const uint64_t *offsets,
                     int64_t n, const int8_t *S, int32_t A, int32_t go,
                     int32_t ge, uint32_t flags, int32_t *score,
                     int32_t *end_ref, int32_t *end_query, int32_t nthreads) {
    return run_batch(0, algo, q, Q, residues, offsets, n, S, A, go, ge, flags,
                     score, end_ref, end_query, nthreads);
}

/* oracle_batch_striped lives in striped.cpp (query state built once). */
