/*
 * oracle.h -- CPU restatement of the MNMGDatalog semi-naive fixpoint path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (libdlg.so, the CLIs) never links or calls it.
 *
 * Tuples are int32 pairs.  Output row layouts follow the reference's
 * one-rank output files:
 *   TC   rows (src, dst) ordered by (dst, src)   (tc.cu:316-333, common/kernels.cu:268-277)
 *   SG   rows (x, y)     ordered by (x, y)       (sg.cu:392-400, common/kernels.cu:279-288)
 *   WCC  rows (label, 0) ordered by label        (wcc.cu:389-485)
 *   SJ   rows (a, c)     ordered by (c, a)       (single_join.cu:290-296; the reference never
 *                                                 writes this file, the rows are still returned)
 * All functions are synchronous, single-threaded unless `threads` says otherwise,
 * and return 0 on success, nonzero on invalid input.  Buffers returned through
 * `int32_t **rows_out` are malloc'd and released with orc_free().
 */
#ifndef DLG_ORACLE_H
#define DLG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_stats {
    int32_t iterations;      /* loop rounds, counted as the reference counts them */
    int64_t output_size;     /* |final relation| (TC, SG, SJ) or #components (WCC) */
    int64_t secondary;       /* WCC: nodes in the largest component; else 0 */
    int64_t raw_join_rows;   /* sum of |J| before dedup over all joins (diagnostic) */
    int64_t nodes;           /* WCC: #distinct nodes; else 0 */
} orc_stats;

/* Transitive closure: tc(x,y) :- e(x,y); tc(x,z) :- tc(x,y), e(y,z). */
int orc_tc(const int32_t *rows, int64_t n, int threads, orc_stats *st, int32_t **rows_out);
/* Same generation: sg(x,y) :- e(p,x), e(p,y), x!=y; sg(x,y) :- e(a,x), sg(a,b), e(b,y). */
int orc_sg(const int32_t *rows, int64_t n, int threads, orc_stats *st, int32_t **rows_out);
/* Weakly connected components by min-label propagation with the reference's
 * termination rule (stop when two consecutive delta sizes are equal). */
int orc_wcc(const int32_t *rows, int64_t n, int threads, orc_stats *st, int32_t **rows_out);
/* Weakly connected components, WCC semantics per iteration but stop when the delta
 * is empty (true fixpoint); used to show where the termination rule matters. */
int orc_wcc_fixpoint(const int32_t *rows, int64_t n, orc_stats *st, int32_t **labels_out);
/* Single self-join: unique{(c,a) : R(a,b), R(b,c)} over deduplicated R. */
int orc_single_join(const int32_t *rows, int64_t n, int threads, orc_stats *st, int32_t **rows_out);
/* The relation the reference's single_join generates on one rank:
 * srand(1); v[i] = rand() % range + 1 for i < 2*rows (common/parallel_io.cu:79-108). */
int orc_generate_rand_rows(int64_t rows, int32_t range, int32_t *out /* 2*rows */);

/* Bounded runs for the CPU baseline: stop after `max_iterations` rounds
 * (0 = unbounded).  Same semantics otherwise. */
int orc_tc_bounded(const int32_t *rows, int64_t n, int threads, int32_t max_iterations,
                   orc_stats *st, int32_t **rows_out);
int orc_sg_bounded(const int32_t *rows, int64_t n, int threads, int32_t max_iterations,
                   orc_stats *st, int32_t **rows_out);

void orc_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
