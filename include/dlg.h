/*
 * dlg.h -- C ABI of the MI355X semi-naive fixpoint engine (libdlg.so).
 *
 * The reference (harp-lab/MNMGDatalog) has no library boundary: each program
 * #includes the .cu files of common/.  This header exports the same ROLES as plain C entry
 * points (plain pointers and sizes, no C++ or torch types), so C/C++ drivers,
 * ctypes harnesses or any FFI can bind them.  Each declaration cites the
 * reference interface it replaces.
 *
 * Tuples.  A tuple is the reference's `struct Entity {int key; int value;}`
 * (common/utils.cu:8-11) stored as one uint64_t:
 *     t = ((uint64_t)((uint32_t)key ^ 0x80000000u) << 32) | ((uint32_t)value ^ 0x80000000u)
 * so unsigned 64-bit order equals the reference's signed lexicographic
 * (key, value) order `set_cmp` (common/utils.cu:115-124).
 *
 * Errors.  Every call returns a dlg_status; the library never exits the
 * process (the reference's checkCuda() exits, common/error_handler.cu:1-8).
 * dlg_last_error() gives the message of the last failure on a context.
 *
 * Threading.  A context owns one HIP stream and a device memory pool.  A
 * context is used by one host thread at a time; several contexts may share a
 * device.  Entry points that return sizes synchronise their stream.
 */
#ifndef DLG_H
#define DLG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dlg_status {
    DLG_OK = 0,
    DLG_INVALID_INPUT = 1,   /* bad argument, malformed file, reserved id            */
    DLG_OUT_OF_MEMORY = 2,   /* device allocation failed or HBM budget exceeded      */
    DLG_OVERFLOW = 3,        /* a count exceeds what the engine represents (2^32-2 rows per hash
                                index, 2^32-1 tuples per sort with 4..7-bit digits) */
    DLG_DEVICE_ERROR = 4,    /* HIP runtime error                                    */
    DLG_IO_ERROR = 5,        /* file open/read/write failure                         */
    DLG_REFUSED = 6,         /* serving: admission refused (predicted need > budget) */
    DLG_PENDING = 7          /* serving: the request has not finished yet            */
} dlg_status;

typedef struct dlg_ctx dlg_ctx;
typedef struct dlg_hash dlg_hash;

/* ---------------------------------------------------------------- context */
/* Replaces the per-program set-up in benchmark() (tc.cu:42-60: device query,
 * warm_up_kernel, block 512 / grid 32*SMs). */
dlg_status dlg_ctx_create(int device, dlg_ctx **out);
void dlg_ctx_destroy(dlg_ctx *ctx);
const char *dlg_last_error(const dlg_ctx *ctx);
dlg_status dlg_ctx_sync(dlg_ctx *ctx);
/* Device and library facts: CU count, HBM bytes, LDS bytes per block, arch name. */
dlg_status dlg_device_info(dlg_ctx *ctx, int32_t *num_cus, int64_t *hbm_bytes,
                           int32_t *lds_per_block, char *arch_name, int32_t arch_len);

/* The context's device as a PCI address "dddd:bb:dd.f" (for matching power and
 * energy monitors to the device); `len` >= 13. */
dlg_status dlg_device_pci_bus_id(dlg_ctx *ctx, char *bdf, int32_t len);

/* Per-kernel event timing on the context's stream (off by default).  While on,
 * every launch is bracketed by hipEvents and its algorithmic bytes recorded. */
dlg_status dlg_profiling(dlg_ctx *ctx, int enable);
/* Kernel ids for dlg_kernel_stats(); DLG_K_COUNT entries. */
enum {
    DLG_K_PACK = 0, DLG_K_HASH_BUILD, DLG_K_JOIN_COUNT, DLG_K_SCAN, DLG_K_JOIN_EXPAND,
    DLG_K_SORT_HIST, DLG_K_SORT_UPSWEEP, DLG_K_SORT_DOWNSWEEP, DLG_K_SELECT, DLG_K_MERGE,
    DLG_K_WCC_RELAX, DLG_K_WCC_FINAL, DLG_K_HSET_INSERT, DLG_K_BITMAP_INSERT, DLG_K_MISC, DLG_K_COUNT
};
const char *dlg_kernel_name(int kernel_id);
/* Accumulated since the last reset: launches, summed duration (ms), summed
 * algorithmic bytes.  Returns DLG_INVALID_INPUT for a bad id. */
dlg_status dlg_kernel_stats(dlg_ctx *ctx, int kernel_id, int64_t *launches, double *total_ms,
                            double *alg_bytes);
dlg_status dlg_kernel_stats_reset(dlg_ctx *ctx);

/* ---------------------------------------------------------- device memory */
/* Buffers handed across the ABI.  The reference's callee-allocates /
 * caller-frees convention (get_join returns a cudaMalloc'd buffer, freed by
 * the caller at tc.cu:226) maps to dlg_dev_alloc / dlg_dev_free. */
dlg_status dlg_dev_alloc(dlg_ctx *ctx, size_t bytes, void **d_ptr);
dlg_status dlg_dev_free(dlg_ctx *ctx, void *d_ptr);
dlg_status dlg_copy_h2d(dlg_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
dlg_status dlg_copy_d2h(dlg_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);

/* ------------------------------------------------------- tuple conversion */
typedef enum dlg_orient {
    DLG_ORIENT_KV = 0,   /* tuple (col0, col1): create_entity_ar, get_int_ar_from_entity_ar   */
    DLG_ORIENT_VK = 1    /* tuple (col1, col0): create_entity_ar_reverse, reverse_t_full,
                            get_reverse_int_ar_from_entity_ar                                   */
} dlg_orient;
/* int32 rows -> tuples.  Replaces create_entity_ar / create_entity_ar_reverse
 * (common/kernels.cu:245-265). */
dlg_status dlg_pack_rows(dlg_ctx *ctx, const int32_t *d_rows, int64_t n, int orient,
                         uint64_t *d_tuples);
/* tuples -> int32 rows.  Replaces reverse_t_full / get_int_ar_from_entity_ar /
 * get_reverse_int_ar_from_entity_ar (common/kernels.cu:268-299). */
dlg_status dlg_unpack_rows(dlg_ctx *ctx, const uint64_t *d_tuples, int64_t n, int orient,
                           int32_t *d_rows);

/* ------------------------------------------------------------- hash table */
/* Replaces get_hash_table(grid, block, relation, n, &rows, &secs)
 * (common/hash_table.cu:24-55).  Builds an open-addressed table of packed
 * 64-bit (key, row-group) slots over a copy of the relation grouped by key,
 * linear probing, slots = next_pow2(ceil(#distinct keys / load_factor)).
 * Multimap semantics are kept: every row of `d_rel` (duplicates included)
 * joins once.  load_factor <= 0 selects the tuned default.  n == 0 gives an
 * empty table whose joins are empty, as in the reference (hash_table.cu:30-35). */
dlg_status dlg_hash_build(dlg_ctx *ctx, const uint64_t *d_rel, int64_t n, double load_factor,
                          dlg_hash **out);
dlg_status dlg_hash_info(const dlg_hash *h, int64_t *slots, int64_t *distinct_keys,
                         int64_t *rows);
void dlg_hash_free(dlg_ctx *ctx, dlg_hash *h);

/* ------------------------------------------------------------------- join */
typedef enum dlg_proj {
    DLG_PROJ_HV_PV = 0,  /* (ht.value, probe.value): get_join_result_entity (common/join.cu:38-41) */
    DLG_PROJ_PV_HV = 1,  /* (probe.value, ht.value): the same join followed by reverse_entity_ar
                            (sg.cu:252-253, 300-301) */
    DLG_PROJ_HV_PV_DROP_EQ = 2,  /* DLG_PROJ_HV_PV, then the rows with key == value removed:
                                    SG's base rule, get_join + remove_if(is_key_equal_value)
                                    (sg.cu:166-182; common/utils.cu:90-95) */
    DLG_PROJ_PV_HV_DROP_EQ = 3   /* DLG_PROJ_PV_HV with the same removal */
} dlg_proj;
/* Replaces get_join(grid, block, ht, ht_rows, probe, n, &out_n, &secs)
 * (common/join.cu:51-87): count pass, wavefront/block prefix scan, then a
 * load-balanced materialize pass.  *d_out is allocated by the library (NULL
 * when *out_n == 0) and released with dlg_dev_free.  The DROP_EQ projections
 * keep the remaining rows in join order, as the stable remove_if does. */
dlg_status dlg_join(dlg_ctx *ctx, const dlg_hash *h, const uint64_t *d_probe, int64_t n,
                    int proj, uint64_t **d_out, int64_t *out_n);

/* --------------------------------------------------------- WCC min-label round */
/* One round of WCC's loop body (wcc.cu:255-339) on tuples:
 *   J   = {(m, l) : (n, l) in delta, (n, m) in E}   get_join over the edge index
 *                                                    (wcc.cu:256-259)
 *   cc' = per node, the least label of cc u J        sort/unique of J, merge with cc,
 *                                                    unique(is_equal_key) (wcc.cu:285-323)
 *   Δ'  = cc' minus cc                               set_difference (wcc.cu:333-336)
 * `edges` is dlg_hash_build of the symmetric edge relation E u E^-1
 * (wcc.cu:140-157, 247).  d_cc holds cc: one (node, label) tuple per distinct
 * key of the index, in key order (the reference's cc, wcc.cu:160-175); it is
 * updated in place to cc'.  Δ' goes to d_next (room for n_cc tuples), sorted;
 * *n_next = |Δ'|.  d_delta need not be sorted.  A join row whose node is not a
 * key of the index (E not symmetric), or n_cc different from the index's key
 * count, gives DLG_INVALID_INPUT (d_cc is then unspecified). */
dlg_status dlg_wcc_relax(dlg_ctx *ctx, const dlg_hash *edges, const uint64_t *d_delta, int64_t n_delta,
                         uint64_t *d_cc, int64_t n_cc, uint64_t *d_next, int64_t *n_next);

/* ------------------------------------------------- relational operators */
/* thrust::sort(set_cmp) + thrust::unique(is_equal) (tc.cu:250-253 and every
 * site listed in the header of engine.hip).  In place; *out_n = distinct count. */
dlg_status dlg_sort_unique(dlg_ctx *ctx, uint64_t *d_data, int64_t n, int64_t *out_n);
/* thrust::sort(set_cmp) alone, in place. */
dlg_status dlg_sort(dlg_ctx *ctx, uint64_t *d_data, int64_t n);
/* thrust::set_difference(a, b, set_cmp) of sorted duplicate-free inputs
 * (tc.cu:260-263).  d_out may alias d_a. */
dlg_status dlg_set_difference(dlg_ctx *ctx, const uint64_t *d_a, int64_t na, const uint64_t *d_b,
                              int64_t nb, uint64_t *d_out, int64_t *out_n);
/* thrust::merge(a, b, set_cmp) of sorted inputs into d_out (na + nb slots)
 * (tc.cu:275-279).  Equal elements: a's copy first, as a stable merge does. */
dlg_status dlg_merge(dlg_ctx *ctx, const uint64_t *d_a, int64_t na, const uint64_t *d_b,
                     int64_t nb, uint64_t *d_out);
/* The same merge keeping one copy of each value (sorted inputs, duplicates
 * allowed): thrust::sort + thrust::unique over the concatenation of two
 * sorted arrays, as WCC dedups E ++ E^-1 (wcc.cu:163-167), in one merge.
 * d_out holds na + nb slots; *out_n = the count written. */
dlg_status dlg_merge_unique(dlg_ctx *ctx, const uint64_t *d_a, int64_t na, const uint64_t *d_b,
                            int64_t nb, uint64_t *d_out, int64_t *out_n);
/* Exclusive prefix sum of uint32 counts into uint64 offsets; *total = sum.
 * Replaces thrust::reduce + thrust::exclusive_scan in get_join (common/join.cu:72-74). */
dlg_status dlg_exclusive_scan(dlg_ctx *ctx, const uint32_t *d_in, int64_t n, uint64_t *d_out,
                              uint64_t *total);

/* ------------------------------------------------------- whole programs */
/* Tuning knobs (the reference hard-codes them: block 512, grid 32*SMs,
 * load factor 0.6 -- tc.cu:58-60, common/hash_table.cu:37).  Zero fields
 * select the tuned defaults. */
typedef struct dlg_opts {
    double load_factor;      /* hash table load factor, (0, 0.95]                      */
    int32_t block_size;      /* threads per block for the streaming kernels (64..1024) */
    int32_t radix_bits;      /* radix digit width 4..8                                 */
    int32_t strategy;        /* 0 auto; 1 the reference's scheme (one sorted full
                                relation merged every round; WCC rounds launched one
                                by one); 2 tiered sorted runs for the full relation and
                                hipGraph-captured WCC rounds; 3 append-only full
                                relation with a device hash set for the set
                                difference, sorted once at the end; 4 (TC, SG) the full
                                relation as a bitmap over the id range (at most
                                370,720 ids wide; automatic up to 2^18); (single join)
                                every output row built in LDS and read out of the
                                map, no raw join or sort (automatic up to 2^18 ids
                                when the map is no larger than the raw join).
                                Single join, 2: each output row's candidates sorted
                                and deduplicated by one wave, rows of more than 2,048
                                candidates sorted together (automatic when the map
                                is not taken and the id range allows dense offsets);
                                otherwise the sorted raw join                          */
    int32_t max_iterations;  /* 0 = run to the fixpoint; >0 stops early (sampling)     */
    int32_t keep_output;     /* 1: copy the output rows to host (dlg_result.rows)      */
    int32_t check_output;    /* 1: fill dlg_result.output_sum / output_order_breaks
                                (one pass over the device output, after the timing) */
    int32_t join_method;     /* 0 hash index (tc.cu, sg.cu); 1 sorted index with a
                                binary search per probe, TC's edges deduplicated first
                                (tc_nl.cu, sg_nl.cu; common/join.cu:89-180).  Same
                                relations and iteration counts either way             */
    int32_t reorder;         /* TC in the bitmap layout: renumber the ids into a
                                locality order (recursive BFS bisection) before the
                                fixpoint and map the result back at read-out, so a
                                round's new ends of one source share words and cache
                                lines of F.  0 auto (plain runs with id ranges of at
                                least 16,384 ids), 1 never, 2 always.  Same output   */
    int32_t relax;           /* WCC rounds launched one by one: 0 auto (round one read
                                off the sorted adjacency lists, later rounds pulled
                                when the frontier's edges are at least half of all
                                edges, pushed otherwise); 1 push every round (one
                                thread per frontier edge, atomic minimum at the far
                                end: the reference's join order); 2 pull every round
                                (each node takes the least label of its neighbours).
                                1 and 2 also select rounds launched one by one under
                                strategy 0.  Same labels and iteration counts        */
    int32_t queue;           /* TC in the bitmap layout, plain runs: rounds queued
                                ahead of the host, each reading its sizes from the
                                device state the round before left and leaving its
                                totals in a ring of mapped host words (a round
                                whose output outgrows its buffer does nothing and
                                is queued again after the buffer grows).  0 auto
                                (4), 1 one round at a time, at most 8            */
    int32_t reserved[1];
} dlg_opts;

/* Phase timers in the reference's CSV order (tc.cu:413-420). */
enum {
    DLG_PH_JOIN = 0, DLG_PH_BUFPREP, DLG_PH_COMM, DLG_PH_DEDUP, DLG_PH_MERGE, DLG_PH_INIT,
    DLG_PH_HASHTABLE, DLG_PH_FINAL, DLG_PH_CLEAR, DLG_PH_FILEIO, DLG_PH_COUNT
};

typedef struct dlg_result {
    int32_t iterations;      /* as the reference counts them                        */
    int64_t output_size;     /* |TC|, |SG|, |single join output|, #components (WCC) */
    int64_t secondary;       /* WCC: nodes in the largest component                 */
    int64_t nodes;           /* WCC: distinct nodes                                  */
    int64_t raw_join_rows;   /* sum of |J| over all joins (before dedup)             */
    int64_t input_rows;
    int32_t *rows;           /* host rows in the reference's file order, or NULL    */
    int64_t n_rows;
    double device_seconds;   /* upload-to-resident-result wall time on the device    */
    double phase_seconds[DLG_PH_COUNT];
    /* With check_output: a digest of the output rows (r0, r1) in file order,
     * sum over rows of r0*0x9E3779B97F4A7C15 + r1*0xC2B2AE3D27D4EB4F + r0*r1
     * + 0x632BE59BD9B4E019 (r0, r1 as uint32, arithmetic mod 2^64;
     * order-independent), and the
     * number of adjacent row pairs not strictly increasing in the output's
     * sort order (0 for a sorted set). */
    uint64_t output_sum;
    int64_t output_order_breaks;
} dlg_result;

/* Whole-program entry points, the benchmark() role of tc.cu:42-423,
 * sg.cu:40-500, wcc.cu:58-581 and single_join.cu:39-396 minus MPI and file
 * I/O.  Input rows are host int32 pairs (the reference's file layout); the
 * *_device variants take rows already resident in device memory and time
 * only device work.  Results go to `res`; free with dlg_result_free. */
dlg_status dlg_run_tc(dlg_ctx *ctx, const int32_t *h_rows, int64_t n, const dlg_opts *o,
                      dlg_result *res);
dlg_status dlg_run_sg(dlg_ctx *ctx, const int32_t *h_rows, int64_t n, const dlg_opts *o,
                      dlg_result *res);
dlg_status dlg_run_wcc(dlg_ctx *ctx, const int32_t *h_rows, int64_t n, const dlg_opts *o,
                       dlg_result *res);
dlg_status dlg_run_single_join(dlg_ctx *ctx, const int32_t *h_rows, int64_t n,
                               const dlg_opts *o, dlg_result *res);
dlg_status dlg_run_tc_device(dlg_ctx *ctx, const int32_t *d_rows, int64_t n, const dlg_opts *o,
                             dlg_result *res);
dlg_status dlg_run_sg_device(dlg_ctx *ctx, const int32_t *d_rows, int64_t n, const dlg_opts *o,
                             dlg_result *res);
dlg_status dlg_run_wcc_device(dlg_ctx *ctx, const int32_t *d_rows, int64_t n, const dlg_opts *o,
                              dlg_result *res);
dlg_status dlg_run_single_join_device(dlg_ctx *ctx, const int32_t *d_rows, int64_t n,
                                      const dlg_opts *o, dlg_result *res);
/* Frees a result's host rows.  Rows of up to 4 GiB come in page-locked host
 * memory (while the library holds at most 32 GiB of it), larger ones in
 * malloc'd memory; free them only with dlg_result_free (or dlg_host_free
 * once taken out of the result). */
void dlg_result_free(dlg_result *res);

/* ------------------------------------------------------- batched form */
/* No counterpart in the reference (one instance per process there).  `count`
 * independent instances of one program advance together: their inputs are
 * placed in disjoint id ranges and run as ONE fixpoint over the disjoint
 * union, so every kernel launch serves every live instance.  Per-instance
 * results are exact: rows, sizes and iteration counts equal the instance's
 * plain run (an instance's iteration count is the first round in which it
 * derived nothing; WCC instances stop by their own termination rule and are
 * then frozen).  The summed id ranges of one batch must fit 2^32 ids
 * (DLG_OVERFLOW otherwise: split the batch).  results[i].raw_join_rows is -1.
 * With check_output, results[i].output_sum / output_order_breaks are the
 * instance's own digest (its rows in its own ids), equal to its plain run's. */
typedef enum dlg_prog { DLG_PROG_TC = 0, DLG_PROG_SG = 1, DLG_PROG_WCC = 2, DLG_PROG_SJ = 3 } dlg_prog;
typedef struct dlg_batch_stats {
    double device_seconds;   /* upload-to-resident wall time of the whole batch      */
    int32_t rounds;          /* fixpoint rounds run (max over instances)             */
    int64_t total_output;    /* sum of output_size                                    */
    int64_t union_rows;      /* input rows of the union                               */
} dlg_batch_stats;
dlg_status dlg_batch_run(dlg_ctx *ctx, int prog, int32_t count, const int32_t *const *h_rows,
                         const int64_t *n, const dlg_opts *o, dlg_result *results,
                         dlg_batch_stats *stats);

/* Autotuner: times the program on `h_rows` for every candidate in the grid
 * load_factor x radix_bits x block_size (zero-terminated lists, NULL = the
 * built-in grid) and writes the fastest setting to *best.  Results of every
 * candidate are checked to be identical (DLG_DEVICE_ERROR otherwise). */
dlg_status dlg_autotune(dlg_ctx *ctx, int prog, const int32_t *h_rows, int64_t n,
                        const double *load_factors, const int32_t *radix_bits,
                        const int32_t *block_sizes, int32_t repeats, dlg_opts *best,
                        double *best_seconds);

/* ------------------------------------------------------- served form */
/* No counterpart in the reference.  A request loop on one device:
 *  - `workers` slots, each with its own HIP stream and memory pool; kernels
 *    of requests in different slots overlap on the device;
 *  - admission in arrival order against an HBM budget: a request reserves
 *    its predicted peak footprint (from its row count and id range) and is
 *    admitted when a slot is free and the reservation fits; a request that
 *    can never fit is refused at submission (DLG_REFUSED);
 *  - continuous batching: a slot that frees up takes the oldest admissible
 *    request and, if it is small (<= small_rows), every other waiting small
 *    request of the same program that fits, and runs them as one batch
 *    (dlg_batch_run semantics, exact per request);
 *  - a request that outgrows its reservation is re-run alone with the whole
 *    budget; resident graphs stay on the device across requests.
 * Every result equals the request's plain run (same rows and counts). */
typedef struct dlg_server dlg_server;
typedef struct dlg_server_opts {
    int32_t device;
    int32_t workers;         /* concurrent slots (streams), default 4                   */
    int64_t hbm_budget;      /* bytes for requests; 0 = 85 % of free device memory       */
    int32_t max_batch;       /* most requests in one batch, default 64                   */
    int64_t small_rows;      /* requests up to this many rows are batched, default 20000 */
    int64_t retain_bytes;    /* pool cache kept per slot between requests, default 256 MiB */
    int64_t max_request_bytes; /* refuse requests predicted above this; 0 = the budget   */
    dlg_opts tune;           /* tuning applied to every request                         */
} dlg_server_opts;
typedef struct dlg_request_info {
    int32_t status;          /* dlg_status of the request                               */
    int32_t batch_size;      /* requests in the batch it ran in                          */
    int32_t attempts;        /* 2 when it was re-run alone                               */
    int32_t slot;
    double submit_s, start_s, end_s;   /* seconds since the server was created           */
    int64_t reserved_bytes;
} dlg_request_info;
dlg_status dlg_server_create(const dlg_server_opts *o, dlg_server **out);
/* Keep a graph resident for WCC queries (graph ids from 0). */
dlg_status dlg_server_add_graph(dlg_server *s, const int32_t *h_rows, int64_t n, int32_t *graph_id);
/* Queue a request: prog on the given rows (copied), or, with graph_id >= 0,
 * WCC on a resident graph (rows ignored). */
dlg_status dlg_server_submit(dlg_server *s, int prog, const int32_t *h_rows, int64_t n,
                             int32_t graph_id, uint64_t *ticket);
/* Wait up to timeout_s (< 0: forever) for a request.  DLG_PENDING on timeout;
 * otherwise DLG_OK and the request's own status in info->status.  The result
 * is handed over once (free with dlg_result_free). */
dlg_status dlg_server_wait(dlg_server *s, uint64_t ticket, double timeout_s, dlg_result *res,
                           dlg_request_info *info);
/* Finishes queued requests, joins the slots and frees device memory. */
void dlg_server_destroy(dlg_server *s);
/* Last server-level error text. */
const char *dlg_server_error(const dlg_server *s);

/* ------------------------------------------------------- input formats */
/* Replaces parallel_read (common/parallel_io.cu:5-41) on one rank: raw
 * little-endian int32 pairs, rows = filesize / 8 (trailing bytes ignored).
 * *h_rows is malloc'd; free with dlg_host_free. */
dlg_status dlg_read_rows(const char *path, int32_t **h_rows, int64_t *n);
/* Replaces parallel_write (common/parallel_io.cu:110-126) on one rank; the
 * file is truncated first (the reference never truncates, so a stale longer
 * file kept its tail). */
dlg_status dlg_write_rows(const char *path, const int32_t *h_rows, int64_t n);
/* Replaces parallel_generate (common/parallel_io.cu:79-108) on rank 0:
 * srand(1); v[i] = rand() % range + 1 for i < 2*rows. */
dlg_status dlg_generate_rand_rows(int64_t rows, int32_t range, int32_t **h_rows);
/* Frees any host block the library handed out (rows of results and of the
 * generators and readers below). */
void dlg_host_free(void *p);

/* No counterpart in the reference.  |TC| of the rows (src, dst) before
 * running it, for admission by predicted output and for output caps, on the
 * host: SCC condensation (Tarjan), then per component C the nodes reachable
 * from it, |TC| = sum over C of |C| * |reach(C)|.  Exact (*exact = 1) when the
 * condensation has at most 16,384 components; otherwise estimated from the
 * reach of up to 4,096 components drawn by size (*exact = 0).  `exact` may be NULL. */
dlg_status dlg_predict_tc_size(const int32_t *h_rows, int64_t n, int64_t *tc_size, int32_t *exact);

/* Synthetic workloads (deterministic: SplitMix64-seeded PCG32, seed per
 * instance).  All return malloc'd int32 rows freed with dlg_host_free. */
/* Directed graph with power-law out-degree P(d) ~ d^-alpha, d in [1, n-1],
 * uniform destinations, no self-loops, exactly m distinct edges. */
dlg_status dlg_gen_powerlaw(int32_t n, int64_t m, double alpha, uint64_t seed,
                            int32_t **h_rows, int64_t *n_rows);
/* side x side grid, each right/down lattice edge kept with probability p and
 * stored in both directions, plus round(shortcut_frac * kept) shortcut edges
 * between vertices within Manhattan distance max_dist (both directions). */
dlg_status dlg_gen_grid(int32_t side, double p, double shortcut_frac, int32_t max_dist,
                        uint64_t seed, int32_t **h_rows, int64_t *n_rows);
/* cells^3 structured hex mesh, each cell split into 5 tetrahedra with
 * parity-alternating orientation; tetrahedron edges directed low -> high id. */
dlg_status dlg_gen_tetmesh(int32_t cells, int32_t **h_rows, int64_t *n_rows);
/* Erdos-Renyi directed graph: n vertices, m distinct edges, no self-loops. */
dlg_status dlg_gen_erdos_renyi(int32_t n, int64_t m, uint64_t seed, int32_t **h_rows,
                               int64_t *n_rows);
/* R-MAT(a, b, c, 1-a-b-c) on 2^scale vertices, m distinct edges, no self-loops
 * (when `undirected`, each kept edge is also stored reversed). */
dlg_status dlg_gen_rmat(int32_t scale, int64_t m, double a, double b, double c, uint64_t seed,
                        int undirected, int32_t **h_rows, int64_t *n_rows);

/* ------------------------------------------------- no-allocation join */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): dlg_join with the output in a caller buffer and every
 * scratch word (per-probe match counts and row starts, their scan and its
 * partials) in one static device array of DLG_NOALLOC_SCRATCH_BYTES that each
 * GPU the library runs on holds until the process ends.  The call makes no
 * hipMalloc / hipFree and no other allocating library call, and records no
 * profiling events.  Limits: at most DLG_JOIN_NOALLOC_MAX_PROBES probes
 * (DLG_INVALID_INPUT beyond, with a message); projections DLG_PROJ_HV_PV and
 * DLG_PROJ_PV_HV.  *out_n = |J|; when |J| > out_cap nothing is written and
 * DLG_OVERFLOW is returned with *out_n the capacity needed.  All calls on one
 * GPU share its array: calls that overlap on one GPU (several contexts or
 * threads) race and must be serialised.  Synchronises the context's stream.
 * Same rows in the same order as dlg_join. */
#define DLG_NOALLOC_SCRATCH_BYTES (64ll << 20)
#define DLG_JOIN_NOALLOC_MAX_PROBES 4000000ll
dlg_status dlg_join_noalloc(dlg_ctx *ctx, const dlg_hash *h, const uint64_t *d_probe, int64_t n, int proj,
                            uint64_t *d_out, int64_t out_cap, int64_t *out_n);

#ifdef __cplusplus
}
#endif
#endif /* DLG_H */
