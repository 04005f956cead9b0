// single_join -- one self-join of a random relation, drop-in for single_join.cu.
// Usage: single_join <rows> [cuda_aware_mpi] [method] [rand_range=1000000]  (single_join.cu:81-100)
// The relation is glibc srand(1) / rand() % range + 1 (common/parallel_io.cu:79-108);
// output = |unique{(c, a) : R(a, b), R(b, c)}|.  Like the reference (job_run fixed
// to 1, single_join.cu:78) no output file is written.
#include "cli_common.h"

int main(int argc, char **argv) {
    cli::Args a = cli::parse(argc, argv, "hipc_2019.bin", 1000000);
    int rand_range = a.last;
    int64_t total_rows = std::atoi(a.input);
    dlg_ctx *ctx = nullptr;
    double t0 = cli::now_s();
    dlg_status s = dlg_ctx_create(0, &ctx);
    if (s != DLG_OK) cli::die(nullptr, s, "dlg_ctx_create");
    double init_s = cli::now_s() - t0;
    if (total_rows < 0 || rand_range <= 0) {
        std::fprintf(stderr, "invalid rows %lld or rand_range %d\n", (long long)total_rows, rand_range);
        dlg_ctx_destroy(ctx);
        return 10 + (int)DLG_INVALID_INPUT;
    }
    double tg = cli::now_s();
    int32_t *rows = nullptr;
    s = dlg_generate_rand_rows(total_rows, rand_range, &rows);
    if (s != DLG_OK) cli::die(ctx, s, "generate");
    double gen_s = cli::now_s() - tg;
    dlg_opts o = cli::tuning_from_env();
    o.keep_output = 0;
    dlg_result r;
    s = dlg_run_single_join(ctx, rows, total_rows, &o, &r);
    if (s != DLG_OK) cli::die(ctx, s, "single join");
    const double *p = r.phase_seconds;
    double init = p[DLG_PH_INIT] + init_s;
    double total = init + p[DLG_PH_HASHTABLE] + p[DLG_PH_JOIN] + p[DLG_PH_DEDUP] + p[DLG_PH_MERGE] +
                   p[DLG_PH_FINAL] + p[DLG_PH_CLEAR];
    std::printf("# Input,# Process,# Iterations,# Output,Total Time,Join,Buffer preparation (data "
                "distribution),Communication (data distribution),Buffer preparation (join result),"
                "Communication (join result),Deduplication,Clear,Copy,Finalization,Initialization,"
                "File I/O,Hashtable\n");
    std::printf("%d,%d,%d,%lld,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf\n",
                (int)total_rows, 1, r.iterations, (long long)r.output_size, total, p[DLG_PH_JOIN], 0.0,
                0.0, 0.0, 0.0, p[DLG_PH_DEDUP], p[DLG_PH_CLEAR], 0.0, p[DLG_PH_FINAL], init, gen_s,
                p[DLG_PH_HASHTABLE]);
    dlg_result_free(&r);
    dlg_host_free(rows);
    dlg_ctx_destroy(ctx);
    return 0;
}
