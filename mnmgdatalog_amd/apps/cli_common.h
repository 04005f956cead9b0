// cli_common.h -- shared driver code of the tc / sg / wcc / single_join executables.
//
// Command line, stdout and output file follow the reference programs on one
// rank (tc.cu:78-95, 359-366, 413-420; sg.cu:74-91, 490-497; wcc.cu:96-113,
// 571-578; single_join.cu:81-100, 385-393):
//   tc|sg|wcc <file.bin> [cuda_aware_mpi] [method] [job_run]
//   single_join <rows> [cuda_aware_mpi] [method] [rand_range]
// cuda_aware_mpi and method select MPI transports in the reference and are
// accepted and ignored here (one GPU, no exchange).  job_run == 0 writes the
// output file.  Tuning knobs come from the environment (DLG_LOAD_FACTOR,
// DLG_RADIX_BITS, DLG_STRATEGY) or a tuning file named by DLG_TUNING
// (lines "key=value"), never from the command line.
// Errors go to stderr with a non-zero exit code (the reference exits through
// checkCuda / MPI_Abort).
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "dlg.h"

namespace cli {

inline void die(dlg_ctx *ctx, dlg_status s, const char *what) {
    std::fprintf(stderr, "%s failed (status %d): %s\n", what, (int)s,
                 ctx ? dlg_last_error(ctx) : "");
    if (ctx) dlg_ctx_destroy(ctx);
    std::exit(10 + (int)s);
}

inline void apply_kv(dlg_opts &o, const char *k, const char *v) {
    if (!std::strcmp(k, "load_factor")) o.load_factor = std::atof(v);
    else if (!std::strcmp(k, "radix_bits")) o.radix_bits = std::atoi(v);
    else if (!std::strcmp(k, "strategy")) o.strategy = std::atoi(v);
    else if (!std::strcmp(k, "block_size")) o.block_size = std::atoi(v);
    else if (!std::strcmp(k, "join_method")) o.join_method = std::atoi(v);
}

inline dlg_opts tuning_from_env() {
    dlg_opts o;
    std::memset(&o, 0, sizeof(o));
    if (const char *f = std::getenv("DLG_TUNING")) {
        if (FILE *fp = std::fopen(f, "r")) {
            char line[256];
            while (std::fgets(line, sizeof line, fp)) {
                char *eq = std::strchr(line, '=');
                if (!eq || line[0] == '#') continue;
                *eq = 0;
                char *v = eq + 1;
                v[std::strcspn(v, "\r\n")] = 0;
                apply_kv(o, line, v);
            }
            std::fclose(fp);
        }
    }
    if (const char *v = std::getenv("DLG_LOAD_FACTOR")) o.load_factor = std::atof(v);
    if (const char *v = std::getenv("DLG_RADIX_BITS")) o.radix_bits = std::atoi(v);
    if (const char *v = std::getenv("DLG_STRATEGY")) o.strategy = std::atoi(v);
    return o;
}

struct Args {
    const char *input;
    int cuda_aware_mpi = 0, method = 0, job_run = 0, last = 0;
};

// The reference's if-chain on argc (tc.cu:78-94): 2..5 arguments, otherwise
// the default input.  For single_join the fourth argument is rand_range.
inline Args parse(int argc, char **argv, const char *default_input, int job_run_default) {
    Args a;
    a.input = default_input;
    a.job_run = job_run_default;
    a.last = job_run_default;
    if (argc >= 2 && argc <= 5) a.input = argv[1];
    if (argc >= 3 && argc <= 5) a.cuda_aware_mpi = std::atoi(argv[2]);
    if (argc >= 4 && argc <= 5) a.method = std::atoi(argv[3]);
    if (argc == 5) a.last = std::atoi(argv[4]);
    a.job_run = a.last;
    return a;
}

typedef dlg_status (*runner_t)(dlg_ctx *, const int32_t *, int64_t, const dlg_opts *, dlg_result *);

inline double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// tc / sg / wcc driver.  `suffix` names the output file; `wcc` selects the
// WCC CSV columns.
inline int run_graph_app(int argc, char **argv, const char *default_input, const char *suffix,
                         runner_t fn, bool wcc, int join_method = 0) {
    Args a = parse(argc, argv, default_input, 0);
    std::string out_name = std::string(a.input) + suffix;
    dlg_ctx *ctx = nullptr;
    double t_init0 = now_s();
    dlg_status s = dlg_ctx_create(0, &ctx);
    if (s != DLG_OK) die(nullptr, s, "dlg_ctx_create");
    double init_s = now_s() - t_init0;

    double t_io0 = now_s();
    int32_t *rows = nullptr;
    int64_t n = 0;
    s = dlg_read_rows(a.input, &rows, &n);
    if (s != DLG_OK) {
        std::fprintf(stderr, "Error opening file %s\n", a.input);
        dlg_ctx_destroy(ctx);
        return 10 + (int)s;
    }
    double fileio_s = now_s() - t_io0;

    dlg_opts o = tuning_from_env();
    o.keep_output = 1;
    if (join_method) o.join_method = join_method;   // the *_nl executables
    dlg_result r;
    s = fn(ctx, rows, n, &o, &r);
    if (s != DLG_OK) die(ctx, s, "fixpoint");

    if (a.job_run == 0) {
        double t0 = now_s();
        s = dlg_write_rows(out_name.c_str(), r.rows, r.n_rows);
        if (s != DLG_OK) {
            std::fprintf(stderr, "Error writing file %s\n", out_name.c_str());
            dlg_result_free(&r);
            dlg_ctx_destroy(ctx);
            return 10 + (int)s;
        }
        std::printf("Rank 0 wrote local tuples on file: %s\n", out_name.c_str());
        fileio_s += now_s() - t0;
    }
    const double *p = r.phase_seconds;
    double init = p[DLG_PH_INIT] + init_s;
    double total = init + p[DLG_PH_HASHTABLE] + p[DLG_PH_JOIN] + p[DLG_PH_BUFPREP] + p[DLG_PH_COMM] +
                   p[DLG_PH_DEDUP] + p[DLG_PH_MERGE] + p[DLG_PH_FINAL] + p[DLG_PH_CLEAR];
    if (wcc) {
        std::printf("# Input,# Process,# Iterations,# CC,# Nodes in largest WCC,Total Time,Join,"
                    "Buffer preparation,Communication,Deduplication,Merge,Initialization,Hashtable,"
                    "Finalization,Clear,File I/O\n");
        std::printf("%d,%d,%d,%lld,%lld,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf\n",
                    (int)n, 1, r.iterations, (long long)r.output_size, (long long)r.secondary, total,
                    p[DLG_PH_JOIN], p[DLG_PH_BUFPREP], p[DLG_PH_COMM], p[DLG_PH_DEDUP], p[DLG_PH_MERGE],
                    init, p[DLG_PH_HASHTABLE], p[DLG_PH_FINAL], p[DLG_PH_CLEAR], fileio_s);
    } else {
        std::printf("# Input,# Process,# Iterations,# TC,Total Time,Join,Buffer preparation,"
                    "Communication,Deduplication,Merge,Initialization,Hashtable,Finalization,Clear,"
                    "File I/O\n");
        std::printf("%d,%d,%d,%lld,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf,%.4lf\n",
                    (int)n, 1, r.iterations, (long long)r.output_size, total, p[DLG_PH_JOIN],
                    p[DLG_PH_BUFPREP], p[DLG_PH_COMM], p[DLG_PH_DEDUP], p[DLG_PH_MERGE], init,
                    p[DLG_PH_HASHTABLE], p[DLG_PH_FINAL], p[DLG_PH_CLEAR], fileio_s);
    }
    dlg_result_free(&r);
    dlg_host_free(rows);
    dlg_ctx_destroy(ctx);
    return 0;
}

}  // namespace cli
