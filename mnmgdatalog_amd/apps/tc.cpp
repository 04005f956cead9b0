// tc -- transitive closure, drop-in for the reference's tc.cu executable.
//   tc(x, y) :- edge(x, y).   tc(x, z) :- tc(x, y), edge(y, z).
// Usage: tc <file.bin> [cuda_aware_mpi] [method] [job_run]   (tc.cu:78-95)
// Output file <file.bin>_tc.bin: rows (src, dst) ordered by (dst, src) (tc.cu:316-366).
#include "cli_common.h"

int main(int argc, char **argv) {
    return cli::run_graph_app(argc, argv, "hipc_2019.bin", "_tc.bin", dlg_run_tc, false);
}
