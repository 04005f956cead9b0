// wcc -- weakly connected components, drop-in for the reference's wcc.cu executable.
// Usage: wcc <file.bin> [cuda_aware_mpi] [method] [job_run]   (wcc.cu:96-113)
// Output file <file.bin>_cc.bin: rows (label, 0), one per component, by label (wcc.cu:485-523).
#include "cli_common.h"

int main(int argc, char **argv) {
    return cli::run_graph_app(argc, argv, "data/dummy.bin", "_cc.bin", dlg_run_wcc, true);
}
