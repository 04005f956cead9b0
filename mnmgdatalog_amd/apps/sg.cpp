// sg -- same generation, drop-in for the reference's sg.cu executable.
//   sg(x, y) :- edge(p, x), edge(p, y), x != y.   sg(x, y) :- edge(a, x), sg(a, b), edge(b, y).
// Usage: sg <file.bin> [cuda_aware_mpi] [method] [job_run]   (sg.cu:74-91)
// Output file <file.bin>_sg.bin: rows (x, y) ordered by (x, y) (sg.cu:392-442).
// The CSV header says "# TC" as the reference's does (sg.cu:490).
#include "cli_common.h"

int main(int argc, char **argv) {
    return cli::run_graph_app(argc, argv, "hipc_2019.bin", "_sg.bin", dlg_run_sg, false);
}
