// sg_nl -- same generation with the sorted-index join, drop-in for the
// reference's sg_nl.cu executable (binary-search probes, sg_nl.cu:168-330,
// common/join.cu:89-180).  Same argv, output file (<file.bin>_sg.bin) and
// CSV as sg.
#include "cli_common.h"

int main(int argc, char **argv) {
    return cli::run_graph_app(argc, argv, "hipc_2019.bin", "_sg.bin", dlg_run_sg, false, 1);
}
