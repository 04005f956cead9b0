// tc_nl -- transitive closure with the sorted-index join, drop-in for the
// reference's tc_nl.cu executable: the edge relation is sorted and
// deduplicated, each probe finds its matches by binary search
// (tc_nl.cu:182-185, common/join.cu:89-180).  Same argv, output file
// (<file.bin>_tc.bin) and CSV as tc.
#include "cli_common.h"

int main(int argc, char **argv) {
    return cli::run_graph_app(argc, argv, "hipc_2019.bin", "_tc.bin", dlg_run_tc, false, 1);
}
