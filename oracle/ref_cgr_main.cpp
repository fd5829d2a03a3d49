// ref_cgr_main.cpp -- driver that runs the REFERENCE compressor/loader,
// built from its own sources in place (src/compressor.cpp, src/file.cpp),
// to emit golden CGR file sets for small graphs.  Test infrastructure only.
//
// usage: ref_cgr <edge_list.txt> <one_way 0|1> <small_thresh>
// Writes <edge_list.txt>.{graph,offset,vertices.bin,graph.small.bin,
// offset.small.bin,row,clm} next to the input, exactly as
// File::load_graph does on a cache miss (file.cpp:24-74, 76-95).
#include "file.h"
#include <cstdlib>

int main(int argc, char **argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: %s edges.txt one_way small_thresh\n", argv[0]);
        return 2;
    }
    File f(argv[1], true, atoi(argv[2]) != 0, atoi(argv[3]));
    f.load_graph();
    return 0;
}
