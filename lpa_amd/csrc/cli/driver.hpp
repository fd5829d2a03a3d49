// driver.hpp -- command-line drivers with the reference's argv
// (baselines/nova/nova.cu:21-90, baselines/novaplus/novaplus.cu:29-107,
// baselines/cglp/cglp.cu:21-90):
//
//   <method> dataset_path if_compressed if_directed threshold output_file [options]
//
// if_directed follows the reference's code, not its README: 1 means every
// edge is stored once and its reverse arc is added (file.cpp:136-140).
// The total LP seconds are appended to output_file in the reference's
// format.  Options (no reference counterpart):
//   --iters N            iterations (default 20, the reference's run(20))
//   --until-converged    stop when no label changes
//   --cms T,D,W,SEED     count-min sketch for degree > T (depth D, width W)
//   --labels-out PATH    write the final labels (uint32, n*K entries)
//   --device I
// A wrong argument count prints the usage (the reference silently falls
// back to a hard-coded dataset path, cglp.cu:51-57).
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "lpa/file.h"
#include "lpa/graph.h"
#include "lpa/label_propagator.h"

namespace lpa_cli {

inline int usage(const char *prog) {
    fprintf(stderr,
            "usage: %s dataset_path if_compressed if_directed threshold output_file\n"
            "          [--iters N] [--until-converged] [--cms T,D,W,SEED] [--labels-out PATH] [--device I]\n",
            prog);
    return 1;
}

// method: 0 = fused (nova / novaplus), 1 = two-kernel (cglp)
inline int driver_main(int argc, char **argv, const char *tag, int method) {
    using namespace lpa_api;
    if (argc < 6) return usage(argv[0]);
    const std::string input = argv[1];
    const bool compress = atoi(argv[2]) != 0;
    const bool once = atoi(argv[3]) != 0;
    const int thresh = atoi(argv[4]);
    const std::string output = argv[5];
    int iters = 20;
    std::string labels_out;
    LpOptions opt;
    for (int i = 6; i < argc; i++) {
        std::string a = argv[i];
        if (a == "--iters" && i + 1 < argc) iters = atoi(argv[++i]);
        else if (a == "--until-converged") opt.until_converged = true;
        else if (a == "--labels-out" && i + 1 < argc) labels_out = argv[++i];
        else if (a == "--device" && i + 1 < argc) opt.device = atoi(argv[++i]);
        else if (a == "--cms" && i + 1 < argc) {
            unsigned t, d, w;
            unsigned long long seed;
            if (sscanf(argv[++i], "%u,%u,%u,%llu", &t, &d, &w, &seed) != 4) return usage(argv[0]);
            opt.cms_thresh = t;
            opt.cms_depth = d;
            opt.cms_width = w;
            opt.cms_seed = seed;
        } else return usage(argv[0]);
    }
    File f(input, compress, once, thresh);
    f.load_graph();
    std::cout << "clm_size: " << f.clm.size() << " row_size: " << f.row.size() << " graph words: " << f.graph.size()
              << " offsets: " << f.offset.size() << " vertices: " << f.vertices.size() << std::endl;
    using Vertex = uint32_t;
    using Edge = uint32_t;
    auto graph = std::make_shared<CGRGraph<Vertex, Edge>>(f.clm, f.row, f.graph, f.offset, f.small_graph,
                                                         f.small_offset, f.vertices, thresh);
    std::unique_ptr<LabelPropagator<Vertex, Edge>> prop;
    if (method == 0) prop = make_unique_ptr<NovaPlusPropagator<Vertex, Edge>>(graph, opt);
    else prop = make_unique_ptr<CGLPPropagator<Vertex, Edge>>(graph, opt);
    std::pair<double, double> r;
    try {
        r = prop->run(iters);
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 2;
    }
    const double f1 = r.first, f2 = r.second;
    std::cout << "f1: " << f1 / 1000 << "(s)" << ",f2: " << f2 << std::endl;
    {
        std::ofstream w(output, std::ios::app);
        w << tag << " file: " << input << std::endl << " running time: " << f1 / 1000 << "thre: " << thresh << std::endl;
        w << "~~~~~" << std::endl;
    }
    if (!labels_out.empty()) {
        auto lab = prop->get_labels();
        FILE *o = fopen(labels_out.c_str(), "wb");
        if (!o || fwrite(lab.data(), sizeof(Vertex), lab.size(), o) != lab.size()) {
            std::cerr << "cannot write " << labels_out << std::endl;
            return 3;
        }
        fclose(o);
    }
    return 0;
}

}  // namespace lpa_cli
