// file.cpp -- CgrCompressor and File over liblpa_host.so (see lpa/file.h).
#include "lpa/file.h"

#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>

namespace lpa_api {

namespace {

void check(int rc, const char *what) {
    if (rc != 0) {
        std::cerr << what << ": " << lpah_last_error() << std::endl;
        abort();   // the reference aborts on I/O errors (file.cpp:104-108, 119-123)
    }
}

bool exists(const std::string &p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

template <class T> void read_vec(const std::string &p, std::vector<T> &v) {
    FILE *f = fopen(p.c_str(), "rb");
    if (!f) {
        std::cerr << "open: " << p << "  failed!" << std::endl;
        abort();
    }
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    v.resize((size_t)sz / sizeof(T));
    if (!v.empty() && fread(v.data(), sizeof(T), v.size(), f) != v.size()) {
        std::cerr << "short read: " << p << std::endl;
        abort();
    }
    fclose(f);
}

}  // namespace

CgrCompressor::CgrCompressor(int zeta_k, int min_itv_len, int itv_seg_len, int res_seg_len, int one_way)
    : one_way_(one_way != 0) {
    if (zeta_k != 3 || min_itv_len != 4 || itv_seg_len != 256 || res_seg_len != 256)
        throw std::invalid_argument("CgrCompressor: only zeta 3, min interval 4, 256-bit segments are supported");
}

CgrCompressor::~CgrCompressor() {
    if (cgr_) lpah_cgr_free(cgr_);
    if (csr_) lpah_csr_free(csr_);
}

bool CgrCompressor::load_graph(const std::string &file_path) {
    if (csr_) lpah_csr_free(csr_);
    csr_ = nullptr;
    check(lpah_csr_read_edge_list(file_path.c_str(), one_way_ ? 1 : 0, &csr_), "load_graph");
    return true;
}

void CgrCompressor::compress() { (void)compresswithcost(); }

double CgrCompressor::compresswithcost() {
    // encoding is deferred to write_cgr, which knows the thresholds; the
    // cost is measured there with the same parameters
    auto t0 = std::chrono::steady_clock::now();
    if (cgr_) lpah_cgr_free(cgr_);
    cgr_ = nullptr;
    check(lpah_cgr_encode(csr_, 32, 128, &cgr_), "compress");
    small_ = 32;
    large_ = 128;
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

bool CgrCompressor::write_cgr(const std::string &file_path, int small_node_thresh, int large_node_thresh) {
    if (!csr_) throw std::logic_error("write_cgr before load_graph");
    if (!cgr_ || small_ != small_node_thresh || large_ != large_node_thresh) {
        if (cgr_) lpah_cgr_free(cgr_);
        cgr_ = nullptr;
        check(lpah_cgr_encode(csr_, (uint32_t)small_node_thresh, (uint32_t)large_node_thresh, &cgr_), "compress");
        small_ = small_node_thresh;
        large_ = large_node_thresh;
    }
    check(lpah_files_write(cgr_, nullptr, file_path.c_str()), "write_cgr");
    return true;
}

File::File(const std::string _file_name, bool _compress, bool _one_way, int sms)
    : small_node_thresh(sms), file_name(_file_name), compress(_compress), one_way(_one_way),
      has_small(sms != 0) {}

int File::cache_state() const {
    std::ifstream f(file_name + ".lpa_params");
    if (!f) return 0;
    int t = -1, o = -1, c = -1;
    if (!(f >> t >> o >> c)) return -1;
    return t == small_node_thresh && o == (one_way ? 1 : 0) && c == (compress ? 1 : 0) ? 1 : -1;
}

bool File::cache_matches() const { return cache_state() == 1; }

// A file set with no parameter record (compressed runs only: an
// uncompressed reference set has no .graph, which this loader reads): its
// .offset count must be the number of vertices of degree >= the threshold in
// .row, else the set was written for another threshold.
static bool reference_set_consistent(const std::string &base, bool compress, int thresh) {
    if (!compress) return false;
    std::vector<uint32_t> row;
    read_vec(base + ".row", row);
    std::ifstream f(base + ".offset");
    uint64_t nc = 0;
    if (row.empty() || !(f >> nc)) return false;
    uint64_t want = 0;
    for (size_t v = 0; v + 1 < row.size(); v++) want += row[v + 1] - row[v] >= (uint32_t)thresh;
    return nc == want;
}

void File::write_binary() {
    lpah_csr *g = nullptr;
    check(lpah_csr_read_edge_list(file_name.c_str(), one_way ? 1 : 0, &g), "read edge list");
    lpah_cgr *c = nullptr;
    check(lpah_cgr_encode(g, compress ? (uint32_t)small_node_thresh : 0xFFFFFFFFu, 128, &c), "compress");
    check(lpah_files_write(c, g, file_name.c_str()), "write files");
    lpah_cgr_free(c);
    lpah_csr_free(g);
    std::ofstream(file_name + ".lpa_params") << small_node_thresh << " " << (one_way ? 1 : 0) << " "
                                             << (compress ? 1 : 0) << "\n";
}

void File::load_graph() {
    // the files File::load_graph reads (file.cpp:29-60)
    const bool have = exists(file_name + ".row") && exists(file_name + ".clm") && exists(file_name + ".graph") &&
                      exists(file_name + ".offset") && exists(file_name + ".vertices.bin");
    const int st = cache_state();
    const int t_req = !compress ? 1 : (has_small ? small_node_thresh : 0);
    if (!have || st < 0 || (st == 0 && !reference_set_consistent(file_name, compress, t_req))) write_binary();
    lpah_cgr *c = nullptr;
    // uncompressed runs keep every vertex in the small CSR
    check(lpah_files_load(file_name.c_str(), t_req, &c), "load");
    {
        const uint64_t nw = lpah_cgr_nwords(c);
        graph.assign(lpah_cgr_words(c), lpah_cgr_words(c) + nw);
        offset.assign(lpah_cgr_offsets(c), lpah_cgr_offsets(c) + lpah_cgr_nc(c) + 1);
        if (has_small || !compress) {
            const uint64_t ns = lpah_cgr_n(c) - lpah_cgr_nc(c);
            const uint64_t *so = lpah_cgr_small_offsets(c);
            small_offset.assign(so, so + ns + 1);
            small_graph.assign(lpah_cgr_small_nbrs(c), lpah_cgr_small_nbrs(c) + lpah_cgr_small_m(c));
        }
    }
    vertices.assign(lpah_cgr_vertices(c), lpah_cgr_vertices(c) + lpah_cgr_n(c));
    lpah_cgr_free(c);
    read_vec(file_name + ".row", row);
    read_vec(file_name + ".clm", clm);
    vertex_num = compress ? (Uint)vertices.size() : (Uint)(row.size() - 1);
}

}  // namespace lpa_api
