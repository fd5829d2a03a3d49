// lpa/file.h -- the reference's host call surface for graph input:
// CgrCompressor (include/compressor.h:23-138) and File (include/file.h:28-76),
// backed by liblpa_host.so.
//
// Differences from the reference, all deliberate:
//  * File::offset holds 64-bit bit offsets (the reference's 32-bit ones
//    overflow beyond 2^32 bits, include/file.h:26).
//  * Self loops and duplicate arcs are removed when an edge list is read
//    (the reference encodes a duplicate as a gap of -1, compressor.cpp:459).
//  * The preprocessing cache records its threshold and direction in
//    <path>.lpa_params and is rebuilt when they differ (the reference reuses
//    a stale cache, file.cpp:29-52).  A file set without that record (one
//    the reference wrote) is loaded as the reference loads it, without the
//    edge list, once its compressed-vertex count is found consistent with the
//    requested threshold; an inconsistent one is rebuilt from the edge list.
//  * CgrCompressor supports the parameters File uses (zeta 3, minimum
//    interval 4, 256-bit segments); other values throw std::invalid_argument.
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "lpa_host.h"

namespace lpa_api {

using Uint = uint32_t;

class CgrCompressor {
  public:
    explicit CgrCompressor(int zeta_k = 3, int min_itv_len = 4, int itv_seg_len = 256, int res_seg_len = 256,
                           int one_way = true);
    ~CgrCompressor();
    bool load_graph(const std::string &file_path);
    void compress();
    double compresswithcost();   // seconds spent encoding
    bool write_cgr(const std::string &file_path, int small_node_thresh = 32, int large_node_thresh = 128);

    const lpah_csr *csr() const { return csr_; }

  private:
    bool one_way_;
    lpah_csr *csr_ = nullptr;
    lpah_cgr *cgr_ = nullptr;
    int small_ = -1, large_ = -1;
};

class File {
  public:
    File(const std::string file_name, bool compress, bool one_way, int sms = 32);
    void load_graph();

    // csr of all vertices
    std::vector<Uint> clm;
    std::vector<Uint> row;
    // cgr: big-endian 32-bit words, n_c+1 start bit offsets, stream-order vertices
    std::vector<Uint> graph;
    std::vector<uint64_t> offset;
    std::vector<Uint> vertices;
    // csr of the small (deg < sms) vertices
    std::vector<Uint> small_graph;
    std::vector<Uint> small_offset;
    Uint vertex_num = 0;
    int small_node_thresh;

  private:
    std::string file_name;
    bool compress;
    bool one_way;
    bool has_small;
    void write_binary();
    bool cache_matches() const;
    int cache_state() const;   // 1: matches, 0: absent, -1: differs
};

}  // namespace lpa_api
