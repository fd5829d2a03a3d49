// host_internal.hpp -- shared definitions of the host library.
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "lpa_host.h"

struct lpah_csr {
    uint64_t n = 0;
    std::vector<uint64_t> row;   // n + 1
    std::vector<uint32_t> col;   // m
};

struct lpah_cgr {
    uint64_t n = 0, n_c = 0, nbits = 0;
    std::vector<uint64_t> words;      // stream, big-endian bit order per u64, +2 zero words
    std::vector<uint32_t> words32;    // same stream as big-endian u32 words, +4 zero words
    std::vector<uint64_t> offsets;    // n_c + 1 start bit offsets
    std::vector<uint32_t> vertices;   // n
    std::vector<uint64_t> small_offs; // n - n_c + 1
    std::vector<uint32_t> small_nbrs;
};

namespace lpah {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

enum { E_ARG = -1, E_IO = -2, E_FORMAT = -3, E_MEM = -4, E_RANGE = -5 };

void set_error(const std::string &m);

// Runs f and converts exceptions into a status code + last-error string.
template <class F> int guard(F &&f) {
    try {
        f();
        return 0;
    } catch (const Error &e) {
        set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error("out of host memory");
        return E_MEM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return E_ARG;
    }
}

// splitmix64 finaliser and counter-based streams: value i of stream `key`.
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return mix64(seed * 0x9E3779B97F4A7C15ull + stream + 0x632BE59BD9B4E019ull);
}
static inline uint64_t ctr_rand(uint64_t key, uint64_t i) {
    return mix64(key + (i + 1) * 0x9E3779B97F4A7C15ull);
}
static inline double u01(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }

// Phase timing of graph preparation on stderr when LPA_HOST_TIMING is set.
void phase(const char *name);

// Builds a CSR from arcs (src[i] -> dst[i]); removes self loops and
// duplicates and sorts rows.  Consumes the arrays.
void build_csr(uint64_t n, std::vector<uint32_t> &src, std::vector<uint32_t> &dst,
               bool undirected, lpah_csr &out);

}  // namespace lpah
