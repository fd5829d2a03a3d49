// decode_check.hip -- test library: runs the device CGR decoders of
// lpa_amd/csrc/hip/cgr_decode.hpp and the device record encoder of
// encode_kernels.hpp on the host, so the tests can compare them with the
// CPU oracle without a GPU.
#include <cstdint>

#include "../../lpa_amd/csrc/hip/cgr_decode.hpp"
#include "../../lpa_amd/csrc/hip/encode_kernels.hpp"

namespace {
struct HostOr {   // one thread writes the whole stream here
    static void apply(uint64_t *w, uint64_t v) { *w |= v; }
};
}  // namespace

extern "C" {

// words: the stream as 64-bit words (stream bit 0 = MSB of word 0), with at
// least 8 zero words after the last data word.  Returns the degree (ids
// beyond cap are counted, not written); *end receives the bit after the record.
int64_t dc_decode_record(const uint64_t *words, uint64_t pos, uint32_t v, uint32_t *out, uint64_t cap,
                         uint64_t *end) {
    uint64_t d = 0;
    *end = lpa::decode_record(words, pos, v, [&](uint32_t u) {
        if (d < cap) out[d] = u;
        d++;
    });
    return (int64_t)d;
}

// Segment-parallel decode as the group kernels do it, one "lane" after the
// other: interval segments j = 0..ni-1 (each from its own seek), then the
// residual block from where the last interval segment ended.
int64_t dc_decode_segments(const uint64_t *words, uint64_t pos, uint32_t v, uint32_t *out, uint64_t cap) {
    uint64_t d = 0;
    auto put = [&](uint32_t u) {
        if (d < cap) out[d] = u;
        d++;
    };
    lpa::BitReader r;
    r.seek(words, pos);
    const uint32_t ni = r.gamma() + 1;
    const uint64_t base = r.pos;
    uint64_t res_start = 0;
    for (uint32_t j = 0; j < ni; j++) {
        lpa::BitReader s = r;
        s.move_to(base + (uint64_t)lpa::SEG_BITS * j);
        lpa::read_interval_segment(s, v, [&](uint32_t l, uint32_t len) {
            for (uint32_t x = 0; x < len; x++) put(l + x);
        });
        if (j == ni - 1) res_start = s.pos;
    }
    r.seek(words, res_start);
    const uint32_t nr = r.gamma() + 1;
    const uint64_t rbase = r.pos;
    for (uint32_t j = 0; j < nr; j++) {
        lpa::BitReader s = r;
        s.move_to(rbase + (uint64_t)lpa::SEG_BITS * j);
        lpa::read_residual_segment(s, v, put);
    }
    return (int64_t)d;
}

// Bits of vertex v's record (sorted adjacency adj[0..d)), as the device
// encoder measures it.
uint64_t dc_encode_bits(uint32_t v, const uint32_t *adj, uint32_t d) {
    lpa::enc::Meter m{0};
    lpa::enc::encode_record(m, v, adj, d);
    return m.pos;
}

// Writes the record into words (zeroed, 64-bit, bit 0 = MSB of word 0) at
// bits [start, end).
void dc_encode_record(uint32_t v, const uint32_t *adj, uint32_t d, uint64_t *words, uint64_t start, uint64_t end) {
    lpa::enc::BitWriter<HostOr> w(words, start, end);
    lpa::enc::encode_record(w, v, adj, d);
    w.finish();
}

}  // extern "C"
