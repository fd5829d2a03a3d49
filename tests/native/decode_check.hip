// decode_check.hip -- test library: runs the device CGR decoders of
// lpa_amd/csrc/hip/cgr_decode.hpp, the device record encoder of
// encode_kernels.hpp and the phases of the long-row encoder of
// encode_long.hpp (lane by lane) on the host, so the tests can compare them
// with the CPU oracle and with each other without a GPU.
#include <cstdint>
#include <vector>

#include "../../lpa_amd/csrc/hip/cgr_decode.hpp"
#include "../../lpa_amd/csrc/hip/encode_kernels.hpp"
#include "../../lpa_amd/csrc/hip/encode_long.hpp"

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

// The long-row encoder's phases for one row, run for lanes 0..T-1 in turn
// between the barriers of k_encode_long_len / k_encode_long_write.  Returns
// the record's bits (UINT64_MAX for an unsorted row or an id >= n) and, when
// words is not null, ORs the record into it at bit `start`.
uint64_t dc_encode_long(uint32_t v, const uint32_t *adj, uint32_t d, uint32_t n, uint32_t T, uint64_t *words,
                        uint64_t start) {
    using namespace lpa::enc;
    std::vector<uint64_t> ni(T), nr(T), oi(T), orr(T);
    bool ok = true;
    for (uint32_t t = 0; t < T; t++) ok = runs_count(adj, d, n, t, T, ni[t], nr[t]) && ok;
    if (!ok) return UINT64_MAX;
    uint64_t NI = 0, NR = 0;
    for (uint32_t t = 0; t < T; t++) {
        oi[t] = NI;
        orr[t] = NR;
        NI += ni[t];
        NR += nr[t];
    }
    std::vector<uint32_t> left(d + 1), len(d + 1), rid(d + 1), nxt_i(d + 1), nxt_r(d + 1), seg_i(d + 1),
        seg_r(d + 1);
    std::vector<uint64_t> P_i(d + 1), P_r(d + 1);
    for (uint32_t t = 0; t < T; t++) runs_place(adj, d, t, T, oi[t], orr[t], left.data(), len.data(), rid.data());
    LongBlock bi{true, v, left.data(), len.data(), NI, P_i.data(), nxt_i.data(), seg_i.data()};
    LongBlock br{false, v, rid.data(), nullptr, NR, P_r.data(), nxt_r.data(), seg_r.data()};
    uint64_t segs[2], bits[2];
    const LongBlock *blocks[2] = {&bi, &br};
    for (int k = 0; k < 2; k++) {
        const LongBlock &B = *blocks[k];
        std::vector<uint64_t> sums(T);
        for (uint32_t t = 0; t < T; t++) sums[t] = later_sum(B, t, T);
        uint64_t base = 0;
        for (uint32_t t = 0; t < T; t++) {
            later_prefix(B, t, T, base);
            base += sums[t];
        }
        for (uint32_t t = 0; t < T; t++) segment_ends(B, t, T);
        segs[k] = segment_chain(B);
        bits[k] = block_bits(B, segs[k]);
    }
    if (words) {
        for (uint32_t t = 0; t < T; t++) write_segments<HostOr>(bi, segs[0], start, words, t, T);
        for (uint32_t t = 0; t < T; t++) write_segments<HostOr>(br, segs[1], start + bits[0], words, t, T);
    }
    return bits[0] + bits[1];
}

}  // extern "C"
