// cgr_decode.hpp -- device-side CGR bit reader and segment decoders (gfx950).
//
// Restates the reference's CgrReader / IntervalSegmentHelper /
// ResidualSegmentHelper (include/graph_utils.cuh:19-183) for 64-bit
// windows: the stream is held as 64-bit words whose bit order is the
// stream order (bit 0 of the stream is the MSB of word 0), every code of a
// graph with n < 2^30 fits one 64-bit window, and bit offsets are 64-bit
// (the reference's 32-bit global_offset overflows beyond 2^32 bits,
// graph_utils.cuh:22).  A reader keeps a 6-word register window per segment.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// The reader and the serial decoders also build for the host, where a test
// library checks them against the CPU oracle.
#define LPA_HD __host__ __device__

namespace lpa {

constexpr uint32_t SEG_BITS = 256;   // segment stride (graph_utils.cuh:17)
constexpr uint32_t MIN_ITV = 4;      // interval length offset (graph_utils.cuh:172)

struct BitReader {
    // Six words are loaded together at every seek (one memory latency per
    // 256-bit segment); moving to the next word shifts the registers down
    // without a load, and only a segment longer than ~320 bits (the merged
    // last segment of a block) triggers a second batch.  A reader that
    // loaded the next word at every word step would stall on that load at
    // the step itself (the shifted-in value is copied immediately).
    static constexpr int NW = 6;
    const uint64_t *__restrict__ W;
    uint64_t pos;     // absolute bit position
    uint64_t idx;     // word index held in w[0]
    uint32_t left;    // valid words from w[0]
    uint64_t w[NW];

    LPA_HD __forceinline__ void load_at(uint64_t i) {
#pragma unroll
        for (int k = 0; k < NW; k++) w[k] = W[i + k];
        idx = i;
        left = NW;
    }
    LPA_HD __forceinline__ void seek(const uint64_t *__restrict__ words, uint64_t p) {
        W = words;
        pos = p;
        load_at(p >> 6);
    }
    // Moves forward to bit q >= pos: by shifting when q lies in the loaded
    // words, else by a new batch load.
    LPA_HD __forceinline__ void move_to(uint64_t q) {
        const uint64_t d = (q >> 6) - idx;
        if (d + 2 <= left) {
#pragma unroll
            for (int step = 1; step < NW - 1; step++) {
                if ((uint64_t)step > d) break;
#pragma unroll
                for (int k = 0; k < NW - 1; k++) w[k] = w[k + 1];
            }
            idx += d;
            left -= (uint32_t)d;
            pos = q;
        } else {
            pos = q;
            load_at(q >> 6);
        }
    }
    // 64 bits starting at pos; pos - idx*64 < 64 always holds.
    LPA_HD __forceinline__ uint64_t window() const {
        uint32_t s = (uint32_t)(pos & 63);
        return s ? (w[0] << s) | (w[1] >> (64 - s)) : w[0];
    }
    LPA_HD __forceinline__ void advance(uint32_t bits) {
        pos += bits;
        if ((pos >> 6) != idx) {   // codes are < 64 bits: at most one word step
#pragma unroll
            for (int k = 0; k < NW - 1; k++) w[k] = w[k + 1];
            idx++;
            if (--left < 2) load_at(idx);
        }
    }
    // Elias gamma: 0^L then the L+1 bits of x+1 (graph_utils.cuh:63-67)
    LPA_HD __forceinline__ uint32_t gamma() {
        uint64_t x = window();
        uint32_t L = (uint32_t)__builtin_clzll(x | 1ull);   // L <= 31 for n < 2^30
        uint32_t len = 2 * L + 1;
        uint64_t y = x >> (64 - len);
        advance(len);
        return (uint32_t)(y - 1);
    }
    // zeta_3: 0^h 1 then 3(h+1) bits of x+1 (graph_utils.cuh:70-77)
    LPA_HD __forceinline__ uint32_t zeta() {
        uint64_t x = window();
        uint32_t h = (uint32_t)__builtin_clzll(x | 1ull);
        uint32_t body = 3 * (h + 1);
        uint64_t y = (x << (h + 1)) >> (64 - body);
        advance(4 * (h + 1));
        return (uint32_t)(y - 1);
    }
};

// decode_first_num (graph_utils.cuh:34-37)
LPA_HD __forceinline__ uint32_t from_nat(uint32_t v, uint32_t x) {
    return (x & 1) ? v - (x >> 1) - 1 : v + (x >> 1);
}

// Decodes the interval segment the reader stands at (vertex v); calls
// emit_run(left, len) for every interval; leaves the reader after it.
template <class F>
LPA_HD __forceinline__ void read_interval_segment(BitReader &r, uint32_t v, F &&emit_run) {
    const uint32_t c = r.gamma();
    uint32_t stop = 0;
    for (uint32_t t = 0; t < c; t++) {
        const uint32_t left = (t == 0) ? from_nat(v, r.gamma()) : stop + r.gamma() + 1;
        const uint32_t len = r.gamma() + MIN_ITV;
        emit_run(left, len);
        stop = left + len;
    }
}

// Decodes the residual segment the reader stands at; emit(u) per neighbour.
template <class F>
LPA_HD __forceinline__ void read_residual_segment(BitReader &r, uint32_t v, F &&emit) {
    const uint32_t c = r.gamma();
    uint32_t cur = 0;
    for (uint32_t t = 0; t < c; t++) {
        cur = (t == 0) ? from_nat(v, r.zeta()) : cur + r.zeta() + 1;
        emit(cur);
    }
}

// Interval segment at bit q; returns the bit just after it.
template <class F>
LPA_HD __forceinline__ uint64_t decode_interval_segment(const uint64_t *__restrict__ W, uint64_t q,
                                                            uint32_t v, F &&emit_run) {
    BitReader r;
    r.seek(W, q);
    read_interval_segment(r, v, emit_run);
    return r.pos;
}

// Residual segment at bit q; returns the bit just after it.
template <class F>
LPA_HD __forceinline__ uint64_t decode_residual_segment(const uint64_t *__restrict__ W, uint64_t q,
                                                            uint32_t v, F &&emit) {
    BitReader r;
    r.seek(W, q);
    read_residual_segment(r, v, emit);
    return r.pos;
}

// Segment-count header of a block at bit p: returns #segments, *base = first segment.
LPA_HD __forceinline__ uint32_t block_header(const uint64_t *__restrict__ W, uint64_t p, uint64_t *base) {
    BitReader r;
    r.seek(W, p);
    uint32_t s = r.gamma() + 1;
    *base = r.pos;
    return s;
}

// Serial decode of a whole record by one lane with one reader: emit(u) per
// neighbour, in encoded order (intervals, then residuals).  Padded segments
// are reached by shifting the loaded words when they are close.  Returns
// the bit after the record.
template <class F>
LPA_HD __forceinline__ uint64_t decode_record(const uint64_t *__restrict__ W, uint64_t p, uint32_t v,
                                                  F &&emit) {
    BitReader r;
    r.seek(W, p);
    uint32_t ns = r.gamma() + 1;
    uint64_t base = r.pos;
    for (uint32_t j = 0; j < ns; j++) {
        r.move_to(base + (uint64_t)SEG_BITS * j);
        read_interval_segment(r, v, [&](uint32_t l, uint32_t len) {
            for (uint32_t x = 0; x < len; x++) emit(l + x);
        });
    }
    ns = r.gamma() + 1;
    base = r.pos;
    for (uint32_t j = 0; j < ns; j++) {
        r.move_to(base + (uint64_t)SEG_BITS * j);
        read_residual_segment(r, v, emit);
    }
    return r.pos;
}

}  // namespace lpa
