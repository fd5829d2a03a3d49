// cgr_decode.hpp -- device-side CGR bit reader and segment decoders (gfx950).
//
// Restates the reference's CgrReader / IntervalSegmentHelper /
// ResidualSegmentHelper (include/graph_utils.cuh:19-183) for 64-bit
// windows: the stream is held as 64-bit words whose bit order is the
// stream order (bit 0 of the stream is the MSB of word 0), every code of a
// graph with n < 2^30 fits one 64-bit window, and bit offsets are 64-bit
// (the reference's 32-bit global_offset overflows beyond 2^32 bits,
// graph_utils.cuh:22).  A reader keeps a rolling 4-word register buffer so
// a serial decode touches memory once per 64 bits, with three words of
// look-ahead in flight.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lpa {

constexpr uint32_t SEG_BITS = 256;   // segment stride (graph_utils.cuh:17)
constexpr uint32_t MIN_ITV = 4;      // interval length offset (graph_utils.cuh:172)

struct BitReader {
    const uint64_t *__restrict__ W;
    uint64_t pos;     // absolute bit position
    uint64_t idx;     // word index held in w0
    uint64_t w0, w1, w2, w3;

    __device__ __forceinline__ void seek(const uint64_t *__restrict__ words, uint64_t p) {
        W = words;
        pos = p;
        idx = p >> 6;
        w0 = W[idx];
        w1 = W[idx + 1];
        w2 = W[idx + 2];
        w3 = W[idx + 3];
    }
    // 64 bits starting at pos; pos - idx*64 < 64 always holds.
    __device__ __forceinline__ uint64_t window() const {
        uint32_t s = (uint32_t)(pos & 63);
        return s ? (w0 << s) | (w1 >> (64 - s)) : w0;
    }
    __device__ __forceinline__ void advance(uint32_t bits) {
        pos += bits;
        if ((pos >> 6) != idx) {   // codes are < 64 bits: at most one word step
            idx++;
            w0 = w1;
            w1 = w2;
            w2 = w3;
            w3 = W[idx + 3];
        }
    }
    // Elias gamma: 0^L then the L+1 bits of x+1 (graph_utils.cuh:63-67)
    __device__ __forceinline__ uint32_t gamma() {
        uint64_t x = window();
        uint32_t L = (uint32_t)__builtin_clzll(x | 1ull);   // L <= 31 for n < 2^30
        uint32_t len = 2 * L + 1;
        uint64_t y = x >> (64 - len);
        advance(len);
        return (uint32_t)(y - 1);
    }
    // zeta_3: 0^h 1 then 3(h+1) bits of x+1 (graph_utils.cuh:70-77)
    __device__ __forceinline__ uint32_t zeta() {
        uint64_t x = window();
        uint32_t h = (uint32_t)__builtin_clzll(x | 1ull);
        uint32_t body = 3 * (h + 1);
        uint64_t y = (x << (h + 1)) >> (64 - body);
        advance(4 * (h + 1));
        return (uint32_t)(y - 1);
    }
};

// decode_first_num (graph_utils.cuh:34-37)
__device__ __forceinline__ uint32_t from_nat(uint32_t v, uint32_t x) {
    return (x & 1) ? v - (x >> 1) - 1 : v + (x >> 1);
}

// Decodes interval segment at bit q of vertex v; calls emit_run(left, len)
// for every interval; returns the bit just after the segment.
template <class F>
__device__ __forceinline__ uint64_t decode_interval_segment(const uint64_t *__restrict__ W, uint64_t q,
                                                            uint32_t v, F &&emit_run) {
    BitReader r;
    r.seek(W, q);
    uint32_t c = r.gamma();
    uint32_t stop = 0;
    for (uint32_t t = 0; t < c; t++) {
        uint32_t left = (t == 0) ? from_nat(v, r.gamma()) : stop + r.gamma() + 1;
        uint32_t len = r.gamma() + MIN_ITV;
        emit_run(left, len);
        stop = left + len;
    }
    return r.pos;
}

// Decodes residual segment at bit q of vertex v; emit(u) per neighbour.
template <class F>
__device__ __forceinline__ uint64_t decode_residual_segment(const uint64_t *__restrict__ W, uint64_t q,
                                                            uint32_t v, F &&emit) {
    BitReader r;
    r.seek(W, q);
    uint32_t c = r.gamma();
    uint32_t cur = 0;
    for (uint32_t t = 0; t < c; t++) {
        cur = (t == 0) ? from_nat(v, r.zeta()) : cur + r.zeta() + 1;
        emit(cur);
    }
    return r.pos;
}

// Segment-count header of a block at bit p: returns #segments, *base = first segment.
__device__ __forceinline__ uint32_t block_header(const uint64_t *__restrict__ W, uint64_t p, uint64_t *base) {
    BitReader r;
    r.seek(W, p);
    uint32_t s = r.gamma() + 1;
    *base = r.pos;
    return s;
}

// Serial decode of a whole record by one lane: emit(u) per neighbour, in
// encoded order (intervals, then residuals).  Returns the bit after the record.
template <class F>
__device__ __forceinline__ uint64_t decode_record(const uint64_t *__restrict__ W, uint64_t p, uint32_t v,
                                                  F &&emit) {
    uint64_t base;
    uint32_t ns = block_header(W, p, &base);
    uint64_t end = base;
    for (uint32_t j = 0; j < ns; j++)
        end = decode_interval_segment(W, base + (uint64_t)SEG_BITS * j, v, [&](uint32_t l, uint32_t len) {
            for (uint32_t x = 0; x < len; x++) emit(l + x);
        });
    ns = block_header(W, end, &base);
    end = base;
    for (uint32_t j = 0; j < ns; j++)
        end = decode_residual_segment(W, base + (uint64_t)SEG_BITS * j, v, emit);
    return end;
}

}  // namespace lpa
