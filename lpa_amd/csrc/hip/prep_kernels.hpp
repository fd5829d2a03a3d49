// prep_kernels.hpp -- one-time graph preparation on the device: record
// validation with degrees and cost features, and the bin-ordered relayout
// of the CGR stream.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cgr_decode.hpp"

namespace lpa {

constexpr int NFEAT = 8;

// Bounds-checked serial reader: every code must end at or before `lim`.
struct SafeReader {
    const uint64_t *W;
    uint64_t pos, lim;
    bool bad;
    __device__ uint64_t window() const {
        uint64_t i = pos >> 6;
        uint32_t s = (uint32_t)(pos & 63);
        return s ? (W[i] << s) | (W[i + 1] >> (64 - s)) : W[i];
    }
    __device__ uint32_t gamma() {
        if (bad || pos >= lim) { bad = true; return 0; }
        uint64_t x = window();
        uint32_t L = (uint32_t)__builtin_clzll(x | 1ull);
        if (L > 31 || pos + 2 * L + 1 > lim) { bad = true; return 0; }
        uint64_t y = x >> (64 - (2 * L + 1));
        pos += 2 * L + 1;
        return (uint32_t)(y - 1);
    }
    __device__ uint32_t zeta() {
        if (bad || pos >= lim) { bad = true; return 0; }
        uint64_t x = window();
        uint32_t h = (uint32_t)__builtin_clzll(x | 1ull);
        if (h > 10 || pos + 4 * (h + 1) > lim) { bad = true; return 0; }
        uint64_t y = (x << (h + 1)) >> (64 - 3 * (h + 1));
        pos += 4 * (h + 1);
        return (uint32_t)(y - 1);
    }
};

// One lane per compressed position: decodes the record with every bound
// checked (codes inside the record, padded segments inside 256 bits,
// neighbour ids < n, record ends exactly at the next offset).  Writes the
// degree, cost features and segment metadata per position -- {bits from the
// record start to the first interval segment, #interval segments, bits to
// the first residual segment, #residual segments}, which lets a group seek
// every segment at once -- and sets *err to 1 + position on the first
// malformed record it sees (any one of them).
__global__ void __launch_bounds__(256) k_scan_records(const uint64_t *__restrict__ W,
                                                      const uint64_t *__restrict__ offs,
                                                      const uint32_t *__restrict__ vid, uint64_t n_c,
                                                      uint32_t n, uint32_t *__restrict__ deg,
                                                      float *__restrict__ feat, uint4 *__restrict__ meta,
                                                      unsigned long long *__restrict__ err) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_c) return;
    const uint32_t v = vid[p];
    const uint64_t start = offs[p], stop = offs[p + 1];
    SafeReader r{W, start, stop, start > stop || v >= n};
    uint64_t d = 0, n_itv = 0, n_res = 0, itv_len = 0;
    uint32_t si = 0, sr = 0;
    uint64_t ibase = 0, rbase = 0;
    uint64_t block_end = start;
    for (int blk = 0; blk < 2 && !r.bad; blk++) {
        r.pos = block_end;
        r.lim = stop;
        const uint32_t ns = r.gamma() + 1;
        const uint64_t base = r.pos;
        if (r.bad || base + (uint64_t)SEG_BITS * (ns - 1) >= stop + (ns > 1 ? 0 : 1)) { r.bad = true; break; }
        if (blk == 0) { si = ns; ibase = base - start; } else { sr = ns; rbase = base - start; }
        for (uint32_t j = 0; j < ns && !r.bad; j++) {
            r.pos = base + (uint64_t)SEG_BITS * j;
            r.lim = j + 1 < ns ? r.pos + SEG_BITS : stop;
            const uint32_t c = r.gamma();
            if (c > r.lim - r.pos) { r.bad = true; break; }
            if (blk == 0) {
                uint64_t stop_prev = 0;
                for (uint32_t t = 0; t < c && !r.bad; t++) {
                    uint32_t x = r.gamma();
                    uint64_t left = t == 0 ? (uint64_t)(int64_t)((x & 1) ? (int64_t)v - (x >> 1) - 1
                                                                         : (int64_t)v + (x >> 1))
                                           : stop_prev + x + 1;
                    uint64_t len = (uint64_t)r.gamma() + MIN_ITV;
                    if (t == 0 && (x & 1) && (uint64_t)(x >> 1) + 1 > v) r.bad = true;
                    if (left + len > n) r.bad = true;
                    stop_prev = left + len;
                    d += len;
                    n_itv++;
                    itv_len += len;
                }
            } else {
                uint64_t cur = 0;
                for (uint32_t t = 0; t < c && !r.bad; t++) {
                    uint32_t x = r.zeta();
                    if (t == 0) {
                        if ((x & 1) && (uint64_t)(x >> 1) + 1 > v) r.bad = true;
                        cur = (x & 1) ? (uint64_t)v - (x >> 1) - 1 : (uint64_t)v + (x >> 1);
                    } else {
                        cur = cur + x + 1;
                    }
                    if (cur >= n) r.bad = true;
                    d++;
                    n_res++;
                }
            }
            block_end = r.pos;
        }
    }
    if (!r.bad && block_end != stop) r.bad = true;
    if (rbase >= (1ull << 32)) r.bad = true;   // segment metadata holds 32-bit relative offsets
    if (r.bad) {
        atomicCAS(err, 0ull, (unsigned long long)p + 1);
        d = 0;
    }
    deg[p] = (uint32_t)d;
    meta[p] = make_uint4((uint32_t)ibase, si, (uint32_t)rbase, sr);
    if (feat) {
        float *f = feat + p * NFEAT;
        f[0] = (float)d;
        f[1] = (float)n_itv;
        f[2] = (float)n_res;
        f[3] = (float)(stop - start);
        f[4] = (float)si;
        f[5] = (float)sr;
        f[6] = n_itv ? (float)itv_len / (float)n_itv : 0.0f;
        f[7] = __log2f(1.0f + (float)d);
    }
}

// Copies `len` bits from bit `s` of src to bit `t` of dst (dst zeroed
// beforehand; words shared with neighbouring records are ORed atomically).
__device__ void copy_bits(const uint64_t *__restrict__ src, uint64_t s, uint64_t *dst, uint64_t t, uint64_t len) {
    for (uint64_t i = 0; i < len; i += 64) {
        const uint32_t c = (uint32_t)(len - i < 64 ? len - i : 64);
        const uint64_t a = s + i;
        const uint32_t sh = (uint32_t)(a & 63);
        uint64_t x = sh ? (src[a >> 6] << sh) | (src[(a >> 6) + 1] >> (64 - sh)) : src[a >> 6];
        if (c < 64) x &= ~0ull << (64 - c);
        const uint64_t b = t + i;
        const uint32_t o = (uint32_t)(b & 63);
        atomicOr((unsigned long long *)&dst[b >> 6], (unsigned long long)(x >> o));
        if (o + c > 64) atomicOr((unsigned long long *)&dst[(b >> 6) + 1], (unsigned long long)(x << (64 - o)));
    }
}

// new position q takes old position perm[q]: record bits, offset, vertex id.
__global__ void __launch_bounds__(256) k_relayout(const uint64_t *__restrict__ W, const uint64_t *__restrict__ offs,
                                                  const uint32_t *__restrict__ vid, const uint4 *__restrict__ meta,
                                                  const uint32_t *__restrict__ perm,
                                                  const uint64_t *__restrict__ new_offs, uint64_t *new_W,
                                                  uint32_t *__restrict__ new_vid, uint4 *__restrict__ new_meta,
                                                  uint64_t n_c) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_c) return;
    const uint32_t p = perm[q];
    new_vid[q] = vid[p];
    new_meta[q] = meta[p];
    copy_bits(W, offs[p], new_W, new_offs[q], offs[p + 1] - offs[p]);
}

// Big-endian 32-bit words (File::graph) -> 64-bit words of the same bit order.
__global__ void k_pack64(const uint32_t *__restrict__ w32, uint64_t n32, uint64_t *__restrict__ w64, uint64_t n64) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n64) return;
    uint64_t hi = 2 * i < n32 ? w32[2 * i] : 0, lo = 2 * i + 1 < n32 ? w32[2 * i + 1] : 0;
    w64[i] = (hi << 32) | lo;
}

__global__ void k_gather_u4(const uint4 *__restrict__ src, const uint32_t *__restrict__ idx, uint32_t n,
                            uint4 *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = src[idx[i]];
}

// plain copy, 16 bytes per lane (counter calibration)
__global__ void __launch_bounds__(256) k_calibration_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                          uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// random row gathers (measurement): lane i reads `per_lane` rows of
// W 32-bit words at uniformly random row indices of a buffer of nrows rows
// (indices from a counter-based splitmix64 stream, so no load depends on
// another) and stores one folded word, so the loads cannot be dropped; the
// pattern of the label-row gathers of a step, at a rate the kernels can be
// compared with
template <int W>
__global__ void __launch_bounds__(256) k_calibration_gather(const uint32_t *__restrict__ src, uint64_t nrows,
                                                            uint32_t per_lane, uint64_t seed,
                                                            uint32_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t acc = 0;
#pragma unroll 8
    for (uint32_t k = 0; k < per_lane; k++) {
        uint64_t z = seed + (i * per_lane + k + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const uint32_t *row = src + (z % nrows) * W;
        if (W % 4 == 0) {
#pragma unroll
            for (int q = 0; q < W / 4; q++) {
                const uint4 w = reinterpret_cast<const uint4 *>(row)[q];
                acc ^= w.x ^ w.y ^ w.z ^ w.w;
            }
        } else {
#pragma unroll
            for (int q = 0; q < W; q++) acc ^= row[q];
        }
    }
    out[i] = acc;
}

__global__ void k_fill_u64(unsigned long long *p, uint64_t n, unsigned long long v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace lpa
