// encode_kernels.hpp -- CGR encoding on the device (CSR -> bitstream).
//
// Produces the same bitstream as the reference's CgrCompressor with File's
// parameters (zeta 3, minimum interval 4, 256-bit interval and residual
// segments; src/compressor.cpp:306-515, src/file.cpp:84-87): per vertex an
// interval block then a residual block, each "gamma(#segments-1)" followed
// by segments "gamma(count) codes", every segment but the last zero-padded
// to 256 bits, the trailing partial segment merged into the last full one.
//
// One lane encodes one record, streaming over the sorted adjacency with no
// per-record buffers: the interval block's entries are the runs of >= 4
// consecutive ids, the residual block's the ids outside them.  A block is
// walked once to count its entries and segments (the segment rule needs
// only the running segment size and the previous entry), and each
// segment's count is found by a look-ahead walk from its first entry just
// before the segment is written.  k_encode_len measures every record (and
// validates the adjacency), a scan turns lengths into offsets, and
// k_encode_write ORs the bits into 64-bit words (stream bit 0 = MSB of word
// 0); words inside one record are stored, the two it may share with its
// neighbours are combined with atomic ORs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cgr_decode.hpp"   // LPA_HD: the record encoder also builds for the host (tests)

namespace lpa {

namespace enc {

constexpr uint32_t SEG = 256;
constexpr uint32_t MIN_RUN = 4;

LPA_HD __forceinline__ uint32_t flog2(uint64_t y) { return 63u - (uint32_t)__builtin_clzll(y); }
LPA_HD __forceinline__ uint32_t gamma_len(uint64_t x) { return 2u * flog2(x + 1) + 1u; }
LPA_HD __forceinline__ uint32_t zeta_len(uint64_t x) { return (flog2(x + 1) / 3u + 1u) * 4u; }
LPA_HD __forceinline__ uint64_t nat(int64_t d) { return d >= 0 ? (uint64_t)d << 1 : (uint64_t)(-2 * d - 1); }

// Runs of >= MIN_RUN consecutive ids: (left, len).
struct Runs {
    const uint32_t *adj;
    uint32_t d, a;
    LPA_HD bool next(uint32_t &left, uint32_t &len) {
        while (a < d) {
            uint32_t b = a + 1;
            while (b < d && adj[b - 1] + 1 == adj[b]) b++;
            const uint32_t s = a;
            a = b;
            if (b - s >= MIN_RUN) {
                left = adj[s];
                len = b - s;
                return true;
            }
        }
        return false;
    }
};

// Ids outside such runs, in order.
struct Singles {
    const uint32_t *adj;
    uint32_t d, a, i, e;   // pending short run adj[i..e)
    LPA_HD bool next(uint32_t &u) {
        for (;;) {
            if (i < e) {
                u = adj[i++];
                return true;
            }
            if (a >= d) return false;
            uint32_t b = a + 1;
            while (b < d && adj[b - 1] + 1 == adj[b]) b++;
            if (b - a < MIN_RUN) {
                i = a;
                e = b;
            }
            a = b;
        }
    }
};

// Interval entries: code of (left, len) given the previous entry of its segment.
struct ItvBlock {
    Runs g;
    uint32_t v;
    uint64_t prev_end;   // previous left + len in the segment
    uint32_t left, len;
    LPA_HD bool next() { return g.next(left, len); }
    LPA_HD uint64_t code(bool first) const {
        return first ? nat((int64_t)left - (int64_t)v) : (uint64_t)left - prev_end - 1;
    }
    LPA_HD uint32_t bits(bool first) const { return gamma_len(code(first)) + gamma_len(len - MIN_RUN); }
    LPA_HD void advance() { prev_end = (uint64_t)left + len; }
};

struct ResBlock {
    Singles g;
    uint32_t v;
    uint64_t prev;
    uint32_t u;
    LPA_HD bool next() { return g.next(u); }
    LPA_HD uint64_t code(bool first) const { return first ? nat((int64_t)u - (int64_t)v) : (uint64_t)u - prev - 1; }
    LPA_HD uint32_t bits(bool first) const { return zeta_len(code(first)); }
    LPA_HD void advance() { prev = u; }
};

// Sinks: Meter counts bits; BitWriter ORs them into the stream (OrShared
// combines a word the record may share with a neighbouring record).
struct Meter {
    uint64_t pos;
    LPA_HD void put(uint64_t, uint32_t len) { pos += len; }
    LPA_HD void skip_to(uint64_t p) { pos = p; }
};

struct DeviceOr {
    __device__ static void apply(uint64_t *w, uint64_t v) { atomicOr((unsigned long long *)w, (unsigned long long)v); }
};

template <class OrShared> struct BitWriter {
    uint64_t *words;
    uint64_t pos, acc, wi, w_first, w_last;   // record's first and last word
    LPA_HD BitWriter(uint64_t *w, uint64_t start, uint64_t end)
        : words(w), pos(start), acc(0), wi(start >> 6), w_first(start >> 6), w_last((end - 1) >> 6) {}
    LPA_HD void flush() {
        if (wi > w_first && wi < w_last) {
            words[wi] = acc;   // owned by this record alone
        } else if (acc) {
            OrShared::apply(&words[wi], acc);
        }
        acc = 0;
    }
    LPA_HD void move(uint64_t np) {
        pos = np;
        if ((pos >> 6) != wi) {
            flush();
            wi = pos >> 6;
        }
    }
    LPA_HD void put(uint64_t val, uint32_t len) {   // 1 <= len <= 64
        const uint32_t off = (uint32_t)(pos & 63);
        if (off + len <= 64) {
            acc |= len == 64 ? val : val << (64 - off - len);
        } else {
            const uint32_t spill = off + len - 64;
            acc |= val >> spill;
            flush();
            wi++;
            acc = val << (64 - spill);
        }
        move(pos + len);
    }
    LPA_HD void skip_to(uint64_t p) { move(p); }
    LPA_HD void finish() { flush(); }
};

template <class S> LPA_HD __forceinline__ void put_gamma(S &s, uint64_t x) {
    const uint64_t y = x + 1;
    s.put(y, 2 * flog2(y) + 1);
}
template <class S> LPA_HD __forceinline__ void put_zeta(S &s, uint64_t x) {
    const uint64_t y = x + 1;
    const uint32_t h = flog2(y) / 3;
    s.put((1ull << (3 * (h + 1))) | y, 4 * (h + 1));
}

// Entries and segments of a block (the segment rule of compressor.cpp:390-399
// and 462-468: an entry that would overflow the segment, counting the
// segment's own gamma(count + 1) header, starts a new one; the trailing
// partial segment is merged into the last full one, :412-425 and 479-483).
template <class B> LPA_HD void plan(B b, uint64_t &entries, uint64_t &segs) {
    uint64_t cur = 0, cnt = 0, closed = 0;
    entries = 0;
    while (b.next()) {
        uint32_t x = b.bits(cnt == 0);
        if (gamma_len(cnt + 1) + cur + x > SEG) {
            closed++;
            cnt = 0;
            cur = 0;
            x = b.bits(true);
        }
        cnt++;
        cur += x;
        b.advance();
        entries++;
    }
    segs = closed ? closed : 1;
}

// Entries of the segment starting at the block's current entry (look-ahead).
template <class B> LPA_HD uint64_t segment_count(B b) {
    uint64_t cur = 0, cnt = 0;
    while (b.next()) {
        const uint32_t x = b.bits(cnt == 0);
        if (gamma_len(cnt + 1) + cur + x > SEG) return cnt;
        cnt++;
        cur += x;
        b.advance();
    }
    return cnt;
}

template <class B, class S, class Put> LPA_HD void write_block(B b, S &s, Put &&put_entry) {
    uint64_t entries, segs;
    plan(b, entries, segs);
    put_gamma(s, segs - 1);
    uint64_t done = 0;
    for (uint64_t k = 0; k < segs; k++) {
        const uint64_t start = s.pos;
        const uint64_t cnt = k + 1 < segs ? segment_count(b) : entries - done;
        put_gamma(s, cnt);
        for (uint64_t j = 0; j < cnt; j++) {
            b.next();
            put_entry(s, b, j == 0);
            b.advance();
        }
        done += cnt;
        if (k + 1 < segs) s.skip_to(start + SEG);
    }
}

template <class S> LPA_HD void encode_record(S &s, uint32_t v, const uint32_t *adj, uint32_t d) {
    ItvBlock ib{Runs{adj, d, 0}, v, 0, 0, 0};
    write_block(ib, s, [] LPA_HD (S &w, const ItvBlock &e, bool first) {
        put_gamma(w, e.code(first));
        put_gamma(w, e.len - MIN_RUN);
    });
    ResBlock rb{Singles{adj, d, 0, 0, 0}, v, 0, 0};
    write_block(rb, s, [] LPA_HD (S &w, const ResBlock &e, bool first) { put_zeta(w, e.code(first)); });
}

}  // namespace enc

// Record length of every compressed position (vid[p] names its vertex)
// whose row has at most long_min ids (longer rows are measured by the
// workgroup kernels of encode_long.hpp); *err = 1 + vertex of the first row
// that is not strictly increasing or has an id >= n (the format's gaps need
// sorted, duplicate-free rows).
__global__ void __launch_bounds__(256) k_encode_len(const uint64_t *__restrict__ row, const uint32_t *__restrict__ col,
                                                    const uint32_t *__restrict__ vid, uint64_t n_c, uint32_t n,
                                                    uint64_t *__restrict__ len, unsigned long long *__restrict__ err,
                                                    uint32_t long_min) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_c) return;
    const uint32_t v = vid[p];
    const uint64_t r0 = row[v], r1 = row[v + 1];
    if (r1 >= r0 && r1 - r0 > long_min) return;
    const uint32_t d = (uint32_t)(r1 - r0);
    const uint32_t *adj = col + r0;
    bool bad = r1 < r0;
    for (uint32_t i = 0; i < d && !bad; i++) bad = adj[i] >= n || (i && adj[i] <= adj[i - 1]);
    if (bad) {
        atomicCAS(err, 0ull, (unsigned long long)v + 1);
        len[p] = 0;
        return;
    }
    enc::Meter m{0};
    enc::encode_record(m, v, adj, d);
    len[p] = m.pos;
}

__global__ void __launch_bounds__(256) k_encode_write(const uint64_t *__restrict__ row, const uint32_t *__restrict__ col,
                                                      const uint32_t *__restrict__ vid, uint64_t n_c,
                                                      const uint64_t *__restrict__ offs, uint64_t *words,
                                                      uint32_t long_min) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_c) return;
    const uint32_t v = vid[p];
    const uint64_t r0 = row[v];
    if (row[v + 1] - r0 > long_min) return;
    const uint32_t d = (uint32_t)(row[v + 1] - r0);
    enc::BitWriter<enc::DeviceOr> w(words, offs[p], offs[p + 1]);
    enc::encode_record(w, v, col + r0, d);
    w.finish();
}

// Exclusive prefix sum of n 64-bit values in place (out[0] = 0, out[n] =
// total), three launches: per-block sums, a single-block scan of those,
// per-block scans with their offsets.  Blocks of 256 lanes x 16 values.
constexpr uint32_t SCAN_PER_BLOCK = 256 * 16;

__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t x, uint64_t *sh, uint64_t &total) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o) inc += y;
    }
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    uint64_t base = 0;
    total = 0;
    for (uint32_t k = 0; k < blockDim.x / 64; k++) {
        if (k < w) base += sh[k];
        total += sh[k];
    }
    __syncthreads();
    return base + inc - x;
}

__global__ void __launch_bounds__(256) k_scan_sums(const uint64_t *__restrict__ in, uint64_t n,
                                                   uint64_t *__restrict__ sums) {
    __shared__ uint64_t sh[4];
    const uint64_t b0 = (uint64_t)blockIdx.x * SCAN_PER_BLOCK;
    uint64_t s = 0;
    for (uint32_t k = 0; k < 16; k++) {
        const uint64_t i = b0 + (uint64_t)k * 256 + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint64_t total;
    (void)block_exclusive_scan(s, sh, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(256) k_scan_top(uint64_t *__restrict__ sums, uint64_t nb) {
    __shared__ uint64_t sh[4];
    uint64_t carry = 0;
    for (uint64_t c0 = 0; c0 < nb; c0 += 256) {
        const uint64_t i = c0 + threadIdx.x;
        const uint64_t x = i < nb ? sums[i] : 0;
        uint64_t total;
        const uint64_t ex = block_exclusive_scan(x, sh, total);
        if (i < nb) sums[i] = carry + ex;
        carry += total;
    }
}

// out[i] = base of the block + exclusive scan inside it; out[n] = total.
// out may be in (an in-place scan): each lane reads its 16 values before the
// block's barrier and writes the same 16 after it, so in and out are not
// declared restrict.
__global__ void __launch_bounds__(256) k_scan_apply(const uint64_t *in, uint64_t n,
                                                    const uint64_t *__restrict__ sums, uint64_t *out,
                                                    uint64_t nb) {
    __shared__ uint64_t sh[4];
    const uint64_t b0 = (uint64_t)blockIdx.x * SCAN_PER_BLOCK;
    // each lane scans 16 consecutive values
    uint64_t v[16];
    uint64_t s = 0;
    const uint64_t i0 = b0 + (uint64_t)threadIdx.x * 16;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        v[k] = i0 + k < n ? in[i0 + k] : 0;
        s += v[k];
    }
    uint64_t total;
    uint64_t run = sums[blockIdx.x] + block_exclusive_scan(s, sh, total);
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (i0 + k < n) out[i0 + k] = run;
        run += v[k];
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = sums[blockIdx.x] + total;
}

}  // namespace lpa
