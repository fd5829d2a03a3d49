// encode_long.hpp -- CGR records of long rows, encoded by a whole workgroup.
//
// k_encode_len / k_encode_write give a record to one lane, which walks its
// row several times; a hub of millions of ids keeps that lane busy for
// seconds while the rest of the device idles.  Rows longer than LONG_ROW
// are left to the kernels here, one 1024-lane workgroup per row, which
// produce the same bits in data-parallel phases:
//
//  1. runs: id i opens a maximal run of consecutive ids when adj[i-1] + 1 !=
//     adj[i]; along a sorted row adj[j] - j never decreases, so the run ends
//     at the last j with adj[j] - j == adj[i] - i, found by binary search.
//     Runs of >= 4 ids are interval entries, the ids of shorter runs
//     residual entries; every lane takes the runs that open in its slice of
//     the row, and a scan of the lanes' counts places the entries in order.
//  2. per entry, the bits it takes when it opens a segment (coded against
//     the vertex) and when it follows another (a gap from the previous
//     entry); P = the prefix sum of the latter.
//  3. a segment opened by entry s holds the entries that fit before one
//     would overflow it (the segment rule of the record encoder): the first
//     that does not is found by a binary search over P, for every s at once.
//     The segments actually opened are the chain 0, next(0), next(next(0)),
//     ... (one lane follows it), the trailing open segment merged into the
//     last closed one.
//  4. every segment but the last fills exactly 256 bits, so all segments
//     are written at once, one lane each; words shared by two segments (or
//     two records) take atomic ORs.
// The phases are host-device functions of (lane, lanes) so that a host
// build runs them lane by lane and checks them against the record encoder.
#pragma once
#include "encode_kernels.hpp"

namespace lpa {

namespace enc {

constexpr uint32_t LONG_ROW = 1u << 16;   // rows longer than this take the workgroup kernels
constexpr uint32_t LONG_LANES = 1024;

// Last index j >= s of the run of consecutive ids containing adj[s] (adj
// strictly increasing): adj[j] - j == adj[s] - s exactly on that run.
LPA_HD __forceinline__ uint64_t run_last(const uint32_t *adj, uint64_t d, uint64_t s) {
    const int64_t key = (int64_t)adj[s] - (int64_t)s;
    uint64_t lo = s, hi = d - 1;   // adj[lo] - lo == key
    while (lo < hi) {
        const uint64_t mid = hi - (hi - lo) / 2;
        if ((int64_t)adj[mid] - (int64_t)mid == key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

LPA_HD __forceinline__ void lane_slice(uint64_t n, uint32_t t, uint32_t T, uint64_t &a, uint64_t &b) {
    const uint64_t S = (n + T - 1) / T;
    a = (uint64_t)t * S < n ? (uint64_t)t * S : n;
    b = a + S < n ? a + S : n;
}

// Phase 1a: validates the lane's slice of the row and counts the interval
// and residual entries of the runs opening in it.
LPA_HD inline bool runs_count(const uint32_t *adj, uint64_t d, uint32_t n, uint32_t t, uint32_t T, uint64_t &ni,
                              uint64_t &nr) {
    uint64_t a, b;
    lane_slice(d, t, T, a, b);
    bool ok = true;
    for (uint64_t i = a; i < b; i++) ok = ok && adj[i] < n && (i == 0 || adj[i] > adj[i - 1]);
    ni = nr = 0;
    if (!ok) return false;
    for (uint64_t i = a; i < b;) {
        if (i > 0 && adj[i - 1] + 1 == adj[i]) {   // inside a run opened before the slice
            i = run_last(adj, d, i) + 1;
            continue;
        }
        const uint64_t e = run_last(adj, d, i), len = e - i + 1;
        if (len >= MIN_RUN) ni++;
        else nr += len;
        i = e + 1;
    }
    return true;
}

// Phase 1b: the entries of the runs opening in the lane's slice, at the
// lane's offsets oi / orr (exclusive scans of the counts).
LPA_HD inline void runs_place(const uint32_t *adj, uint64_t d, uint32_t t, uint32_t T, uint64_t oi, uint64_t orr,
                              uint32_t *left, uint32_t *len, uint32_t *rid) {
    uint64_t a, b;
    lane_slice(d, t, T, a, b);
    for (uint64_t i = a; i < b;) {
        if (i > 0 && adj[i - 1] + 1 == adj[i]) {
            i = run_last(adj, d, i) + 1;
            continue;
        }
        const uint64_t e = run_last(adj, d, i), l = e - i + 1;
        if (l >= MIN_RUN) {
            left[oi] = adj[i];
            len[oi] = (uint32_t)l;
            oi++;
        } else {
            for (uint64_t j = i; j <= e; j++) rid[orr++] = adj[j];
        }
        i = e + 1;
    }
}

// One block's entries: intervals (a = left ids, l = lengths) or residuals
// (a = ids).
struct LongBlock {
    bool itv;
    uint32_t v;
    const uint32_t *a, *l;
    uint64_t n;
    uint64_t *P;     // n + 1 words: P[k] = sum of later_bits(e) over 1 <= e < k
    uint32_t *nxt;   // n: the entry that does not fit a segment opened by s (n: all fit)
    uint32_t *seg;   // segment starts, then n
    LPA_HD uint64_t first_code(uint64_t e) const { return nat((int64_t)a[e] - (int64_t)v); }
    LPA_HD uint64_t later_code(uint64_t e) const {
        return itv ? (uint64_t)a[e] - ((uint64_t)a[e - 1] + l[e - 1]) - 1 : (uint64_t)a[e] - a[e - 1] - 1;
    }
    LPA_HD uint32_t first_bits(uint64_t e) const {
        return itv ? gamma_len(first_code(e)) + gamma_len(l[e] - MIN_RUN) : zeta_len(first_code(e));
    }
    LPA_HD uint32_t later_bits(uint64_t e) const {
        return itv ? gamma_len(later_code(e)) + gamma_len(l[e] - MIN_RUN) : zeta_len(later_code(e));
    }
    // entry j (> s) still fits a segment opened by entry s
    LPA_HD bool fits(uint64_t s, uint64_t j) const {
        return gamma_len(j - s + 1) + first_bits(s) + (P[j + 1] - P[s + 1]) <= SEG;
    }
    template <class S> LPA_HD void put(S &w, uint64_t e, bool first) const {
        if (itv) {
            put_gamma(w, first ? first_code(e) : later_code(e));
            put_gamma(w, l[e] - MIN_RUN);
        } else {
            put_zeta(w, first ? first_code(e) : later_code(e));
        }
    }
};

// Phase 2: the lane's chunk of entries [a, b) and its sum of later bits.
LPA_HD inline uint64_t later_sum(const LongBlock &B, uint32_t t, uint32_t T) {
    uint64_t a, b, s = 0;
    lane_slice(B.n, t, T, a, b);
    for (uint64_t e = a < 1 ? 1 : a; e < b; e++) s += B.later_bits(e);
    return s;
}

// Phase 2b: P over the lane's chunk, from its offset base.
LPA_HD inline void later_prefix(const LongBlock &B, uint32_t t, uint32_t T, uint64_t base) {
    uint64_t a, b;
    lane_slice(B.n, t, T, a, b);
    if (t == 0) B.P[0] = 0;
    for (uint64_t e = a; e < b; e++) {
        base += e >= 1 ? B.later_bits(e) : 0;
        B.P[e + 1] = base;
    }
}

// Phase 3: next(s) for the lane's chunk (a segment holds at most 255
// entries: every entry takes at least one bit).
LPA_HD inline void segment_ends(const LongBlock &B, uint32_t t, uint32_t T) {
    uint64_t a, b;
    lane_slice(B.n, t, T, a, b);
    for (uint64_t s = a; s < b; s++) {
        uint64_t lo = s + 1, hi = B.n < s + SEG ? B.n : s + SEG;
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (B.fits(s, mid)) lo = mid + 1;
            else hi = mid;
        }
        B.nxt[s] = (uint32_t)lo;
    }
}

// Phase 3b (one lane): the segments opened, seg[0 .. segs) and seg[segs] =
// n; returns segs.
LPA_HD inline uint64_t segment_chain(const LongBlock &B) {
    uint64_t k = 0;
    B.seg[0] = 0;
    for (uint64_t s = 0; B.n && B.nxt[s] < B.n;) {
        s = B.nxt[s];
        B.seg[++k] = (uint32_t)s;
    }
    const uint64_t segs = k ? k : 1;   // the open segment merges into the last closed one
    B.seg[segs] = (uint32_t)B.n;
    return segs;
}

LPA_HD inline uint64_t block_bits(const LongBlock &B, uint64_t segs) {
    const uint64_t s = B.seg[segs - 1], cnt = B.n - s;
    uint64_t bits = gamma_len(segs - 1) + (uint64_t)SEG * (segs - 1) + gamma_len(cnt);
    if (cnt) bits += B.first_bits(s) + (B.P[B.n] - B.P[s + 1]);
    return bits;
}

// Phase 4: the lane's segments (k = t, t + T, ...) and, lane 0, the header.
template <class OrShared> struct OrWriter {
    uint64_t *words;
    uint64_t pos, acc, wi;   // wi: the word acc belongs to
    LPA_HD OrWriter(uint64_t *w, uint64_t p) : words(w), pos(p), acc(0), wi(p >> 6) {}
    LPA_HD void flush() {
        if (acc) OrShared::apply(&words[wi], acc);
        acc = 0;
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
        pos += len;
        if ((pos >> 6) != wi) {
            flush();
            wi = pos >> 6;
        }
    }
};

template <class OrShared>
LPA_HD inline void write_segments(const LongBlock &B, uint64_t segs, uint64_t start, uint64_t *words, uint32_t t,
                                  uint32_t T) {
    const uint64_t body = start + gamma_len(segs - 1);
    if (t == 0) {
        OrWriter<OrShared> w(words, start);
        put_gamma(w, segs - 1);
        w.flush();
    }
    for (uint64_t k = t; k < segs; k += T) {
        OrWriter<OrShared> w(words, body + (uint64_t)SEG * k);
        const uint64_t s = B.seg[k], e = B.seg[k + 1];
        put_gamma(w, e - s);
        for (uint64_t j = s; j < e; j++) B.put(w, j, j == s);
        w.flush();
    }
}

// Per long row: entries and segments of both blocks, kept from the length
// kernel for the write kernel.
struct LongInfo {
    uint64_t ni, nr, segs_i, segs_r, bits_i, bits_r;
};

// Scratch of the long rows: row r's arrays start at base[r] (the sum of
// d + 1 over the rows before it).
struct LongScratch {
    const uint64_t *base;
    uint32_t *left, *len, *rid, *nxt_i, *nxt_r, *seg_i, *seg_r;
    uint64_t *P_i, *P_r;
    LongInfo *info;
    LPA_HD LongBlock block(uint32_t r, bool itv, uint32_t v, uint64_t n) const {
        const uint64_t b = base[r];
        return itv ? LongBlock{true, v, left + b, len + b, n, P_i + b, nxt_i + b, seg_i + b}
                   : LongBlock{false, v, rid + b, nullptr, n, P_r + b, nxt_r + b, seg_r + b};
    }
};

}  // namespace enc

// Exclusive scan of one value per lane over the workgroup (LONG_LANES lanes).
__device__ __forceinline__ uint64_t long_scan(uint64_t x, uint64_t *sh, uint64_t &total) {
    return block_exclusive_scan(x, sh, total);
}

// Length of every long row's record (len[pos[r]]), with the entries,
// segment ends and chains kept in the scratch for k_encode_long_write;
// *err as k_encode_len.
__global__ void __launch_bounds__(enc::LONG_LANES) k_encode_long_len(const uint64_t *__restrict__ row,
                                                                     const uint32_t *__restrict__ col,
                                                                     const uint32_t *__restrict__ vid,
                                                                     const uint32_t *__restrict__ pos, uint32_t n,
                                                                     enc::LongScratch sc, uint64_t *__restrict__ len,
                                                                     unsigned long long *__restrict__ err) {
    __shared__ uint64_t sh[enc::LONG_LANES / 64];
    __shared__ uint32_t bad;
    __shared__ uint64_t segs_sh[2];
    const uint32_t r = blockIdx.x, t = threadIdx.x, T = blockDim.x;
    const uint32_t p = pos[r], v = vid[p];
    const uint64_t r0 = row[v], d = row[v + 1] - r0;
    const uint32_t *adj = col + r0;
    if (t == 0) bad = 0;
    __syncthreads();
    uint64_t ni, nr;
    if (!enc::runs_count(adj, d, n, t, T, ni, nr)) atomicOr(&bad, 1u);
    __syncthreads();
    if (bad) {
        if (t == 0) {
            atomicCAS(err, 0ull, (unsigned long long)v + 1);
            len[p] = 0;
        }
        return;
    }
    uint64_t NI, NR;
    const uint64_t oi = long_scan(ni, sh, NI);
    const uint64_t orr = long_scan(nr, sh, NR);
    const uint64_t b = sc.base[r];
    enc::runs_place(adj, d, t, T, oi, orr, sc.left + b, sc.len + b, sc.rid + b);
    const enc::LongBlock bi = sc.block(r, true, v, NI), br = sc.block(r, false, v, NR);
    __syncthreads();
    uint64_t tot;
    const uint64_t pi = long_scan(enc::later_sum(bi, t, T), sh, tot);
    const uint64_t pr = long_scan(enc::later_sum(br, t, T), sh, tot);
    enc::later_prefix(bi, t, T, pi);
    enc::later_prefix(br, t, T, pr);
    __syncthreads();
    enc::segment_ends(bi, t, T);
    enc::segment_ends(br, t, T);
    __syncthreads();
    if (t == 0) segs_sh[0] = enc::segment_chain(bi);
    if (t == 64) segs_sh[1] = enc::segment_chain(br);
    __syncthreads();
    if (t == 0) {
        enc::LongInfo f{NI, NR, segs_sh[0], segs_sh[1], enc::block_bits(bi, segs_sh[0]), enc::block_bits(br, segs_sh[1])};
        sc.info[r] = f;
        len[p] = f.bits_i + f.bits_r;
    }
}

// Writes every long row's record at its offset (stream words zeroed).
__global__ void __launch_bounds__(enc::LONG_LANES) k_encode_long_write(const uint32_t *__restrict__ vid,
                                                                       const uint32_t *__restrict__ pos,
                                                                       enc::LongScratch sc,
                                                                       const uint64_t *__restrict__ offs,
                                                                       uint64_t *words) {
    const uint32_t r = blockIdx.x, t = threadIdx.x, T = blockDim.x;
    const uint32_t p = pos[r], v = vid[p];
    const enc::LongInfo f = sc.info[r];
    const enc::LongBlock bi = sc.block(r, true, v, f.ni), br = sc.block(r, false, v, f.nr);
    enc::write_segments<enc::DeviceOr>(bi, f.segs_i, offs[p], words, t, T);
    enc::write_segments<enc::DeviceOr>(br, f.segs_r, offs[p] + f.bits_i, words, t, T);
}

}  // namespace lpa
