// lp_kernels.hpp -- gfx950 kernels of one synchronous label-propagation step.
//
// The reference computes each vertex's label with three kernel shapes picked
// by degree (baselines/novaplus/kernel.cuh:14-191): a block per vertex of
// degree >= 128 with a count-min sketch (include/decoder.cuh:103-419), a warp
// per medium vertex with an LDS hash (decoder.cuh:618-800), and warp kernels
// over an uncompressed CSR for small vertices.  It updates labels in place,
// prefers larger labels on ties and drops colliding labels.
//
// Here every shape computes the exact LP-SC result (double-buffered labels,
// most frequent label, ties to the smallest) and shapes differ only in
// speed:
//   k_lane   one lane per vertex (deg <= 32) decodes into an LDS list;
//            then one lane per (vertex, label vector) pair gathers the
//            labels into registers, sorts them with a bitonic network and
//            takes the longest run.
//   k_group  G lanes per vertex (G = 16: a quarter wave, 64: one wave,
//            256: one workgroup): segments (256-bit units) decoded one per
//            lane into an LDS list, then each neighbour's label row gathered
//            once per batch of label vectors and counted, vector by vector,
//            in an open-address LDS hash table with 64-bit atomics on
//            (label << 32 | count) words.
// Vertices above the count-min-sketch threshold and vertices too large for
// a workgroup's LDS table run through the heavy-vertex kernels.
// A lane keeps max((count << 32) | ~label) over every count it produced, so
// the group maximum names the most frequent label and, among equal counts,
// the smallest one, without a pass over the table.
//
// Labels: layout L[v*K + k]; bit 31 marks a clamped seed (semi-supervised),
// 0xFFFFFFFF an unlabelled entry.  Bins are contiguous ranges of the
// device layout (records reordered by bin, then degree).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cgr_decode.hpp"

namespace lpa {

constexpr uint32_t UNL = 0xFFFFFFFFu;
constexpr uint32_t SEED_BIT = 0x80000000u;
constexpr uint64_t EMPTY_SLOT = 0xFFFFFFFF00000000ull;
constexpr int BLOCK = 256;

struct StepArgs {
    // compressed part (device layout)
    const uint64_t *__restrict__ words;
    const uint64_t *__restrict__ offs;     // start bit offset per position
    const uint32_t *__restrict__ vid;      // vertex id per position
    const uint4 *__restrict__ meta;        // segment metadata per position (see k_scan_records)
    // CSR part (small vertices, or the CGLP decoded CSR)
    const uint64_t *__restrict__ row;      // per position: [row[p], row[p+1])
    const uint32_t *__restrict__ nbr;
    const uint32_t *__restrict__ cvid;     // vertex id per CSR position
    // labels
    const uint32_t *__restrict__ lin;
    uint32_t *__restrict__ lout;
    uint32_t K;
    // CMS (row hashes a_r, b_r)
    uint32_t cms_depth, cms_logw;
    uint64_t cms_ab[16];
    unsigned long long *__restrict__ changed;
    uint32_t begin, end;                   // position range of this bin
    // work recorder (null: off): lane-cycles spent per vertex id
    unsigned long long *__restrict__ rec;
    uint32_t xcds;                         // > 1: tiles dealt per XCD (xcd_tile), else in workgroup order
};

// Workgroup b of a launch runs on XCD b mod X (the dispatcher deals
// workgroups to the XCDs round robin, and each XCD has its own L2).  The
// tile a workgroup takes: XCD x's workgroups take the contiguous tiles
// [x q + min(x, r), ...) of the n = X q + r tiles, so the workgroups sharing
// an L2 work on neighbouring vertices (a bijection of [0, n)).
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n, uint32_t X) {
    if (X <= 1) return b;
    const uint32_t x = b % X, i = b / X, q = n / X, r = n % X;
    return x * q + (x < r ? x : r) + i;
}

// shader clock (cycles)
__device__ __forceinline__ uint64_t cycles() { return __builtin_readcyclecounter(); }

__device__ __forceinline__ uint32_t clean_label(uint32_t raw) {
    return raw == UNL ? UNL : (raw & ~SEED_BIT);
}

__device__ __forceinline__ uint32_t pick(uint64_t best, uint32_t old) {
    // best == 0: no labelled neighbour -> keep; seeds stay clamped
    if ((old != UNL && (old & SEED_BIT)) || best == 0) return old;
    return ~(uint32_t)best;
}

// Adds the block's changed labels to the counter with one atomic: n is the
// calling lane's count; every thread of the block must call it.
__device__ __forceinline__ void flush_changes(uint32_t n, unsigned long long *counter) {
    __shared__ uint32_t total;
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o, 64);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(&total, n);
    __syncthreads();
    if (threadIdx.x == 0 && total) atomicAdd(counter, (unsigned long long)total);
}

__device__ __forceinline__ void count_changes(bool ch, unsigned long long *counter) {
    uint64_t b = __ballot(ch);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(counter, (unsigned long long)__popcll(b));
}

// ------------------------------------------------------------ lane kernel

template <int N>
__device__ __forceinline__ void bitonic_sort(uint32_t (&x)[N]) {
#pragma unroll
    for (int k = 2; k <= N; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < N; i++) {
                const int l = i ^ j;
                if (l > i) {
                    uint32_t a = x[i], b = x[l];
                    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
                    if ((i & k) == 0) { x[i] = lo; x[l] = hi; }
                    else { x[i] = hi; x[l] = lo; }
                }
            }
}

// longest run of a sorted array; UNL entries (sorted last) never win
template <int N>
__device__ __forceinline__ uint64_t longest_run(const uint32_t (&x)[N]) {
    uint32_t best = 0, bestc = 0, run = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        run = (i > 0 && x[i] == x[i - 1]) ? run + 1 : 1;
        bool win = x[i] != UNL && run > bestc;
        bestc = win ? run : bestc;
        best = win ? x[i] : best;
    }
    return bestc ? ((uint64_t)bestc << 32) | (uint32_t)~best : 0;
}

// Loads the labels of vectors k0 .. k0+kb-1 (kb <= KB) of vertex u, cleaned
// (UNL beyond kb): the K labels of a vertex are contiguous, so one gather
// brings a whole row slice (16-byte loads when K and k0 are multiples of 4).
template <int KB>
__device__ __forceinline__ void load_row(const StepArgs &a, uint32_t u, uint32_t k0, uint32_t kb,
                                         uint32_t (&x)[KB]) {
    const uint32_t *src = a.lin + (uint64_t)u * a.K + k0;
    if (KB % 4 == 0 && kb == (uint32_t)KB && ((a.K | k0) & 3) == 0) {
#pragma unroll
        for (int q = 0; q < KB / 4; q++) {
            const uint4 w = reinterpret_cast<const uint4 *>(src)[q];
            x[4 * q] = clean_label(w.x);
            x[4 * q + 1] = clean_label(w.y);
            x[4 * q + 2] = clean_label(w.z);
            x[4 * q + 3] = clean_label(w.w);
        }
    } else {
#pragma unroll
        for (int t = 0; t < KB; t++) x[t] = (uint32_t)t < kb ? clean_label(src[t]) : UNL;
    }
}

// One lane per vertex decodes the record into LDS (CSR=false) or looks up
// its CSR row (CSR=true); then the block's (vertex, label vector) pairs are
// dealt out to the lanes with the vector index fastest, so the lanes that
// gather neighbour u's labels for consecutive vectors read consecutive
// words of u's label row in one access.  Each lane sorts its pair's labels
// with a bitonic network and takes the longest run.
template <int N, bool CSR>
__global__ void __launch_bounds__(BLOCK) k_lane(StepArgs a) {
    // CSR rows follow the layout back to back, so the block's rows are one
    // range of nbr: the workgroup copies it to LDS with coalesced loads
    // instead of every lane walking its own row (N <= 32: at most 32 KB)
    constexpr bool STAGE = CSR && N <= 32;
    __shared__ uint32_t list[CSR ? 1 : N][BLOCK];
    __shared__ uint32_t rows[STAGE ? N * BLOCK : 1];
    __shared__ uint32_t sv[BLOCK], sd[BLOCK];
    __shared__ uint64_t sbase[CSR ? BLOCK : 1];
    const uint32_t first = a.begin + xcd_tile(blockIdx.x, gridDim.x, a.xcds) * BLOCK;
    const uint32_t p = first + threadIdx.x;
    const uint32_t nact = a.end - first < (uint32_t)BLOCK ? a.end - first : (uint32_t)BLOCK;
    const uint64_t t0 = a.rec ? cycles() : 0;
    const uint64_t b0 = STAGE ? a.row[first] : 0;
    if (p < a.end) {
        uint32_t d = 0;
        if (CSR) {
            sv[threadIdx.x] = a.cvid[p];
            sbase[threadIdx.x] = a.row[p];
            d = (uint32_t)(a.row[p + 1] - a.row[p]);
        } else {
            const uint32_t v = a.vid[p];
            sv[threadIdx.x] = v;
            decode_record(a.words, a.offs[p], v, [&](uint32_t u) {
                if (d < N) list[d][threadIdx.x] = u;
                d++;
            });
        }
        sd[threadIdx.x] = d;
        if (a.rec) atomicAdd(&a.rec[sv[threadIdx.x]], cycles() - t0);
    }
    if (STAGE) {
        // every row of a k_lane<N> bin holds at most N ids; each lane issues
        // a round's 8 loads before its LDS stores
        uint64_t tot = a.row[first + nact] - b0;
        tot = tot < (uint64_t)N * BLOCK ? tot : (uint64_t)N * BLOCK;
        constexpr int U = 8;
        for (uint32_t x0 = 0; x0 < (uint32_t)tot; x0 += U * BLOCK) {
            uint32_t r[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t x = x0 + u * BLOCK + threadIdx.x;
                r[u] = a.nbr[b0 + (x < (uint32_t)tot ? x : 0)];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint32_t x = x0 + u * BLOCK + threadIdx.x;
                if (x < (uint32_t)tot) rows[x] = r[u];
            }
        }
    }
    __syncthreads();
    const uint32_t K = a.K, pairs = nact * K;
    uint32_t nch = 0;
    for (uint32_t f0 = 0; f0 < pairs; f0 += BLOCK) {
        const uint64_t t1 = a.rec ? cycles() : 0;
        const uint32_t f = f0 + threadIdx.x;
        const bool active = f < pairs;
        const uint32_t vb = active ? f / K : 0, k = active ? f - vb * K : 0;
        const uint32_t d = active ? sd[vb] : 0;
        const uint64_t idx = (uint64_t)sv[vb] * K + k;
        const uint32_t old = active ? a.lin[idx] : UNL, own = clean_label(old);
        // the neighbours' labels: a lane issues all its loads before using
        // any (a slot beyond its row reads its own label word), and slots
        // beyond every row of the wave are skipped by a wave-uniform test
        uint32_t dw = d;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const uint32_t y = __shfl_xor(dw, o, 64);
            dw = y > dw ? y : dw;
        }
        dw = __builtin_amdgcn_readfirstlane(dw);
        const uint64_t rb = CSR ? (STAGE ? sbase[CSR ? vb : 0] - b0 : sbase[CSR ? vb : 0]) : 0;
        uint32_t x[N];
#pragma unroll
        for (int j = 0; j < N; j++) {
            x[j] = UNL;
            if ((uint32_t)j < dw) {
                const uint32_t jj = (uint32_t)j < d ? (uint32_t)j : 0u;
                const uint32_t u = STAGE ? rows[(uint32_t)rb + jj] : CSR ? a.nbr[rb + jj] : list[CSR ? 0 : j][vb];
                x[j] = a.lin[(uint32_t)j < d ? (uint64_t)u * K + k : idx];
            }
        }
#pragma unroll
        for (int j = 0; j < N; j++) x[j] = (uint32_t)j < d ? clean_label(x[j]) : UNL;
        // own label a strict majority of the labelled neighbours -> it is
        // the mode; otherwise sort and take the longest run
        uint32_t oc = 0, lc = 0;
#pragma unroll
        for (int j = 0; j < N; j++) {
            oc += x[j] == own && own != UNL;
            lc += x[j] != UNL;
        }
        uint64_t best;
        if (own != UNL && 2 * oc > lc) {
            best = ((uint64_t)oc << 32) | (uint32_t)~own;
        } else {
            bitonic_sort<N>(x);
            best = longest_run<N>(x);
        }
        if (active) {
            const uint32_t v = sv[vb];
            const uint32_t nl = pick(best, old);
            a.lout[idx] = nl;
            nch += nl != old;
            if (a.rec) atomicAdd(&a.rec[v], cycles() - t1);
        }
    }
    flush_changes(nch, a.changed);
}

// ------------------------------------------------------------ group kernels

// Groups of G lanes: G = 16 (four vertices per wave), 64 (one wave) or 256
// (one workgroup).  Sub-wave groups synchronise with the wave.
template <int G>
__device__ __forceinline__ void group_sync() {
    if (G == BLOCK) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// lanes of the wave that belong to this lane's group
template <int G>
__device__ __forceinline__ uint64_t group_mask() {
    if (G >= 64) return ~0ull;
    const uint32_t lane = __lane_id();
    return (((1ull << G) - 1) << (lane & ~(uint32_t)(G - 1)));
}

template <int W>
__device__ __forceinline__ uint64_t max_u64_over(uint64_t x) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) {
        uint64_t y = __shfl_xor(x, o, 64);
        x = y > x ? y : x;
    }
    return x;
}

// Add `cnt` occurrences of lab to the table (2^logT slots); returns the
// label's count after the addition.  A slot is one 64-bit word (label << 32
// | count), cleared to EMPTY_SLOT; the label half is claimed with a 32-bit
// compare-and-swap (which also finds a slot already holding the label) and
// the count half takes a 32-bit atomic add, so every probe is one CAS and
// the hit one more add, with no separate read.
__device__ __forceinline__ uint32_t tab_add(uint64_t *tab, uint32_t logT, uint32_t lab, uint32_t cnt) {
    const uint32_t mask = (1u << logT) - 1;
    // the counting kernels are VALU-bound: an xor-fold of the label (three
    // full-rate operations, no quarter-rate multiply) measured ~1 % faster
    // than a multiplicative hash; the result never depends on the slot a
    // label takes
    uint32_t h = (lab ^ (lab >> 16)) & mask;
    uint32_t *w = reinterpret_cast<uint32_t *>(tab);   // little endian: w[2h] count, w[2h+1] label
    for (;;) {
        const uint32_t old = atomicCAS(&w[2 * h + 1], UNL, lab);
        if (old == UNL || old == lab) return atomicAdd(&w[2 * h], cnt) + cnt;
        h = (h + 1) & mask;
    }
}

// Pre-aggregation inside a group: the active lanes of the group holding the
// same label as its first active lane are folded into one update by the
// lowest of them.  Returns how many occurrences this lane applies (0: folded).
template <int G>
__device__ __forceinline__ uint32_t fold_first_label(uint32_t lab) {
    const uint64_t gm = group_mask<G>();
    const uint64_t act = __ballot(1) & gm;
    const int leader = __builtin_ctzll(act);
    const uint32_t first = G >= 64 ? __builtin_amdgcn_readfirstlane(lab) : (uint32_t)__shfl((int)lab, leader, 64);
    const uint64_t m = __ballot(lab == first) & gm;
    if (lab != first) return 1;
    return __lane_id() == (uint32_t)__builtin_ctzll(m) ? (uint32_t)__popcll(m) : 0;
}

__device__ __forceinline__ uint32_t cms_col(const StepArgs &a, uint32_t r, uint32_t lab) {
    uint64_t h = a.cms_ab[2 * r] * (uint64_t)lab + a.cms_ab[2 * r + 1];
    return a.cms_logw ? (uint32_t)(h >> (64 - a.cms_logw)) : 0;
}

// Intervals at least this long are written by the whole group.
constexpr uint32_t LONG_RUN = 16;

template <int G> struct GroupShared {
    static constexpr int QCAP = G == BLOCK ? 512 : G == 64 ? 64 : G == 32 ? 32 : 16;
    uint64_t res_start;
    uint32_t cnt;                 // ids written to the list so far
    uint32_t qn;
    uint32_t qleft[QCAP];
    uint32_t qlen[QCAP];
    uint32_t qpos[QCAP];
};

// Phase A: the G lanes of a group decode the record at position p into
// `list` (order irrelevant to the mode).  Segments (256-bit units) go one
// per lane: the interval block first, then the residual block, which starts
// where the last interval segment ends.  Each run claims its slots with one
// LDS atomic, each residual segment with one atomic for its whole count;
// runs of LONG_RUN or more ids are written by all lanes together so a hub
// whose neighbours form one long run is not written by a single lane.
// Returns the degree (valid in every lane of the group).
template <int G>
__device__ __forceinline__ uint32_t group_decode(const StepArgs &a, uint32_t p, uint32_t v, uint32_t gl,
                                                 GroupShared<G> *sh, uint32_t *list) {
    if (gl == 0) {
        sh->cnt = 0;
        sh->qn = 0;
    }
    group_sync<G>();
    // segment j of the record (intervals first, then residuals) is at a
    // position known from the metadata: every lane seeks its own segment
    const uint64_t start = a.offs[p];
    const uint4 m = a.meta[p];
    const uint32_t ni = m.y, ns = m.y + m.w;
    auto emit_run = [&](uint32_t l, uint32_t len) {
        const uint32_t o = atomicAdd(&sh->cnt, len);
        if (len >= LONG_RUN) {
            const uint32_t q = atomicAdd(&sh->qn, 1u);
            if (q < (uint32_t)GroupShared<G>::QCAP) {
                sh->qleft[q] = l;
                sh->qlen[q] = len;
                sh->qpos[q] = o;
                return;
            }
        }
        for (uint32_t x = 0; x < len; x++) list[o + x] = l + x;
    };
    for (uint32_t j = gl; j < ns; j += G) {
        BitReader r;
        if (j < ni) {
            r.seek(a.words, start + m.x + (uint64_t)SEG_BITS * j);
            read_interval_segment(r, v, emit_run);
        } else {
            r.seek(a.words, start + m.z + (uint64_t)SEG_BITS * (j - ni));
            BitReader peek = r;   // the segment's count claims its list slots
            uint32_t o = atomicAdd(&sh->cnt, peek.gamma());
            read_residual_segment(r, v, [&](uint32_t u) { list[o++] = u; });
        }
    }
    group_sync<G>();
    const uint32_t nq = sh->qn < (uint32_t)GroupShared<G>::QCAP ? sh->qn : (uint32_t)GroupShared<G>::QCAP;
    for (uint32_t q = 0; q < nq; q++) {
        const uint32_t l = sh->qleft[q], len = sh->qlen[q], o = sh->qpos[q];
        for (uint32_t x = gl; x < len; x += G) list[o + x] = l + x;
    }
    if (G == BLOCK) __threadfence_block();
    group_sync<G>();
    return sh->cnt;
}

// Aggregated counting rounds (per wave, or per sub-wave group): the group's
// first lane still holding an uncounted label offers it, every copy of it
// among the group's staged labels is found with one ballot per register and
// the total is added by that lane with a single table update.  Rounds stop
// when the wave has nothing left, after MATCH_ROUNDS, or when no group of
// the wave matched MATCH_MIN copies (diverse neighbourhoods go one label at
// a time instead).  A round costs a compare, a ballot and a mask update per
// staged register, and the group kernels are VALU-bound with 8 vectors
// staged: one round (the most common label of a settled neighbourhood is
// usually the first offered) measured fastest -- 3.5 % on R-MAT 22 and 11 %
// on R-MAT 24 with 64 semi-supervised vectors against eight rounds, and
// faster than none.
#ifndef LPA_MATCH_ROUNDS
#define LPA_MATCH_ROUNDS 1
#endif
constexpr int MATCH_ROUNDS = LPA_MATCH_ROUNDS;
constexpr uint32_t MATCH_MIN = 4;

template <int G, int IPL, int KB, class Agg>
__device__ __forceinline__ void match_rounds(const uint32_t (&labs)[IPL][KB], int s, uint32_t &pm, Agg &&agg) {
    const uint64_t gm = group_mask<G>();
    const uint32_t lane = __lane_id();
    for (int r = 0; r < MATCH_ROUNDS; r++) {
        const uint64_t wact = __ballot(pm != 0);
        if (wact == 0) break;
        const uint64_t act = wact & gm;
        uint32_t mine = UNL;
#pragma unroll
        for (int t = IPL - 1; t >= 0; t--)
            if (pm & (1u << t)) mine = labs[t][s];
        const uint32_t leader = act ? (uint32_t)__builtin_ctzll(act) : lane;
        const uint32_t lead = (uint32_t)__shfl((int)mine, (int)leader, 64);
        uint32_t c = 0;
#pragma unroll
        for (int t = 0; t < IPL; t++) {
            const bool m = (pm & (1u << t)) && labs[t][s] == lead;
            c += (uint32_t)__popcll(__ballot(m) & gm);
            if (m) pm &= ~(1u << t);
        }
        const bool lead_lane = act != 0 && lane == leader;
        if (lead_lane) agg(s, lead, c);
        if (__ballot(lead_lane && c >= MATCH_MIN) == 0) break;
    }
}

// Cross-wave sums of a workgroup (every thread calls it): lane 0 of each
// wave puts its KB words in xw (KB * waves words), threads 0..KB-1 add them
// up in xs (KB words), and every thread returns with the totals in x.  Two
// barriers; xw and xs can be reused by the next call without another.
template <int KB>
__device__ __forceinline__ void block_sum_words(uint32_t (&x)[KB], uint32_t *xw, uint32_t *xs) {
    const uint32_t w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int s = 0; s < KB; s++) xw[s * nw + w] = x[s];
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)KB) {
        uint32_t t = 0;
        for (uint32_t ww = 0; ww < nw; ww++) t += xw[threadIdx.x * nw + ww];
        xs[threadIdx.x] = t;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < KB; s++) x[s] = xs[s];
}

// Own-label count and labelled count of a vector, packed as (own << 16 |
// labelled): every list the own-label check sees is shorter than 2^16.
__device__ __forceinline__ uint32_t own_count(uint32_t packed) { return packed >> 16; }
__device__ __forceinline__ uint32_t lab_count(uint32_t packed) { return packed & 0xFFFFu; }
__device__ __forceinline__ bool own_majority_of(uint32_t own, uint32_t packed) {
    return own != UNL && 2 * own_count(packed) > lab_count(packed);
}

// The vertex's own label decides a vector when it is a strict majority of
// the labelled neighbours: its count then exceeds the sum of all the others,
// so it is the most frequent label, with no tie (LP-SC exact mode).  In a
// settled neighbourhood this holds for nearly every vertex and vector, and
// the check costs a compare and a ballot per staged label instead of a
// table.  own[s] is the clean own label of vector k0+s (UNL: never decides);
// every lane of a group holds the same own[] and reaches this together.
// xw, xs: LDS scratch of block_sum_words when G > 64 (the group's counts are
// summed across its waves), unused otherwise.  Returns the mask of decided
// vectors; pc[s] = the packed counts.
template <int G, int IPL, int KB>
__device__ __forceinline__ uint32_t own_majority(const uint32_t (&labs)[IPL][KB], const uint32_t (&own)[KB],
                                                 uint32_t kb, uint32_t *xw, uint32_t *xs, uint32_t (&pc)[KB]) {
    const uint64_t gm = group_mask<G>();
#pragma unroll
    for (int s = 0; s < KB; s++) {
        pc[s] = 0;
#pragma unroll
        for (int t = 0; t < IPL; t++) {
            pc[s] += (uint32_t)__popcll(__ballot(own[s] != UNL && labs[t][s] == own[s]) & gm) << 16;
            pc[s] += (uint32_t)__popcll(__ballot(labs[t][s] != UNL) & gm);
        }
    }
    if (G > 64) block_sum_words<KB>(pc, xw, xs);
    uint32_t mask = 0;
#pragma unroll
    for (int s = 0; s < KB; s++)
        if ((uint32_t)s < kb && own_majority_of(own[s], pc[s])) mask |= 1u << s;
    return mask;
}

// Phase B: counts label vectors k0 .. k0+kb-1 (kb <= KB, the same in every
// lane of the group) of the neighbours ids[0..d), one vector at a time.
// Lane gl first gathers the label-row slices of neighbours gl, gl+G, ...,
// gl+(IPL-1)G into registers (IPL*KB values, all loads in flight together),
// so a vertex's neighbour rows are read once per batch of KB vectors;
// neighbours beyond IPL*G are gathered per vector.  With use_own (the list
// is the vertex's whole neighbourhood and all of it is staged), a vector
// whose own label is a strict majority is settled by decide(s, key) with
// no table work; for every other vector s: begin(s) prepares the table (and
// synchronises the group); staged labels go through aggregated rounds
// (agg(s, lab, count) by one lane per match) and then use(s, lab) one at a
// time; end(s) finishes the vector.  st (or null, KB a multiple of 4): the
// labels come from a row stash instead of the label rows -- clean labels
// in groups of four vectors, vector k0+s of neighbour i at
// st[((s/4)*sn + i)*4 + s%4] -- written by the lane that reads them (lane
// i % G), so no synchronisation is needed.
template <int G, int IPL, int KB, class Begin, class Agg, class Use, class End, class Decide>
__device__ __forceinline__ void count_staged(const StepArgs &a, const uint32_t *__restrict__ ids, uint32_t d,
                                             uint32_t gl, uint32_t k0, uint32_t kb, const uint32_t (&own)[KB],
                                             bool use_own, uint32_t *xw, uint32_t *xs, Begin &&begin, Agg &&agg,
                                             Use &&use, End &&end, Decide &&decide,
                                             const uint32_t *__restrict__ st = nullptr, uint32_t sn = 0) {
    uint32_t labs[IPL][KB];
#pragma unroll
    for (int t = 0; t < IPL; t++) {
        const uint32_t i = t * G + gl;
        if (KB % 4 == 0 && i < d && st) {
#pragma unroll
            for (int q = 0; q < KB / 4; q++) {
                const uint4 w = 4u * q < kb ? reinterpret_cast<const uint4 *>(st)[(uint64_t)q * sn + i]
                                            : make_uint4(UNL, UNL, UNL, UNL);
                labs[t][4 * q] = w.x;
                labs[t][4 * q + 1] = 4u * q + 1 < kb ? w.y : UNL;
                labs[t][4 * q + 2] = 4u * q + 2 < kb ? w.z : UNL;
                labs[t][4 * q + 3] = 4u * q + 3 < kb ? w.w : UNL;
            }
        } else if (i < d) {
            load_row<KB>(a, ids[i], k0, kb, labs[t]);
        } else {
#pragma unroll
            for (int s = 0; s < KB; s++) labs[t][s] = UNL;
        }
    }
    uint32_t decided = 0, pc[KB];
    if (use_own && d <= (uint32_t)(IPL * G)) decided = own_majority<G, IPL, KB>(labs, own, kb, xw, xs, pc);
#pragma unroll
    for (int s = 0; s < KB; s++) {
        if ((uint32_t)s < kb) {
            if (decided & (1u << s)) {
                decide(s, ((uint64_t)own_count(pc[s]) << 32) | (uint32_t)~own[s]);
                continue;
            }
            begin(s);
            uint32_t pm = 0;
#pragma unroll
            for (int t = 0; t < IPL; t++) pm |= (labs[t][s] != UNL ? 1u : 0u) << t;
            match_rounds<G, IPL, KB>(labs, s, pm, agg);
#pragma unroll
            for (int t = 0; t < IPL; t++)
                if (pm & (1u << t)) use(s, labs[t][s]);
            for (uint32_t i = IPL * G + gl; i < d; i += G) {
                const uint32_t lab = KB % 4 == 0 && st ? st[((uint64_t)(s >> 2) * sn + i) * 4 + (s & 3)]
                                                       : clean_label(a.lin[(uint64_t)ids[i] * a.K + k0 + s]);
                if (lab != UNL) use(s, lab);
            }
            end(s);
        }
    }
}

// One group of G lanes per vertex; dynamic LDS holds, per group, 2^logT
// table slots followed by the neighbour list (listcap ids).  The record is
// decoded once; label vectors are counted KB at a time from label rows
// staged in registers, IPL neighbours per lane (the launch picks the
// smallest IPL with IPL * G >= listcap, so the staged labels, IPL * KB
// registers, cost no more occupancy than the bin needs).
template <int G, bool CSR, int KB, int IPL>
__global__ void __launch_bounds__(BLOCK) k_group(StepArgs a, uint32_t logT, uint32_t listcap) {
    constexpr int NG = BLOCK / G;
    extern __shared__ uint64_t lds[];
    __shared__ uint64_t red[BLOCK / 64];
    __shared__ uint32_t xw[G > 64 ? KB * (G / 64) : 1], xs[KB];
    __shared__ GroupShared<G> gsh[NG];
    const uint32_t gid = threadIdx.x / G, gl = threadIdx.x % G;
    const uint32_t T = 1u << logT;
    uint64_t *tab = lds + (uint64_t)gid * (T + (listcap + 1) / 2);
    uint32_t *list = (uint32_t *)(tab + T);
    GroupShared<G> *sh = &gsh[gid];
    const uint32_t stride = gridDim.x * NG;
    uint32_t nch = 0;
    for (uint32_t p0 = a.begin + xcd_tile(blockIdx.x, gridDim.x, a.xcds) * NG; p0 < a.end; p0 += stride) {
        // all groups of a wave (or workgroup) run the same trip count
        const uint32_t p = p0 + gid;
        const bool live = p < a.end;
        const uint64_t t0 = a.rec ? cycles() : 0;
        const uint32_t v = live ? (CSR ? a.cvid[p] : a.vid[p]) : 0;
        const uint32_t *ids = list;
        uint32_t d = 0;
        if (CSR) {
            if (live) {
                ids = a.nbr + a.row[p];
                d = (uint32_t)(a.row[p + 1] - a.row[p]);
            }
        } else if (live) {
            d = group_decode<G>(a, p, v, gl, sh, list);
        }
        for (uint32_t k0 = 0; k0 < a.K; k0 += KB) {
            const uint32_t kb = a.K - k0 < (uint32_t)KB ? a.K - k0 : (uint32_t)KB;
            uint64_t best = 0, mine = 0;   // lane gl < kb keeps vector k0+gl's result
            uint32_t own[KB];
#pragma unroll
            for (int s = 0; s < KB; s++)
                own[s] = live && (uint32_t)s < kb ? clean_label(a.lin[(uint64_t)v * a.K + k0 + s]) : UNL;
            count_staged<G, IPL, KB>(
                a, ids, d, gl, k0, kb, own, true, xw, xs,
                [&](int) {
                    for (uint32_t i = gl; i < T; i += G) tab[i] = EMPTY_SLOT;
                    best = 0;
                    group_sync<G>();
                },
                [&](int, uint32_t lab, uint32_t cnt) {
                    const uint32_t c = tab_add(tab, logT, lab, cnt);
                    const uint64_t key = ((uint64_t)c << 32) | (uint32_t)~lab;
                    best = key > best ? key : best;
                },
                [&](int, uint32_t lab) {
                    const uint32_t cnt = fold_first_label<G>(lab);
                    if (cnt == 0) return;
                    const uint32_t c = tab_add(tab, logT, lab, cnt);
                    const uint64_t key = ((uint64_t)c << 32) | (uint32_t)~lab;
                    best = key > best ? key : best;
                },
                [&](int s) {
                    best = max_u64_over<(G < 64 ? G : 64)>(best);
                    if (G == BLOCK) {
                        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
                        __syncthreads();
#pragma unroll
                        for (int w = 0; w < BLOCK / 64; w++) best = red[w] > best ? red[w] : best;
                    }
                    if (gl == (uint32_t)s) mine = best;
                    group_sync<G>();   // table / red reused next
                },
                [&](int s, uint64_t key) {
                    if (gl == (uint32_t)s) mine = key;
                });
            // the batch's labels are picked and written by kb lanes at once
            // (one memory round trip per batch, not per vector)
            if (gl < kb && live) {
                const uint64_t idx = (uint64_t)v * a.K + k0 + gl;
                const uint32_t old = a.lin[idx];
                const uint32_t nl = pick(mine, old);
                a.lout[idx] = nl;
                nch += nl != old;
            }
        }
        if (a.rec && live && gl == 0) a.rec[v] = (cycles() - t0) * G;
    }
    flush_changes(nch, a.changed);
}

// --------------------------------------------------------- CGLP kernel 1

// Decode the records of positions [begin,end) into CSR rows crow[p]..crow[p+1]
// (rows follow the device layout; one lane per record; the two-kernel
// method's decode step).
__global__ void __launch_bounds__(BLOCK) k_decode_csr(const uint64_t *__restrict__ words,
                                                      const uint64_t *__restrict__ offs,
                                                      const uint32_t *__restrict__ vid,
                                                      const uint64_t *__restrict__ crow,
                                                      uint32_t *__restrict__ ccol, uint32_t begin, uint32_t end) {
    const uint32_t p = begin + blockIdx.x * BLOCK + threadIdx.x;
    if (p >= end) return;
    const uint32_t v = vid[p];
    uint64_t o = crow[p];
    decode_record(words, offs[p], v, [&](uint32_t u) { ccol[o++] = u; });
}

// Same for long records: a workgroup per record, one lane per segment;
// output slots are claimed with an LDS cursor, and intervals of LONG_RUN or
// more ids are written by all lanes together.
__global__ void __launch_bounds__(BLOCK) k_decode_csr_group(const uint64_t *__restrict__ words,
                                                            const uint64_t *__restrict__ offs,
                                                            const uint32_t *__restrict__ vid,
                                                            const uint64_t *__restrict__ crow,
                                                            uint32_t *__restrict__ ccol, uint32_t begin,
                                                            uint32_t end) {
    constexpr int QCAP = 512;
    __shared__ uint64_t res_start;
    __shared__ unsigned long long cursor;
    __shared__ uint32_t qn;
    __shared__ uint32_t qleft[QCAP], qlen[QCAP];
    __shared__ uint64_t qout[QCAP];
    for (uint32_t p = begin + blockIdx.x; p < end; p += gridDim.x) {
        const uint32_t v = vid[p];
        if (threadIdx.x == 0) {
            cursor = crow[p];
            qn = 0;
        }
        __syncthreads();
        uint64_t base;
        const uint32_t ni = block_header(words, offs[p], &base);
        uint64_t my_end = 0;
        for (uint32_t j = threadIdx.x; j < ni; j += BLOCK) {
            uint64_t e = decode_interval_segment(words, base + (uint64_t)SEG_BITS * j, v,
                                                 [&](uint32_t l, uint32_t len) {
                                                     uint64_t o = atomicAdd(&cursor, (unsigned long long)len);
                                                     if (len >= LONG_RUN) {
                                                         uint32_t q = atomicAdd(&qn, 1u);
                                                         if (q < QCAP) {
                                                             qleft[q] = l;
                                                             qlen[q] = len;
                                                             qout[q] = o;
                                                             return;
                                                         }
                                                     }
                                                     for (uint32_t x = 0; x < len; x++) ccol[o + x] = l + x;
                                                 });
            if (j == ni - 1) my_end = e;
        }
        if (threadIdx.x == ((ni - 1) & (BLOCK - 1))) res_start = my_end;
        __syncthreads();
        const uint32_t nq = qn < (uint32_t)QCAP ? qn : (uint32_t)QCAP;
        for (uint32_t q = 0; q < nq; q++)
            for (uint32_t x = threadIdx.x; x < qlen[q]; x += BLOCK) ccol[qout[q] + x] = qleft[q] + x;
        const uint32_t nr = block_header(words, res_start, &base);
        for (uint32_t j = threadIdx.x; j < nr; j += BLOCK) {
            const uint64_t q = base + (uint64_t)SEG_BITS * j;
            BitReader r;
            r.seek(words, q);
            const uint32_t c = r.gamma();
            uint64_t o = atomicAdd(&cursor, (unsigned long long)c);
            decode_residual_segment(words, q, v, [&](uint32_t u) { ccol[o++] = u; });
        }
        __syncthreads();
    }
}

}  // namespace lpa
