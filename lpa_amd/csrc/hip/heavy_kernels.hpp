// heavy_kernels.hpp -- the heaviest vertices (count-min-sketch vertices and
// vertices whose hash table does not fit LDS), spread over many workgroups.
//
// One workgroup per vertex cannot keep enough label gathers in flight for a
// hub (R-MAT scale 22 has a vertex of degree 163,396; scale 27 hubs reach
// millions), so a step over these "heavy" vertices runs as
//   k_heavy_decode  one workgroup per run of up to 256 segments of a heavy
//                   record: decode into the vertex's neighbour list in HBM,
//   k_heavy_clear   clear the sketches / tables / result words of a batch of
//                   label vectors [k0, k0+kb) (their memory is reused by the
//                   next batch when K is large),
//   k_heavy_chunks<1 or 0>  one workgroup per chunk of CHUNK list entries:
//                   gather the labels and add them to the vertex's count-min
//                   sketch (1), or count them in its HBM hash table (0, exact
//                   mode; the running max of (count << 32 | ~label) goes to
//                   the result word),
//   k_heavy_chunks<2>  (CMS only) the same chunks again: estimate every
//                   label from the completed sketch, max into the result word,
//   k_heavy_finish  one lane per heavy vertex and label vector: pick the
//                   label, write it, count changes.
// Small-vertex CSR rows (and the CGLP CSR) skip the decode: their list is
// the CSR row itself.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lp_kernels.hpp"

namespace lpa {

constexpr uint32_t CHUNK = 4096;
// workgroup size of the chunk and one-workgroup heavy kernels (large LDS
// tables and sketches; many waves share them)
constexpr int HBLOCK = 1024;
// CMS vertices up to this degree are counted by one workgroup (LDS
// sketches; with several label vectors their cells are 16-bit, two vectors
// to a word, and no cell of a list this short can reach 2^16).
constexpr uint32_t CMS_LOCAL = 65535;

struct HeavyArgs {
    StepArgs s;                                // words/offs/vid, row/nbr/cvid, labels, CMS hashes
    uint32_t nh;                               // heavy vertices in this launch
    const uint32_t *__restrict__ pos;          // position (device layout) of each heavy vertex
    const uint64_t *__restrict__ loff;         // start of each heavy vertex's list (nh + 1)
    uint32_t *__restrict__ list;               // decoded ids (unused when csr)
    bool csr;                                  // lists are CSR rows: ids = nbr + row[pos]
    const uint64_t *__restrict__ toff;         // exact mode: table slots prefix per heavy vertex (nh + 1);
                                               // vector k of the batch uses tab[toff[i]*kb + (k-k0)*T_i ...]
    uint64_t *__restrict__ tab;
    uint32_t *__restrict__ sketch;             // CMS: (cms_depth << cms_logw) words per heavy vertex and vector
    unsigned long long *__restrict__ result;   // nh * kb
    uint32_t nchunks;
    const uint32_t *__restrict__ chunk_h;      // heavy vertex of each chunk
    const uint64_t *__restrict__ chunk_b;      // first list entry of each chunk (relative to the vertex)
    int cms;                                   // 1: CMS mode, 0: exact
    uint32_t k0, kb;                           // label vectors of this batch
    uint32_t kbs;                              // CMS: label vectors per sketch pass (sketches side by side in LDS)
    uint32_t kbl;                              // CMS one-workgroup kernel: vectors per pass (packed 16-bit cells)
    const uint32_t *__restrict__ slot;         // chunked vertex i -> its sketch / result slot
    uint32_t *stash;                           // CMS, several vectors: per-workgroup label stash; exact
                                               // one-workgroup kernel: per-workgroup row stash (or NULL)
    uint64_t stash_stride;                     // words per workgroup
    uint32_t *epoch;                           // exact mode, tagged tables: passes completed (device word)
    int tagged;                                // exact mode: HBM table slots carry the pass's generation
    uint32_t kb_alloc;                         // exact mode: vectors the table buffer is laid out for
    uint32_t *wq;                              // exact one-workgroup kernel with a row stash: work
                                               // counter and exit counter (zero between launches)
    uint32_t *hcnt;                            // CMS, chunked hubs: per (slot, vector of the batch) 4 words:
                                               // own-label count, labelled count, full-collision flag, unused
};

// Exact-mode HBM tables with generation tags.  A slot's count word holds
// (generation << 24 | count) and a slot whose generation is not the current
// pass's counts as empty, so a pass needs no clearing of its tables: only
// when the 8-bit generation wraps (every TABLE_GENS passes) are the tables
// cleared in full.  Used when every chunked list is shorter than 2^24.
// The generation is derived on the device from the pass counter, which the
// pass's last kernel advances, so a captured iteration replays correctly.
constexpr uint32_t TABLE_GENS = 255;
constexpr uint32_t COUNT_BITS = 24;
constexpr uint32_t COUNT_MASK = (1u << COUNT_BITS) - 1;

__device__ __forceinline__ uint32_t table_gen(const HeavyArgs &h) { return *h.epoch % TABLE_GENS + 1; }

// tab_add for a tagged HBM table: claims a stale or empty slot with one
// 64-bit CAS, else adds to the matching label's count word.  Returns the
// label's count after the addition.
__device__ __forceinline__ uint32_t tab_add_tagged(uint64_t *tab, uint32_t logT, uint32_t lab, uint32_t cnt,
                                                   uint32_t gen) {
    const uint64_t mask = (1ull << logT) - 1;
    uint64_t hh = logT ? (lab * 2654435769u) >> (32 - logT) : 0;
    uint64_t cur = __hip_atomic_load(&tab[hh], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        const uint32_t slab = (uint32_t)(cur >> 32), cw = (uint32_t)cur;
        if (slab != UNL && (cw >> COUNT_BITS) == gen) {
            if (slab == lab) {
                uint32_t *w = reinterpret_cast<uint32_t *>(&tab[hh]);   // little endian: w[0] count word
                return ((atomicAdd(w, cnt) + cnt) & COUNT_MASK);
            }
            hh = (hh + 1) & mask;
            cur = __hip_atomic_load(&tab[hh], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            continue;
        }
        const uint64_t want = ((uint64_t)lab << 32) | ((uint64_t)gen << COUNT_BITS) | cnt;
        const uint64_t old = atomicCAS((unsigned long long *)&tab[hh], (unsigned long long)cur,
                                       (unsigned long long)want);
        if (old == cur) return cnt;
        cur = old;   // taken meanwhile: look at what is there now
    }
}

__device__ __forceinline__ const uint32_t *heavy_ids(const HeavyArgs &h, uint32_t i) {
    return h.csr ? h.s.nbr + h.s.row[h.pos[i]] : h.list + h.loff[i];
}

__device__ __forceinline__ uint32_t heavy_deg(const HeavyArgs &h, uint32_t i) {
    return (uint32_t)(h.loff[i + 1] - h.loff[i]);
}

// Decode tasks: up to BLOCK consecutive segments (one per lane; intervals
// first, then residuals) of one heavy record, so a hub's record is decoded
// by many workgroups at once.  A task's descriptor is fixed for a layout:
// where its record's interval and residual segments start, the vertex, its
// first segment, and where its ids go in the vertex's list and how many
// there are (counted once, by k_heavy_task_count, when the layout is built;
// the order of a list is irrelevant to the counts).
struct DecodeTask {
    uint64_t ibase, rbase;   // absolute bit of interval segment 0 / residual segment 0
    uint64_t out;            // first list slot of the task's ids
    uint32_t v, j0, ni, ns;  // vertex, first segment, #interval segments, #segments
    uint32_t count, pad;     // ids the task writes
};
constexpr uint32_t DECODE_STAGE = 8192;

__device__ __forceinline__ uint32_t task_segment(const StepArgs &a, const DecodeTask &T, uint32_t j, BitReader &r) {
    r.seek(a.words, j < T.ni ? T.ibase + (uint64_t)SEG_BITS * j : T.rbase + (uint64_t)SEG_BITS * (j - T.ni));
    return j < T.ni;
}

// One-time: ids of every task (pass of the decoder that only counts).
__global__ void __launch_bounds__(BLOCK) k_heavy_task_count(StepArgs a, DecodeTask *tasks, uint32_t ntasks) {
    __shared__ uint32_t total;
    const uint32_t t = blockIdx.x;
    if (t >= ntasks) return;
    const DecodeTask T = tasks[t];
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
    const uint32_t j = T.j0 + threadIdx.x;
    uint32_t mine = 0;
    if (j < T.ns) {
        BitReader r;
        if (task_segment(a, T, j, r)) read_interval_segment(r, T.v, [&](uint32_t, uint32_t len) { mine += len; });
        else mine = r.gamma();
    }
    if (mine) atomicAdd(&total, mine);
    __syncthreads();
    if (threadIdx.x == 0) tasks[t].count = total;
}

// One workgroup per task: each lane decodes its segment, claiming output
// slots with one LDS atomic per run (runs of LONG_RUN or more ids are
// written by the whole workgroup) or per residual segment; the task's ids
// are assembled in LDS and stored to the list with coalesced stores, or
// written in place when they exceed the staging buffer.
// CMS: the heavy set's mode (names the launch in profiles; same code)
template <int CMS>
__global__ void __launch_bounds__(BLOCK) k_heavy_decode(HeavyArgs h, const DecodeTask *__restrict__ tasks,
                                                        uint32_t ntasks) {
    __shared__ GroupShared<BLOCK> sh;
    __shared__ uint32_t stage[DECODE_STAGE];
    const uint32_t t = blockIdx.x;
    if (t >= ntasks || h.csr) return;
    const uint64_t t0 = h.s.rec ? cycles() : 0;
    const DecodeTask T = tasks[t];
    if (threadIdx.x == 0) {
        sh.cnt = 0;
        sh.qn = 0;
    }
    __syncthreads();
    uint32_t *out = h.list + T.out;
    const bool staged = T.count <= DECODE_STAGE;
    uint32_t *dst = staged ? stage : out;
    const uint32_t j = T.j0 + threadIdx.x;
    if (j < T.ns) {
        BitReader r;
        if (task_segment(h.s, T, j, r)) {
            read_interval_segment(r, T.v, [&](uint32_t l, uint32_t len) {
                const uint32_t o = atomicAdd(&sh.cnt, len);
                if (len >= LONG_RUN) {
                    const uint32_t q = atomicAdd(&sh.qn, 1u);
                    if (q < (uint32_t)GroupShared<BLOCK>::QCAP) {
                        sh.qleft[q] = l;
                        sh.qlen[q] = len;
                        sh.qpos[q] = o;
                        return;
                    }
                }
                for (uint32_t x = 0; x < len; x++) dst[o + x] = l + x;
            });
        } else {
            BitReader peek = r;   // the segment's count claims its slots
            uint32_t o = atomicAdd(&sh.cnt, peek.gamma());
            read_residual_segment(r, T.v, [&](uint32_t u) { dst[o++] = u; });
        }
    }
    __syncthreads();
    const uint32_t nq = sh.qn < (uint32_t)GroupShared<BLOCK>::QCAP ? sh.qn : (uint32_t)GroupShared<BLOCK>::QCAP;
    for (uint32_t q = 0; q < nq; q++) {
        const uint32_t l = sh.qleft[q], len = sh.qlen[q], qo = sh.qpos[q];
        for (uint32_t x = threadIdx.x; x < len; x += BLOCK) dst[qo + x] = l + x;
    }
    if (staged) {
        __syncthreads();
        for (uint32_t x = threadIdx.x; x < T.count; x += BLOCK) out[x] = stage[x];
    }
    if (h.s.rec && threadIdx.x == 0) atomicAdd(&h.s.rec[T.v], (cycles() - t0) * BLOCK);
}

// chunked[] lists the heavy-vertex indices handled by the chunk kernels.
template <int CMS>
__global__ void __launch_bounds__(BLOCK) k_heavy_clear(HeavyArgs h, const uint32_t *__restrict__ chunked,
                                                       uint32_t nchunked) {
    if (blockIdx.x >= nchunked) return;
    const uint32_t i = chunked[blockIdx.x], sl = h.slot[i];
    for (uint32_t k = threadIdx.x; k < h.kb; k += BLOCK) h.result[(uint64_t)sl * h.kb + k] = 0;
    if (h.cms) {
        const uint64_t cells = ((uint64_t)h.s.cms_depth << h.s.cms_logw) * h.kb;
        uint32_t *sk = h.sketch + (uint64_t)sl * cells;
        for (uint64_t c = threadIdx.x; c < cells; c += BLOCK) sk[c] = 0;
        for (uint32_t k = threadIdx.x; k < 4 * h.kb; k += BLOCK) h.hcnt[(uint64_t)sl * h.kb * 4 + k] = 0;
    } else if (!h.tagged) {
        for (uint64_t c = h.toff[i] * h.kb + threadIdx.x; c < h.toff[i + 1] * h.kb; c += BLOCK) h.tab[c] = EMPTY_SLOT;
    } else if (*h.epoch % TABLE_GENS == 0) {
        // generation 1 again: clear the vertex's share of the whole buffer
        // (laid out for kb_alloc vectors, whatever this pass's kb)
        for (uint64_t c = h.toff[i] * h.kb_alloc + threadIdx.x; c < h.toff[i + 1] * h.kb_alloc; c += BLOCK)
            h.tab[c] = EMPTY_SLOT;
    }
}

// Block-wide maximum (every lane gets it); red[] is free again on return.
__device__ __forceinline__ uint64_t block_max(uint64_t x, uint64_t *red) {
    x = max_u64_over<64>(x);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
#pragma unroll
    for (int w = 0; w < HBLOCK / 64; w++) x = red[w] > x ? red[w] : x;
    __syncthreads();
    return x;
}

// Exact counting stages the label rows of up to CHUNK neighbours per
// workgroup in registers (HEAVY_IPL per lane), KB vectors at a time.
constexpr int HEAVY_IPL = CHUNK / HBLOCK;
// Sketch passes walk the list in rounds of CMS_R neighbours per lane and
// serve up to KB label vectors (kbs sketches side by side) per round.
constexpr int CMS_R = 4;
// Label vectors per staged batch of the heavy kernels (template KB): 1 for
// a single vector, else 8 for the exact chunk kernel, 4 for the exact
// one-workgroup kernel with 64 KB tables (16 staged labels per lane fit the
// 64 registers of eight waves per SIMD without spilling) and 8 for the one
// with 128 KB tables (one workgroup per CU: 128 registers).
constexpr int HEAVY_KB_MAX = 8;
constexpr int HEAVY_KB_LOCAL = 4;
// The sketch kernels (one-workgroup and chunk) pack two vectors' 16-bit LDS
// cells in a word, so eight sketches take the LDS of four and a row is
// gathered once per 8 vectors.
constexpr int HEAVY_KB_CMS_LOCAL = 8;

// One sketch pass over ids[0..n) for vectors k0 .. k0+kb-1 (kb <= KB;
// sketch s at sk + s*cells, or with PACK the 16-bit half s&1 of the words
// at sk + (s>>1)*cells).  ADD: add every label to its sketch (LDS,
// pre-aggregated per wave); else: estimate every label from its sketch and
// keep the best (count << 32 | ~label) per vector in best[s].
// stash (HBM, n*KB words, or NULL): the ADD pass stores every gathered
// label there ([id][KB]) and the estimating pass reads the labels back in
// order instead of gathering the rows again (a second random gather of a
// graph whose labels do not fit the caches costs a full memory access per
// arc).  With one label vector the decoded list itself is the stash.
template <bool ADD, int KB, bool PACK = false>
__device__ __forceinline__ void cms_rounds(const StepArgs &a, const uint32_t *__restrict__ ids, uint32_t n,
                                           uint32_t k0, uint32_t kb, uint32_t *sk, uint32_t cells,
                                           uint64_t (&best)[KB], uint32_t *stash = nullptr, uint32_t qmask = ~0u) {
    // neighbours per lane and round: fewer when a row carries 8 labels
    constexpr int R = KB > 4 ? 1 : CMS_R;
    // the row hashes stay in scalar registers for the whole pass (a runtime
    // row index into the argument block would reload them per label)
    constexpr int MAXD = 8;
    uint64_t ha[MAXD], hb[MAXD];
#pragma unroll
    for (int r = 0; r < MAXD; r++) {
        ha[r] = a.cms_ab[2 * r];
        hb[r] = a.cms_ab[2 * r + 1];
    }
    const uint32_t depth = a.cms_depth, logw = a.cms_logw, sh = 64 - logw;
    auto col = [&](int r, uint32_t lab) -> uint32_t {
        return logw ? (uint32_t)((ha[r] * (uint64_t)lab + hb[r]) >> sh) : 0u;
    };
    for (uint32_t i0 = 0; i0 < n; i0 += R * HBLOCK) {
        uint32_t labs[R][KB];
#pragma unroll
        for (int t = 0; t < R; t++) {
            const uint32_t i = i0 + t * HBLOCK + threadIdx.x;
            if (i < n) {
                if (stash && !ADD) {
                    if (KB == 8) {
                        const uint4 w0 = reinterpret_cast<const uint4 *>(stash)[2 * i];
                        const uint4 w1 = reinterpret_cast<const uint4 *>(stash)[2 * i + 1];
                        const uint32_t v8[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
                        for (int q = 0; q < KB; q++) labs[t][q] = v8[q & 7];
                    } else if (KB == 4) {
                        const uint4 w = reinterpret_cast<const uint4 *>(stash)[i];
                        labs[t][0] = w.x;
                        labs[t][KB > 1 ? 1 : 0] = w.y;
                        labs[t][KB > 2 ? 2 : 0] = w.z;
                        labs[t][KB > 3 ? 3 : 0] = w.w;
                    } else {
#pragma unroll
                        for (int q = 0; q < KB; q++) labs[t][q] = stash[(uint64_t)i * KB + q];
                    }
                } else {
                    load_row<KB>(a, ids[i], k0, kb, labs[t]);
                    if (stash) {
                        if (KB == 8) {
                            uint4 *st = reinterpret_cast<uint4 *>(stash) + 2 * i;
                            st[0] = make_uint4(labs[t][0], labs[t][1 % KB], labs[t][2 % KB], labs[t][3 % KB]);
                            st[1] = make_uint4(labs[t][4 % KB], labs[t][5 % KB], labs[t][6 % KB], labs[t][7 % KB]);
                        } else if (KB == 4) {
                            reinterpret_cast<uint4 *>(stash)[i] =
                                make_uint4(labs[t][0], labs[t][KB > 1 ? 1 : 0], labs[t][KB > 2 ? 2 : 0],
                                           labs[t][KB > 3 ? 3 : 0]);
                        } else {
#pragma unroll
                            for (int q = 0; q < KB; q++) stash[(uint64_t)i * KB + q] = labs[t][q];
                        }
                    }
                }
            } else {
#pragma unroll
                for (int q = 0; q < KB; q++) labs[t][q] = UNL;
            }
        }
#pragma unroll
        for (int q = 0; q < KB; q++) {
            if ((uint32_t)q >= kb || !((qmask >> q) & 1)) continue;
            uint32_t *skq = sk + (uint64_t)(PACK ? q >> 1 : q) * cells;
            const uint32_t half = PACK ? 16u * (q & 1) : 0u;   // a constant once unrolled
#pragma unroll
            for (int t = 0; t < R; t++) {
                const uint32_t lab = labs[t][q];
                if (lab == UNL) continue;
                if (ADD) {
                    const uint32_t cnt = fold_first_label<HBLOCK>(lab);
                    if (cnt == 0) continue;
#pragma unroll
                    for (int r = 0; r < MAXD; r++) {
                        if ((uint32_t)r >= depth) break;
                        atomicAdd(&skq[((uint32_t)r << logw) + col(r, lab)], cnt << half);
                    }
                } else {
                    uint32_t est = UNL;
#pragma unroll
                    for (int r = 0; r < MAXD; r++) {
                        if ((uint32_t)r >= depth) break;
                        const uint32_t x = PACK ? (skq[((uint32_t)r << logw) + col(r, lab)] >> half) & 0xFFFFu
                                                : skq[((uint32_t)r << logw) + col(r, lab)];
                        est = x < est ? x : est;
                    }
                    const uint64_t key = ((uint64_t)est << 32) | (uint32_t)~lab;
                    best[q] = key > best[q] ? key : best[q];
                }
            }
        }
    }
}

// LDS a one-workgroup CMS vertex (or a chunk) needs beside its sketches: per
// vector the compacted stash cursor, the collision bit and the cross-wave
// sums of the packed own-label / labelled counts.
struct CmsLocalShared {
    uint32_t cur[HEAVY_KB_CMS_LOCAL];
    uint32_t coll;
    uint32_t xw[HEAVY_KB_CMS_LOCAL * (HBLOCK / 64)], xs[HEAVY_KB_CMS_LOCAL];   // block_sum_words
};

// The sketch's row hashes h_r(x) = (a_r x + b_r mod 2^64) >> (64 - log2 w),
// held in registers for a whole pass (indexing the kernel-argument block with
// a runtime row would reload them per label).
struct CmsRows {
    static constexpr int MAXD = 8;
    uint64_t ha[MAXD], hb[MAXD];
    uint32_t depth, logw;
    __device__ __forceinline__ explicit CmsRows(const StepArgs &a) : depth(a.cms_depth), logw(a.cms_logw) {
#pragma unroll
        for (int r = 0; r < MAXD; r++) {
            ha[r] = a.cms_ab[2 * r];
            hb[r] = a.cms_ab[2 * r + 1];
        }
    }
    __device__ __forceinline__ uint32_t col(int r, uint32_t lab) const {
        return logw ? (uint32_t)((ha[r] * (uint64_t)lab + hb[r]) >> (64 - logw)) : 0u;
    }
    // estimate of lab in sketch q (16-bit half q & 1 of the packed words of
    // pair q >> 1 with PACK, else the 32-bit cells of sketch q)
    template <bool PACK>
    __device__ __forceinline__ uint32_t estimate(const uint32_t *sk, uint32_t cells, uint32_t q, uint32_t lab) const {
        const uint32_t *skq = sk + (uint64_t)(PACK ? q >> 1 : q) * cells;
        const uint32_t half = PACK ? 16u * (q & 1) : 0u;
        uint32_t est = UNL;
#pragma unroll
        for (int r = 0; r < MAXD; r++) {
            if ((uint32_t)r >= depth) break;
            const uint32_t w = skq[((uint32_t)r << logw) + col(r, lab)];
            const uint32_t x = PACK ? (w >> half) & 0xFFFFu : w;
            est = x < est ? x : est;
        }
        return est;
    }
};

// Own labels of vectors [k0, k0+kb) of vertex v (clean; UNL beyond kb) and
// their row-0 columns.
template <int KB>
__device__ __forceinline__ void cms_own(const StepArgs &a, const CmsRows &H, uint32_t v, uint32_t k0, uint32_t kb,
                                        uint32_t (&own)[KB], uint32_t (&own0)[KB]) {
#pragma unroll
    for (int q = 0; q < KB; q++) {
        own[q] = (uint32_t)q < kb ? clean_label(a.lin[(uint64_t)v * a.K + k0 + q]) : UNL;
        own0[q] = H.col(0, own[q]);
    }
}

// Pass 1 of a count-min step over ids[0..n) (n < 2^16) for vectors [k0,
// k0+kb) (LP-SC rule 4), sketches sk in LDS (PACK: 16-bit cells, two vectors
// to a word; else 32-bit cells).  Every neighbour's label row is gathered
// once.  Per vector, copies of the vertex's own label are only counted (pc,
// packed own / labelled counts of the wave); every other labelled entry is
// added to the sketch (a wave's copies of one label folded into one update)
// and checked for sharing the own label's cell in every row (cs->coll bit
// q).  stash (or null): one entry per folded group is appended to
// stash + q * cap (cursor cs->cur[q]), so the stash holds every distinct
// non-own label present.  inplace (or null, one vector only): every
// gathered label is written over its id.
template <bool PACK, int KB>
__device__ __forceinline__ void cms_add_pass(const StepArgs &a, const CmsRows &H, const uint32_t *ids, uint32_t n,
                                             uint32_t k0, uint32_t kb, const uint32_t (&own)[KB],
                                             const uint32_t (&own0)[KB], uint32_t *sk, uint32_t cells,
                                             CmsLocalShared *cs, uint32_t *stash, uint32_t cap, uint32_t *inplace,
                                             uint32_t (&pc)[KB]) {
    const uint32_t lane = __lane_id();
    const uint64_t lt = (1ull << lane) - 1;   // lanes below this one
#pragma unroll
    for (int q = 0; q < KB; q++) pc[q] = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += HBLOCK) {
        const uint32_t i = i0 + threadIdx.x;
        uint32_t labs[KB];
        if (i < n) {
            load_row<KB>(a, ids[i], k0, kb, labs);
            if (KB == 1 && inplace) inplace[i] = labs[0];
        } else {
#pragma unroll
            for (int q = 0; q < KB; q++) labs[q] = UNL;
        }
#pragma unroll
        for (int q = 0; q < KB; q++) {
            if ((uint32_t)q >= kb) continue;
            const uint32_t lab = labs[q];
            const bool is_own = own[q] != UNL && lab == own[q];
            pc[q] += ((uint32_t)__popcll(__ballot(is_own)) << 16) + (uint32_t)__popcll(__ballot(lab != UNL));
            if (lab == UNL || is_own) continue;
            const uint32_t cnt = fold_first_label<HBLOCK>(lab);
            if (cnt == 0) continue;
            uint32_t *skq = sk + (uint64_t)(PACK ? q >> 1 : q) * cells;
            const uint32_t half = PACK ? 16u * (q & 1) : 0u;
            // full: lab shares the own label's cell in every row so far
            bool full = own[q] != UNL;
#pragma unroll
            for (int r = 0; r < CmsRows::MAXD; r++) {
                if ((uint32_t)r >= H.depth) break;
                const uint32_t c = H.col(r, lab);
                atomicAdd(&skq[((uint32_t)r << H.logw) + c], cnt << half);
                if (full) full = c == (r == 0 ? own0[q] : H.col(r, own[q]));
            }
            if (full) atomicOr(&cs->coll, 1u << q);
            if (stash) {
                // one stash entry per folded group, compacted per wave
                const uint64_t m = __ballot(1);
                const uint32_t leader = (uint32_t)__builtin_ctzll(m);
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(&cs->cur[q], (uint32_t)__popcll(m));
                base = (uint32_t)__shfl((int)base, (int)leader, 64);
                stash[(uint64_t)q * cap + base + (uint32_t)__popcll(m & lt)] = lab;
            }
        }
    }
}

// Workgroup totals of the packed counts (every thread gets them), then the
// own label's copies enter the sketch as one update per vector.
template <bool PACK, int KB>
__device__ __forceinline__ void cms_own_totals(const CmsRows &H, uint32_t kb, const uint32_t (&own)[KB],
                                               uint32_t *sk, uint32_t cells, CmsLocalShared *cs, uint32_t (&pc)[KB]) {
    block_sum_words<KB>(pc, cs->xw, cs->xs);
    if (threadIdx.x < kb) {
#pragma unroll
        for (int q = 0; q < KB; q++) {
            const uint32_t oc = own_count(pc[q]);
            if ((uint32_t)q != threadIdx.x || own[q] == UNL || oc == 0) continue;
            uint32_t *skq = sk + (uint64_t)(PACK ? q >> 1 : q) * cells;
            const uint32_t half = PACK ? 16u * (q & 1) : 0u;
#pragma unroll
            for (int r = 0; r < CmsRows::MAXD; r++) {
                if ((uint32_t)r >= H.depth) break;
                atomicAdd(&skq[((uint32_t)r << H.logw) + H.col(r, own[q])], oc << half);
            }
        }
    }
    __syncthreads();
}

// A vector is settled by its own label when that label is a strict majority
// of the labelled neighbours and no present label shares its cell in every
// row: any other label's estimate is then bounded by a cell holding none of
// the own label's copies, so by fewer than half of the labelled neighbours,
// while the own label's estimate is at least its count.
__device__ __forceinline__ bool cms_settled(uint32_t own, uint32_t packed, bool coll) {
    return own_majority_of(own, packed) && !coll;
}

// The same test from a chunked hub's counters (own count, labelled count,
// full-collision flag; 32-bit words, a hub's list may exceed 2^16).
__device__ __forceinline__ bool hub_settled(uint32_t own, const uint32_t *hc) {
    return own != UNL && 2ull * hc[0] > hc[1] && hc[2] == 0;
}

// PASS 1: sketch increments, 2: sketch queries, 0: exact counting, for the
// label vectors [k0, k0+kb) of one chunk.  Passes 1 and 0 first aggregate
// the chunk in LDS (chunk-local sketches, or a chunk-local hash table of
// 2*CHUNK slots) and then merge only the non-zero cells / occupied slots
// into the vertex's HBM sketch or table, so a hub whose neighbours share a
// few labels does not funnel every arc into the same HBM atomic.  Dynamic
// LDS: kbs local sketches (pass 1) or the local table (pass 0).
template <int PASS, int KB>
__global__ void __launch_bounds__(HBLOCK) k_heavy_chunks(HeavyArgs h) {
    extern __shared__ uint64_t lds[];
    __shared__ uint64_t red[HBLOCK / 64];
    __shared__ CmsLocalShared cs[1];
    const uint32_t c = blockIdx.x;
    if (c >= h.nchunks) return;
    const uint32_t i = h.chunk_h[c];
    const uint32_t d = heavy_deg(h, i);
    const uint64_t b = h.chunk_b[c];
    const uint32_t n = (uint32_t)((d - b) < CHUNK ? (d - b) : CHUNK);
    const uint32_t *ids = heavy_ids(h, i) + b;
    const StepArgs &a = h.s;
    const uint64_t t0 = a.rec ? cycles() : 0;
    const uint32_t sl = h.slot[i];
    const uint32_t v = h.csr ? a.cvid[h.pos[i]] : a.vid[h.pos[i]];
    if (PASS == 0) {
        const uint32_t gen = h.tagged ? table_gen(h) : 0;
        const uint64_t T = h.toff[i + 1] - h.toff[i];
        const uint32_t logT = 63 - __builtin_clzll(T);
        constexpr uint32_t LT = 2 * CHUNK, LOG_LT = 31 - __builtin_clz(LT);
        uint64_t *ltab = lds;
        for (uint32_t kq = 0; kq < h.kb; kq += KB) {
            const uint32_t kb = h.kb - kq < (uint32_t)KB ? h.kb - kq : (uint32_t)KB;
            uint64_t best = 0;
            // a chunk holds part of the hub's list: no own-label shortcut
            uint32_t no_own[KB];
#pragma unroll
            for (int s = 0; s < KB; s++) no_own[s] = UNL;
            count_staged<HBLOCK, HEAVY_IPL, KB>(
                a, ids, n, threadIdx.x, h.k0 + kq, kb, no_own, false, nullptr, nullptr,
                [&](int) {
                    for (uint32_t x = threadIdx.x; x < LT; x += HBLOCK) ltab[x] = EMPTY_SLOT;
                    best = 0;
                    __syncthreads();
                },
                [&](int, uint32_t lab, uint32_t cnt) { (void)tab_add(ltab, LOG_LT, lab, cnt); },
                [&](int, uint32_t lab) {
                    const uint32_t cnt = fold_first_label<HBLOCK>(lab);
                    if (cnt) (void)tab_add(ltab, LOG_LT, lab, cnt);
                },
                [&](int s) {
                    __syncthreads();
                    uint64_t *tab = h.tab + h.toff[i] * h.kb + (uint64_t)(kq + s) * T;
                    for (uint32_t x = threadIdx.x; x < LT; x += HBLOCK) {
                        const uint64_t e = ltab[x];
                        if ((uint32_t)(e >> 32) == UNL) continue;
                        const uint32_t lab = (uint32_t)(e >> 32);
                        const uint32_t cc = h.tagged ? tab_add_tagged(tab, logT, lab, (uint32_t)e, gen)
                                                     : tab_add(tab, logT, lab, (uint32_t)e);
                        const uint64_t key = ((uint64_t)cc << 32) | (uint32_t)~lab;
                        best = key > best ? key : best;
                    }
                    best = block_max(best, red);
                    if (threadIdx.x == 0 && best)
                        atomicMax(&h.result[(uint64_t)sl * h.kb + kq + s], (unsigned long long)best);
                },
                [&](int, uint64_t) {});
        }
    } else {
        // several vectors: the chunk-local sketches take 16-bit cells, two
        // vectors to a word (a chunk of at most CHUNK ids cannot overflow
        // one); the hub's HBM sketches keep 32-bit cells
        constexpr bool PACK = KB > 1;
        const uint32_t cells = a.cms_depth << a.cms_logw;
        for (uint32_t kq = 0; kq < h.kb; kq += h.kbs) {
            const uint32_t kb = h.kb - kq < h.kbs ? h.kb - kq : h.kbs;
            uint32_t *hsk = h.sketch + ((uint64_t)sl * h.kb + kq) * cells;
            uint64_t best[KB] = {};
            if (PASS == 1) {
                // own-label counts go to the hub's counters; the other
                // labels, and the chunk's own copies as one update, to the
                // chunk-local sketches, merged into the hub's
                uint32_t *lsk = (uint32_t *)lds;
                const uint32_t words = (PACK ? (kb + 1) / 2 : kb) * cells;
                for (uint32_t x = threadIdx.x; x < words; x += HBLOCK) lsk[x] = 0;
                if (threadIdx.x == 0) cs->coll = 0;
                __syncthreads();
                const CmsRows H(a);
                uint32_t own[KB], own0[KB], pc[KB];
                cms_own<KB>(a, H, v, h.k0 + kq, kb, own, own0);
                cms_add_pass<PACK, KB>(a, H, ids, n, h.k0 + kq, kb, own, own0, lsk, cells, cs, nullptr, 0,
                                       KB == 1 && a.K == 1 && !h.csr ? const_cast<uint32_t *>(ids) : nullptr, pc);
                cms_own_totals<PACK, KB>(H, kb, own, lsk, cells, cs, pc);
                for (uint32_t x = threadIdx.x; x < words; x += HBLOCK) {
                    const uint32_t w = lsk[x];
                    if (!w) continue;
                    if (PACK) {
                        const uint32_t pair = x / cells, c = x - pair * cells;
                        if (w & 0xFFFFu) atomicAdd(&hsk[(uint64_t)(2 * pair) * cells + c], w & 0xFFFFu);
                        if (w >> 16) atomicAdd(&hsk[(uint64_t)(2 * pair + 1) * cells + c], w >> 16);
                    } else {
                        atomicAdd(&hsk[x], w);
                    }
                }
                const uint32_t coll = cs->coll;
#pragma unroll
                for (int q = 0; q < KB; q++) {
                    if ((uint32_t)q >= kb || threadIdx.x != (uint32_t)q) continue;
                    uint32_t *hc = h.hcnt + ((uint64_t)sl * h.kb + kq + q) * 4;
                    if (own_count(pc[q])) atomicAdd(&hc[0], own_count(pc[q]));
                    if (lab_count(pc[q])) atomicAdd(&hc[1], lab_count(pc[q]));
                    if ((coll >> q) & 1) atomicOr(&hc[2], 1u);
                }
                __syncthreads();
            } else {
                // vectors settled by the hub's own label need no estimates
                uint32_t undecided = 0;
#pragma unroll
                for (int q = 0; q < KB; q++) {
                    if ((uint32_t)q >= kb) continue;
                    const uint32_t *hc = h.hcnt + ((uint64_t)sl * h.kb + kq + q) * 4;
                    const uint32_t own = clean_label(a.lin[(uint64_t)v * a.K + h.k0 + kq + q]);
                    if (!hub_settled(own, hc)) undecided |= 1u << q;
                }
                if (!undecided) continue;
                cms_rounds<false, KB>(a, ids, n, h.k0 + kq, kb, hsk, cells, best,
                                      KB == 1 && a.K == 1 && !h.csr ? const_cast<uint32_t *>(ids) : nullptr, undecided);
#pragma unroll
                for (int q = 0; q < KB; q++) {
                    if ((uint32_t)q >= kb || !((undecided >> q) & 1)) continue;
                    const uint64_t m = block_max(best[q], red);
                    if (threadIdx.x == 0 && m)
                        atomicMax(&h.result[(uint64_t)sl * h.kb + kq + q], (unsigned long long)m);
                }
            }
        }
    }
    if (a.rec && threadIdx.x == 0) {
        const uint32_t p = h.pos[i];
        atomicAdd(&a.rec[h.csr ? a.cvid[p] : a.vid[p]], (cycles() - t0) * HBLOCK);
    }
}

// One CMS vertex (neighbours ids[0..d)), label vectors [k0, k0+kb), by one
// workgroup with kb sketches in LDS: pass 1 (cms_add_pass with the
// compacted stash); vectors settled by their own label take it; for the
// others, pass 2 estimates the own label and the stashed entries (every
// distinct label present) and keeps the largest estimate, ties to the
// smallest label.  result(q, key) is called by thread 0 with the vector's
// (count << 32 | ~label).
template <bool PACK, int KB, class Result>
__device__ __forceinline__ void cms_local_own(const StepArgs &a, const uint32_t *__restrict__ ids, uint32_t d,
                                              uint32_t v, uint32_t k0, uint32_t kb, uint32_t *sk, uint32_t cells,
                                              uint32_t *stash, uint32_t cap, CmsLocalShared *cs, uint64_t *red,
                                              Result &&result) {
    const CmsRows H(a);
    uint32_t own[KB], own0[KB], pc[KB];
    cms_own<KB>(a, H, v, k0, kb, own, own0);
    const uint32_t words = (PACK ? (kb + 1) / 2 : kb) * cells;
    for (uint32_t x = threadIdx.x; x < words; x += HBLOCK) sk[x] = 0;
    if (threadIdx.x < (uint32_t)KB) cs->cur[threadIdx.x] = 0;
    if (threadIdx.x == 0) cs->coll = 0;
    __syncthreads();
    cms_add_pass<PACK, KB>(a, H, ids, d, k0, kb, own, own0, sk, cells, cs, stash, cap, nullptr, pc);
    cms_own_totals<PACK, KB>(H, kb, own, sk, cells, cs, pc);
    const uint32_t coll = cs->coll;
#pragma unroll
    for (int q = 0; q < KB; q++) {
        if ((uint32_t)q >= kb) continue;
        if (cms_settled(own[q], pc[q], (coll >> q) & 1)) {
            if (threadIdx.x == 0) result(q, ((uint64_t)own_count(pc[q]) << 32) | (uint32_t)~own[q]);
            continue;
        }
        uint64_t best = 0;
        if (threadIdx.x == 0 && own[q] != UNL && own_count(pc[q]))
            best = ((uint64_t)H.estimate<PACK>(sk, cells, q, own[q]) << 32) | (uint32_t)~own[q];
        const uint32_t n = cs->cur[q];
        for (uint32_t j = threadIdx.x; j < n; j += HBLOCK) {
            const uint32_t lab = stash[(uint64_t)q * cap + j];
            const uint64_t key = ((uint64_t)H.estimate<PACK>(sk, cells, q, lab) << 32) | (uint32_t)~lab;
            best = key > best ? key : best;
        }
        best = block_max(best, red);
        if (threadIdx.x == 0) result(q, best);
    }
    __syncthreads();   // the sketches, cursors and counts are reused next
}

// Row stash of an exact one-workgroup vertex: the clean labels of all K
// vectors of neighbours ids[0..n) in groups of four vectors (vector k of
// neighbour i at st[((k/4)*n + i)*4 + k%4]; ceil(K/4)*4*n words).  Each
// neighbour's label row is read whole, once (its two 128-byte lines at
// K = 64), instead of once per batch of KB vectors, where every batch
// fetches the lines again for a KB-label slice; the batches then read the
// stash with coalesced 16-byte loads.  Lane i % HBLOCK writes neighbour i,
// the lane that count_staged reads it in.
__device__ __forceinline__ void stash_rows(const StepArgs &a, const uint32_t *__restrict__ ids, uint32_t n,
                                           uint32_t *__restrict__ st) {
    const uint32_t K = a.K;
    uint4 *st4 = reinterpret_cast<uint4 *>(st);
    for (uint32_t i = threadIdx.x; i < n; i += HBLOCK) {
        const uint32_t *src = a.lin + (uint64_t)ids[i] * K;
        if ((K & 15) == 0) {
            for (uint32_t k = 0; k < K; k += 16) {
                uint4 w[4];
#pragma unroll
                for (int q = 0; q < 4; q++) w[q] = reinterpret_cast<const uint4 *>(src + k)[q];
#pragma unroll
                for (int q = 0; q < 4; q++)
                    st4[(uint64_t)(k / 4 + q) * n + i] = make_uint4(clean_label(w[q].x), clean_label(w[q].y),
                                                                    clean_label(w[q].z), clean_label(w[q].w));
            }
        } else {
            for (uint32_t k = 0; k < K; k++) st[((uint64_t)(k >> 2) * n + i) * 4 + (k & 3)] = clean_label(src[k]);
        }
    }
}

template <int CMS, int KB, int IPL, bool ST>
__device__ __forceinline__ void heavy_local_vertex(const HeavyArgs &h, uint32_t i, uint64_t *lds, uint64_t *red,
                                                   uint32_t *olds, uint32_t *xw, uint32_t *xs, CmsLocalShared *cs) {
    static_assert(KB <= HEAVY_KB_MAX, "olds[] holds one batch");
    const uint32_t d = heavy_deg(h, i);
    const uint32_t *ids = heavy_ids(h, i);
    const StepArgs &a = h.s;
    const uint64_t t0 = a.rec ? cycles() : 0;
    const uint32_t p = h.pos[i];
    const uint32_t v = h.csr ? a.cvid[p] : a.vid[p];
    uint32_t nch = 0;
    // a batch's current labels are read into olds[] (LDS) by its first kb
    // lanes before the counting, so picking a vector's label costs thread 0
    // no memory round trip
    auto prefetch = [&](uint32_t k0, uint32_t kb) {
        if (threadIdx.x < kb) olds[threadIdx.x] = a.lin[(uint64_t)v * a.K + k0 + threadIdx.x];
    };
    auto finish = [&](uint32_t k0, uint32_t q, uint64_t best) {   // thread 0
        const uint64_t idx = (uint64_t)v * a.K + k0 + q;
        const uint32_t old = olds[q];
        const uint32_t nl = pick(best, old);
        a.lout[idx] = nl;
        nch += nl != old;
    };
    if (CMS) {
        // several vectors: 16-bit cells, two vectors to a word
        constexpr bool PACK = KB > 1;
        const uint32_t cells = a.cms_depth << a.cms_logw;
        uint32_t *stash = h.stash + (uint64_t)blockIdx.x * h.stash_stride;
        for (uint32_t k0 = 0; k0 < a.K; k0 += h.kbl) {
            const uint32_t kb = a.K - k0 < h.kbl ? a.K - k0 : h.kbl;
            prefetch(k0, kb);
            cms_local_own<PACK, KB>(a, ids, d, v, k0, kb, (uint32_t *)lds, cells, stash, CMS_LOCAL, cs, red,
                                    [&](uint32_t q, uint64_t key) { finish(k0, q, key); });
        }
    } else {
        const uint32_t T = d ? 1u << (32 - __builtin_clz(2 * d - 1)) : 1u;
        const uint32_t logT = 31 - __builtin_clz(T);
        uint64_t *tab = lds;
        // with a row stash (many vectors) every row is read once for all K
        uint32_t *st = ST ? h.stash + (uint64_t)blockIdx.x * h.stash_stride : nullptr;
        if (ST) {
            __syncthreads();   // the previous vertex's reads of the region are done
            stash_rows(a, ids, d, st);
        }
        for (uint32_t k0 = 0; k0 < a.K; k0 += KB) {
            const uint32_t kb = a.K - k0 < (uint32_t)KB ? a.K - k0 : (uint32_t)KB;
            uint64_t best = 0;
            prefetch(k0, kb);
            uint32_t own[KB];
#pragma unroll
            for (int s = 0; s < KB; s++)
                own[s] = (uint32_t)s < kb ? clean_label(a.lin[(uint64_t)v * a.K + k0 + s]) : UNL;
            count_staged<HBLOCK, IPL, KB>(
                a, ids, d, threadIdx.x, k0, kb, own, true, xw, xs,
                [&](int) {
                    for (uint32_t x = threadIdx.x; x < T; x += HBLOCK) tab[x] = EMPTY_SLOT;
                    best = 0;
                    __syncthreads();
                },
                [&](int, uint32_t lab, uint32_t cnt) {
                    const uint32_t c = tab_add(tab, logT, lab, cnt);
                    const uint64_t key = ((uint64_t)c << 32) | (uint32_t)~lab;
                    best = key > best ? key : best;
                },
                [&](int, uint32_t lab) {
                    const uint32_t cnt = fold_first_label<HBLOCK>(lab);
                    if (cnt == 0) return;
                    const uint32_t c = tab_add(tab, logT, lab, cnt);
                    const uint64_t key = ((uint64_t)c << 32) | (uint32_t)~lab;
                    best = key > best ? key : best;
                },
                [&](int s) {
                    const uint64_t m = block_max(best, red);
                    if (threadIdx.x == 0) finish(k0, s, m);
                },
                [&](int s, uint64_t key) {
                    if (threadIdx.x == 0) finish(k0, s, key);
                },
                ST ? st + (uint64_t)k0 * d : nullptr, d);   // k0 is a multiple of 4
        }
    }
    if (threadIdx.x == 0 && nch) atomicAdd(a.changed, (unsigned long long)nch);
    if (a.rec && threadIdx.x == 0) atomicAdd(&a.rec[v], (cycles() - t0) * HBLOCK);
}

// Heavy vertices small enough for one workgroup's LDS (exact: degree <=
// 2*CHUNK, table of 2^ceil(log2 2d) slots -- up to 128 KB; CMS: any degree
// up to the local limit, kbs sketches in LDS): one workgroup per vertex
// counts its whole list -- both CMS passes, or the exact table -- and writes
// the labels itself, with no HBM table, sketch or merge.  local[] lists the
// heavy-vertex indices; IPL = staged ids per lane (exact).
// (Eight waves per SIMD requested -- two 1024-lane workgroups per CU -- when
// the tables leave LDS for two: IPL = 4; the 128 KB tables of IPL = 8 admit
// one workgroup per CU, so those instances take four waves per SIMD and the
// registers that frees.)  ST (exact mode, IPL = 4): with a row stash, the
// grid is the workgroups the CUs hold at once, each taking vertices one by
// one into its own stash region (h.stash); without one, a workgroup per
// vertex.
template <int CMS, int KB, int IPL, bool ST = false>
__global__ void __launch_bounds__(HBLOCK, IPL == 8 || (CMS && KB > 1) ? 4 : 8)
    k_heavy_local(HeavyArgs h, const uint32_t *__restrict__ local, uint32_t nlocal) {
    static_assert(!ST || (!CMS && IPL * HBLOCK == CHUNK), "row stash regions hold CHUNK neighbours");
    extern __shared__ uint64_t lds[];
    __shared__ uint64_t red[HBLOCK / 64];
    __shared__ uint32_t olds[HEAVY_KB_MAX];
    __shared__ uint32_t xw[CMS ? 1 : HEAVY_KB_MAX * (HBLOCK / 64)], xs[HEAVY_KB_MAX];   // block_sum_words
    __shared__ CmsLocalShared cs[1];
    if (ST) {
        // workgroups take the next vertex from the work counter h.wq[0]
        // (vertices of very different degrees: a fixed share per workgroup
        // leaves some with far more work); the last workgroup to leave
        // resets the counters for the next launch, when no workgroup of
        // this one reads them any more
        __shared__ uint32_t next;
        for (;;) {
            if (threadIdx.x == 0) next = atomicAdd(&h.wq[0], 1u);
            __syncthreads();
            const uint32_t j = next;
            if (j >= nlocal) break;
            // the vertex synchronises the workgroup before thread 0 writes
            // `next` again
            heavy_local_vertex<CMS, KB, IPL, ST>(h, local[j], lds, red, olds, xw, xs, cs);
        }
        if (threadIdx.x == 0 && atomicAdd(&h.wq[1], 1u) == gridDim.x - 1) {
            atomicExch(&h.wq[0], 0u);
            atomicExch(&h.wq[1], 0u);
        }
    } else if (CMS) {
        // workgroups loop over the vertices, each with its stash region
        for (uint32_t j = blockIdx.x; j < nlocal; j += gridDim.x)
            heavy_local_vertex<CMS, KB, IPL, ST>(h, local[j], lds, red, olds, xw, xs, cs);
    } else if (blockIdx.x < nlocal) {
        heavy_local_vertex<CMS, KB, IPL, ST>(h, local[blockIdx.x], lds, red, olds, xw, xs, cs);
    }
}

template <int CMS>
__global__ void __launch_bounds__(BLOCK) k_heavy_finish(HeavyArgs h, const uint32_t *__restrict__ chunked,
                                                        uint32_t nchunked) {
    const uint64_t t0 = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
    bool ch = false;
    if (t0 < (uint64_t)nchunked * h.kb) {
        const uint32_t i = chunked[t0 / h.kb], k = h.k0 + (uint32_t)(t0 % h.kb);
        const uint64_t t = (uint64_t)h.slot[i] * h.kb + (t0 % h.kb);
        const uint32_t p = h.pos[i];
        const uint32_t v = h.csr ? h.s.cvid[p] : h.s.vid[p];
        const uint64_t idx = (uint64_t)v * h.s.K + k;
        const uint32_t old = h.s.lin[idx];
        uint64_t key = h.result[t];
        if (CMS) {
            // settled by the hub's own label (the chunk passes skipped it)
            const uint32_t *hc = h.hcnt + t * 4, own = clean_label(old);
            if (hub_settled(own, hc)) key = ((uint64_t)hc[0] << 32) | (uint32_t)~own;
        }
        const uint32_t nl = pick(key, old);
        h.s.lout[idx] = nl;
        ch = nl != old;
    }
    count_changes(ch, h.s.changed);
    // the pass is complete: the next one counts under the next generation
    if (!CMS && h.tagged && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(h.epoch, 1u);
}

}  // namespace lpa
