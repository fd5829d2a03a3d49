// heavy_kernels.hpp -- the heaviest vertices (count-min-sketch vertices and
// vertices whose hash table does not fit LDS), spread over many workgroups.
//
// One workgroup per vertex cannot keep enough label gathers in flight for a
// hub (R-MAT scale 22 has a vertex of degree 163,396; scale 27 hubs reach
// millions), so a step over these "heavy" vertices runs as
//   k_heavy_decode  one workgroup per heavy vertex: decode the record into a
//                   neighbour list in HBM (long runs written by every lane),
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
// CMS vertices up to this degree are counted by one workgroup (LDS sketch).
constexpr uint32_t CMS_LOCAL = 65536;

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
    const uint32_t *__restrict__ slot;         // chunked vertex i -> its sketch / result slot
};

__device__ __forceinline__ const uint32_t *heavy_ids(const HeavyArgs &h, uint32_t i) {
    return h.csr ? h.s.nbr + h.s.row[h.pos[i]] : h.list + h.loff[i];
}

__device__ __forceinline__ uint32_t heavy_deg(const HeavyArgs &h, uint32_t i) {
    return (uint32_t)(h.loff[i + 1] - h.loff[i]);
}

__global__ void __launch_bounds__(BLOCK) k_heavy_decode(HeavyArgs h) {
    __shared__ GroupShared<BLOCK> sh;
    const uint32_t i = blockIdx.x;
    if (i >= h.nh || h.csr) return;
    const uint64_t t0 = h.s.rec ? cycles() : 0;
    const uint32_t p = h.pos[i];
    group_decode<BLOCK>(h.s, p, h.s.vid[p], threadIdx.x, &sh, h.list + h.loff[i]);
    if (h.s.rec && threadIdx.x == 0) atomicAdd(&h.s.rec[h.s.vid[p]], (cycles() - t0) * BLOCK);
}

// chunked[] lists the heavy-vertex indices handled by the chunk kernels.
__global__ void __launch_bounds__(BLOCK) k_heavy_clear(HeavyArgs h, const uint32_t *__restrict__ chunked,
                                                       uint32_t nchunked) {
    if (blockIdx.x >= nchunked) return;
    const uint32_t i = chunked[blockIdx.x], sl = h.slot[i];
    for (uint32_t k = threadIdx.x; k < h.kb; k += BLOCK) h.result[(uint64_t)sl * h.kb + k] = 0;
    if (h.cms) {
        const uint64_t cells = ((uint64_t)h.s.cms_depth << h.s.cms_logw) * h.kb;
        uint32_t *sk = h.sketch + (uint64_t)sl * cells;
        for (uint64_t c = threadIdx.x; c < cells; c += BLOCK) sk[c] = 0;
    } else {
        for (uint64_t c = h.toff[i] * h.kb + threadIdx.x; c < h.toff[i + 1] * h.kb; c += BLOCK) h.tab[c] = EMPTY_SLOT;
    }
}

// PASS 1: sketch increments, 2: sketch queries, 0: exact counting.
// Passes 1 and 0 first aggregate the chunk in LDS (a chunk-local sketch, or
// a chunk-local hash table of 2*CHUNK slots) and then merge only the
// non-zero cells / occupied slots into the vertex's HBM sketch or table, so
// a hub whose neighbours share a few labels does not funnel every arc into
// the same HBM atomic.  Dynamic LDS: the local sketch or table.
template <int PASS>
__global__ void __launch_bounds__(BLOCK) k_heavy_chunks(HeavyArgs h) {
    extern __shared__ uint64_t lds[];
    __shared__ uint64_t red[BLOCK / 64];
    const uint32_t c = blockIdx.x;
    if (c >= h.nchunks) return;
    const uint32_t i = h.chunk_h[c];
    const uint32_t d = heavy_deg(h, i);
    const uint64_t b = h.chunk_b[c];
    const uint32_t n = (uint32_t)((d - b) < CHUNK ? (d - b) : CHUNK);
    const uint32_t *ids = heavy_ids(h, i) + b;
    const StepArgs &a = h.s;
    const uint64_t T = h.cms ? 0 : h.toff[i + 1] - h.toff[i];
    const uint32_t logT = h.cms ? 0 : 63 - __builtin_clzll(T);
    const uint32_t cells = a.cms_depth << a.cms_logw;
    constexpr uint32_t LT = 2 * CHUNK, LOG_LT = 31 - __builtin_clz(LT);
    uint32_t *lsk = (uint32_t *)lds;   // PASS 1
    uint64_t *ltab = lds;              // PASS 0
    const uint64_t t0 = a.rec ? cycles() : 0;
    for (uint32_t kk = 0; kk < h.kb; kk++) {
        const uint32_t k = h.k0 + kk;
        uint64_t best = 0;
        uint32_t *sk = h.cms ? h.sketch + ((uint64_t)h.slot[i] * h.kb + kk) * cells : nullptr;
        uint64_t *tab = h.cms ? nullptr : h.tab + h.toff[i] * h.kb + kk * T;
        if (PASS == 1) {
            for (uint32_t x = threadIdx.x; x < cells; x += BLOCK) lsk[x] = 0;
        } else if (PASS == 0) {
            for (uint32_t x = threadIdx.x; x < LT; x += BLOCK) ltab[x] = EMPTY_SLOT;
        }
        __syncthreads();
        group_count<BLOCK>(a, ids, n, threadIdx.x, k, [&](uint32_t lab) {
            if (PASS == 2) {
                uint32_t est = UNL;
                for (uint32_t r = 0; r < a.cms_depth; r++) {
                    uint32_t x = sk[(r << a.cms_logw) + cms_col(a, r, lab)];
                    est = x < est ? x : est;
                }
                uint64_t key = ((uint64_t)est << 32) | (uint32_t)~lab;
                best = key > best ? key : best;
                return;
            }
            const uint32_t cnt = fold_first_label<BLOCK>(lab);
            if (cnt == 0) return;
            if (PASS == 1) {
                for (uint32_t r = 0; r < a.cms_depth; r++)
                    atomicAdd(&lsk[(r << a.cms_logw) + cms_col(a, r, lab)], cnt);
            } else {
                (void)tab_add(ltab, LOG_LT, lab, cnt);
            }
        });
        __syncthreads();
        if (PASS == 1) {
            for (uint32_t x = threadIdx.x; x < cells; x += BLOCK)
                if (lsk[x]) atomicAdd(&sk[x], lsk[x]);
        } else if (PASS == 0) {
            for (uint32_t x = threadIdx.x; x < LT; x += BLOCK) {
                const uint64_t e = ltab[x];
                if ((uint32_t)(e >> 32) == UNL) continue;
                const uint32_t lab = (uint32_t)(e >> 32);
                const uint32_t cc = tab_add(tab, logT, lab, (uint32_t)e);
                uint64_t key = ((uint64_t)cc << 32) | (uint32_t)~lab;
                best = key > best ? key : best;
            }
        }
        if (PASS != 1) {
            best = max_u64_over<64>(best);
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
            __syncthreads();
            if (threadIdx.x == 0) {
                for (int w = 0; w < BLOCK / 64; w++) best = red[w] > best ? red[w] : best;
                if (best) atomicMax(&h.result[(uint64_t)h.slot[i] * h.kb + kk], (unsigned long long)best);
            }
        }
        __syncthreads();
    }
    if (a.rec && threadIdx.x == 0) {
        const uint32_t p = h.pos[i];
        atomicAdd(&a.rec[h.csr ? a.cvid[p] : a.vid[p]], (cycles() - t0) * BLOCK);
    }
}

// Heavy vertices small enough for one workgroup's LDS (exact: degree <=
// CHUNK, table of 2^ceil(log2 2d) slots; CMS: any degree up to the local
// limit, sketch in LDS): one workgroup per vertex counts its whole list --
// both CMS passes, or the exact table -- and writes the label itself, with
// no HBM table, sketch or merge.  local[] lists the heavy-vertex indices.
template <int CMS>
__global__ void __launch_bounds__(BLOCK) k_heavy_local(HeavyArgs h, const uint32_t *__restrict__ local,
                                                       uint32_t nlocal) {
    extern __shared__ uint64_t lds[];
    __shared__ uint64_t red[BLOCK / 64];
    if (blockIdx.x >= nlocal) return;
    const uint32_t i = local[blockIdx.x];
    const uint32_t d = heavy_deg(h, i);
    const uint32_t *ids = heavy_ids(h, i);
    const StepArgs &a = h.s;
    const uint64_t t0 = a.rec ? cycles() : 0;
    const uint32_t p = h.pos[i];
    const uint32_t v = h.csr ? a.cvid[p] : a.vid[p];
    const uint32_t cells = a.cms_depth << a.cms_logw;
    const uint32_t T = d ? 1u << (32 - __builtin_clz(2 * d - 1)) : 1u;
    const uint32_t logT = 31 - __builtin_clz(T);
    uint32_t *sk = (uint32_t *)lds;
    uint64_t *tab = lds;
    for (uint32_t k = 0; k < a.K; k++) {
        uint64_t best = 0;
        if (CMS) {
            for (uint32_t x = threadIdx.x; x < cells; x += BLOCK) sk[x] = 0;
            __syncthreads();
            group_count<BLOCK>(a, ids, d, threadIdx.x, k, [&](uint32_t lab) {
                const uint32_t cnt = fold_first_label<BLOCK>(lab);
                if (cnt == 0) return;
                for (uint32_t r = 0; r < a.cms_depth; r++) atomicAdd(&sk[(r << a.cms_logw) + cms_col(a, r, lab)], cnt);
            });
            __syncthreads();
            group_count<BLOCK>(a, ids, d, threadIdx.x, k, [&](uint32_t lab) {
                uint32_t est = UNL;
                for (uint32_t r = 0; r < a.cms_depth; r++) {
                    const uint32_t x = sk[(r << a.cms_logw) + cms_col(a, r, lab)];
                    est = x < est ? x : est;
                }
                const uint64_t key = ((uint64_t)est << 32) | (uint32_t)~lab;
                best = key > best ? key : best;
            });
        } else {
            for (uint32_t x = threadIdx.x; x < T; x += BLOCK) tab[x] = EMPTY_SLOT;
            __syncthreads();
            group_count<BLOCK>(a, ids, d, threadIdx.x, k, [&](uint32_t lab) {
                const uint32_t cnt = fold_first_label<BLOCK>(lab);
                if (cnt == 0) return;
                const uint32_t c = tab_add(tab, logT, lab, cnt);
                const uint64_t key = ((uint64_t)c << 32) | (uint32_t)~lab;
                best = key > best ? key : best;
            });
        }
        best = max_u64_over<64>(best);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 0; w < BLOCK / 64; w++) best = red[w] > best ? red[w] : best;
            const uint64_t idx = (uint64_t)v * a.K + k;
            const uint32_t old = a.lin[idx];
            const uint32_t nl = pick(best, old);
            a.lout[idx] = nl;
            if (nl != old) atomicAdd(a.changed, 1ull);
        }
        __syncthreads();
    }
    if (a.rec && threadIdx.x == 0) atomicAdd(&a.rec[v], (cycles() - t0) * BLOCK);
}

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
        const uint32_t nl = pick(h.result[t], old);
        h.s.lout[idx] = nl;
        ch = nl != old;
    }
    count_changes(ch, h.s.changed);
}

}  // namespace lpa
