// bfs_relabel.hpp -- BFS relabelling of a CSR on the device (gfx950), the device
// counterpart of the host's lpah_csr_bfs_relabel: vertices get new ids in
// the order a sequential queue-based BFS from `root` visits them, taking
// neighbours in ascending id order, and unreached vertices follow in id
// order; every row of the relabelled CSR is sorted.  The result is the
// host function's, bit for bit.
//
// Level-synchronous: a vertex found at level L+1 belongs to the frontier
// entry of smallest position that has it as a neighbour (its first parent,
// an atomic minimum over the frontier), and the new level is ordered by
// (parent position, the parent's adjacency order).  Each pass gives one
// wave to a frontier entry, whose 64 lanes take its neighbours 64 at a time,
// so a hub's row is spread over a wave rather than walked by one lane; the
// parent's children are placed in adjacency order with ballot prefix counts.
// The relabelled rows are sorted with rocPRIM's segmented radix sort
// (preparation, not the label-propagation path).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "encode_kernels.hpp"   // the exclusive-scan kernels
#include "lpa_amd.h"

namespace lpa {

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr int WPB = 4;   // waves per 256-lane workgroup

#define BFS_HIPCHK(x)                                                                                    \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T> struct Buf {
    T *p = nullptr;
    Buf() = default;
    Buf(const Buf &) = delete;
    Buf &operator=(const Buf &) = delete;
    explicit Buf(uint64_t n) { BFS_HIPCHK(hipMalloc((void **)&p, std::max<uint64_t>(n, 1) * sizeof(T))); }
    ~Buf() {
        if (p) (void)hipFree(p);
    }
};

// global wave index and wave count of a launch of 256-lane workgroups
__device__ __forceinline__ uint64_t wave_id() { return ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; }
__device__ __forceinline__ uint64_t wave_count() { return ((uint64_t)gridDim.x * blockDim.x) >> 6; }

__global__ void __launch_bounds__(256) k_bfs_fill(uint32_t *a, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// *bad = 1 when a neighbour id is >= n (checked before any id indexes an array)
__global__ void __launch_bounds__(256) k_bfs_check_ids(const uint32_t *__restrict__ col, uint64_t m, uint64_t n,
                                                       uint32_t *bad) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x)
        if (col[e] >= n) *bad = 1;
}

__global__ void k_bfs_root(uint32_t *order, uint32_t *nid, uint32_t root) {
    order[0] = root;
    nid[root] = 0;
}

// pass 1: every undiscovered neighbour of frontier entry p keeps the
// smallest p that reaches it
__global__ void __launch_bounds__(256) k_bfs_claim(const uint32_t *__restrict__ order, uint64_t lo, uint64_t fsz,
                                                   const uint64_t *__restrict__ row, const uint32_t *__restrict__ col,
                                                   const uint32_t *__restrict__ nid, uint32_t *claim) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint64_t f = wave_id(); f < fsz; f += wave_count()) {
        const uint32_t p = (uint32_t)(lo + f), u = order[p];
        const uint64_t e0 = row[u], e1 = row[u + 1];
        for (uint64_t e = e0 + lane; e < e1; e += 64) {
            const uint32_t w = col[e];
            if (nid[w] == NONE) atomicMin(&claim[w], p);
        }
    }
}

// pass 2: children per frontier entry (neighbours it claimed)
__global__ void __launch_bounds__(256) k_bfs_count(const uint32_t *__restrict__ order, uint64_t lo, uint64_t fsz,
                                                   const uint64_t *__restrict__ row, const uint32_t *__restrict__ col,
                                                   const uint32_t *__restrict__ nid, const uint32_t *__restrict__ claim,
                                                   uint64_t *__restrict__ cnt) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint64_t f = wave_id(); f < fsz; f += wave_count()) {
        const uint32_t p = (uint32_t)(lo + f), u = order[p];
        const uint64_t e0 = row[u], e1 = row[u + 1];
        uint64_t c = 0;
        for (uint64_t e = e0 + lane; e < e1; e += 64) {
            const uint32_t w = col[e];
            c += nid[w] == NONE && claim[w] == p;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if (lane == 0) cnt[f] = c;
    }
}

// pass 3: the children of entry p take positions hi + off[f] ..., in the
// order of p's (ascending) adjacency
__global__ void __launch_bounds__(256) k_bfs_place(uint32_t *order, uint64_t lo, uint64_t fsz, uint64_t hi,
                                                   const uint64_t *__restrict__ row, const uint32_t *__restrict__ col,
                                                   const uint32_t *__restrict__ nid, const uint32_t *__restrict__ claim,
                                                   const uint64_t *__restrict__ off) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1;
    for (uint64_t f = wave_id(); f < fsz; f += wave_count()) {
        const uint32_t p = (uint32_t)(lo + f), u = order[p];
        const uint64_t e0 = row[u], e1 = row[u + 1];
        uint64_t o = hi + off[f];
        for (uint64_t b = e0; b < e1; b += 64) {
            const uint64_t e = b + lane;
            uint32_t w = 0;
            bool take = false;
            if (e < e1) {
                w = col[e];
                take = nid[w] == NONE && claim[w] == p;
            }
            const uint64_t m = __ballot(take);
            if (take) order[o + (uint64_t)__popcll(m & below)] = w;
            o += (uint64_t)__popcll(m);
        }
    }
}

__global__ void __launch_bounds__(256) k_bfs_mark(const uint32_t *__restrict__ order, uint64_t q0, uint64_t q1,
                                                  uint32_t *nid) {
    for (uint64_t q = q0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < q1;
         q += (uint64_t)gridDim.x * blockDim.x)
        nid[order[q]] = (uint32_t)q;
}

// unreached vertices, by id: 1 where nid is NONE
__global__ void __launch_bounds__(256) k_bfs_unreached_flag(const uint32_t *__restrict__ nid, uint64_t n,
                                                            uint64_t *__restrict__ flag) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        flag[v] = nid[v] == NONE;
}

__global__ void __launch_bounds__(256) k_bfs_unreached_place(uint32_t *order, uint32_t *nid, uint64_t n, uint64_t base,
                                                             const uint64_t *__restrict__ rank) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        if (nid[v] == NONE) {
            const uint64_t q = base + rank[v];
            order[q] = (uint32_t)v;
            nid[v] = (uint32_t)q;
        }
}

// degree of new vertex i (its old vertex order[i])
__global__ void __launch_bounds__(256) k_bfs_new_degree(const uint32_t *__restrict__ order, uint64_t n,
                                                        const uint64_t *__restrict__ row, uint64_t *__restrict__ deg) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = order[i];
        deg[i] = row[u + 1] - row[u];
    }
}

// new row i = new ids of the old row order[i] (unsorted); a wave per row
__global__ void __launch_bounds__(256) k_bfs_new_rows(const uint32_t *__restrict__ order, uint64_t n,
                                                      const uint64_t *__restrict__ row, const uint32_t *__restrict__ col,
                                                      const uint32_t *__restrict__ nid,
                                                      const uint64_t *__restrict__ new_row, uint32_t *__restrict__ out) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint64_t i = wave_id(); i < n; i += wave_count()) {
        const uint32_t u = order[i];
        const uint64_t e0 = row[u], d = row[u + 1] - e0, o = new_row[i];
        for (uint64_t j = lane; j < d; j += 64) out[o + j] = nid[col[e0 + j]];
    }
}

// offsets of rows [s0, s1) relative to the first
__global__ void __launch_bounds__(256) k_bfs_rel_offsets(const uint64_t *__restrict__ new_row, uint64_t s0, uint64_t s1,
                                                         uint32_t *__restrict__ rel) {
    for (uint64_t i = s0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= s1;
         i += (uint64_t)gridDim.x * blockDim.x)
        rel[i - s0] = (uint32_t)(new_row[i] - new_row[s0]);
}

unsigned blocks_for(uint64_t items, uint64_t per_block) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((items + per_block - 1) / per_block, 1ull << 20));
}

// exclusive scan of n values in place (out[n] = total; out has n + 1 entries)
void scan_in_place(hipStream_t s, uint64_t *a, uint64_t n) {
    const uint64_t nb = std::max<uint64_t>(1, (n + SCAN_PER_BLOCK - 1) / SCAN_PER_BLOCK);
    Buf<uint64_t> sums(nb);
    hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nb), dim3(256), 0, s, a, n, sums.p);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, s, sums.p, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(256), 0, s, a, n, sums.p, a, nb);
    BFS_HIPCHK(hipGetLastError());
    BFS_HIPCHK(hipStreamSynchronize(s));   // sums is freed on return
}

uint64_t read_u64(hipStream_t s, const uint64_t *p) {
    uint64_t x = 0;
    BFS_HIPCHK(hipMemcpyAsync(&x, p, sizeof x, hipMemcpyDeviceToHost, s));
    BFS_HIPCHK(hipStreamSynchronize(s));
    return x;
}

}  // namespace

// Host arrays in and out: row[n+1], col[row[n]] (any CSR with ids < n);
// new_row[n+1], new_col[row[n]], perm[n] (perm[v] = new id of v).
void bfs_relabel_device(hipStream_t s, uint64_t n, const uint64_t *h_row, const uint32_t *h_col, uint32_t root,
                        uint64_t *h_new_row, uint32_t *h_new_col, uint32_t *h_perm) {
    if (n == 0) {
        h_new_row[0] = 0;
        return;
    }
    if (root >= n) throw std::invalid_argument("root >= n");
    const uint64_t m = h_row[n];
    Buf<uint64_t> row(n + 1), new_row(n + 1), cnt(n + 1);
    Buf<uint32_t> col(m), nid(n), claim(n), order(n);
    BFS_HIPCHK(hipMemcpyAsync(row.p, h_row, (n + 1) * 8, hipMemcpyHostToDevice, s));
    if (m) BFS_HIPCHK(hipMemcpyAsync(col.p, h_col, m * 4, hipMemcpyHostToDevice, s));
    {
        Buf<uint32_t> bad(1);
        hipLaunchKernelGGL(k_bfs_fill, dim3(1), dim3(64), 0, s, bad.p, (uint64_t)1, 0u);
        if (m) hipLaunchKernelGGL(k_bfs_check_ids, dim3(blocks_for(m, 256)), dim3(256), 0, s, col.p, m, n, bad.p);
        BFS_HIPCHK(hipGetLastError());
        uint32_t b = 0;
        BFS_HIPCHK(hipMemcpyAsync(&b, bad.p, 4, hipMemcpyDeviceToHost, s));
        BFS_HIPCHK(hipStreamSynchronize(s));
        if (b) throw std::invalid_argument("neighbour id >= n");
    }
    hipLaunchKernelGGL(k_bfs_fill, dim3(blocks_for(n, 256)), dim3(256), 0, s, nid.p, n, NONE);
    hipLaunchKernelGGL(k_bfs_fill, dim3(blocks_for(n, 256)), dim3(256), 0, s, claim.p, n, NONE);
    hipLaunchKernelGGL(k_bfs_root, dim3(1), dim3(1), 0, s, order.p, nid.p, root);
    BFS_HIPCHK(hipGetLastError());
    uint64_t lo = 0, hi = 1;
    while (lo < hi) {
        const uint64_t fsz = hi - lo;
        const dim3 g(blocks_for(fsz, WPB)), b(256);
        hipLaunchKernelGGL(k_bfs_claim, g, b, 0, s, order.p, lo, fsz, row.p, col.p, nid.p, claim.p);
        hipLaunchKernelGGL(k_bfs_count, g, b, 0, s, order.p, lo, fsz, row.p, col.p, nid.p, claim.p, cnt.p);
        BFS_HIPCHK(hipGetLastError());
        scan_in_place(s, cnt.p, fsz);
        const uint64_t total = read_u64(s, cnt.p + fsz);
        if (total) {
            hipLaunchKernelGGL(k_bfs_place, g, b, 0, s, order.p, lo, fsz, hi, row.p, col.p, nid.p, claim.p, cnt.p);
            hipLaunchKernelGGL(k_bfs_mark, dim3(blocks_for(total, 256)), b, 0, s, order.p, hi, hi + total, nid.p);
            BFS_HIPCHK(hipGetLastError());
        }
        lo = hi;
        hi += total;
    }
    if (hi < n) {
        // unreached vertices follow in id order (cnt holds n + 1 words)
        hipLaunchKernelGGL(k_bfs_unreached_flag, dim3(blocks_for(n, 256)), dim3(256), 0, s, nid.p, n, cnt.p);
        BFS_HIPCHK(hipGetLastError());
        scan_in_place(s, cnt.p, n);
        hipLaunchKernelGGL(k_bfs_unreached_place, dim3(blocks_for(n, 256)), dim3(256), 0, s, order.p, nid.p, n, hi,
                           cnt.p);
        BFS_HIPCHK(hipGetLastError());
    }
    // relabelled CSR: rows in the new order, then every row sorted
    hipLaunchKernelGGL(k_bfs_new_degree, dim3(blocks_for(n, 256)), dim3(256), 0, s, order.p, n, row.p, new_row.p);
    BFS_HIPCHK(hipGetLastError());
    scan_in_place(s, new_row.p, n);
    Buf<uint32_t> unsorted(m);
    hipLaunchKernelGGL(k_bfs_new_rows, dim3(blocks_for(n, WPB)), dim3(256), 0, s, order.p, n, row.p, col.p, nid.p,
                       new_row.p, unsorted.p);
    BFS_HIPCHK(hipGetLastError());
    std::vector<uint64_t> nr(n + 1);
    BFS_HIPCHK(hipMemcpyAsync(nr.data(), new_row.p, (n + 1) * 8, hipMemcpyDeviceToHost, s));
    BFS_HIPCHK(hipStreamSynchronize(s));
    // the old col buffer receives the sorted rows; batches of rows below 2^31 keys
    const uint64_t batch_keys = 1ull << 31;
    Buf<uint32_t> rel(n + 1);
    size_t temp_bytes = 0;
    void *temp = nullptr;
    try {
        for (uint64_t s0 = 0; s0 < n;) {
            uint64_t s1 = s0 + 1;
            // as many rows as fit the batch (a single longer row is still one batch)
            const uint64_t lim = nr[s0] + batch_keys;
            s1 = (uint64_t)(std::upper_bound(nr.begin() + s0 + 1, nr.end(), lim) - nr.begin()) - 1;
            if (s1 <= s0) s1 = s0 + 1;
            s1 = std::min<uint64_t>(s1, std::min<uint64_t>(n, s0 + (1ull << 31)));
            const uint64_t keys = nr[s1] - nr[s0];
            if (keys) {
                hipLaunchKernelGGL(k_bfs_rel_offsets, dim3(blocks_for(s1 - s0 + 1, 256)), dim3(256), 0, s, new_row.p,
                                   s0, s1, rel.p);
                BFS_HIPCHK(hipGetLastError());
                size_t need = 0;
                BFS_HIPCHK(rocprim::segmented_radix_sort_keys(nullptr, need, unsorted.p + nr[s0], col.p + nr[s0],
                                                              (unsigned)keys, (unsigned)(s1 - s0), rel.p, rel.p + 1,
                                                              0, 32, s));
                if (need > temp_bytes) {
                    if (temp) BFS_HIPCHK(hipFree(temp));
                    temp = nullptr;
                    BFS_HIPCHK(hipMalloc(&temp, need));
                    temp_bytes = need;
                }
                BFS_HIPCHK(rocprim::segmented_radix_sort_keys(temp, need, unsorted.p + nr[s0], col.p + nr[s0],
                                                              (unsigned)keys, (unsigned)(s1 - s0), rel.p, rel.p + 1,
                                                              0, 32, s));
                BFS_HIPCHK(hipStreamSynchronize(s));
            }
            s0 = s1;
        }
    } catch (...) {
        if (temp) (void)hipFree(temp);
        throw;
    }
    if (temp) BFS_HIPCHK(hipFree(temp));
    std::copy(nr.begin(), nr.end(), h_new_row);
    if (m) BFS_HIPCHK(hipMemcpyAsync(h_new_col, col.p, m * 4, hipMemcpyDeviceToHost, s));
    BFS_HIPCHK(hipMemcpyAsync(h_perm, nid.p, n * 4, hipMemcpyDeviceToHost, s));
    BFS_HIPCHK(hipStreamSynchronize(s));
}

}  // namespace lpa
