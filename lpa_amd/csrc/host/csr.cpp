// csr.cpp -- CSR construction, BFS relabelling, text edge lists.
//
// Replaces File::read_from_edge_list (src/file.cpp:115-195) and
// CgrCompressor::load_graph (src/compressor.cpp:22-68): both build sorted
// adjacency lists from a text edge list, single threaded, with 8-byte ids
// and without removing duplicates.  Here: 32-bit ids, 64-bit offsets,
// counting sort by source over OpenMP threads, duplicates and self loops
// removed.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstdio>
#include <thread>

#include "host_internal.hpp"

namespace lpah {

static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }

void phase(const char *name) {
    static const bool on = getenv("LPA_HOST_TIMING") != nullptr;
    static auto last = std::chrono::steady_clock::now();
    if (!on) return;
    auto now = std::chrono::steady_clock::now();
    fprintf(stderr, "[lpa host] %-28s %8.3f s\n", name, std::chrono::duration<double>(now - last).count());
    last = now;
}

static void exclusive_scan(std::vector<uint64_t> &a) {
    uint64_t s = 0;
    for (auto &x : a) { uint64_t t = x; x = s; s += t; }
}

// Arcs are distributed without atomics: per-thread counts of arcs per
// source bucket (a range of 2^shift source ids), a prefix sum over
// (bucket, thread), a scatter of (src, dst) pairs into bucket order, then one
// thread per bucket counts, places and sorts the rows of its id range in a
// cache-sized window; rows are then sorted over all threads.  Atomic
// counters would serialise every thread on the cache lines of R-MAT hubs.
void build_csr(uint64_t n, std::vector<uint32_t> &src, std::vector<uint32_t> &dst,
               bool undirected, lpah_csr &out) {
    const uint64_t m = src.size();
    if (dst.size() != m) throw Error(E_ARG, "src/dst length mismatch");
    for (uint64_t i = 0; i < m; i++)
        if (src[i] >= n || dst[i] >= n) throw Error(E_RANGE, "edge endpoint >= n");
    uint32_t lg = 0;
    while ((1ull << lg) < std::max<uint64_t>(n, 1)) lg++;
    const uint32_t shift = lg > 10 ? lg - 10 : 0;
    const uint64_t nb = ((std::max<uint64_t>(n, 1) - 1) >> shift) + 1;
    const int T = omp_get_max_threads();
    std::vector<uint64_t> hist((uint64_t)T * nb, 0);
    const uint64_t chunk = (m + T - 1) / std::max(T, 1);
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        uint64_t *h = &hist[(uint64_t)t * nb];
        const uint64_t b0 = std::min(m, t * chunk), b1 = std::min(m, b0 + chunk);
        for (uint64_t i = b0; i < b1; i++) {
            const uint32_t u = src[i], v = dst[i];
            if (u == v) continue;
            h[u >> shift]++;
            if (undirected) h[v >> shift]++;
        }
    }
    // bucket-major prefix: bstart[b] = first slot of bucket b
    phase("csr: bucket counts");
    std::vector<uint64_t> bstart(nb + 1, 0);
    uint64_t total = 0;
    for (uint64_t b = 0; b < nb; b++) {
        bstart[b] = total;
        for (int t = 0; t < T; t++) {
            uint64_t c = hist[(uint64_t)t * nb + b];
            hist[(uint64_t)t * nb + b] = total;
            total += c;
        }
    }
    bstart[nb] = total;
    std::vector<uint32_t> as(total), ad(total);
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        uint64_t *h = &hist[(uint64_t)t * nb];
        const uint64_t b0 = std::min(m, t * chunk), b1 = std::min(m, b0 + chunk);
        for (uint64_t i = b0; i < b1; i++) {
            const uint32_t u = src[i], v = dst[i];
            if (u == v) continue;
            uint64_t o = h[u >> shift]++;
            as[o] = u;
            ad[o] = v;
            if (undirected) {
                o = h[v >> shift]++;
                as[o] = v;
                ad[o] = u;
            }
        }
    }
    std::vector<uint32_t>().swap(src);
    std::vector<uint32_t>().swap(dst);
    std::vector<uint64_t>().swap(hist);
    phase("csr: scatter to buckets");
    // per bucket: degrees, then rows placed, sorted and deduplicated in place
    std::vector<uint64_t> deg(n + 1, 0);
    std::vector<uint32_t> col0(total);
    std::vector<uint64_t> start(n + 1, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (uint64_t b = 0; b < nb; b++) {
        const uint64_t v0 = b << shift, v1 = std::min<uint64_t>(n, (b + 1) << shift);
        std::vector<uint64_t> cur(v1 - v0 + 1, 0);
        for (uint64_t e = bstart[b]; e < bstart[b + 1]; e++) cur[as[e] - v0 + 1]++;
        for (uint64_t k = 0; k < v1 - v0; k++) cur[k + 1] += cur[k];
        for (uint64_t k = 0; k < v1 - v0; k++) start[v0 + k] = bstart[b] + cur[k];
        for (uint64_t e = bstart[b]; e < bstart[b + 1]; e++) col0[bstart[b] + cur[as[e] - v0]++] = ad[e];
    }
    start[n] = total;
    phase("csr: place rows");
    std::vector<uint32_t>().swap(as);
    std::vector<uint32_t>().swap(ad);
    // rows sorted over all threads (hub rows would serialise a bucket)
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint64_t v = 0; v < n; v++) {
        auto it0 = col0.begin() + start[v], it1 = col0.begin() + start[v + 1];
        std::sort(it0, it1);
        deg[v] = (uint64_t)(std::unique(it0, it1) - it0);
    }
    phase("csr: sort rows");
    exclusive_scan(deg);
    out.n = n;
    out.col.assign(deg[n], 0);
#pragma omp parallel for schedule(dynamic, 4096)
    for (uint64_t v = 0; v < n; v++)
        std::copy(col0.begin() + start[v], col0.begin() + start[v] + (deg[v + 1] - deg[v]), out.col.begin() + deg[v]);
    out.row.swap(deg);
    phase("csr: compact");
}

}  // namespace lpah

using namespace lpah;

extern "C" {

const char *lpah_last_error(void) { return lpah::g_err.c_str(); }

int lpah_threads(void) { return omp_get_max_threads(); }

int lpah_csr_from_edges(uint64_t n, const uint32_t *src, const uint32_t *dst, uint64_t m,
                        int undirected, lpah_csr **out) {
    return guard([&] {
        if (!out || (m && (!src || !dst))) throw Error(E_ARG, "null argument");
        if (n >= (1ull << 30)) throw Error(E_RANGE, "n must be < 2^30");
        std::vector<uint32_t> s(src, src + m), d(dst, dst + m);
        auto *g = new lpah_csr;
        try { build_csr(n, s, d, undirected != 0, *g); } catch (...) { delete g; throw; }
        *out = g;
    });
}

int lpah_csr_from_arrays(uint64_t n, const uint64_t *row, const uint32_t *col, lpah_csr **out) {
    return guard([&] {
        if (!out || !row) throw Error(E_ARG, "null argument");
        if (row[0] != 0) throw Error(E_FORMAT, "row[0] != 0");
        for (uint64_t v = 0; v < n; v++)
            if (row[v + 1] < row[v]) throw Error(E_FORMAT, "row offsets decrease");
        int bad = 0;   // 1: id >= n, 2: row not strictly ascending
#pragma omp parallel for schedule(dynamic, 4096) reduction(max : bad)
        for (uint64_t v = 0; v < n; v++)
            for (uint64_t e = row[v]; e < row[v + 1]; e++) {
                if (col[e] >= n) bad = std::max(bad, 1);
                else if (e > row[v] && col[e] <= col[e - 1]) bad = std::max(bad, 2);
            }
        if (bad == 1) throw Error(E_RANGE, "neighbour id >= n");
        if (bad == 2) throw Error(E_FORMAT, "row not strictly ascending");
        auto *g = new lpah_csr;
        g->n = n;
        g->row.assign(row, row + n + 1);
        g->col.resize(row[n]);
#pragma omp parallel for schedule(static)
        for (uint64_t b = 0; b < row[n]; b += 1 << 20)
            std::copy(col + b, col + std::min<uint64_t>(row[n], b + (1 << 20)), g->col.begin() + b);
        *out = g;
    });
}

int lpah_csr_read_edge_list(const char *path, int undirected, lpah_csr **out) {
    return guard([&] {
        FILE *f = fopen(path, "rb");
        if (!f) throw Error(E_IO, std::string("cannot open ") + path);
        std::vector<uint32_t> s, d;
        std::vector<char> buf(1 << 20);
        uint64_t val = 0, maxid = 0;
        bool in_num = false;
        int field = 0;
        uint32_t u = 0;
        size_t got;
        while ((got = fread(buf.data(), 1, buf.size(), f)) > 0) {
            for (size_t i = 0; i < got; i++) {
                char ch = buf[i];
                if (ch >= '0' && ch <= '9') {
                    val = val * 10 + (uint64_t)(ch - '0');
                    if (val >= (1ull << 30)) { fclose(f); throw Error(E_RANGE, "vertex id >= 2^30"); }
                    in_num = true;
                } else if (in_num) {
                    if (field == 0) { u = (uint32_t)val; field = 1; }
                    else { s.push_back(u); d.push_back((uint32_t)val); field = 0; }
                    maxid = std::max<uint64_t>(maxid, val);
                    val = 0; in_num = false;
                }
            }
        }
        fclose(f);
        if (in_num && field == 1) { s.push_back(u); d.push_back((uint32_t)val); maxid = std::max<uint64_t>(maxid, val); }
        uint64_t n = s.empty() ? 0 : maxid + 1;
        auto *g = new lpah_csr;
        try { build_csr(n, s, d, undirected != 0, *g); } catch (...) { delete g; throw; }
        *out = g;
    });
}

uint64_t lpah_csr_n(const lpah_csr *g) { return g->n; }
uint64_t lpah_csr_m(const lpah_csr *g) { return g->col.size(); }
const uint64_t *lpah_csr_row(const lpah_csr *g) { return g->row.data(); }
const uint32_t *lpah_csr_col(const lpah_csr *g) { return g->col.data(); }
uint32_t lpah_csr_max_degree(const lpah_csr *g) {
    uint64_t mx = 0;
#pragma omp parallel for reduction(max : mx)
    for (uint64_t v = 0; v < g->n; v++) mx = std::max(mx, g->row[v + 1] - g->row[v]);
    return (uint32_t)mx;
}
void lpah_csr_free(lpah_csr *g) { delete g; }

// BFS order from `root` with neighbours taken in ascending order, computed
// level by level in parallel: a vertex discovered at level L+1 belongs to
// the frontier vertex of smallest position that has it as a neighbour, and
// the new level is ordered by (that position, vertex id) -- exactly the
// order a sequential queue-based BFS produces.
static void bfs_order(const lpah_csr *g, uint32_t root, std::vector<uint32_t> &order, std::vector<uint32_t> &nid) {
    const uint64_t n = g->n;
    const uint32_t NONE = 0xFFFFFFFFu;
    nid.assign(n, NONE);
    std::vector<uint32_t> claim(n, NONE);
    order.clear();
    order.reserve(n);
    order.push_back(root);
    nid[root] = 0;
    uint64_t lo = 0, hi = 1;   // current frontier = order[lo, hi)
    std::vector<uint64_t> cnt;
    while (lo < hi) {
        const uint64_t fsz = hi - lo;
        // pass 1: every undiscovered neighbour remembers its first parent
#pragma omp parallel for schedule(dynamic, 256)
        for (uint64_t i = 0; i < fsz; i++) {
            const uint32_t u = order[lo + i];
            for (uint64_t e = g->row[u]; e < g->row[u + 1]; e++) {
                const uint32_t w = g->col[e];
                if (nid[w] != NONE) continue;
                uint32_t cur = __atomic_load_n(&claim[w], __ATOMIC_RELAXED);
                while ((uint32_t)i < cur &&
                       !__atomic_compare_exchange_n(&claim[w], &cur, (uint32_t)i, true, __ATOMIC_RELAXED,
                                                    __ATOMIC_RELAXED)) {
                }
            }
        }
        // pass 2: each parent counts its children, in adjacency order
        cnt.assign(fsz + 1, 0);
#pragma omp parallel for schedule(dynamic, 256)
        for (uint64_t i = 0; i < fsz; i++) {
            const uint32_t u = order[lo + i];
            uint64_t c = 0;
            for (uint64_t e = g->row[u]; e < g->row[u + 1]; e++) {
                const uint32_t w = g->col[e];
                if (nid[w] == NONE && claim[w] == (uint32_t)i) c++;
            }
            cnt[i + 1] = c;
        }
        for (uint64_t i = 0; i < fsz; i++) cnt[i + 1] += cnt[i];
        order.resize(hi + cnt[fsz]);
        // pass 3: children take consecutive positions
#pragma omp parallel for schedule(dynamic, 256)
        for (uint64_t i = 0; i < fsz; i++) {
            const uint32_t u = order[lo + i];
            uint64_t o = hi + cnt[i];
            for (uint64_t e = g->row[u]; e < g->row[u + 1]; e++) {
                const uint32_t w = g->col[e];
                if (nid[w] == NONE && claim[w] == (uint32_t)i) order[o++] = w;
            }
        }
        const uint64_t nhi = order.size();
#pragma omp parallel for schedule(static)
        for (uint64_t q = hi; q < nhi; q++) nid[order[q]] = (uint32_t)q;
        lo = hi;
        hi = nhi;
    }
    for (uint64_t v = 0; v < n; v++)
        if (nid[v] == NONE) {
            nid[v] = (uint32_t)order.size();
            order.push_back((uint32_t)v);
        }
}

int lpah_csr_bfs_relabel(const lpah_csr *g, uint32_t root, uint32_t *perm_out, lpah_csr **out) {
    return guard([&] {
        const uint64_t n = g->n;
        if (n == 0) { *out = new lpah_csr(*g); return; }
        if (root >= n) throw Error(E_RANGE, "root >= n");
        std::vector<uint32_t> nid, order;
        bfs_order(g, root, order, nid);
        auto *r = new lpah_csr;
        r->n = n;
        r->row.assign(n + 1, 0);
        for (uint64_t i = 0; i < n; i++) r->row[i + 1] = r->row[i] + (g->row[order[i] + 1] - g->row[order[i]]);
        r->col.resize(g->col.size());
#pragma omp parallel for schedule(dynamic, 4096)
        for (uint64_t i = 0; i < n; i++) {
            uint32_t u = order[i];
            uint64_t o = r->row[i];
            for (uint64_t e = g->row[u]; e < g->row[u + 1]; e++) r->col[o++] = nid[g->col[e]];
            std::sort(r->col.begin() + r->row[i], r->col.begin() + r->row[i + 1]);
        }
        if (perm_out) std::copy(nid.begin(), nid.end(), perm_out);
        *out = r;
    });
}

}  // extern "C"
