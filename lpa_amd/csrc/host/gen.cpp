// gen.cpp -- deterministic synthetic graphs (no reference counterpart).
//
// Every random decision is value i of a counter-based splitmix64 stream
// (ctr_rand(stream_key(seed, stream), i)), so a graph depends only on its
// parameters and seed, never on the thread count.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <unordered_set>

#include "host_internal.hpp"

using namespace lpah;

namespace {

// lattice edge id of (r,c)->(r,c+1) is 2v, of (r,c)->(r+1,c) is 2v+1, v = r*side+c
void gen_lattice(uint32_t side, double p_drop, double shortcut_frac, uint64_t seed,
                 lpah_csr &out) {
    const uint64_t n = (uint64_t)side * side;
    const uint64_t kdrop = stream_key(seed, 1), ksc = stream_key(seed, 2);
    std::vector<uint8_t> keep(2 * n, 0);
#pragma omp parallel for schedule(static)
    for (uint64_t v = 0; v < n; v++) {
        uint64_t r = v / side, c = v % side;
        keep[2 * v] = c + 1 < side && u01(ctr_rand(kdrop, 2 * v)) >= p_drop;
        keep[2 * v + 1] = r + 1 < side && u01(ctr_rand(kdrop, 2 * v + 1)) >= p_drop;
    }
    std::vector<uint32_t> s, d;
    for (uint64_t v = 0; v < n; v++) {
        if (keep[2 * v]) { s.push_back((uint32_t)v); d.push_back((uint32_t)(v + 1)); }
        if (keep[2 * v + 1]) { s.push_back((uint32_t)v); d.push_back((uint32_t)(v + side)); }
    }
    const uint64_t want = (uint64_t)std::floor(shortcut_frac * (double)n);
    std::unordered_set<uint64_t> added;
    for (uint64_t j = 0; added.size() < want && n > 1; j++) {
        uint64_t a = ctr_rand(ksc, 2 * j) % n, b = ctr_rand(ksc, 2 * j + 1) % n;
        if (a == b) continue;
        uint64_t lo = std::min(a, b), hi = std::max(a, b);
        bool lattice = (hi == lo + 1 && lo % side + 1 < side && keep[2 * lo]) ||
                       (hi == lo + side && keep[2 * lo + 1]);
        if (lattice || !added.insert(lo * n + hi).second) continue;
        s.push_back((uint32_t)lo);
        d.push_back((uint32_t)hi);
    }
    build_csr(n, s, d, true, out);
}

// Draw i takes its quadrant choices from 21-bit fields of the random words
// ctr_rand(key, i * ceil(scale / 3) + j): level l uses field l % 3 of word
// l / 3, compared with the thresholds a, a+b, a+b+c scaled to 2^21.
void gen_rmat(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed,
              lpah_csr &out) {
    const uint64_t n = 1ull << scale, m = (uint64_t)ef << scale;
    const uint64_t key = stream_key(seed, 3);
    const uint32_t F = 1u << 21;
    const uint32_t ta = (uint32_t)std::llround(a * F), tab = (uint32_t)std::llround((a + b) * F),
                   tabc = (uint32_t)std::llround((a + b + c) * F);
    const uint32_t words = (scale + 2) / 3;
    phase("rmat: start");
    std::vector<uint32_t> s(m), d(m);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < m; i++) {
        uint32_t u = 0, v = 0;
        for (uint32_t w = 0; w < words; w++) {
            const uint64_t r = ctr_rand(key, i * words + w);
            for (uint32_t f = 0; f < 3 && 3 * w + f < scale; f++) {
                const uint32_t x = (uint32_t)(r >> (21 * f)) & (F - 1);
                const uint32_t bu = x >= tab, bv = (x >= ta && x < tab) || x >= tabc;
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
        }
        s[i] = u;
        d[i] = v;
    }
    phase("rmat: draws");
    build_csr(n, s, d, true, out);
}

// out-degree: min(cap, floor(xmin * (1-u)^(-1/(alpha-1)))), u ~ U[0,1)
uint32_t pl_degree(double u, double xmin, double alpha, uint32_t cap) {
    double x = xmin * std::pow(1.0 - u, -1.0 / (alpha - 1.0));
    return x >= (double)cap ? cap : (uint32_t)x;
}

double pl_mean(double xmin, double alpha, uint32_t cap) {
    const int S = 1 << 20;   // stratified quadrature over u
    double s = 0;
    for (int i = 0; i < S; i++) s += pl_degree((i + 0.5) / S, xmin, alpha, cap);
    return s / S;
}

// Copying model (directed).  Frozen rules, every draw from a counter-based
// stream of the seed:
//  * out-degree d(v) = min(cap, floor(x_min (1-u)^(-1/(alpha-1)))), u ~ U[0,1)
//    (stream 4, index v), x_min set by bisection so that the mean of this
//    degree over u (stratified quadrature) is mean_deg;
//  * base list of u: B(u)[j] = uniform id (stream 8, index u * 2^17 + j),
//    j < d(u);
//  * prototype p(v) = v + U{-window..window} (stream 5, index v), clamped to
//    [0, n-1], and v + 1 (mod n) when that is v itself;
//  * v copies c(v) = min(round(copy_frac * d(v)), d(p)) entries of B(p(v)) at
//    distinct positions -- a partial Fisher-Yates shuffle of the positions
//    0..d(p)-1 driven by stream 7, index v * 2^17 + j -- so copies never
//    repeat a position;
//  * the other d(v) - c(v) targets are B(v)[c(v) .. d(v)): uniform ids;
//  * self loops and repeated targets are removed by the CSR build (rare:
//    both need a collision among uniform ids).
void gen_copying(uint64_t n, double mean_deg, double alpha, uint32_t cap, double copy_frac,
                 uint32_t window, uint64_t seed, lpah_csr &out) {
    if (n < 2) throw Error(E_ARG, "copying model needs n >= 2");
    if (cap >= (1u << 17)) throw Error(E_RANGE, "degree cap must be < 2^17");
    if (copy_frac < 0 || copy_frac > 1) throw Error(E_ARG, "copy fraction must be in [0, 1]");
    double lo = 0.05, hi = 1000.0;   // bisection on xmin for the requested mean
    for (int it = 0; it < 60; it++) {
        double mid = 0.5 * (lo + hi);
        (pl_mean(mid, alpha, cap) < mean_deg ? lo : hi) = mid;
    }
    const double xmin = 0.5 * (lo + hi);
    const uint64_t kdeg = stream_key(seed, 4), kproto = stream_key(seed, 5), kpick = stream_key(seed, 7),
                   krand = stream_key(seed, 8);
    std::vector<uint32_t> deg(n);
#pragma omp parallel for schedule(static)
    for (uint64_t v = 0; v < n; v++) deg[v] = pl_degree(u01(ctr_rand(kdeg, v)), xmin, alpha, cap);
    std::vector<uint64_t> off(n + 1, 0);
    for (uint64_t v = 0; v < n; v++) off[v + 1] = off[v] + deg[v];
    std::vector<uint32_t> s(off[n]), d(off[n]);
    auto base = [&](uint64_t u, uint64_t j) { return (uint32_t)(ctr_rand(krand, (u << 17) | j) % n); };
#pragma omp parallel
    {
        std::vector<uint32_t> pos;   // positions of the prototype's list not taken yet
#pragma omp for schedule(dynamic, 1024)
        for (uint64_t v = 0; v < n; v++) {
            int64_t p = (int64_t)v + (int64_t)(ctr_rand(kproto, v) % (2ull * window + 1)) - (int64_t)window;
            p = std::min<int64_t>(std::max<int64_t>(p, 0), (int64_t)n - 1);
            if ((uint64_t)p == v) p = (int64_t)((v + 1) % n);
            const uint64_t dv = deg[v], dp = deg[p];
            const uint64_t c = std::min<uint64_t>((uint64_t)std::llround(copy_frac * (double)dv), dp);
            pos.resize(dp);
            for (uint64_t k = 0; k < dp; k++) pos[k] = (uint32_t)k;
            uint32_t *sv = &s[off[v]], *tv = &d[off[v]];
            for (uint64_t j = 0; j < c; j++) {
                const uint64_t r = j + ctr_rand(kpick, (v << 17) | j) % (dp - j);
                std::swap(pos[j], pos[r]);
                sv[j] = (uint32_t)v;
                tv[j] = base((uint64_t)p, pos[j]);
            }
            for (uint64_t j = c; j < dv; j++) {
                sv[j] = (uint32_t)v;
                tv[j] = base(v, j);
            }
        }
    }
    build_csr(n, s, d, false, out);
}

}  // namespace

extern "C" {

int lpah_gen_lattice(uint32_t side, double p_drop, double shortcut_frac, uint64_t seed,
                     lpah_csr **out) {
    return guard([&] {
        if (!out || side == 0 || (uint64_t)side * side >= (1ull << 30)) throw Error(E_ARG, "bad lattice size");
        auto *g = new lpah_csr;
        try { gen_lattice(side, p_drop, shortcut_frac, seed, *g); } catch (...) { delete g; throw; }
        *out = g;
    });
}

int lpah_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                  uint64_t seed, int bfs, lpah_csr **out) {
    return guard([&] {
        if (!out || scale == 0 || scale > 29) throw Error(E_ARG, "scale must be in 1..29");
        if (a < 0 || b < 0 || c < 0 || a + b + c > 1) throw Error(E_ARG, "bad R-MAT probabilities");
        auto *g = new lpah_csr;
        try { gen_rmat(scale, edge_factor, a, b, c, seed, *g); } catch (...) { delete g; throw; }
        if (bfs) {
            lpah_csr *r = nullptr;
            int rc = lpah_csr_bfs_relabel(g, 0, nullptr, &r);
            delete g;
            if (rc) throw Error(rc, lpah_last_error());
            g = r;
        }
        *out = g;
    });
}

int lpah_gen_copying(uint64_t n, double mean_deg, double exponent, uint32_t cap,
                     double copy_frac, uint32_t window, uint64_t seed, lpah_csr **out) {
    return guard([&] {
        if (!out || n >= (1ull << 30) || exponent <= 1.0) throw Error(E_ARG, "bad copying-model parameters");
        auto *g = new lpah_csr;
        try { gen_copying(n, mean_deg, exponent, cap, copy_frac, window, seed, *g); } catch (...) { delete g; throw; }
        *out = g;
    });
}

}  // extern "C"
