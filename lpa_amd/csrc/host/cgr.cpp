// cgr.cpp -- parallel CGR encoder and the reference's on-disk file set.
//
// Produces the same bitstream as CgrCompressor with File's parameters
// (zeta 3, minimum interval 4, 256-bit interval and residual segments;
// src/file.cpp:84-87): per vertex an interval block then a residual block,
// each "gamma(#segments-1)" followed by segments "gamma(count) codes",
// every segment but the last zero-padded to 256 bits and the trailing
// partial segment merged into the last full one (src/compressor.cpp:364-499).
//
// The reference builds each record in a std::vector<bool> through O(n) code
// tables, one vertex at a time.  Here each record is planned (segment
// boundaries from code lengths), its length summed in a first parallel pass,
// and its bits written straight into 64-bit words in a second pass; words
// shared by two records are combined with an atomic OR.
#include <omp.h>

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <sys/stat.h>

#include "host_internal.hpp"

using namespace lpah;

namespace {

constexpr uint64_t SEG = 256;   // segment bits (file.cpp:84)
constexpr uint64_t MIN_ITV = 4; // minimum interval length (file.cpp:84)

inline int flog2(uint64_t y) { return 63 - __builtin_clzll(y); }
inline uint64_t gamma_len(uint64_t x) { return 2ull * flog2(x + 1) + 1; }
inline uint64_t zeta_len(uint64_t x) { return (uint64_t)(flog2(x + 1) / 3 + 1) * 4; }
inline uint64_t nat(int64_t d) { return d >= 0 ? (uint64_t)d << 1 : (uint64_t)(-2 * d - 1); }

// Sink that only measures.
struct Counter {
    uint64_t pos = 0;
    void put(uint64_t, int len) { pos += (uint64_t)len; }
    void skip(uint64_t len) { pos += len; }
};

// Sink that ORs bits, MSB first, into big-endian-ordered 64-bit words.
struct Writer {
    uint64_t *words;
    uint64_t pos;
    uint64_t acc = 0;
    uint64_t wi;
    Writer(uint64_t *w, uint64_t start) : words(w), pos(start), wi(start >> 6) {}
    void flush() {
        if (acc) __atomic_fetch_or(&words[wi], acc, __ATOMIC_RELAXED);
        acc = 0;
    }
    void advance_to(uint64_t np) {
        pos = np;
        if ((pos >> 6) != wi) { flush(); wi = pos >> 6; }
    }
    void put(uint64_t val, int len) {   // 1 <= len <= 64, val < 2^len
        int off = (int)(pos & 63);
        if (off + len <= 64) {
            acc |= (len == 64) ? val : val << (64 - off - len);
        } else {
            int spill = off + len - 64;
            acc |= val >> spill;
            flush();
            wi++;
            acc = val << (64 - spill);
        }
        advance_to(pos + (uint64_t)len);
    }
    void skip(uint64_t len) { advance_to(pos + len); }
    ~Writer() { flush(); }
};

template <class S> inline void put_gamma(S &s, uint64_t x) {
    uint64_t y = x + 1;
    s.put(y, 2 * flog2(y) + 1);   // L zeros then the L+1 bits of y
}
template <class S> inline void put_zeta(S &s, uint64_t x) {
    uint64_t y = x + 1;
    int h = flog2(y) / 3;
    s.put((1ull << (3 * (h + 1))) | y, 4 * (h + 1));   // 0^h 1 then 3(h+1) bits of y
}

struct Scratch {
    std::vector<uint64_t> il, ilen, res;
    std::vector<std::pair<uint64_t, uint64_t>> segs;   // (first index, count)
};

// Segment plan of one block: `code(i, first)` is the code of entry i when it
// opens a segment (first) or continues one; `len(code)` its size.
template <class CodeLen>
void plan_segments(uint64_t count, CodeLen &&entry_bits, std::vector<std::pair<uint64_t, uint64_t>> &segs) {
    segs.clear();
    uint64_t cur = 0, cnt = 0, first = 0;
    for (uint64_t i = 0; i < count; i++) {
        uint64_t b = entry_bits(i, cnt == 0);
        if (gamma_len(cnt + 1) + cur + b > SEG) {          // compressor.cpp:390-399, 462-468
            segs.emplace_back(first, cnt);
            cnt = 0;
            cur = 0;
            first = i;
            b = entry_bits(i, true);
        }
        cnt++;
        cur += b;
    }
    if (segs.empty()) segs.emplace_back(first, cnt);
    else segs.back().second += cnt;                        // merge, compressor.cpp:412-425, 479-483
}

template <class S>
void encode_vertex(S &s, uint64_t v, const uint32_t *adj, uint64_t d, Scratch &t) {
    t.il.clear(); t.ilen.clear(); t.res.clear();
    for (uint64_t a = 0; a < d;) {                         // intervalize, compressor.cpp:330-362
        uint64_t b = a + 1;
        while (b < d && (uint64_t)adj[b - 1] + 1 == adj[b]) b++;
        if (b - a >= MIN_ITV) { t.il.push_back(adj[a]); t.ilen.push_back(b - a); }
        else for (uint64_t k = a; k < b; k++) t.res.push_back(adj[k]);
        a = b;
    }
    auto itv_left_code = [&](uint64_t i, bool first) {
        return first ? nat((int64_t)t.il[i] - (int64_t)v) : t.il[i] - t.il[i - 1] - t.ilen[i - 1] - 1;
    };
    plan_segments(t.il.size(), [&](uint64_t i, bool first) {
        return gamma_len(itv_left_code(i, first)) + gamma_len(t.ilen[i] - MIN_ITV);
    }, t.segs);
    put_gamma(s, t.segs.size() - 1);
    for (size_t k = 0; k < t.segs.size(); k++) {
        uint64_t start = s.pos;
        auto [first, cnt] = t.segs[k];
        put_gamma(s, cnt);
        for (uint64_t i = first; i < first + cnt; i++) {
            put_gamma(s, itv_left_code(i, i == first));
            put_gamma(s, t.ilen[i] - MIN_ITV);
        }
        if (k + 1 < t.segs.size()) s.skip(start + SEG - s.pos);
    }
    auto res_code = [&](uint64_t i, bool first) {
        return first ? nat((int64_t)t.res[i] - (int64_t)v) : t.res[i] - t.res[i - 1] - 1;
    };
    plan_segments(t.res.size(), [&](uint64_t i, bool first) { return zeta_len(res_code(i, first)); }, t.segs);
    put_gamma(s, t.segs.size() - 1);
    for (size_t k = 0; k < t.segs.size(); k++) {
        uint64_t start = s.pos;
        auto [first, cnt] = t.segs[k];
        put_gamma(s, cnt);
        for (uint64_t i = first; i < first + cnt; i++) put_zeta(s, res_code(i, i == first));
        if (k + 1 < t.segs.size()) s.skip(start + SEG - s.pos);
    }
}

void encode(const lpah_csr &g, uint32_t small_t, uint32_t large_t, lpah_cgr &c) {
    const uint64_t n = g.n;
    c.n = n;
    c.vertices.clear();
    c.vertices.reserve(n);
    auto deg = [&](uint64_t v) { return g.row[v + 1] - g.row[v]; };
    for (uint64_t v = 0; v < n; v++) if (deg(v) >= large_t) c.vertices.push_back((uint32_t)v);
    for (uint64_t v = 0; v < n; v++) if (deg(v) >= small_t && deg(v) < large_t) c.vertices.push_back((uint32_t)v);
    c.n_c = c.vertices.size();
    for (uint64_t v = 0; v < n; v++) if (deg(v) < small_t) c.vertices.push_back((uint32_t)v);

    std::vector<uint64_t> len(c.n_c + 1, 0);
#pragma omp parallel
    {
        Scratch t;
#pragma omp for schedule(dynamic, 1024)
        for (uint64_t i = 0; i < c.n_c; i++) {
            uint64_t v = c.vertices[i];
            Counter k;
            encode_vertex(k, v, g.col.data() + g.row[v], deg(v), t);
            len[i] = k.pos;
        }
    }
    c.offsets.assign(c.n_c + 1, 0);
    for (uint64_t i = 0; i < c.n_c; i++) c.offsets[i + 1] = c.offsets[i] + len[i];
    c.nbits = c.offsets[c.n_c];
    c.words.assign((c.nbits + 63) / 64 + 2, 0);
#pragma omp parallel
    {
        Scratch t;
#pragma omp for schedule(dynamic, 1024)
        for (uint64_t i = 0; i < c.n_c; i++) {
            uint64_t v = c.vertices[i];
            Writer w(c.words.data(), c.offsets[i]);
            encode_vertex(w, v, g.col.data() + g.row[v], deg(v), t);
        }
    }
    const uint64_t nbytes = (c.nbits + 7) / 8, nw32 = (nbytes + 3) / 4;
    c.words32.assign(nw32 + 4, 0);
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < nw32; i++)
        c.words32[i] = (uint32_t)(c.words[i >> 1] >> ((i & 1) ? 0 : 32));
    c.small_offs.assign(1, 0);
    c.small_nbrs.clear();
    for (uint64_t i = c.n_c; i < n; i++) {
        uint64_t v = c.vertices[i];
        c.small_nbrs.insert(c.small_nbrs.end(), g.col.begin() + g.row[v], g.col.begin() + g.row[v + 1]);
        c.small_offs.push_back(c.small_nbrs.size());
    }
}

bool file_exists(const std::string &p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

template <class T> void write_bin(const std::string &p, const T *data, uint64_t n) {
    FILE *f = fopen(p.c_str(), "wb");
    if (!f) throw Error(E_IO, "cannot create " + p);
    if (n && fwrite(data, sizeof(T), n, f) != n) { fclose(f); throw Error(E_IO, "short write " + p); }
    fclose(f);
}

template <class T> std::vector<T> read_bin(const std::string &p) {
    FILE *f = fopen(p.c_str(), "rb");
    if (!f) throw Error(E_IO, "cannot open " + p);
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (sz < 0 || sz % (long)sizeof(T)) { fclose(f); throw Error(E_FORMAT, "bad size " + p); }
    std::vector<T> v((size_t)sz / sizeof(T));
    if (!v.empty() && fread(v.data(), sizeof(T), v.size(), f) != v.size()) { fclose(f); throw Error(E_IO, "short read " + p); }
    fclose(f);
    return v;
}

}  // namespace

extern "C" {

int lpah_cgr_encode(const lpah_csr *g, uint32_t small_thresh, uint32_t large_thresh, lpah_cgr **out) {
    return guard([&] {
        if (!g || !out) throw Error(E_ARG, "null argument");
        if (g->n >= (1ull << 30)) throw Error(E_RANGE, "n must be < 2^30");
        auto *c = new lpah_cgr;
        try { encode(*g, small_thresh, large_thresh, *c); } catch (...) { delete c; throw; }
        *out = c;
    });
}

uint64_t lpah_cgr_n(const lpah_cgr *c) { return c->n; }
uint64_t lpah_cgr_nc(const lpah_cgr *c) { return c->n_c; }
uint64_t lpah_cgr_nbits(const lpah_cgr *c) { return c->nbits; }
const uint32_t *lpah_cgr_words(const lpah_cgr *c) { return c->words32.data(); }
uint64_t lpah_cgr_nwords(const lpah_cgr *c) { return c->words32.size() - 4; }
const uint64_t *lpah_cgr_offsets(const lpah_cgr *c) { return c->offsets.data(); }
const uint32_t *lpah_cgr_vertices(const lpah_cgr *c) { return c->vertices.data(); }
const uint64_t *lpah_cgr_small_offsets(const lpah_cgr *c) { return c->small_offs.data(); }
const uint32_t *lpah_cgr_small_nbrs(const lpah_cgr *c) { return c->small_nbrs.data(); }
uint64_t lpah_cgr_small_m(const lpah_cgr *c) { return c->small_nbrs.size(); }
void lpah_cgr_free(lpah_cgr *c) { delete c; }

int lpah_files_write(const lpah_cgr *c, const lpah_csr *csr, const char *path) {
    return guard([&] {
        if (!c || !path) throw Error(E_ARG, "null argument");
        std::string p(path);
        const uint64_t nbytes = (c->nbits + 7) / 8;
        std::vector<uint8_t> bytes(nbytes);
        for (uint64_t i = 0; i < nbytes; i++) bytes[i] = (uint8_t)(c->words[i >> 3] >> (56 - 8 * (i & 7)));
        write_bin(p + ".graph", bytes.data(), nbytes);
        {   // count line, then cumulative END offsets (compressor.cpp:190, 219, 248)
            FILE *f = fopen((p + ".offset").c_str(), "w");
            if (!f) throw Error(E_IO, "cannot create " + p + ".offset");
            fprintf(f, "%llu\n", (unsigned long long)c->n_c);
            for (uint64_t i = 1; i <= c->n_c; i++) fprintf(f, "%llu\n", (unsigned long long)c->offsets[i]);
            fclose(f);
        }
        write_bin(p + ".offset64.bin", c->offsets.data(), c->offsets.size());
        write_bin(p + ".vertices.bin", c->vertices.data(), c->vertices.size());
        if (!c->small_nbrs.empty()) {   // written only when non-empty (compressor.cpp:284-288)
            write_bin(p + ".graph.small.bin", c->small_nbrs.data(), c->small_nbrs.size());
            std::vector<uint32_t> so(c->small_offs.begin(), c->small_offs.end());
            write_bin(p + ".offset.small.bin", so.data(), so.size());
        }
        if (csr && csr->col.size() < (1ull << 32)) {   // file.cpp:92-93
            std::vector<uint32_t> row(csr->row.begin(), csr->row.end());
            write_bin(p + ".row", row.data(), row.size());
            write_bin(p + ".clm", csr->col.data(), csr->col.size());
        }
    });
}

int lpah_files_load(const char *path, int small_thresh, lpah_cgr **out) {
    return guard([&] {
        if (!path || !out) throw Error(E_ARG, "null argument");
        std::string p(path);
        auto *c = new lpah_cgr;
        try {
            auto bytes = read_bin<uint8_t>(p + ".graph");
            c->vertices = read_bin<uint32_t>(p + ".vertices.bin");
            c->n = c->vertices.size();
            if (file_exists(p + ".offset64.bin")) {
                c->offsets = read_bin<uint64_t>(p + ".offset64.bin");
                if (c->offsets.empty() || c->offsets[0] != 0) throw Error(E_FORMAT, "bad .offset64.bin");
                c->n_c = c->offsets.size() - 1;
            } else {
                std::ifstream f(p + ".offset");
                if (!f) throw Error(E_IO, "cannot open " + p + ".offset");
                unsigned long long cnt = 0, x = 0;
                if (!(f >> cnt)) throw Error(E_FORMAT, "empty .offset");
                c->offsets.assign(1, 0);                    // the loader prepends 0 (file.cpp:243)
                for (unsigned long long i = 0; i < cnt; i++) {
                    if (!(f >> x)) throw Error(E_FORMAT, ".offset has fewer lines than its count");
                    c->offsets.push_back(x);
                }
                c->n_c = cnt;
            }
            if (c->n_c > c->n) throw Error(E_FORMAT, "more offsets than vertices");
            for (uint64_t i = 0; i < c->n_c; i++)
                if (c->offsets[i + 1] < c->offsets[i]) throw Error(E_FORMAT, "offsets decrease");
            c->nbits = c->offsets[c->n_c];
            if ((c->nbits + 7) / 8 > bytes.size()) throw Error(E_FORMAT, ".graph shorter than the offsets say");
            c->words.assign(bytes.size() / 8 + 3, 0);
            for (uint64_t i = 0; i < bytes.size(); i++) c->words[i >> 3] |= (uint64_t)bytes[i] << (56 - 8 * (i & 7));
            const uint64_t nw32 = (bytes.size() + 3) / 4;
            c->words32.assign(nw32 + 4, 0);
            for (uint64_t i = 0; i < nw32; i++) c->words32[i] = (uint32_t)(c->words[i >> 1] >> ((i & 1) ? 0 : 32));
            for (uint32_t v : c->vertices) if (v >= c->n) throw Error(E_FORMAT, "vertex id >= n in .vertices.bin");
            c->small_offs.assign(1, 0);
            if (c->n_c < c->n) {
                if (small_thresh <= 0) throw Error(E_FORMAT, "graph has small vertices but small_thresh == 0");
                if (file_exists(p + ".graph.small.bin")) {
                    c->small_nbrs = read_bin<uint32_t>(p + ".graph.small.bin");
                    auto so = read_bin<uint32_t>(p + ".offset.small.bin");
                    if (so.size() != c->n - c->n_c + 1) throw Error(E_FORMAT, "bad .offset.small.bin length");
                    c->small_offs.assign(so.begin(), so.end());
                    if (c->small_offs.back() != c->small_nbrs.size()) throw Error(E_FORMAT, "small CSR size mismatch");
                    for (uint32_t v : c->small_nbrs) if (v >= c->n) throw Error(E_FORMAT, "small neighbour >= n");
                } else {
                    c->small_offs.assign(c->n - c->n_c + 1, 0);   // every small vertex isolated
                }
            }
        } catch (...) { delete c; throw; }
        *out = c;
    });
}

}  // extern "C"
