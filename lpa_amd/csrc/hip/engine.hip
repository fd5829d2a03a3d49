// engine.hip -- liblpa_hip.so: the C ABI of include/lpa_amd.h over the
// gfx950 kernels of lp_kernels.hpp / prep_kernels.hpp.
//
// Device layout after scheduling (all HBM-resident, 64-bit offsets):
//   words   CGR stream as 64-bit words, records reordered so every bin is a
//           contiguous run of positions, by degree (descending; or in
//           vertex-id order) inside the bin; 8 zero words of padding for
//           the readers' look-ahead
//   offs    n_c+1 start bit offsets of the reordered records
//   vid     vertex id of each position
//   srow/snbr/svid   small-vertex CSR (deg < small threshold), same order rule
//   L[2]    two label buffers, n*K uint32 each, layout [v][K]
//   crow/ccol        CGLP: the compressed part decoded to CSR (positions order)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "bfs_relabel.hpp"
#include "encode_kernels.hpp"
#include "encode_long.hpp"
#include "heavy_kernels.hpp"
#include "mlp_kernels.hpp"
#include "lp_kernels.hpp"
#include "lpa_amd.h"
#include "prep_kernels.hpp"

using namespace lpa;

namespace {

thread_local std::string g_err;

struct Fail : std::runtime_error {
    int code;
    Fail(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess)                                                                  \
            throw Fail(LPA_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));           \
    } while (0)

template <class F> int guard(F &&f) {
    try {
        f();
        return LPA_OK;
    } catch (const Fail &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_err = "out of host memory";
        return LPA_ERR_RANGE;
    } catch (const std::exception &e) {
        g_err = e.what();
        return LPA_ERR_ARG;
    }
}

template <class T> struct DevBuf {
    T *p = nullptr;
    uint64_t n = 0;
    bool host = false;   // pinned host memory, mapped into the device's address space
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)(host ? hipHostFree(p) : hipFree(p));
        p = nullptr;
        n = 0;
        host = false;
    }
    void alloc(uint64_t cnt) {
        release();
        if (cnt == 0) cnt = 1;
        HIPCHK(hipMalloc((void **)&p, cnt * sizeof(T)));
        n = cnt;
    }
    // pinned host memory that kernels can read in place (same address on
    // the device) and copies can stream from at full link speed
    void alloc_host(uint64_t cnt) {
        release();
        if (cnt == 0) cnt = 1;
        HIPCHK(hipHostMalloc((void **)&p, cnt * sizeof(T), hipHostMallocDefault));
        host = true;
        n = cnt;
        void *d = nullptr;
        HIPCHK(hipHostGetDevicePointer(&d, p, 0));
        if (d != (void *)p) throw Fail(LPA_ERR_HIP, "pinned host memory is mapped at another device address");
    }
    void swap(DevBuf &o) {
        std::swap(p, o.p);
        std::swap(n, o.n);
        std::swap(host, o.host);
    }
    void upload(const T *h, uint64_t cnt, hipStream_t s) {
        alloc(cnt);
        if (cnt) HIPCHK(hipMemcpyAsync(p, h, cnt * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

// Heavy vertices of one part (compressed or small CSR) and one mode (CMS or
// exact HBM hash), processed by the chunked kernels of heavy_kernels.hpp.
struct HeavySet {
    bool csr_part = false;
    bool cms = false;
    uint32_t nh = 0, nchunks = 0, kb = 1;
    uint32_t nlocal = 0, nlocal_big = 0, nchunked = 0;
    // heavy-vertex indices: one workgroup each (the last nlocal_big of them
    // exact with degree in (CHUNK, 2*CHUNK]) / chunked; chunk slots
    DevBuf<uint32_t> local, chunked, slot;
    DevBuf<uint32_t> pos, chunk_h;
    DevBuf<uint64_t> loff, chunk_b, toff;
    DevBuf<uint32_t> list;
    DevBuf<uint64_t> tab;
    DevBuf<uint32_t> sketch;
    DevBuf<unsigned long long> result;
    DevBuf<uint32_t> hcnt;                   // CMS: own-label counters per (slot, vector of a batch)
    uint32_t ntasks = 0;
    DevBuf<DecodeTask> tasks;                // decode tasks (fixed per layout)
    DevBuf<uint32_t> stash;                  // CMS, several vectors: label stash per resident workgroup;
                                             // exact, many vectors: row stash per resident workgroup
    uint32_t stash_blocks = 0;
    uint32_t rstash_small = 0;               // exact: grid of the one-workgroup kernel of degrees up
                                             // to CHUNK with a row stash (0: none, a workgroup per vertex)
    DevBuf<uint32_t> wq;                     // its work and exit counters
    bool tagged = false;                     // exact: generation-tagged tables (see TABLE_GENS)
    DevBuf<uint32_t> epoch;                  // exact: passes completed
    std::vector<uint32_t> task_pos;          // stream position of every decode task (ascending)
};

// One batch of a host-resident stream: stream positions [p0, p1) of the
// compressed part, whose records (and the readers' 8 look-ahead words) lie
// in words [w0, w0 + nw), and per heavy set the decode tasks [first,
// second) of those records.
struct StreamBatch {
    uint32_t p0 = 0, p1 = 0;
    uint64_t w0 = 0, nw = 0;
    std::vector<std::pair<uint32_t, uint32_t>> tasks;
};

// default bytes of one staging buffer of a host-resident stream
constexpr uint64_t DEFAULT_BATCH_BYTES = 1ull << 30;

// XCDs of an MI355X (gfx950): workgroups are dealt to them round robin
constexpr uint32_t XCDS = 8;

struct Bin {
    int kind;          // LPA_BIN_*
    int param;         // LANE: network size N; WAVE/BLOCK: log2 table slots
    bool csr;          // small-vertex CSR part
    uint32_t begin, end;
    uint32_t maxdeg;
};

inline uint32_t ceil_log2(uint64_t x) {
    uint32_t l = 0;
    while ((1ull << l) < x) l++;
    return l;
}

uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

// Device scratch of the allocation-free encoder (lpa_encode_csr_device): word
// 0 is the error flag, words 1.. the scan's block sums.  One copy per GPU,
// held until the process ends; its address on each device comes from
// hipGetSymbolAddress (the host-side name is only a stand-in).
constexpr uint64_t NOALLOC_SCRATCH_WORDS = 1ull << 17;   // 1 MiB
constexpr uint64_t NOALLOC_MAX_NC = (NOALLOC_SCRATCH_WORDS - 1) * SCAN_PER_BLOCK;
static __device__ __align__(256) uint64_t scratch[NOALLOC_SCRATCH_WORDS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint64_t *g_scratch_ptr[MAX_DEVICES] = {};

// The scratch address on `device`, resolved once per device; the device
// must be current.
uint64_t *device_scratch(int device) {
    if (device < 0 || device >= MAX_DEVICES) throw Fail(LPA_ERR_RANGE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw Fail(LPA_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[device] = static_cast<uint64_t *>(ptr);
    }
    return g_scratch_ptr[device];
}
}  // namespace

struct lpa_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t *scratch_dev = nullptr;   // this device's copy of `scratch`
    int num_cu = 0;
    size_t lds_max = 65536;

    // graph as uploaded (file order)
    uint64_t n = 0, n_c = 0, n_s = 0, nbits = 0, m_c = 0, m_s = 0;
    DevBuf<uint64_t> words0, offs0;
    DevBuf<uint32_t> vid0;
    DevBuf<uint4> meta0;                    // segment metadata per file-order compressed position
    std::vector<uint64_t> h_offs;           // file-order start offsets
    std::vector<uint64_t> h_loffs;          // start offsets of the scheduled (relayouted) stream
    std::vector<uint32_t> h_vertices;       // file order
    std::vector<uint64_t> h_soffs;          // small CSR (file order)
    std::vector<uint32_t> h_snbr;
    std::vector<uint32_t> deg_c;            // degree per file-order compressed position
    DevBuf<float> feat0;                    // features per file-order compressed position
    bool have_graph = false;

    // scheduled device layout
    DevBuf<uint64_t> words, offs, srow;
    DevBuf<uint32_t> vid, snbr, svid;
    DevBuf<uint4> meta;
    std::vector<Bin> bins;
    std::vector<uint32_t> order_c;          // device position -> file-order position
    std::vector<uint32_t> order_s;          // same for the small-vertex CSR
    uint64_t bin_counts[LPA_NBIN_KINDS] = {0};
    bool scheduled = false;
    std::vector<std::unique_ptr<HeavySet>> heavy;

    // LP state
    uint32_t K = 1;
    int semi = 0;
    uint32_t cms_thresh = 0, cms_depth = 4, cms_width = 1024;
    uint64_t cms_seed = 0;
    uint64_t cms_ab[16] = {0};
    DevBuf<uint32_t> L[2];
    int cur = 0;
    bool have_labels = false;
    DevBuf<unsigned long long> counter;
    unsigned long long *h_counter = nullptr;   // pinned

    // CGLP
    DevBuf<uint64_t> crow;
    DevBuf<uint32_t> ccol;
    bool have_csr = false;

    // host-resident stream (lpa_stream_residence): words0 and words in pinned
    // host memory; the fused and two-kernel steps copy the scheduled stream
    // batch by batch into stage[0] / stage[1] on copy_stream
    bool host_stream = false;
    uint64_t batch_bytes = DEFAULT_BATCH_BYTES;
    std::vector<StreamBatch> batches;
    DevBuf<uint64_t> stage[2];
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_copied[2] = {nullptr, nullptr}, ev_free[2] = {nullptr, nullptr};

    // timing of the last timed run, per bin kind
    float kind_ms[LPA_NBIN_KINDS] = {0};

    // kind of every vertex under the current schedule (by vertex id)
    std::vector<uint8_t> kind_of;
    lpa_sched_params sched_used{32, 512, 0, 1, 128};
    // work recorder (lane-cycles per vertex id), null when off
    unsigned long long *rec = nullptr;

    // captured steps: graphs[method][cur]
    bool graph_mode = false;
    hipGraphExec_t graphs[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
    void drop_graphs() {
        for (auto &m : graphs)
            for (auto &g : m)
                if (g) {
                    (void)hipGraphExecDestroy(g);
                    g = nullptr;
                }
    }

    ~lpa_engine() {
        drop_graphs();
        if (h_counter) (void)hipHostFree(h_counter);
        for (hipEvent_t x : {ev_copied[0], ev_copied[1], ev_free[0], ev_free[1]})
            if (x) (void)hipEventDestroy(x);
        if (copy_stream) (void)hipStreamDestroy(copy_stream);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void sync() { HIPCHK(hipStreamSynchronize(stream)); }

    StepArgs base_args(bool cglp) const {
        StepArgs a{};
        a.words = words.p;
        a.offs = offs.p;
        a.vid = vid.p;
        a.meta = meta.p;
        a.row = srow.p;
        a.nbr = snbr.p;
        a.cvid = svid.p;
        if (cglp) {
            a.row = crow.p;
            a.nbr = ccol.p;
            a.cvid = vid.p;
        }
        a.lin = L[cur].p;
        a.lout = L[cur ^ 1].p;
        a.K = K;
        a.cms_depth = cms_depth;
        a.cms_logw = ceil_log2(cms_width);
        std::memcpy(a.cms_ab, cms_ab, sizeof cms_ab);
        a.changed = counter.p;
        a.rec = rec;
        a.xcds = sched_used.xcd_tiles == 1 ? XCDS : 1;
        return a;
    }

    void launch_bin(const Bin &b, StepArgs a, bool as_csr);
    HeavyArgs heavy_args(const HeavySet &hs, const StepArgs &a, bool as_csr) const;
    void launch_heavy_decode(const HeavySet &hs, const StepArgs &a, uint32_t t0, uint32_t t1);
    void launch_heavy(HeavySet &hs, const StepArgs &a, bool as_csr, bool decode = true);
    template <class F> void for_each_batch(F &&fn);
    void step_host_stream();
    void step(bool cglp, hipEvent_t *launch_events);
    void run_heavy(bool cglp, hipEvent_t *launch_events);
    void step_bin_times(std::vector<float> &bin_ms, std::vector<float> &heavy_ms);
};

// ---------------------------------------------------------------- launches

template <bool CSR> static void launch_lane(int N, StepArgs a, uint32_t count, hipStream_t s) {
    dim3 grid((count + BLOCK - 1) / BLOCK);
    switch (N) {
    case 4: hipLaunchKernelGGL((k_lane<4, CSR>), grid, dim3(BLOCK), 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_lane<8, CSR>), grid, dim3(BLOCK), 0, s, a); break;
    case 16: hipLaunchKernelGGL((k_lane<16, CSR>), grid, dim3(BLOCK), 0, s, a); break;
    case 32: hipLaunchKernelGGL((k_lane<32, CSR>), grid, dim3(BLOCK), 0, s, a); break;
    default: hipLaunchKernelGGL((k_lane<64, CSR>), grid, dim3(BLOCK), 0, s, a); break;
    }
}

// LDS bytes of the WAVE / BLOCK kernels' per-group table (2^logT slots) and
// neighbour list (2^logT / 2 ids), for NG groups per workgroup.
static size_t group_lds(int ng, uint32_t logT) { return (size_t)ng * 10ull * (1ull << logT); }
static int wave_groups(uint32_t logT, uint32_t quarter_max = 128) {
    return (1u << (logT - 1)) <= quarter_max ? 16 : 4;
}

// label vectors staged per row gather by the group kernels: 8 from K = 5
// (one gather per neighbour and step at K = 8; with one aggregated match
// round the kernels stay fast enough at the lower occupancy that even label
// rows held by the 256 MB last-level cache gain from it), else K rounded
// down to 4 / 2 / 1.
static int group_kb(uint32_t K, uint64_t n) {
    (void)n;
    return K >= 5 ? 8 : K >= 3 ? 4 : K == 2 ? 2 : 1;
}

// instance of the group kernel for a batch size and staged ids per lane
template <int G, bool CSR> static const void *group_fn(int kb, uint32_t ipl) {
    switch (kb) {
    case 1: return (const void *)k_group<G, CSR, 1, 8>;
    case 2: return (const void *)k_group<G, CSR, 2, 8>;
    case 4:
        return ipl <= 2 ? (const void *)k_group<G, CSR, 4, 2>
             : ipl <= 4 ? (const void *)k_group<G, CSR, 4, 4> : (const void *)k_group<G, CSR, 4, 8>;
    default:
        return ipl <= 2 ? (const void *)k_group<G, CSR, 8, 2>
             : ipl <= 4 ? (const void *)k_group<G, CSR, 8, 4> : (const void *)k_group<G, CSR, 8, 8>;
    }
}

template <int G>
static void launch_group(bool as_csr, uint32_t K, uint64_t n, dim3 grid, size_t lds, hipStream_t s,
                         const StepArgs &a, uint32_t logT, uint32_t cap) {
    const uint32_t ipl = (cap + G - 1) / G;
    // quarter-wave and wave groups hold few labels per lane: all 8 vectors
    // of a batch fit their registers, and one gather per row wins
    const int kb = G <= 64 && K >= 5 ? 8 : group_kb(K, n);
    const void *fn = as_csr ? group_fn<G, true>(kb, ipl) : group_fn<G, false>(kb, ipl);
    StepArgs args = a;
    void *params[] = {&args, &logT, &cap};
    (void)hipLaunchKernel(fn, grid, dim3(BLOCK), params, lds, s);
}

template <int G> static void allow_group_lds(int dyn) {
    for (int kb : {1, 2, 4, 8})
        for (uint32_t ipl : {2u, 4u, 8u})
            for (const void *f : {group_fn<G, false>(kb, ipl), group_fn<G, true>(kb, ipl)})
                (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
}

void lpa_engine::launch_bin(const Bin &b, StepArgs a, bool as_csr) {
    a.begin = b.begin;
    a.end = b.end;
    const uint32_t count = b.end - b.begin;
    if (count == 0) return;
    switch (b.kind) {
    case LPA_BIN_LANE:
        if (as_csr) launch_lane<true>(b.param, a, count, stream);
        else launch_lane<false>(b.param, a, count, stream);
        break;
    case LPA_BIN_WAVE: {
        // a quarter wave per vertex up to degree 128, a whole wave above
        const uint32_t logT = (uint32_t)b.param, cap = 1u << (logT - 1);
        int ng = wave_groups(logT, sched_used.quarter_max ? sched_used.quarter_max : 128);
        // with 8 vectors per staged gather, lists of 65..128 ids take half
        // waves (4 ids per lane) rather than quarter waves (8 per lane):
        // half the staged registers, twice the waves resident
        if (ng == 16 && cap > 64 && K >= 5) ng = 8;
        const size_t lds = group_lds(ng, logT);
        dim3 grid((count + ng - 1) / ng);
        if (ng == 16) launch_group<16>(as_csr, K, n, grid, lds, stream, a, logT, cap);
        else if (ng == 8) launch_group<32>(as_csr, K, n, grid, lds, stream, a, logT, cap);
        else launch_group<64>(as_csr, K, n, grid, lds, stream, a, logT, cap);
        break;
    }
    case LPA_BIN_BLOCK: {
        const uint32_t logT = (uint32_t)b.param, cap = 1u << (logT - 1);
        launch_group<BLOCK>(as_csr, K, n, dim3(count), group_lds(1, logT), stream, a, logT, cap);
        break;
    }
    default: break;   // CMS and GLOBAL bins run through launch_heavy
    }
}

// One-workgroup CMS kernel: label vectors per pass and its LDS bytes.  One
// vector keeps 32-bit cells; several take 16-bit cells, two vectors to a
// word, as many pairs as the LDS holds (at most HEAVY_KB_CMS_LOCAL vectors).
static uint32_t cms_local_kb(uint32_t K, size_t cells, size_t lds) {
    if (K == 1) return 1;
    const size_t pairs = std::max<size_t>(1, lds / (cells * sizeof(uint32_t)));
    return (uint32_t)std::min<size_t>({(size_t)K, (size_t)HEAVY_KB_CMS_LOCAL, 2 * pairs});
}

// Row stashes for the exact one-workgroup kernel (degrees up to CHUNK) from
// this many label vectors up (below it a row is a line or less and the
// batches are few); LPA_ROW_STASH=0 turns them off.
constexpr uint32_t ROW_STASH_MIN_K = 16;

static bool row_stash_on(const lpa_engine *e) {
    const char *env = std::getenv("LPA_ROW_STASH");
    return e->K >= ROW_STASH_MIN_K && !(env && std::strcmp(env, "0") == 0);
}

// words of a row stash per neighbour: K rounded up to whole groups of four
static uint64_t row_stash_k(const lpa_engine *e) { return (e->K + 3ull) / 4 * 4; }

static size_t cms_local_lds(uint32_t kbl, size_t cells) {
    return (kbl == 1 ? 1 : (kbl + 1) / 2) * cells * sizeof(uint32_t);
}

HeavyArgs lpa_engine::heavy_args(const HeavySet &hs, const StepArgs &a, bool as_csr) const {
    HeavyArgs h{};
    h.s = a;
    h.nh = hs.nh;
    h.pos = hs.pos.p;
    h.loff = hs.loff.p;
    h.list = hs.list.p;
    h.csr = as_csr;
    h.toff = hs.toff.p;
    h.tab = hs.tab.p;
    h.sketch = hs.sketch.p;
    h.result = hs.result.p;
    h.hcnt = hs.hcnt.p;
    h.nchunks = hs.nchunks;
    h.chunk_h = hs.chunk_h.p;
    h.chunk_b = hs.chunk_b.p;
    h.cms = hs.cms ? 1 : 0;
    h.slot = hs.slot.p;
    h.epoch = hs.epoch.p;
    h.tagged = hs.tagged ? 1 : 0;
    h.kb_alloc = hs.kb;
    // CMS: sketches of up to KB label vectors side by side in LDS; both CMS
    // kernels pack two vectors' 16-bit LDS cells in a word when K > 1 (a
    // list of <= CMS_LOCAL ids, or a chunk, cannot overflow one)
    h.kbs = cms_local_kb(K, (size_t)cms_depth * cms_width, lds_max - 20480);
    h.kbl = h.kbs;
    return h;
}

// Decode tasks [t0, t1) of a heavy set (compressed part) into its lists.
void lpa_engine::launch_heavy_decode(const HeavySet &hs, const StepArgs &a, uint32_t t0, uint32_t t1) {
    if (t1 <= t0) return;
    const HeavyArgs h = heavy_args(hs, a, false);
    if (hs.cms) hipLaunchKernelGGL(k_heavy_decode<1>, dim3(t1 - t0), dim3(BLOCK), 0, stream, h, hs.tasks.p + t0, t1 - t0);
    else hipLaunchKernelGGL(k_heavy_decode<0>, dim3(t1 - t0), dim3(BLOCK), 0, stream, h, hs.tasks.p + t0, t1 - t0);
}

// A heavy set's step: decode its records into the lists (unless `decode` is
// off: a host-resident stream decodes them batch by batch), then count.
void lpa_engine::launch_heavy(HeavySet &hs, const StepArgs &a, bool as_csr, bool decode) {
    if (hs.nh == 0) return;
    HeavyArgs h = heavy_args(hs, a, as_csr);
    const bool one = K == 1;
    const size_t cells = (size_t)cms_depth * cms_width;
    if (!as_csr && decode) launch_heavy_decode(hs, a, 0, hs.ntasks);
    if (hs.nlocal) {
        const dim3 b(HBLOCK);
        const size_t sk = cms_local_lds(h.kbl, cells);
        const uint32_t nsmall = hs.nlocal - hs.nlocal_big;
        const uint32_t *lsmall = hs.local.p, *lbig = hs.local.p + nsmall;
        if (hs.cms) {
            // resident workgroups loop over the vertices, each with its stash
            h.stash = hs.stash.p;
            h.stash_stride = (uint64_t)CMS_LOCAL * (one ? 1 : HEAVY_KB_CMS_LOCAL);
            const dim3 g(std::min(hs.nlocal, hs.stash_blocks));
            if (one) hipLaunchKernelGGL((k_heavy_local<1, 1, 4>), g, b, sk, stream, h, hs.local.p, hs.nlocal);
            else hipLaunchKernelGGL((k_heavy_local<1, HEAVY_KB_CMS_LOCAL, 4>), g, b, sk, stream, h, hs.local.p, hs.nlocal);
        } else {
            // tables of up to 2*CHUNK (8192 ids: 4 per lane) or 4*CHUNK slots (8 per lane);
            // with a row stash, resident workgroups take the vertices one by one
            const size_t t1 = 2 * CHUNK * sizeof(uint64_t), t2 = 4 * CHUNK * sizeof(uint64_t);
            if (nsmall) {
                h.stash = hs.rstash_small ? hs.stash.p : nullptr;
                h.stash_stride = (uint64_t)CHUNK * row_stash_k(this);
                h.wq = hs.wq.p;
                const dim3 g(hs.rstash_small ? hs.rstash_small : nsmall);
                if (one) hipLaunchKernelGGL((k_heavy_local<0, 1, 4>), g, b, t1, stream, h, lsmall, nsmall);
                else if (h.stash) hipLaunchKernelGGL((k_heavy_local<0, HEAVY_KB_LOCAL, 4, true>), g, b, t1, stream, h, lsmall, nsmall);
                else hipLaunchKernelGGL((k_heavy_local<0, HEAVY_KB_LOCAL, 4>), g, b, t1, stream, h, lsmall, nsmall);
            }
            h.stash = nullptr;
            if (hs.nlocal_big) {
                const dim3 g(hs.nlocal_big);
                if (one) hipLaunchKernelGGL((k_heavy_local<0, 1, 8>), g, b, t2, stream, h, lbig, hs.nlocal_big);
                else hipLaunchKernelGGL((k_heavy_local<0, HEAVY_KB_MAX, 8>), g, b, t2, stream, h, lbig, hs.nlocal_big);
            }
        }
    }
    if (hs.nchunked == 0) return;
    for (uint32_t k0 = 0; k0 < K; k0 += hs.kb) {
        h.k0 = k0;
        h.kb = std::min(hs.kb, K - k0);
        if (hs.cms) hipLaunchKernelGGL(k_heavy_clear<1>, dim3(hs.nchunked), dim3(BLOCK), 0, stream, h, hs.chunked.p, hs.nchunked);
        else hipLaunchKernelGGL(k_heavy_clear<0>, dim3(hs.nchunked), dim3(BLOCK), 0, stream, h, hs.chunked.p, hs.nchunked);
        const dim3 g(hs.nchunks), b(HBLOCK);
        if (hs.cms) {
            const size_t sk = cms_local_lds(std::min(h.kbs, h.kb), cells);
            if (one) {
                hipLaunchKernelGGL((k_heavy_chunks<1, 1>), g, b, sk, stream, h);
                hipLaunchKernelGGL((k_heavy_chunks<2, 1>), g, b, 0, stream, h);
            } else {
                hipLaunchKernelGGL((k_heavy_chunks<1, HEAVY_KB_CMS_LOCAL>), g, b, sk, stream, h);
                hipLaunchKernelGGL((k_heavy_chunks<2, HEAVY_KB_CMS_LOCAL>), g, b, 0, stream, h);
            }
        } else {
            const size_t tb = 2 * CHUNK * sizeof(uint64_t);
            if (one) hipLaunchKernelGGL((k_heavy_chunks<0, 1>), g, b, tb, stream, h);
            else hipLaunchKernelGGL((k_heavy_chunks<0, HEAVY_KB_MAX>), g, b, tb, stream, h);
        }
        const uint64_t fin = (uint64_t)hs.nchunked * h.kb;
        const dim3 fg((unsigned)((fin + BLOCK - 1) / BLOCK));
        if (hs.cms) hipLaunchKernelGGL(k_heavy_finish<1>, fg, dim3(BLOCK), 0, stream, h, hs.chunked.p, hs.nchunked);
        else hipLaunchKernelGGL(k_heavy_finish<0>, fg, dim3(BLOCK), 0, stream, h, hs.chunked.p, hs.nchunked);
    }
}

// One synchronous step: clear the counter, every bin reads L[cur] and writes
// L[cur^1].  Vertices without neighbours are never written (both buffers
// hold their initial label).  launch_events (or null): 2 * (bins + heavy
// sets) events, recorded around bin i at [2i, 2i+1] and around heavy set j
// at [2(nbins + j), ...], so a family's time is the sum over its launches.
void lpa_engine::step(bool cglp, hipEvent_t *launch_events) {
    if (host_stream && !cglp) {
        step_host_stream();
        return;
    }
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, stream, counter.p, (uint64_t)1, 0ull);
    StepArgs a = base_args(cglp);
    for (size_t i = 0; i < bins.size(); i++) {
        const Bin &b = bins[i];
        if (b.kind == LPA_BIN_SKIP || b.kind == LPA_BIN_CMS || b.kind == LPA_BIN_GLOBAL || b.end == b.begin) continue;
        if (launch_events) HIPCHK(hipEventRecord(launch_events[2 * i], stream));
        launch_bin(b, a, b.csr || cglp);
        if (launch_events) HIPCHK(hipEventRecord(launch_events[2 * i + 1], stream));
    }
    run_heavy(cglp, launch_events);
    HIPCHK(hipGetLastError());
}

void lpa_engine::run_heavy(bool cglp, hipEvent_t *launch_events) {
    for (size_t j = 0; j < heavy.size(); j++) {
        HeavySet &hs = *heavy[j];
        if (hs.nh == 0) continue;
        hipEvent_t *ev = launch_events ? launch_events + 2 * (bins.size() + j) : nullptr;
        if (ev) HIPCHK(hipEventRecord(ev[0], stream));
        const StepArgs a = base_args(cglp && !hs.csr_part);
        launch_heavy(hs, a, hs.csr_part || cglp);
        if (ev) HIPCHK(hipEventRecord(ev[1], stream));
    }
}

// Runs fn(batch, words) for every batch of a host-resident stream, in
// stream order on the engine stream.  Batch i is copied into stage[i % 2]
// on the copy stream once the kernels of batch i - 2 (the buffer's last
// user) are done, so the copy of batch i + 1 overlaps batch i's kernels;
// `words` is the staging buffer rebased by the batch's first word, so the
// kernels address the batch's records by their absolute stream positions.
template <class F> void lpa_engine::for_each_batch(F &&fn) {
    for (size_t i = 0; i < batches.size(); i++) {
        const StreamBatch &b = batches[i];
        const int s = (int)(i & 1);
        HIPCHK(hipStreamWaitEvent(copy_stream, ev_free[s], 0));
        HIPCHK(hipMemcpyAsync(stage[s].p, words.p + b.w0, b.nw * sizeof(uint64_t), hipMemcpyHostToDevice,
                              copy_stream));
        HIPCHK(hipEventRecord(ev_copied[s], copy_stream));
        HIPCHK(hipStreamWaitEvent(stream, ev_copied[s], 0));
        const auto *w = reinterpret_cast<const uint64_t *>(reinterpret_cast<uintptr_t>(stage[s].p) -
                                                           b.w0 * sizeof(uint64_t));
        fn(b, w);
        HIPCHK(hipEventRecord(ev_free[s], stream));
    }
}

// The fused step over a host-resident stream: per batch, every bin's share
// of the batch's positions and the decode tasks of its heavy records; then
// the small CSR's bins and the heavy counting, which read only HBM.  Every
// launch reads L[cur] and writes L[cur ^ 1], so the order of the launches
// changes nothing.
void lpa_engine::step_host_stream() {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, stream, counter.p, (uint64_t)1, 0ull);
    const StepArgs a = base_args(false);
    for_each_batch([&](const StreamBatch &sb, const uint64_t *w) {
        StepArgs ab = a;
        ab.words = w;
        for (const Bin &b : bins) {
            if (b.csr || b.kind == LPA_BIN_SKIP || b.kind == LPA_BIN_CMS || b.kind == LPA_BIN_GLOBAL) continue;
            Bin c = b;
            c.begin = std::max(b.begin, sb.p0);
            c.end = std::min(b.end, sb.p1);
            if (c.begin < c.end) launch_bin(c, ab, false);
        }
        for (size_t j = 0; j < heavy.size(); j++)
            if (!heavy[j]->csr_part) launch_heavy_decode(*heavy[j], ab, sb.tasks[j].first, sb.tasks[j].second);
    });
    for (const Bin &b : bins)
        if (b.csr && b.kind != LPA_BIN_SKIP && b.kind != LPA_BIN_CMS && b.kind != LPA_BIN_GLOBAL) launch_bin(b, a, true);
    for (auto &hp : heavy) launch_heavy(*hp, a, hp->csr_part, false);
    HIPCHK(hipGetLastError());
}

// One fused step like step(false, ...), timing every bin launch and every
// heavy set with HIP events (ms; 0 where nothing ran).
void lpa_engine::step_bin_times(std::vector<float> &bin_ms, std::vector<float> &heavy_ms) {
    const size_t nb = bins.size(), nhs = heavy.size();
    bin_ms.assign(nb, 0.0f);
    heavy_ms.assign(nhs, 0.0f);
    std::vector<hipEvent_t> ev(2 * (nb + nhs), nullptr);
    std::vector<char> used(nb + nhs, 0);
    try {
        for (auto &x : ev) HIPCHK(hipEventCreate(&x));
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, stream, counter.p, (uint64_t)1, 0ull);
        const StepArgs a = base_args(false);
        for (size_t i = 0; i < nb; i++) {
            const Bin &b = bins[i];
            if (b.kind == LPA_BIN_SKIP || b.kind == LPA_BIN_CMS || b.kind == LPA_BIN_GLOBAL || b.end == b.begin) continue;
            HIPCHK(hipEventRecord(ev[2 * i], stream));
            launch_bin(b, a, b.csr);
            HIPCHK(hipEventRecord(ev[2 * i + 1], stream));
            used[i] = 1;
        }
        for (size_t j = 0; j < nhs; j++) {
            HeavySet &hs = *heavy[j];
            if (hs.nh == 0) continue;
            HIPCHK(hipEventRecord(ev[2 * (nb + j)], stream));
            launch_heavy(hs, base_args(false), hs.csr_part);
            HIPCHK(hipEventRecord(ev[2 * (nb + j) + 1], stream));
            used[nb + j] = 1;
        }
        HIPCHK(hipGetLastError());
        sync();
        for (size_t i = 0; i < nb + nhs; i++) {
            if (!used[i]) continue;
            float t = 0;
            HIPCHK(hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]));
            (i < nb ? bin_ms[i] : heavy_ms[i - nb]) = t;
        }
    } catch (...) {
        for (auto &x : ev)
            if (x) (void)hipEventDestroy(x);
        throw;
    }
    for (auto &x : ev) (void)hipEventDestroy(x);
}

// Validation pass over the device stream (words0, offs0, vid0 and the host
// copies set): degrees, features, segment metadata; rejects malformed
// records.
static void finish_graph(lpa_engine *e) {
    const uint64_t n_c = e->n_c;
    DevBuf<uint32_t> ddeg;
    ddeg.alloc(n_c);
    e->feat0.alloc(n_c * NFEAT);
    e->meta0.alloc(n_c);
    e->counter.alloc(1);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, e->stream, e->counter.p, (uint64_t)1, 0ull);
    if (n_c)
        hipLaunchKernelGGL(k_scan_records, dim3((unsigned)((n_c + 255) / 256)), dim3(256), 0, e->stream, e->words0.p,
                           e->offs0.p, e->vid0.p, n_c, (uint32_t)e->n, ddeg.p, e->feat0.p, e->meta0.p, e->counter.p);
    HIPCHK(hipGetLastError());
    unsigned long long bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, e->counter.p, sizeof bad, hipMemcpyDeviceToHost, e->stream));
    e->deg_c.assign(n_c, 0);
    if (n_c) HIPCHK(hipMemcpyAsync(e->deg_c.data(), ddeg.p, n_c * 4, hipMemcpyDeviceToHost, e->stream));
    e->sync();
    if (bad) throw Fail(LPA_ERR_FORMAT, "malformed CGR record at stream position " + std::to_string(bad - 1));
    e->m_c = 0;
    for (uint32_t d : e->deg_c) e->m_c += d;
    e->have_graph = true;
}

// ------------------------------------------------------------------ C ABI

static uint64_t lds_budget(const lpa_engine *e);
static uint32_t wave_cap(const lpa_engine *e);
static uint32_t block_cap(const lpa_engine *e);
static void move_stream(lpa_engine *e, DevBuf<uint64_t> &b, bool to_host);
static void plan_batches(lpa_engine *e);

extern "C" {

const char *lpa_last_error(void) { return g_err.c_str(); }

int lpa_device_count(int *count) {
    return guard([&] {
        if (!count) throw Fail(LPA_ERR_ARG, "null argument");
        HIPCHK(hipGetDeviceCount(count));
    });
}

int lpa_engine_create(int device, lpa_engine **out) {
    return guard([&] {
        if (!out) throw Fail(LPA_ERR_ARG, "null argument");
        auto *e = new lpa_engine;
        try {
            e->device = device;
            HIPCHK(hipSetDevice(device));
            hipDeviceProp_t prop;
            HIPCHK(hipGetDeviceProperties(&prop, device));
            e->num_cu = prop.multiProcessorCount;
            e->lds_max = prop.sharedMemPerBlock;
            HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
            e->scratch_dev = device_scratch(device);
            e->counter.alloc(1);
            HIPCHK(hipHostMalloc((void **)&e->h_counter, sizeof(unsigned long long), hipHostMallocDefault));
            const int dyn = (int)(e->lds_max - 20480);
            allow_group_lds<16>(dyn);
            allow_group_lds<32>(dyn);
            allow_group_lds<64>(dyn);
            allow_group_lds<BLOCK>(dyn);
            const void *heavy_fns[] = {(const void *)k_heavy_chunks<0, 1>, (const void *)k_heavy_chunks<0, HEAVY_KB_MAX>,
                                       (const void *)k_heavy_chunks<1, 1>, (const void *)k_heavy_chunks<1, HEAVY_KB_CMS_LOCAL>,
                                       (const void *)k_heavy_local<0, 1, 4>, (const void *)k_heavy_local<0, HEAVY_KB_LOCAL, 4>,
                                       (const void *)k_heavy_local<0, 1, 8>, (const void *)k_heavy_local<0, HEAVY_KB_MAX, 8>,
                                       (const void *)k_heavy_local<0, HEAVY_KB_LOCAL, 4, true>,
                                       (const void *)k_heavy_local<1, 1, 4>, (const void *)k_heavy_local<1, HEAVY_KB_CMS_LOCAL, 4>};
            for (const void *f : heavy_fns) (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
            (void)hipGetLastError();
        } catch (...) {
            delete e;
            throw;
        }
        *out = e;
    });
}

int lpa_engine_info(lpa_engine *e, uint64_t *info) {
    return guard([&] {
        if (!e || !info) throw Fail(LPA_ERR_ARG, "null argument");
        info[0] = (uint64_t)e->num_cu;
        info[1] = (uint64_t)e->lds_max;
        info[2] = wave_cap(e);
        info[3] = block_cap(e);
    });
}

int lpa_engine_destroy(lpa_engine *e) {
    return guard([&] {
        if (!e) return;
        (void)hipSetDevice(e->device);
        (void)hipStreamSynchronize(e->stream);
        delete e;
    });
}

int lpa_graph_upload(lpa_engine *e, uint64_t n, uint64_t n_c, const uint32_t *words, uint64_t nwords,
                     const uint64_t *offsets, const uint32_t *vertices, const uint64_t *small_offsets,
                     const uint32_t *small_nbrs) {
    return guard([&] {
        if (!e || !offsets || !vertices || (nwords && !words)) throw Fail(LPA_ERR_ARG, "null argument");
        if (n >= (1ull << 30)) throw Fail(LPA_ERR_RANGE, "n must be < 2^30");
        if (n_c > n) throw Fail(LPA_ERR_ARG, "n_c > n");
        if (n_c >= (1ull << 32) - 1) throw Fail(LPA_ERR_RANGE, "too many compressed vertices");
        HIPCHK(hipSetDevice(e->device));
        e->drop_graphs();
        e->have_graph = e->scheduled = e->have_labels = e->have_csr = false;
        if (offsets[0] != 0) throw Fail(LPA_ERR_FORMAT, "offsets[0] != 0");
        for (uint64_t i = 0; i < n_c; i++)
            if (offsets[i + 1] < offsets[i]) throw Fail(LPA_ERR_FORMAT, "offsets decrease");
        const uint64_t nbits = offsets[n_c];
        if ((nbits + 31) / 32 > nwords) throw Fail(LPA_ERR_FORMAT, "stream shorter than the offsets");
        std::vector<uint8_t> seen(n, 0);
        for (uint64_t i = 0; i < n; i++) {
            if (vertices[i] >= n || seen[vertices[i]]) throw Fail(LPA_ERR_FORMAT, "vertices is not a permutation of 0..n-1");
            seen[vertices[i]] = 1;
        }
        const uint64_t n_s = n - n_c;
        e->h_soffs.assign(n_s + 1, 0);
        e->h_snbr.clear();
        if (n_s) {
            if (!small_offsets) throw Fail(LPA_ERR_ARG, "small vertices present but no small CSR");
            if (small_offsets[0] != 0) throw Fail(LPA_ERR_FORMAT, "small_offsets[0] != 0");
            for (uint64_t i = 0; i < n_s; i++)
                if (small_offsets[i + 1] < small_offsets[i]) throw Fail(LPA_ERR_FORMAT, "small offsets decrease");
            const uint64_t ms = small_offsets[n_s];
            if (ms && !small_nbrs) throw Fail(LPA_ERR_ARG, "small_nbrs is null");
            for (uint64_t i = 0; i < ms; i++)
                if (small_nbrs[i] >= n) throw Fail(LPA_ERR_FORMAT, "small neighbour id >= n");
            e->h_soffs.assign(small_offsets, small_offsets + n_s + 1);
            e->h_snbr.assign(small_nbrs, small_nbrs + ms);
        }
        e->n = n;
        e->n_c = n_c;
        e->n_s = n_s;
        e->nbits = nbits;
        e->m_s = e->h_snbr.size();
        e->h_offs.assign(offsets, offsets + n_c + 1);
        e->h_vertices.assign(vertices, vertices + n);

        const uint64_t n64 = (nbits + 63) / 64 + 8;
        const uint64_t n32 = std::min<uint64_t>(nwords, (nbits + 31) / 32);
        if (e->host_stream) {
            // packed on the host straight into pinned memory: the stream
            // never occupies HBM
            e->words0.alloc_host(n64);
            for (uint64_t i = 0; i < n64; i++) {
                const uint64_t hi = 2 * i < n32 ? words[2 * i] : 0, lo = 2 * i + 1 < n32 ? words[2 * i + 1] : 0;
                e->words0.p[i] = (hi << 32) | lo;
            }
        } else {
            DevBuf<uint32_t> w32;
            w32.upload(words, nwords, e->stream);
            e->words0.alloc(n64);
            hipLaunchKernelGGL(k_pack64, dim3((unsigned)((n64 + 255) / 256)), dim3(256), 0, e->stream, w32.p, n32,
                               e->words0.p, n64);
            e->sync();
        }
        e->offs0.upload(offsets, n_c + 1, e->stream);
        e->vid0.upload(vertices, n_c, e->stream);
        finish_graph(e);
    });
}

int lpa_graph_encode_csr(lpa_engine *e, uint64_t n, const uint64_t *row, const uint32_t *col,
                         uint32_t small_thresh) {
    return guard([&] {
        if (!e || !row || (row[n] && !col)) throw Fail(LPA_ERR_ARG, "null argument");
        if (n >= (1ull << 30)) throw Fail(LPA_ERR_RANGE, "n must be < 2^30");
        if (row[0] != 0) throw Fail(LPA_ERR_FORMAT, "row[0] != 0");
        for (uint64_t v = 0; v < n; v++)
            if (row[v + 1] < row[v]) throw Fail(LPA_ERR_FORMAT, "row offsets decrease");
        HIPCHK(hipSetDevice(e->device));
        e->drop_graphs();
        e->have_graph = e->scheduled = e->have_labels = e->have_csr = false;
        const uint64_t m = row[n];
        // stream order (compressor.cpp:157-290): deg >= 128 by id, then
        // small_thresh <= deg < 128 by id; deg < small_thresh -> small CSR
        const uint32_t large = 128;
        std::vector<uint32_t> vert;
        vert.reserve(n);
        auto deg = [&](uint64_t v) { return row[v + 1] - row[v]; };
        for (uint64_t v = 0; v < n; v++) if (deg(v) >= large) vert.push_back((uint32_t)v);
        for (uint64_t v = 0; v < n; v++) if (deg(v) >= small_thresh && deg(v) < large) vert.push_back((uint32_t)v);
        const uint64_t n_c = vert.size();
        for (uint64_t v = 0; v < n; v++) if (deg(v) < small_thresh) vert.push_back((uint32_t)v);
        const uint64_t n_s = n - n_c;
        e->h_soffs.assign(1, 0);
        e->h_snbr.clear();
        for (uint64_t j = n_c; j < n; j++) {
            const uint32_t v = vert[j];
            for (uint64_t k = row[v]; k < row[v + 1]; k++) {
                if (col[k] >= n) throw Fail(LPA_ERR_FORMAT, "neighbour id >= n");
                e->h_snbr.push_back(col[k]);
            }
            e->h_soffs.push_back(e->h_snbr.size());
        }
        e->n = n;
        e->n_c = n_c;
        e->n_s = n_s;
        e->m_s = e->h_snbr.size();
        e->h_vertices = vert;
        // encode on the device
        DevBuf<uint64_t> drow, len, offs;
        DevBuf<uint32_t> dcol;
        drow.upload(row, n + 1, e->stream);
        dcol.upload(col, m, e->stream);
        e->vid0.upload(vert.data(), n_c, e->stream);
        len.alloc(n_c);
        e->offs0.alloc(n_c + 1);
        e->counter.alloc(1);
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, e->stream, e->counter.p, (uint64_t)1, 0ull);
        if (n_c)
            hipLaunchKernelGGL(k_encode_len, dim3((unsigned)((n_c + 255) / 256)), dim3(256), 0, e->stream, drow.p,
                               dcol.p, e->vid0.p, n_c, (uint32_t)n, len.p, e->counter.p, enc::LONG_ROW);
        // rows longer than LONG_ROW: one workgroup each, with scratch of
        // d + 1 entries per row for every per-entry array (44 bytes each),
        // in groups of rows whose scratch stays within 2^28 entries
        // (LPA_LONG_SCRATCH overrides the limit: tests force several groups)
        uint64_t LONG_SCRATCH = 1ull << 28;
        if (const char *ls = std::getenv("LPA_LONG_SCRATCH")) LONG_SCRATCH = std::max<uint64_t>(1, std::strtoull(ls, nullptr, 10));
        struct LongGroup {
            std::vector<uint32_t> pos;
            std::vector<uint64_t> base{0};
        };
        std::vector<LongGroup> lgroups;
        for (uint64_t p = 0; p < n_c; p++) {
            const uint64_t d = deg(vert[p]);
            if (d <= enc::LONG_ROW) continue;
            if (lgroups.empty() || (lgroups.back().base.back() + d + 1 > LONG_SCRATCH && !lgroups.back().pos.empty()))
                lgroups.emplace_back();
            lgroups.back().pos.push_back((uint32_t)p);
            lgroups.back().base.push_back(lgroups.back().base.back() + d + 1);
        }
        DevBuf<uint32_t> dlpos, lleft, llen, lrid, lnxt_i, lnxt_r, lseg_i, lseg_r;
        DevBuf<uint64_t> dlbase, lP_i, lP_r;
        DevBuf<enc::LongInfo> linfo;
        enc::LongScratch lsc{};
        // the scratch of group g: its rows' entries, segment ends and chains
        auto long_group = [&](const LongGroup &G) {
            dlpos.upload(G.pos.data(), G.pos.size(), e->stream);
            dlbase.upload(G.base.data(), G.base.size(), e->stream);
            lsc = enc::LongScratch{dlbase.p, lleft.p, llen.p, lrid.p, lnxt_i.p, lnxt_r.p, lseg_i.p, lseg_r.p,
                                   lP_i.p, lP_r.p, linfo.p};
            hipLaunchKernelGGL(k_encode_long_len, dim3((unsigned)G.pos.size()), dim3(enc::LONG_LANES), 0, e->stream,
                               drow.p, dcol.p, e->vid0.p, dlpos.p, (uint32_t)n, lsc, len.p, e->counter.p);
        };
        if (!lgroups.empty()) {
            uint64_t tot = 0, rows = 0;
            for (const auto &G : lgroups) {
                tot = std::max(tot, G.base.back());
                rows = std::max<uint64_t>(rows, G.pos.size());
            }
            for (DevBuf<uint32_t> *b : {&lleft, &llen, &lrid, &lnxt_i, &lnxt_r, &lseg_i, &lseg_r}) b->alloc(tot);
            lP_i.alloc(tot);
            lP_r.alloc(tot);
            linfo.alloc(rows);
            for (const auto &G : lgroups) {
                long_group(G);
                e->sync();   // the next group reuses the scratch
            }
        }
        const uint64_t nb = std::max<uint64_t>(1, (n_c + SCAN_PER_BLOCK - 1) / SCAN_PER_BLOCK);
        DevBuf<uint64_t> sums;
        sums.alloc(nb);
        hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nb), dim3(256), 0, e->stream, len.p, n_c, sums.p);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, e->stream, sums.p, nb);
        hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(256), 0, e->stream, len.p, n_c, sums.p,
                           e->offs0.p, nb);
        HIPCHK(hipGetLastError());
        unsigned long long bad = 0;
        HIPCHK(hipMemcpyAsync(&bad, e->counter.p, sizeof bad, hipMemcpyDeviceToHost, e->stream));
        e->h_offs.assign(n_c + 1, 0);
        HIPCHK(hipMemcpyAsync(e->h_offs.data(), e->offs0.p, (n_c + 1) * 8, hipMemcpyDeviceToHost, e->stream));
        e->sync();
        if (bad) throw Fail(LPA_ERR_FORMAT, "row of vertex " + std::to_string(bad - 1) +
                                                " is not strictly increasing or has an id >= n");
        e->nbits = e->h_offs[n_c];
        const uint64_t n64 = (e->nbits + 63) / 64 + 8;
        e->words0.alloc(n64);
        hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((n64 + 255) / 256)), dim3(256), 0, e->stream,
                           (unsigned long long *)e->words0.p, n64, 0ull);
        if (n_c)
            hipLaunchKernelGGL(k_encode_write, dim3((unsigned)((n_c + 255) / 256)), dim3(256), 0, e->stream, drow.p,
                               dcol.p, e->vid0.p, n_c, e->offs0.p, e->words0.p, enc::LONG_ROW);
        for (const auto &G : lgroups) {
            // one group: its scratch is still in place; several: each is
            // measured again before it is written (the same lengths result)
            if (lgroups.size() > 1) long_group(G);
            hipLaunchKernelGGL(k_encode_long_write, dim3((unsigned)G.pos.size()), dim3(enc::LONG_LANES), 0,
                               e->stream, e->vid0.p, dlpos.p, lsc, e->offs0.p, e->words0.p);
            if (lgroups.size() > 1) e->sync();
        }
        HIPCHK(hipGetLastError());
        if (e->host_stream) move_stream(e, e->words0, true);
        finish_graph(e);
    });
}

int lpa_encode_csr_device(lpa_engine *e, uint64_t n, uint64_t n_c, const uint64_t *row, const uint32_t *col,
                          const uint32_t *vid, uint64_t *offsets, uint64_t *words, uint64_t words_cap,
                          uint64_t *nbits) {
    return guard([&] {
        if (!e || !row || !vid || !offsets || !words || !nbits) throw Fail(LPA_ERR_ARG, "null argument");
        if (n >= (1ull << 30)) throw Fail(LPA_ERR_RANGE, "n must be < 2^30");
        if (n_c > n) throw Fail(LPA_ERR_ARG, "n_c > n");
        if (n_c > NOALLOC_MAX_NC)
            throw Fail(LPA_ERR_RANGE, "n_c exceeds the allocation-free encoder's limit of " +
                                          std::to_string(NOALLOC_MAX_NC) + " records");
        HIPCHK(hipSetDevice(e->device));
        uint64_t *scr = e->scratch_dev;   // [0] error flag, [1..nb] block sums
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, e->stream, (unsigned long long *)scr, (uint64_t)1, 0ull);
        if (n_c == 0) {
            hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, e->stream, (unsigned long long *)offsets, (uint64_t)1,
                               0ull);
            HIPCHK(hipGetLastError());
            e->sync();
            *nbits = 0;
            return;
        }
        // record lengths into offsets[0..n_c), then an in-place exclusive scan
        hipLaunchKernelGGL(k_encode_len, dim3((unsigned)((n_c + 255) / 256)), dim3(256), 0, e->stream, row, col, vid,
                           n_c, (uint32_t)n, offsets, (unsigned long long *)scr, UINT32_MAX);
        const uint64_t nb = (n_c + SCAN_PER_BLOCK - 1) / SCAN_PER_BLOCK;
        hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nb), dim3(256), 0, e->stream, offsets, n_c, scr + 1);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, e->stream, scr + 1, nb);
        hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(256), 0, e->stream, offsets, n_c, scr + 1, offsets,
                           nb);
        HIPCHK(hipGetLastError());
        // error flag and total length to the host (the engine's pinned word)
        HIPCHK(hipMemcpyAsync(e->h_counter, scr, sizeof(uint64_t), hipMemcpyDeviceToHost, e->stream));
        e->sync();
        if (*e->h_counter)
            throw Fail(LPA_ERR_FORMAT, "row of vertex " + std::to_string(*e->h_counter - 1) +
                                           " is not strictly increasing or has an id >= n");
        HIPCHK(hipMemcpyAsync(e->h_counter, offsets + n_c, sizeof(uint64_t), hipMemcpyDeviceToHost, e->stream));
        e->sync();
        const uint64_t total = *e->h_counter, nw = (total + 63) / 64;
        if (nw > words_cap)
            throw Fail(LPA_ERR_RANGE, "the stream needs " + std::to_string(nw) + " words, words_cap is " +
                                          std::to_string(words_cap));
        if (nw)
            hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, e->stream,
                               (unsigned long long *)words, nw, 0ull);
        hipLaunchKernelGGL(k_encode_write, dim3((unsigned)((n_c + 255) / 256)), dim3(256), 0, e->stream, row, col, vid,
                           n_c, offsets, words, UINT32_MAX);
        HIPCHK(hipGetLastError());
        e->sync();
        *nbits = total;
    });
}

int lpa_csr_bfs_relabel(lpa_engine *e, uint64_t n, const uint64_t *row, const uint32_t *col, uint32_t root,
                        uint64_t *new_row, uint32_t *new_col, uint32_t *perm) {
    return guard([&] {
        if (!e || !row || !new_row || (n && (!perm || (row[n] && (!col || !new_col)))))
            throw Fail(LPA_ERR_ARG, "null argument");
        if (n >= (1ull << 30)) throw Fail(LPA_ERR_RANGE, "n must be < 2^30");
        if (row[0] != 0) throw Fail(LPA_ERR_FORMAT, "row[0] != 0");
        for (uint64_t v = 0; v < n; v++)
            if (row[v + 1] < row[v]) throw Fail(LPA_ERR_FORMAT, "row offsets decrease");
        if (n && root >= n) throw Fail(LPA_ERR_RANGE, "root >= n");
        HIPCHK(hipSetDevice(e->device));
        try {
            bfs_relabel_device(e->stream, n, row, col, root, new_row, new_col, perm);
        } catch (const std::invalid_argument &x) {
            throw Fail(LPA_ERR_FORMAT, x.what());
        } catch (const std::runtime_error &x) {
            throw Fail(LPA_ERR_HIP, x.what());
        }
    });
}

int lpa_graph_stream(lpa_engine *e, uint64_t *n_c, uint64_t *nbits, uint64_t *words, uint64_t *offsets,
                     uint32_t *vertices) {
    return guard([&] {
        if (!e || !n_c || !nbits) throw Fail(LPA_ERR_ARG, "null argument");
        if (!e->have_graph) throw Fail(LPA_ERR_STATE, "no graph uploaded");
        HIPCHK(hipSetDevice(e->device));
        *n_c = e->n_c;
        *nbits = e->nbits;
        if (words && e->nbits)
            HIPCHK(hipMemcpyAsync(words, e->words0.p, (e->nbits + 63) / 64 * 8, hipMemcpyDefault, e->stream));
        if (offsets) std::memcpy(offsets, e->h_offs.data(), (e->n_c + 1) * 8);
        if (vertices) std::memcpy(vertices, e->h_vertices.data(), e->n * 4);
        e->sync();
    });
}

int lpa_graph_degrees(lpa_engine *e, uint32_t *deg) {
    return guard([&] {
        if (!e || !deg) throw Fail(LPA_ERR_ARG, "null argument");
        if (!e->have_graph) throw Fail(LPA_ERR_STATE, "no graph uploaded");
        for (uint64_t i = 0; i < e->n_c; i++) deg[e->h_vertices[i]] = e->deg_c[i];
        for (uint64_t j = 0; j < e->n_s; j++)
            deg[e->h_vertices[e->n_c + j]] = (uint32_t)(e->h_soffs[j + 1] - e->h_soffs[j]);
    });
}

int lpa_graph_features(lpa_engine *e, float *feat) {
    return guard([&] {
        if (!e || !feat) throw Fail(LPA_ERR_ARG, "null argument");
        if (!e->have_graph) throw Fail(LPA_ERR_STATE, "no graph uploaded");
        HIPCHK(hipSetDevice(e->device));
        std::vector<float> f(e->n_c * NFEAT);
        if (e->n_c) HIPCHK(hipMemcpy(f.data(), e->feat0.p, f.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < e->n_c; i++)
            std::memcpy(feat + (uint64_t)e->h_vertices[i] * NFEAT, &f[i * NFEAT], NFEAT * sizeof(float));
        for (uint64_t j = 0; j < e->n_s; j++) {
            float *o = feat + (uint64_t)e->h_vertices[e->n_c + j] * NFEAT;
            const float d = (float)(e->h_soffs[j + 1] - e->h_soffs[j]);
            const float z[NFEAT] = {d, 0, d, 32.0f * d, 0, 0, 0, std::log2(1.0f + d)};
            std::memcpy(o, z, sizeof z);
        }
    });
}

int lpa_lp_config(lpa_engine *e, uint32_t K, int semi, uint32_t cms_thresh, uint32_t cms_depth,
                  uint32_t cms_width, uint64_t cms_seed) {
    return guard([&] {
        if (!e) throw Fail(LPA_ERR_ARG, "null argument");
        if (K == 0 || K > 1024) throw Fail(LPA_ERR_RANGE, "K must be in 1..1024");
        if (cms_thresh) {
            if (cms_depth == 0 || cms_depth > 8) throw Fail(LPA_ERR_RANGE, "cms_depth must be in 1..8");
            if (cms_width == 0 || (cms_width & (cms_width - 1))) throw Fail(LPA_ERR_RANGE, "cms_width must be a power of two");
            if ((uint64_t)cms_depth * cms_width * 4 > lds_budget(e)) throw Fail(LPA_ERR_RANGE, "CMS sketch exceeds LDS");
        }
        // the schedule depends on K and the CMS parameters (bins, heavy-vertex buffers)
        const bool changed_cms = cms_thresh != e->cms_thresh || K != e->K || cms_depth != e->cms_depth ||
                                 cms_width != e->cms_width;
        e->K = K;
        e->semi = semi ? 1 : 0;
        e->cms_thresh = cms_thresh;
        e->cms_depth = cms_depth ? cms_depth : 4;
        e->cms_width = cms_width ? cms_width : 1024;
        e->cms_seed = cms_seed;
        uint64_t s = cms_seed;
        std::memset(e->cms_ab, 0, sizeof e->cms_ab);
        for (uint32_t r = 0; r < e->cms_depth && r < 8; r++) {
            e->cms_ab[2 * r] = splitmix64(s) | 1ull;
            e->cms_ab[2 * r + 1] = splitmix64(s);
        }
        e->have_labels = false;
        if (changed_cms) e->scheduled = false;   // call lpa_schedule_* again
    });
}

// Heavy sets from the CMS / GLOBAL bins: per heavy vertex its position,
// list offsets (degree prefix), chunk table and, in exact mode, hash-table
// offsets; tables of several label vectors share memory when K is large.
static void build_heavy(lpa_engine *e) {
    e->heavy.clear();
    // bytes of HBM hash tables for the label vectors counted at once: 40 %
    // of the free device memory (the old layout's tables already freed), at
    // least 16 GiB; LPA_TABLE_GB overrides.  Every batch gathers the chunked
    // hubs' rows again, so fewer, larger batches pay (R-MAT 24, K = 64: 95.7
    // ms per step with 6 batches, 86.1 ms with one).
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    uint64_t table_budget = std::max<uint64_t>(16ull << 30, (uint64_t)(0.4 * (double)free_b));
    if (const char *tb = std::getenv("LPA_TABLE_GB")) {
        const double gb = std::atof(tb);
        if (gb > 0) table_budget = (uint64_t)(gb * (1ull << 30));
    }
    for (int part = 0; part < 2; part++) {
        for (int cms = 0; cms < 2; cms++) {
            auto hs = std::make_unique<HeavySet>();
            hs->csr_part = part == 1;
            hs->cms = cms == 1;
            std::vector<uint32_t> pos, local, local_big, chunked, slot;
            std::vector<uint64_t> loff(1, 0), toff(1, 0), cb;
            std::vector<uint32_t> ch;
            uint64_t max_chunked = 0;
            // one workgroup holds the whole count: exact up to 2*CHUNK ids
            // (LDS table of up to 4*CHUNK slots, when the LDS holds it), CMS
            // up to CMS_LOCAL ids (LDS sketches)
            const bool big_ok = 4ull * CHUNK * 8 <= lds_budget(e);
            const uint64_t local_max = cms ? CMS_LOCAL : big_ok ? 2 * CHUNK : CHUNK;
            for (const Bin &b : e->bins) {
                if (b.csr != hs->csr_part) continue;
                if (b.kind != (cms ? LPA_BIN_CMS : LPA_BIN_GLOBAL)) continue;
                for (uint32_t q = b.begin; q < b.end; q++) {
                    const uint64_t d = hs->csr_part ? e->h_soffs[e->order_s[q] + 1] - e->h_soffs[e->order_s[q]]
                                                    : e->deg_c[e->order_c[q]];
                    const uint32_t i = (uint32_t)pos.size();
                    pos.push_back(q);
                    const bool is_local = d <= local_max;
                    slot.push_back(is_local ? 0u : (uint32_t)chunked.size());
                    if (!is_local) {
                        chunked.push_back(i);
                        max_chunked = std::max(max_chunked, d);
                    }
                    else if (!cms && d > CHUNK) local_big.push_back(i);
                    else local.push_back(i);
                    if (!is_local)
                        for (uint64_t c0 = 0; c0 < d; c0 += CHUNK) {
                            ch.push_back(i);
                            cb.push_back(c0);
                        }
                    loff.push_back(loff.back() + d);
                    toff.push_back(toff.back() +
                                   (cms || is_local ? 0 : (1ull << ceil_log2(std::max<uint64_t>(2 * d, 2)))));
                }
            }
            if (pos.empty()) continue;
            hs->nlocal_big = (uint32_t)local_big.size();
            local.insert(local.end(), local_big.begin(), local_big.end());
            hs->nh = (uint32_t)pos.size();
            hs->nchunks = (uint32_t)ch.size();
            hs->nlocal = (uint32_t)local.size();
            hs->nchunked = (uint32_t)chunked.size();
            hs->pos.upload(pos.data(), pos.size(), e->stream);
            hs->loff.upload(loff.data(), loff.size(), e->stream);
            hs->local.upload(local.data(), local.size(), e->stream);
            hs->chunked.upload(chunked.data(), chunked.size(), e->stream);
            hs->slot.upload(slot.data(), slot.size(), e->stream);
            hs->chunk_h.upload(ch.data(), ch.size(), e->stream);
            hs->chunk_b.upload(cb.data(), cb.size(), e->stream);
            if (!hs->csr_part) {
                hs->list.alloc(loff.back());
                // decode tasks of <= BLOCK segments per record, with their
                // seek positions (device metadata) and output ranges
                DevBuf<uint4> m;
                m.alloc(hs->nh);
                hipLaunchKernelGGL(k_gather_u4, dim3((hs->nh + 255) / 256), dim3(256), 0, e->stream, e->meta.p,
                                   hs->pos.p, hs->nh, m.p);
                std::vector<uint4> hm(hs->nh);
                HIPCHK(hipMemcpyAsync(hm.data(), m.p, hs->nh * sizeof(uint4), hipMemcpyDeviceToHost, e->stream));
                e->sync();
                std::vector<DecodeTask> tk;
                std::vector<uint32_t> owner;
                for (uint32_t i = 0; i < hs->nh; i++) {
                    const uint32_t q = pos[i];
                    const uint64_t start = e->h_loffs[q];   // the scheduled stream
                    const uint32_t segs = hm[i].y + hm[i].w;
                    for (uint32_t j0 = 0; j0 < segs; j0 += BLOCK) {
                        DecodeTask t{};
                        t.ibase = start + hm[i].x;
                        t.rbase = start + hm[i].z;
                        t.v = e->h_vertices[e->order_c[q]];
                        t.j0 = j0;
                        t.ni = hm[i].y;
                        t.ns = segs;
                        tk.push_back(t);
                        owner.push_back(i);
                        hs->task_pos.push_back(q);
                    }
                }
                hs->ntasks = (uint32_t)tk.size();
                hs->tasks.upload(tk.data(), tk.size(), e->stream);
                if (hs->ntasks) {
                    StepArgs a = e->base_args(false);
                    hipLaunchKernelGGL(k_heavy_task_count, dim3(hs->ntasks), dim3(BLOCK), 0, e->stream, a, hs->tasks.p,
                                       hs->ntasks);
                    HIPCHK(hipGetLastError());
                    HIPCHK(hipMemcpyAsync(tk.data(), hs->tasks.p, tk.size() * sizeof(DecodeTask), hipMemcpyDeviceToHost,
                                          e->stream));
                    e->sync();
                    // tasks of one vertex fill its list back to back
                    for (size_t t = 0; t < tk.size(); t++) {
                        const uint32_t i = owner[t];
                        tk[t].out = (t > 0 && owner[t - 1] == i) ? tk[t - 1].out + tk[t - 1].count : loff[i];
                    }
                    for (uint32_t i = 0, t = 0; i < hs->nh; i++) {
                        uint64_t sum = 0;
                        for (; t < tk.size() && owner[t] == i; t++) sum += tk[t].count;
                        if (sum != loff[i + 1] - loff[i]) throw Fail(LPA_ERR_FORMAT, "heavy record decodes to a wrong degree");
                    }
                    hs->tasks.upload(tk.data(), tk.size(), e->stream);
                }
            }
            if (hs->cms) {
                hs->kb = e->K;
                hs->sketch.alloc(((uint64_t)e->cms_depth * e->cms_width) * std::max<uint32_t>(hs->nchunked, 1) * hs->kb);
                hs->hcnt.alloc(4ull * std::max<uint32_t>(hs->nchunked, 1) * hs->kb);
            } else {
                const uint64_t per_vec = toff.back() * 8;
                hs->kb = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(e->K, table_budget / std::max<uint64_t>(per_vec, 1)));
                hs->toff.upload(toff.data(), toff.size(), e->stream);
                hs->tab.alloc(std::max<uint64_t>(toff.back(), 1) * hs->kb);
                // counts share the slot's count word with the generation
                const char *tags = std::getenv("LPA_TABLE_TAGS");
                hs->tagged = max_chunked <= COUNT_MASK && !(tags && std::strcmp(tags, "0") == 0);
                hs->epoch.alloc(1);
                HIPCHK(hipMemsetAsync(hs->epoch.p, 0, sizeof(uint32_t), e->stream));
            }
            hs->result.alloc((uint64_t)std::max<uint32_t>(hs->nchunked, 1) * hs->kb);
            if (hs->cms && hs->nlocal) {
                // one compacted stash (CMS_LOCAL entries per label vector of a
                // batch) per workgroup the CUs hold at once
                const size_t cells = (size_t)e->cms_depth * e->cms_width;
                const bool one = e->K == 1;
                const uint32_t kbl = cms_local_kb(e->K, cells, lds_budget(e));
                const void *fn = one ? (const void *)k_heavy_local<1, 1, 4>
                                     : (const void *)k_heavy_local<1, HEAVY_KB_CMS_LOCAL, 4>;
                int per_cu = 0;
                HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, HBLOCK, cms_local_lds(kbl, cells)));
                hs->stash_blocks = (uint32_t)std::min<uint64_t>(hs->nlocal, (uint64_t)std::max(per_cu, 1) * e->num_cu);
                hs->stash.alloc((uint64_t)hs->stash_blocks * CMS_LOCAL * (one ? 1 : HEAVY_KB_CMS_LOCAL));
            }
            const uint32_t nsmall = hs->nlocal - hs->nlocal_big;
            if (!hs->cms && nsmall && row_stash_on(e)) {
                // exact one-workgroup vertices of degree <= CHUNK with many
                // vectors: a row stash (CHUNK x K words) per workgroup the CUs
                // hold at once.  (Degrees up to 2*CHUNK keep a workgroup per
                // vertex: their stashes, 2 MB each at K = 64, overflow the
                // caches and the step measured slower with them.)
                int occ = 0;
                HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &occ, (const void *)k_heavy_local<0, HEAVY_KB_LOCAL, 4, true>, HBLOCK, 2 * CHUNK * sizeof(uint64_t)));
                hs->rstash_small = (uint32_t)std::min<uint64_t>(nsmall, (uint64_t)std::max(occ, 1) * e->num_cu);
                hs->stash.alloc((uint64_t)hs->rstash_small * CHUNK * row_stash_k(e));
                hs->wq.alloc(2);
                HIPCHK(hipMemsetAsync(hs->wq.p, 0, 2 * sizeof(uint32_t), e->stream));
            }
            e->heavy.push_back(std::move(hs));
        }
    }
    e->sync();
}

// ------------------------------------------------- host-resident stream

static uint64_t *rebased(uint64_t *stage, uint64_t w0) {
    return reinterpret_cast<uint64_t *>(reinterpret_cast<uintptr_t>(stage) - w0 * sizeof(uint64_t));
}

// Position batches of the scheduled stream (h_loffs): each takes whole
// records while its words, the readers' 8 look-ahead words included, fit in
// `cap` words; a longer record is a batch of its own.
static std::vector<StreamBatch> stream_batches(const lpa_engine *e, uint64_t cap) {
    std::vector<StreamBatch> out;
    auto span_end = [&](uint64_t r) { return ((e->h_loffs[r] + 63) >> 6) + 8; };
    for (uint64_t q = 0; q < e->n_c;) {
        StreamBatch b;
        b.p0 = (uint32_t)q;
        b.w0 = e->h_loffs[q] >> 6;
        uint64_t r = q + 1;
        while (r < e->n_c && span_end(r + 1) - b.w0 <= cap) r++;
        b.p1 = (uint32_t)r;
        b.nw = span_end(r) - b.w0;
        out.push_back(std::move(b));
        q = r;
    }
    return out;
}

// Staging buffers (HBM) for the largest batch, the copy stream and its
// events; both buffers start free.
static void ensure_staging(lpa_engine *e, const std::vector<StreamBatch> &bs) {
    uint64_t maxw = 1;
    for (const auto &b : bs) maxw = std::max(maxw, b.nw);
    e->sync();
    if (e->copy_stream) HIPCHK(hipStreamSynchronize(e->copy_stream));
    for (auto &s : e->stage)
        if (s.n < maxw) s.alloc(maxw);
    if (!e->copy_stream) {
        HIPCHK(hipStreamCreateWithFlags(&e->copy_stream, hipStreamNonBlocking));
        for (int s = 0; s < 2; s++) {
            HIPCHK(hipEventCreateWithFlags(&e->ev_copied[s], hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&e->ev_free[s], hipEventDisableTiming));
            HIPCHK(hipEventRecord(e->ev_free[s], e->stream));
        }
    }
}

// The batches of the current layout, with each heavy set's decode tasks.
static void plan_batches(lpa_engine *e) {
    auto bs = stream_batches(e, std::max<uint64_t>(e->batch_bytes / sizeof(uint64_t), 16));
    for (auto &b : bs) {
        b.tasks.resize(e->heavy.size());
        for (size_t j = 0; j < e->heavy.size(); j++) {
            const auto &tp = e->heavy[j]->task_pos;
            b.tasks[j] = {(uint32_t)(std::lower_bound(tp.begin(), tp.end(), b.p0) - tp.begin()),
                          (uint32_t)(std::lower_bound(tp.begin(), tp.end(), b.p1) - tp.begin())};
        }
    }
    ensure_staging(e, bs);
    e->batches = std::move(bs);
}

// Relayout into a host-resident stream, batch by batch: k_relayout ORs the
// records into a zeroed staging buffer (atomically, for words shared by two
// records; atomics are kept off the host link), then the batch's words are
// copied back.  The first word of a batch may hold the previous batch's
// last bits: it is ORed in on the host.
static void relayout_to_host(lpa_engine *e, const uint32_t *perm) {
    const uint64_t n64 = (e->nbits + 63) / 64 + 8;
    e->words.alloc_host(n64);
    const auto bs = stream_batches(e, std::max<uint64_t>(e->batch_bytes / sizeof(uint64_t), 16));
    ensure_staging(e, bs);
    uint64_t *st = e->stage[0].p;
    uint64_t written = 0;   // host words [0, written) are filled; the last may be shared
    for (const auto &b : bs) {
        const uint64_t nw = ((e->h_loffs[b.p1] + 63) >> 6) - b.w0;   // words holding the batch's bits
        const uint32_t cnt = b.p1 - b.p0;
        if (nw == 0) continue;
        HIPCHK(hipMemsetAsync(st, 0, nw * sizeof(uint64_t), e->stream));
        hipLaunchKernelGGL(k_relayout, dim3((cnt + 255) / 256), dim3(256), 0, e->stream, e->words0.p, e->offs0.p,
                           e->vid0.p, e->meta0.p, perm + b.p0, e->offs.p + b.p0, rebased(st, b.w0), e->vid.p + b.p0,
                           e->meta.p + b.p0, (uint64_t)cnt);
        HIPCHK(hipGetLastError());
        uint64_t first = 0;
        HIPCHK(hipMemcpyAsync(&first, st, sizeof first, hipMemcpyDeviceToHost, e->stream));
        if (nw > 1)
            HIPCHK(hipMemcpyAsync(e->words.p + b.w0 + 1, st + 1, (nw - 1) * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                  e->stream));
        e->sync();
        e->words.p[b.w0] = (b.w0 < written ? e->words.p[b.w0] : 0) | first;
        written = b.w0 + nw;
    }
    std::memset(e->words.p + written, 0, (n64 - written) * sizeof(uint64_t));
}

// Moves a stream buffer between HBM and pinned host memory.
static void move_stream(lpa_engine *e, DevBuf<uint64_t> &b, bool to_host) {
    if (!b.p || b.host == to_host) return;
    e->sync();
    DevBuf<uint64_t> t;
    if (to_host) t.alloc_host(b.n);
    else t.alloc(b.n);
    HIPCHK(hipMemcpy(t.p, b.p, b.n * sizeof(uint64_t), hipMemcpyDefault));
    b.swap(t);
}

// Builds the device layout from a kind per file-order position (compressed
// positions first, then small positions).
static void build_layout(lpa_engine *e, const std::vector<uint8_t> &kind_c, const std::vector<uint8_t> &kind_s,
                         const lpa_sched_params &p) {
    e->drop_graphs();
    e->kind_of.assign(e->n, LPA_BIN_SKIP);
    for (uint64_t i = 0; i < e->n_c; i++) e->kind_of[e->h_vertices[i]] = kind_c[i];
    for (uint64_t j = 0; j < e->n_s; j++) e->kind_of[e->h_vertices[e->n_c + j]] = kind_s[j];
    e->sched_used = p;
    auto lane_n = [](uint32_t d) { return d <= 4 ? 4 : d <= 8 ? 8 : d <= 16 ? 16 : d <= 32 ? 32 : 64; };
    auto sub_of = [&](int kind, uint32_t d) -> int {
        switch (kind) {
        case LPA_BIN_LANE: return lane_n(d);
        case LPA_BIN_WAVE: return (int)std::max<uint32_t>(6, ceil_log2(2ull * std::max<uint32_t>(d, 1)));
        case LPA_BIN_BLOCK: return (int)std::max<uint32_t>(8, ceil_log2(2ull * std::max<uint32_t>(d, 1)));
        default: return 0;
        }
    };
    e->bins.clear();
    std::memset(e->bin_counts, 0, sizeof e->bin_counts);
    for (int part = 0; part < 2; part++) {
        const bool csr = part == 1;
        const uint64_t cnt = csr ? e->n_s : e->n_c;
        const auto &kinds = csr ? kind_s : kind_c;
        auto deg = [&](uint64_t i) -> uint32_t {
            return csr ? (uint32_t)(e->h_soffs[i + 1] - e->h_soffs[i]) : e->deg_c[i];
        };
        std::vector<uint32_t> order(cnt);
        std::iota(order.begin(), order.end(), 0u);
        std::vector<int> sub(cnt);
        for (uint64_t i = 0; i < cnt; i++) sub[i] = sub_of(kinds[i], deg(i));
        // inside a bin: degree descending (the default), or file order
        // (vertex ids); measured alike on R-MAT 27 (within 1.5 %)
        const bool by_degree = p.bin_order != 1;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
            if (kinds[x] != kinds[y]) return kinds[x] < kinds[y];
            if (sub[x] != sub[y]) return sub[x] < sub[y];
            return by_degree && deg(x) > deg(y);
        });
        for (uint64_t q = 0; q < cnt;) {
            uint64_t r = q;
            const uint32_t x = order[q];
            while (r < cnt && kinds[order[r]] == kinds[x] && sub[order[r]] == sub[x]) r++;
            Bin b{kinds[x], sub[x], csr, (uint32_t)q, (uint32_t)r, deg(x)};
            e->bins.push_back(b);
            e->bin_counts[b.kind] += r - q;
            q = r;
        }
        if (!csr) {
            e->order_c = order;
            // reorder the stream on the device
            std::vector<uint64_t> noffs(cnt + 1, 0);
            for (uint64_t q = 0; q < cnt; q++)
                noffs[q + 1] = noffs[q] + (e->h_offs[order[q] + 1] - e->h_offs[order[q]]);
            const uint64_t n64 = (e->nbits + 63) / 64 + 8;
            e->offs.upload(noffs.data(), cnt + 1, e->stream);
            e->h_loffs = noffs;
            e->vid.alloc(std::max<uint64_t>(cnt, 1));
            e->meta.alloc(std::max<uint64_t>(cnt, 1));
            DevBuf<uint32_t> perm;
            perm.upload(order.data(), cnt, e->stream);
            if (e->host_stream) {
                relayout_to_host(e, perm.p);
            } else {
                e->words.alloc(n64);
                HIPCHK(hipMemsetAsync(e->words.p, 0, n64 * 8, e->stream));
                if (cnt)
                    hipLaunchKernelGGL(k_relayout, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, e->stream,
                                       e->words0.p, e->offs0.p, e->vid0.p, e->meta0.p, perm.p, e->offs.p,
                                       e->words.p, e->vid.p, e->meta.p, cnt);
            }
            HIPCHK(hipGetLastError());
            e->sync();
        } else {
            e->order_s = order;
            std::vector<uint64_t> row(cnt + 1, 0);
            std::vector<uint32_t> nb;
            std::vector<uint32_t> sv(cnt);
            nb.reserve(e->h_snbr.size());
            for (uint64_t q = 0; q < cnt; q++) {
                const uint32_t i = order[q];
                nb.insert(nb.end(), e->h_snbr.begin() + e->h_soffs[i], e->h_snbr.begin() + e->h_soffs[i + 1]);
                row[q + 1] = nb.size();
                sv[q] = e->h_vertices[e->n_c + i];
            }
            e->srow.upload(row.data(), cnt + 1, e->stream);
            e->snbr.upload(nb.data(), nb.size(), e->stream);
            e->svid.upload(sv.data(), cnt, e->stream);
            e->sync();
        }
    }
    build_heavy(e);
    if (e->host_stream) plan_batches(e);
    e->have_csr = false;
    e->scheduled = true;
    (void)p;
}

// LDS left for tables and lists after the kernels' static shared memory
static uint64_t lds_budget(const lpa_engine *e) { return e->lds_max - 20480; }
// largest degree the WAVE / BLOCK kernels' LDS can hold (table at load <= 1/2)
static uint32_t wave_cap(const lpa_engine *e) {
    uint32_t best = 0;
    for (uint32_t l = 6; l <= 11; l++)
        if (group_lds(wave_groups(l), l) <= lds_budget(e)) best = 1u << (l - 1);
    return best;
}
static uint32_t block_cap(const lpa_engine *e) {
    // up to 4096 slots (40 KB of table and list): several workgroups per CU
    uint32_t best = 0;
    for (uint32_t l = 8; l <= 12; l++)
        if (group_lds(1, l) <= lds_budget(e)) best = 1u << (l - 1);
    return best;
}

static uint8_t legal_kind(const lpa_engine *e, int want, uint32_t d, const lpa_sched_params &p) {
    if (d == 0) return LPA_BIN_SKIP;
    if (e->cms_thresh && d > e->cms_thresh) return LPA_BIN_CMS;
    // promote to the first kind able to hold the degree
    if (want <= LPA_BIN_LANE && d <= std::max<uint32_t>(32, std::min<uint32_t>(p.lane_max, 64))) return LPA_BIN_LANE;
    if (want <= LPA_BIN_WAVE && d <= std::min<uint32_t>(p.wave_max ? p.wave_max : wave_cap(e), wave_cap(e))) return LPA_BIN_WAVE;
    if (want <= LPA_BIN_BLOCK && d <= block_cap(e)) return LPA_BIN_BLOCK;
    return LPA_BIN_GLOBAL;
}

int lpa_schedule_degree(lpa_engine *e, const lpa_sched_params *p) {
    return guard([&] {
        if (!e || !p) throw Fail(LPA_ERR_ARG, "null argument");
        if (!e->have_graph) throw Fail(LPA_ERR_STATE, "no graph uploaded");
        if (p->lane_max > 64) throw Fail(LPA_ERR_RANGE, "lane_max must be <= 64");
        if (p->bin_order > 2 || p->xcd_tiles > 2) throw Fail(LPA_ERR_RANGE, "bin_order and xcd_tiles must be 0, 1 or 2");
        HIPCHK(hipSetDevice(e->device));
        lpa_sched_params q = *p;
        if (q.wave_max == 0) q.wave_max = 512;
        if (q.block_max == 0) q.block_max = block_cap(e);
        auto kind = [&](uint32_t d) -> uint8_t {
            int want = d <= q.lane_max ? LPA_BIN_LANE : d <= q.wave_max ? LPA_BIN_WAVE
                     : d <= q.block_max ? LPA_BIN_BLOCK : LPA_BIN_GLOBAL;
            return legal_kind(e, want, d, q);
        };
        std::vector<uint8_t> kc(e->n_c), ks(e->n_s);
        for (uint64_t i = 0; i < e->n_c; i++) kc[i] = kind(e->deg_c[i]);
        for (uint64_t j = 0; j < e->n_s; j++) ks[j] = kind((uint32_t)(e->h_soffs[j + 1] - e->h_soffs[j]));
        build_layout(e, kc, ks, q);
    });
}

int lpa_schedule_bins(lpa_engine *e, const uint8_t *bins, uint32_t relayout) {
    return guard([&] {
        if (!e || !bins) throw Fail(LPA_ERR_ARG, "null argument");
        if (!e->have_graph) throw Fail(LPA_ERR_STATE, "no graph uploaded");
        HIPCHK(hipSetDevice(e->device));
        lpa_sched_params q{32, 1024, 0, relayout};
        std::vector<uint8_t> kc(e->n_c), ks(e->n_s);
        for (uint64_t i = 0; i < e->n_c; i++) {
            const int want = bins[e->h_vertices[i]];
            if (want >= LPA_NBIN_KINDS) throw Fail(LPA_ERR_ARG, "bin kind out of range");
            kc[i] = legal_kind(e, want, e->deg_c[i], q);
        }
        for (uint64_t j = 0; j < e->n_s; j++) {
            const int want = bins[e->h_vertices[e->n_c + j]];
            if (want >= LPA_NBIN_KINDS) throw Fail(LPA_ERR_ARG, "bin kind out of range");
            ks[j] = legal_kind(e, want, (uint32_t)(e->h_soffs[j + 1] - e->h_soffs[j]), q);
        }
        build_layout(e, kc, ks, q);
    });
}

int lpa_schedule_stats(lpa_engine *e, uint64_t *stats) {
    return guard([&] {
        if (!e || !stats) throw Fail(LPA_ERR_ARG, "null argument");
        if (!e->scheduled) throw Fail(LPA_ERR_STATE, "not scheduled");
        std::memset(stats, 0, 4 * LPA_NBIN_KINDS * sizeof(uint64_t));
        for (const Bin &b : e->bins) {
            uint64_t *s = stats + 4 * b.kind;
            s[0] += b.end - b.begin;
            for (uint32_t q = b.begin; q < b.end; q++) {
                if (b.csr) {
                    const uint32_t j = e->order_s[q];
                    s[1] += e->h_soffs[j + 1] - e->h_soffs[j];
                    s[3] += e->h_soffs[j + 1] - e->h_soffs[j];
                } else {
                    const uint32_t i = e->order_c[q];
                    s[1] += e->deg_c[i];
                    s[2] += e->h_offs[i + 1] - e->h_offs[i];
                }
            }
        }
    });
}

int lpa_schedule_counts(lpa_engine *e, uint64_t *counts) {
    return guard([&] {
        if (!e || !counts) throw Fail(LPA_ERR_ARG, "null argument");
        if (!e->scheduled) throw Fail(LPA_ERR_STATE, "not scheduled");
        std::memcpy(counts, e->bin_counts, sizeof e->bin_counts);
    });
}

int lpa_labels_init(lpa_engine *e, const uint32_t *labels) {
    return guard([&] {
        if (!e) throw Fail(LPA_ERR_ARG, "null argument");
        if (!e->have_graph) throw Fail(LPA_ERR_STATE, "no graph uploaded");
        HIPCHK(hipSetDevice(e->device));
        e->drop_graphs();
        const uint64_t tot = e->n * e->K;
        std::vector<uint32_t> h(tot);
        if (labels) {
            for (uint64_t i = 0; i < tot; i++) {
                uint32_t x = labels[i];
                if (x != UNL && (x & SEED_BIT)) throw Fail(LPA_ERR_RANGE, "labels must be < 2^31 or 0xFFFFFFFF");
                h[i] = (e->semi && x != UNL) ? (x | SEED_BIT) : x;
            }
        } else {
            for (uint64_t v = 0; v < e->n; v++)
                for (uint32_t k = 0; k < e->K; k++) h[v * e->K + k] = (uint32_t)v | (e->semi ? SEED_BIT : 0u);
        }
        e->L[0].upload(h.data(), tot, e->stream);
        e->L[1].upload(h.data(), tot, e->stream);
        e->cur = 0;
        e->sync();
        e->have_labels = true;
    });
}

int lpa_labels_get(lpa_engine *e, uint32_t *labels) {
    return guard([&] {
        if (!e || !labels) throw Fail(LPA_ERR_ARG, "null argument");
        if (!e->have_labels) throw Fail(LPA_ERR_STATE, "no labels");
        HIPCHK(hipSetDevice(e->device));
        const uint64_t tot = e->n * e->K;
        HIPCHK(hipMemcpyAsync(labels, e->L[e->cur].p, tot * 4, hipMemcpyDeviceToHost, e->stream));
        e->sync();
        for (uint64_t i = 0; i < tot; i++)
            if (labels[i] != UNL) labels[i] &= ~SEED_BIT;
    });
}

static void require_scheduled(lpa_engine *e) {
    if (!e) throw Fail(LPA_ERR_ARG, "null argument");
    if (!e->have_graph) throw Fail(LPA_ERR_STATE, "no graph uploaded");
    if (!e->scheduled) throw Fail(LPA_ERR_STATE, "no schedule (call lpa_schedule_degree)");
    HIPCHK(hipSetDevice(e->device));
}

static void require_ready(lpa_engine *e) {
    require_scheduled(e);
    if (!e->have_labels) throw Fail(LPA_ERR_STATE, "labels not initialised");
}

static void cglp_step(lpa_engine *e);
static void cglp_decode(lpa_engine *e);
static void csr_lp(lpa_engine *e);

static void method_step(lpa_engine *e, int method) {
    if (method == 0) e->step(false, nullptr);
    else if (method == 1) cglp_step(e);
    else csr_lp(e);
}

// Enqueues one step (method 0 = fused, 1 = two-kernel, 2 = LP over the
// uncompressed CSR) on the engine stream: eagerly, or by replaying the step
// captured as a HIP graph.
static void enqueue_step(lpa_engine *e, int method) {
    if (method >= 1 && !e->have_csr) {
        // the first decode allocates the CSR: never inside a capture
        if (method == 1) {
            cglp_step(e);
            return;
        }
        cglp_decode(e);
    }
    if (!e->graph_mode) {
        method_step(e, method);
        return;
    }
    hipGraphExec_t &ge = e->graphs[method][e->cur];
    if (!ge) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        try {
            method_step(e, method);
        } catch (...) {
            (void)hipStreamEndCapture(e->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCHK(hipStreamEndCapture(e->stream, &g));
        hipError_t err = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (err != hipSuccess) {
            ge = nullptr;
            throw Fail(LPA_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(err));
        }
    }
    HIPCHK(hipGraphLaunch(ge, e->stream));
}

static void finish_step(lpa_engine *e, uint64_t *changed) {
    HIPCHK(hipMemcpyAsync(e->h_counter, e->counter.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, e->stream));
    e->sync();
    e->cur ^= 1;
    if (changed) *changed = *e->h_counter;
}

int lpa_nova_iterate(lpa_engine *e, uint64_t *changed) {
    return guard([&] {
        require_ready(e);
        enqueue_step(e, 0);
        finish_step(e, changed);
    });
}

int lpa_set_graph_mode(lpa_engine *e, int on) {
    return guard([&] {
        if (!e) throw Fail(LPA_ERR_ARG, "null argument");
        if (on && e->host_stream) throw Fail(LPA_ERR_STATE, "the stream is host-resident: steps are not captured");
        e->graph_mode = on != 0;
        if (!e->graph_mode) e->drop_graphs();
    });
}

// CGLP's first kernel over stream positions [p0, p1): decode every
// compressed record there into its CSR row.
static void decode_range(lpa_engine *e, const uint64_t *words, uint32_t p0, uint32_t p1) {
    for (const Bin &b : e->bins) {
        if (b.csr || b.kind == LPA_BIN_SKIP) continue;
        const uint32_t lo = std::max(b.begin, p0), hi = std::min(b.end, p1);
        if (lo >= hi) continue;
        const uint32_t count = hi - lo;
        if (b.maxdeg <= 256)
            hipLaunchKernelGGL(k_decode_csr, dim3((count + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, e->stream, words,
                               e->offs.p, e->vid.p, e->crow.p, e->ccol.p, lo, hi);
        else
            hipLaunchKernelGGL(k_decode_csr_group, dim3(std::min<uint32_t>(count, 4096)), dim3(BLOCK), 0, e->stream,
                               words, e->offs.p, e->vid.p, e->crow.p, e->ccol.p, lo, hi);
    }
}

// Decodes the whole compressed part to the CSR (batch by batch when the
// stream is host-resident).
static void decode_all(lpa_engine *e) {
    if (e->host_stream)
        e->for_each_batch([&](const StreamBatch &sb, const uint64_t *w) { decode_range(e, w, sb.p0, sb.p1); });
    else decode_range(e, e->words.p, 0, (uint32_t)e->n_c);
    HIPCHK(hipGetLastError());
}

int lpa_stream_residence(lpa_engine *e, int where, uint64_t batch_bytes) {
    return guard([&] {
        if (!e) throw Fail(LPA_ERR_ARG, "null argument");
        if (where != LPA_STREAM_DEVICE && where != LPA_STREAM_HOST)
            throw Fail(LPA_ERR_ARG, "where: LPA_STREAM_DEVICE or LPA_STREAM_HOST");
        const bool host = where == LPA_STREAM_HOST;
        if (host && e->graph_mode) throw Fail(LPA_ERR_STATE, "graph mode is on: a host-resident stream is not captured");
        HIPCHK(hipSetDevice(e->device));
        e->sync();
        if (e->copy_stream) HIPCHK(hipStreamSynchronize(e->copy_stream));
        e->host_stream = host;
        e->batch_bytes = batch_bytes ? batch_bytes : DEFAULT_BATCH_BYTES;
        move_stream(e, e->words0, host);
        move_stream(e, e->words, host);
        if (host && e->scheduled) plan_batches(e);
        if (!host) {
            e->batches.clear();
            for (auto &st : e->stage) st.release();
        }
    });
}

int lpa_stream_batches(lpa_engine *e, uint64_t *nbatches, uint64_t *stage_bytes) {
    return guard([&] {
        if (!e || !nbatches || !stage_bytes) throw Fail(LPA_ERR_ARG, "null argument");
        *nbatches = e->host_stream ? e->batches.size() : 0;
        *stage_bytes = e->host_stream ? 2 * e->stage[0].n * sizeof(uint64_t) : 0;
    });
}

static void cglp_decode(lpa_engine *e) {
    // rows in device-layout order
    std::vector<uint64_t> row(e->n_c + 1, 0);
    for (uint64_t q = 0; q < e->n_c; q++) row[q + 1] = row[q] + e->deg_c[e->order_c[q]];
    e->crow.upload(row.data(), e->n_c + 1, e->stream);
    e->ccol.alloc(std::max<uint64_t>(row[e->n_c], 1));
    decode_all(e);
    e->have_csr = true;
}

int lpa_cglp_decode(lpa_engine *e) {
    return guard([&] {
        require_scheduled(e);
        cglp_decode(e);
        e->sync();
    });
}

// The two-kernel step: decode the compressed part to CSR, then run the LP
// kernels over the CSR (small vertices read their own CSR as in NOVA).
static void cglp_step(lpa_engine *e) {
    if (!e->have_csr) cglp_decode(e);
    else decode_all(e);
    csr_lp(e);
}

// The LP kernels over the CSR of every vertex (the decoded CSR of the
// compressed part, which equals its uncompressed adjacency, and the small
// CSR): CGLP's second kernel, and on its own the uncompressed comparator.
static void csr_lp(lpa_engine *e) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, e->stream, e->counter.p, (uint64_t)1, 0ull);
    StepArgs a = e->base_args(false);
    StepArgs ac = e->base_args(true);
    for (const Bin &b : e->bins) {
        if (b.kind == LPA_BIN_SKIP || b.kind == LPA_BIN_CMS || b.kind == LPA_BIN_GLOBAL) continue;
        if (b.csr) e->launch_bin(b, a, true);
        else e->launch_bin(b, ac, true);
    }
    e->run_heavy(true, nullptr);
    HIPCHK(hipGetLastError());
}

int lpa_cglp_iterate(lpa_engine *e, uint64_t *changed) {
    return guard([&] {
        require_ready(e);
        enqueue_step(e, 1);
        finish_step(e, changed);
    });
}

int lpa_csr_iterate(lpa_engine *e, uint64_t *changed) {
    return guard([&] {
        require_ready(e);
        enqueue_step(e, 2);
        finish_step(e, changed);
    });
}

int lpa_cglp_csr_get(lpa_engine *e, uint64_t *row, uint32_t *col, uint64_t *m) {
    return guard([&] {
        require_scheduled(e);
        if (!row || !m) throw Fail(LPA_ERR_ARG, "null argument");
        if (!e->have_csr) cglp_decode(e);
        e->sync();
        std::vector<uint64_t> crow(e->n_c + 1);
        std::vector<uint32_t> vid(e->n_c);
        HIPCHK(hipMemcpy(crow.data(), e->crow.p, crow.size() * 8, hipMemcpyDeviceToHost));
        if (e->n_c) HIPCHK(hipMemcpy(vid.data(), e->vid.p, e->n_c * 4, hipMemcpyDeviceToHost));
        std::vector<uint32_t> ccol(crow[e->n_c]);
        if (!ccol.empty()) HIPCHK(hipMemcpy(ccol.data(), e->ccol.p, ccol.size() * 4, hipMemcpyDeviceToHost));
        // by vertex id: compressed rows from the device, small rows from the host CSR
        std::vector<uint64_t> deg(e->n, 0), src(e->n, 0);
        std::vector<int8_t> kind(e->n, -1);
        for (uint64_t q = 0; q < e->n_c; q++) {
            deg[vid[q]] = crow[q + 1] - crow[q];
            src[vid[q]] = q;
            kind[vid[q]] = 0;
        }
        for (uint64_t j = 0; j < e->n_s; j++) {
            uint32_t v = e->h_vertices[e->n_c + j];
            deg[v] = e->h_soffs[j + 1] - e->h_soffs[j];
            src[v] = j;
            kind[v] = 1;
        }
        row[0] = 0;
        for (uint64_t v = 0; v < e->n; v++) row[v + 1] = row[v] + deg[v];
        *m = row[e->n];
        if (!col) return;
        for (uint64_t v = 0; v < e->n; v++) {
            const uint32_t *s = kind[v] == 0 ? &ccol[crow[src[v]]] : &e->h_snbr[e->h_soffs[src[v]]];
            std::memcpy(col + row[v], s, deg[v] * 4);
        }
    });
}

// ----------------------------------------------------- learned schedule

struct lpa_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    DevBuf<float> params, grad, m1, m2, loss;
    int t = 0;
    ~lpa_mlp() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static int family_of(int kind) {
    switch (kind) {
    case LPA_BIN_LANE:
    case LPA_BIN_WAVE: return 0;
    case LPA_BIN_BLOCK: return 1;
    case LPA_BIN_GLOBAL: return 2;
    default: return -1;
    }
}

static uint32_t degree_of_vertex(const lpa_engine *e, std::vector<uint32_t> &deg) {
    deg.assign(e->n, 0);
    for (uint64_t i = 0; i < e->n_c; i++) deg[e->h_vertices[i]] = e->deg_c[i];
    for (uint64_t j = 0; j < e->n_s; j++) deg[e->h_vertices[e->n_c + j]] = (uint32_t)(e->h_soffs[j + 1] - e->h_soffs[j]);
    return 0;
}

// kinds per file-order position from kinds per vertex id
static void split_kinds(const lpa_engine *e, const std::vector<uint8_t> &by_v, std::vector<uint8_t> &kc,
                        std::vector<uint8_t> &ks) {
    kc.resize(e->n_c);
    ks.resize(e->n_s);
    for (uint64_t i = 0; i < e->n_c; i++) kc[i] = by_v[e->h_vertices[i]];
    for (uint64_t j = 0; j < e->n_s; j++) ks[j] = by_v[e->h_vertices[e->n_c + j]];
}

int lpa_model_features(lpa_engine *e, float *x) {
    return guard([&] {
        if (!e || !x) throw Fail(LPA_ERR_ARG, "null argument");
        if (!e->have_graph) throw Fail(LPA_ERR_STATE, "no graph uploaded");
        HIPCHK(hipSetDevice(e->device));
        std::vector<float> raw(e->n * NFEAT);
        {
            int rc = lpa_graph_features(e, raw.data());
            if (rc) throw Fail(rc, g_err);
        }
        std::vector<uint8_t> flag(e->n, 0);
        if (e->have_labels) {
            const uint64_t tot = e->n * e->K;
            std::vector<uint32_t> a(tot), b(tot);
            HIPCHK(hipMemcpy(a.data(), e->L[e->cur].p, tot * 4, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(b.data(), e->L[e->cur ^ 1].p, tot * 4, hipMemcpyDeviceToHost));
            for (uint64_t v = 0; v < e->n; v++) flag[v] = a[v * e->K] != b[v * e->K];
        }
        for (uint64_t v = 0; v < e->n; v++) {
            float *o = x + v * NFEAT;
            const float *r = &raw[v * NFEAT];
            for (int f = 0; f < 7; f++) o[f] = std::log2(1.0f + std::max(0.0f, r[f]));
            o[7] = flag[v];
        }
    });
}

int lpa_record_family(lpa_engine *e, int family, uint64_t *cost, uint8_t *measured) {
    return guard([&] {
        require_ready(e);
        if (family < 0 || family > 2 || !cost || !measured) throw Fail(LPA_ERR_ARG, "bad arguments");
        const std::vector<uint8_t> saved_kind = e->kind_of;
        const lpa_sched_params saved_p = e->sched_used;
        const uint64_t tot = e->n * e->K;
        std::vector<uint32_t> saved_lab(tot);
        HIPCHK(hipMemcpy(saved_lab.data(), e->L[e->cur].p, tot * 4, hipMemcpyDeviceToHost));
        const int saved_cur = e->cur;
        // force the family where the kernel can hold the vertex
        std::vector<uint32_t> deg;
        degree_of_vertex(e, deg);
        const int want = family == 0 ? LPA_BIN_LANE : family == 1 ? LPA_BIN_BLOCK : LPA_BIN_GLOBAL;
        lpa_sched_params q{32, 0, 0, 1};
        std::vector<uint8_t> byv(e->n);
        for (uint64_t v = 0; v < e->n; v++) byv[v] = legal_kind(e, want, deg[v], q);
        std::vector<uint8_t> kc, ks;
        split_kinds(e, byv, kc, ks);
        DevBuf<unsigned long long> rec;
        try {
            build_layout(e, kc, ks, q);
            rec.alloc(e->n);
            HIPCHK(hipMemsetAsync(rec.p, 0, e->n * 8, e->stream));
            e->rec = rec.p;
            std::vector<float> bin_ms, heavy_ms;
            e->step_bin_times(bin_ms, heavy_ms);
            e->rec = nullptr;
            std::vector<unsigned long long> h(e->n);
            HIPCHK(hipMemcpyAsync(h.data(), rec.p, e->n * 8, hipMemcpyDeviceToHost, e->stream));
            e->sync();
            // Lane-cycles measure a vertex's latency, not what it costs the
            // device (waves resident per CU differ between kernels): every
            // bin's (and heavy set's) measured time is shared among its
            // vertices in proportion to their lane-cycles.  cost = ps.
            auto vertex_at = [&](bool csr, uint32_t q) -> uint32_t {
                return csr ? e->h_vertices[e->n_c + e->order_s[q]] : e->h_vertices[e->order_c[q]];
            };
            auto share = [&](float ms, const std::vector<const Bin *> &bs) {
                long double sum = 0;
                for (const Bin *b : bs)
                    for (uint32_t q = b->begin; q < b->end; q++) sum += h[vertex_at(b->csr, q)];
                if (sum <= 0) return;
                const long double scale = (long double)ms * 1e9L / sum;
                for (const Bin *b : bs)
                    for (uint32_t q = b->begin; q < b->end; q++) {
                        const uint32_t v = vertex_at(b->csr, q);
                        h[v] = (unsigned long long)std::llround((long double)h[v] * scale);
                    }
            };
            for (size_t i = 0; i < e->bins.size(); i++)
                if (bin_ms[i] > 0) share(bin_ms[i], {&e->bins[i]});
            for (size_t j = 0; j < e->heavy.size(); j++) {
                if (heavy_ms[j] <= 0) continue;
                const HeavySet &hs = *e->heavy[j];
                std::vector<const Bin *> bs;
                for (const Bin &b : e->bins)
                    if (b.csr == hs.csr_part && b.kind == (hs.cms ? LPA_BIN_CMS : LPA_BIN_GLOBAL)) bs.push_back(&b);
                share(heavy_ms[j], bs);
            }
            for (uint64_t v = 0; v < e->n; v++) {
                cost[v] = h[v];
                measured[v] = family_of(byv[v]) == family;
            }
        } catch (...) {
            e->rec = nullptr;
            throw;
        }
        // restore labels and schedule
        e->cur = saved_cur;
        HIPCHK(hipMemcpy(e->L[e->cur].p, saved_lab.data(), tot * 4, hipMemcpyHostToDevice));
        split_kinds(e, saved_kind, kc, ks);
        build_layout(e, kc, ks, saved_p);
    });
}

int lpa_mlp_create(int device, uint64_t seed, lpa_mlp **out) {
    return guard([&] {
        if (!out) throw Fail(LPA_ERR_ARG, "null argument");
        auto *m = new lpa_mlp;
        try {
            m->device = device;
            HIPCHK(hipSetDevice(device));
            HIPCHK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
            // He initialisation (Box-Muller over splitmix64), zero biases
            std::vector<float> p(MLP_NPARAM, 0.0f);
            uint64_t s = seed;
            auto normal = [&]() {
                double u1 = ((splitmix64(s) >> 11) + 1) * 0x1.0p-53, u2 = (splitmix64(s) >> 11) * 0x1.0p-53;
                return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
            };
            for (int i = 0; i < MLP_H * MLP_IN; i++) p[MLP_W1 + i] = (float)(normal() * std::sqrt(2.0 / MLP_IN));
            for (int i = 0; i < MLP_H * MLP_H; i++) p[MLP_W2 + i] = (float)(normal() * std::sqrt(2.0 / MLP_H));
            for (int i = 0; i < MLP_OUT * MLP_H; i++) p[MLP_W3 + i] = (float)(normal() * std::sqrt(2.0 / MLP_H));
            m->params.upload(p.data(), MLP_NPARAM, m->stream);
            m->grad.alloc(MLP_NPARAM);
            m->m1.alloc(MLP_NPARAM);
            m->m2.alloc(MLP_NPARAM);
            m->loss.alloc(1);
            HIPCHK(hipMemsetAsync(m->m1.p, 0, MLP_NPARAM * 4, m->stream));
            HIPCHK(hipMemsetAsync(m->m2.p, 0, MLP_NPARAM * 4, m->stream));
            HIPCHK(hipStreamSynchronize(m->stream));
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

int lpa_mlp_destroy(lpa_mlp *m) {
    return guard([&] {
        if (!m) return;
        (void)hipSetDevice(m->device);
        delete m;
    });
}

int lpa_mlp_get_params(lpa_mlp *m, float *params) {
    return guard([&] {
        if (!m || !params) throw Fail(LPA_ERR_ARG, "null argument");
        HIPCHK(hipSetDevice(m->device));
        HIPCHK(hipMemcpy(params, m->params.p, MLP_NPARAM * 4, hipMemcpyDeviceToHost));
    });
}

int lpa_mlp_set_params(lpa_mlp *m, const float *params) {
    return guard([&] {
        if (!m || !params) throw Fail(LPA_ERR_ARG, "null argument");
        HIPCHK(hipSetDevice(m->device));
        HIPCHK(hipMemcpy(m->params.p, params, MLP_NPARAM * 4, hipMemcpyHostToDevice));
        m->t = 0;
        HIPCHK(hipMemset(m->m1.p, 0, MLP_NPARAM * 4));
        HIPCHK(hipMemset(m->m2.p, 0, MLP_NPARAM * 4));
    });
}

int lpa_mlp_forward(lpa_mlp *m, const float *x, uint64_t n, float *y) {
    return guard([&] {
        if (!m || (n && (!x || !y))) throw Fail(LPA_ERR_ARG, "null argument");
        if (n == 0) return;
        HIPCHK(hipSetDevice(m->device));
        DevBuf<float> dx, dy;
        dx.upload(x, n * MLP_IN, m->stream);
        dy.alloc(n * MLP_OUT);
        hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream,
                           m->params.p, dx.p, dy.p, n);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(y, dy.p, n * MLP_OUT * 4, hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
    });
}

int lpa_mlp_loss_grad(lpa_mlp *m, const float *x, const float *target, const uint8_t *mask, uint64_t n,
                      float *loss, float *grad) {
    return guard([&] {
        if (!m || !x || !target || !mask || !loss || !grad || n == 0) throw Fail(LPA_ERR_ARG, "bad arguments");
        HIPCHK(hipSetDevice(m->device));
        DevBuf<float> dx, dt;
        DevBuf<uint8_t> dm;
        dx.upload(x, n * MLP_IN, m->stream);
        dt.upload(target, n * MLP_OUT, m->stream);
        dm.upload(mask, n * MLP_OUT, m->stream);
        hipLaunchKernelGGL(k_zero_f32, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->stream, m->grad.p, MLP_NPARAM);
        hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, m->stream, m->loss.p, 1);
        hipLaunchKernelGGL(k_mlp_backward, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream, m->params.p,
                           dx.p, dt.p, dm.p, (const uint32_t *)nullptr, n, 1.0f / (float)n, m->grad.p, m->loss.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(grad, m->grad.p, MLP_NPARAM * 4, hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipMemcpyAsync(loss, m->loss.p, 4, hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
    });
}

int lpa_mlp_train(lpa_mlp *m, const float *x, const float *target, const uint8_t *mask, uint64_t n, int epochs,
                  int batch, float lr, uint64_t seed, float *loss_per_epoch) {
    return guard([&] {
        if (!m || !x || !target || !mask || n == 0 || epochs <= 0 || batch <= 0) throw Fail(LPA_ERR_ARG, "bad arguments");
        if (n >= (1ull << 32)) throw Fail(LPA_ERR_RANGE, "too many samples");
        HIPCHK(hipSetDevice(m->device));
        DevBuf<float> dx, dt;
        DevBuf<uint8_t> dm;
        DevBuf<uint32_t> didx;
        dx.upload(x, n * MLP_IN, m->stream);
        dt.upload(target, n * MLP_OUT, m->stream);
        dm.upload(mask, n * MLP_OUT, m->stream);
        didx.alloc(n);
        std::vector<uint32_t> idx(n);
        std::iota(idx.begin(), idx.end(), 0u);
        uint64_t s = seed;
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        for (int ep = 0; ep < epochs; ep++) {
            for (uint64_t i = n - 1; i > 0; i--) std::swap(idx[i], idx[splitmix64(s) % (i + 1)]);
            HIPCHK(hipMemcpyAsync(didx.p, idx.data(), n * 4, hipMemcpyHostToDevice, m->stream));
            hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, m->stream, m->loss.p, 1);
            int nb = 0;
            for (uint64_t b0 = 0; b0 < n; b0 += (uint64_t)batch, nb++) {
                const uint64_t bn = std::min<uint64_t>(batch, n - b0);
                hipLaunchKernelGGL(k_zero_f32, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->stream, m->grad.p,
                                   MLP_NPARAM);
                hipLaunchKernelGGL(k_mlp_backward, dim3((unsigned)((bn + 255) / 256)), dim3(256), 0, m->stream,
                                   m->params.p, dx.p, dt.p, dm.p, didx.p + b0, bn, 1.0f / (float)bn, m->grad.p,
                                   m->loss.p);
                m->t++;
                const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, m->t)));
                const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, m->t)));
                hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->stream, m->params.p,
                                   m->grad.p, m->m1.p, m->m2.p, lr, b1, b2, eps, c1, c2);
            }
            HIPCHK(hipGetLastError());
            float l = 0;
            HIPCHK(hipMemcpyAsync(&l, m->loss.p, 4, hipMemcpyDeviceToHost, m->stream));
            HIPCHK(hipStreamSynchronize(m->stream));
            if (loss_per_epoch) loss_per_epoch[ep] = l / (float)nb;
        }
    });
}

int lpa_schedule_learned(lpa_engine *e, lpa_mlp *m) {
    return guard([&] {
        if (!e || !m) throw Fail(LPA_ERR_ARG, "null argument");
        if (!e->have_graph) throw Fail(LPA_ERR_STATE, "no graph uploaded");
        std::vector<float> x(e->n * NFEAT), y(e->n * MLP_OUT);
        int rc = lpa_model_features(e, x.data());
        if (rc) throw Fail(rc, g_err);
        rc = lpa_mlp_forward(m, x.data(), e->n, y.data());
        if (rc) throw Fail(rc, g_err);
        HIPCHK(hipSetDevice(e->device));
        std::vector<uint32_t> deg;
        degree_of_vertex(e, deg);
        lpa_sched_params q{32, 512, 0, 1, 128};
        const uint32_t wc = wave_cap(e), bc = block_cap(e);
        std::vector<uint8_t> byv(e->n);
        // A vertex leaves the family the default degree schedule gives it
        // only when the model predicts another one cheaper by more than
        // LEARNED_MARGIN (log2 of the cost): near-ties stay put, so noise in
        // the recorded costs does not reshuffle the bins from run to run.
        // Family 0 stops at the default wave limit: above it the wave
        // kernel's lists outgrow its staged registers, which the recorded
        // costs (every vertex of the graph in one family) understate.
        constexpr float LEARNED_MARGIN = 0.5f;   // ~41 %
        const uint32_t def_wave = std::min<uint32_t>(512, wc);
        for (uint64_t v = 0; v < e->n; v++) {
            const uint32_t d = deg[v];
            int fam = 2;
            float best = y[v * MLP_OUT + 2];
            if (d <= bc && y[v * MLP_OUT + 1] < best) { fam = 1; best = y[v * MLP_OUT + 1]; }
            if (d <= def_wave && y[v * MLP_OUT + 0] < best) { fam = 0; best = y[v * MLP_OUT + 0]; }
            const int def = d <= def_wave ? 0 : d <= bc ? 1 : 2;
            if (fam != def && y[v * MLP_OUT + def] - best < LEARNED_MARGIN) fam = def;
            const int want = fam == 0 ? LPA_BIN_LANE : fam == 1 ? LPA_BIN_BLOCK : LPA_BIN_GLOBAL;
            byv[v] = legal_kind(e, want, d, q);
        }
        // the small CSR part has no CGR features to predict from: its
        // vertices keep the default degree family
        for (uint64_t j = e->n_c; j < e->n; j++) {
            const uint32_t v = e->h_vertices[j], d = deg[v];
            byv[v] = legal_kind(e, d <= 32 ? LPA_BIN_LANE : d <= def_wave ? LPA_BIN_WAVE : LPA_BIN_BLOCK, d, q);
        }
        std::vector<uint8_t> kc, ks;
        split_kinds(e, byv, kc, ks);
        build_layout(e, kc, ks, q);
    });
}

int lpa_schedule_learned_judged(lpa_engine *e, lpa_mlp *m, int iters, float *learned_ms, float *degree_ms,
                                int *kept_learned) {
    return guard([&] {
        // builds both layouts itself: needs the graph and the labels only
        if (!e || !m || iters <= 0) throw Fail(LPA_ERR_ARG, "bad arguments");
        if (!e->have_graph) throw Fail(LPA_ERR_STATE, "no graph uploaded");
        if (!e->have_labels) throw Fail(LPA_ERR_STATE, "labels not initialised");
        HIPCHK(hipSetDevice(e->device));
        // both layouts are timed from the same labels, which are put back
        const uint64_t tot = e->n * e->K;
        std::vector<uint32_t> saved(tot);
        HIPCHK(hipMemcpy(saved.data(), e->L[e->cur].p, tot * 4, hipMemcpyDeviceToHost));
        const int saved_cur = e->cur;
        std::vector<float> ms(iters);
        // steps of one layout, from the saved labels (put back afterwards)
        auto run_steps = [&](std::vector<float> &all) {
            int rc = lpa_time_iterations(e, 0, iters, 0, ms.data(), nullptr);
            if (rc) throw Fail(rc, g_err);
            e->cur = saved_cur;
            HIPCHK(hipMemcpy(e->L[e->cur].p, saved.data(), tot * 4, hipMemcpyHostToDevice));
            all.insert(all.end(), ms.begin(), ms.end());
        };
        auto median = [](std::vector<float> v) {
            std::sort(v.begin(), v.end());
            return v[v.size() / 2];
        };
        // order degree, learned, degree: neither layout is always measured
        // first, and the degree layout's median covers both ends
        std::vector<float> dl, ll;
        lpa_sched_params q{32, 512, 0, 1, 128};
        int rc = lpa_schedule_degree(e, &q);
        if (rc) throw Fail(rc, g_err);
        run_steps(dl);
        rc = lpa_schedule_learned(e, m);
        if (rc) throw Fail(rc, g_err);
        run_steps(ll);
        const std::vector<uint8_t> learned_kind = e->kind_of;
        const lpa_sched_params learned_p = e->sched_used;
        rc = lpa_schedule_degree(e, &q);
        if (rc) throw Fail(rc, g_err);
        run_steps(dl);
        const float tl = median(ll), td = median(dl);
        const bool keep = tl <= td;
        if (keep) {
            std::vector<uint8_t> kc, ks;
            split_kinds(e, learned_kind, kc, ks);
            build_layout(e, kc, ks, learned_p);
        }
        if (learned_ms) *learned_ms = tl;
        if (degree_ms) *degree_ms = td;
        if (kept_learned) *kept_learned = keep ? 1 : 0;
    });
}

int lpa_time_iterations(lpa_engine *e, int method, int iters, int use_graph, float *ms, uint64_t *changed) {
    return guard([&] {
        require_ready(e);
        if (iters <= 0 || !ms) throw Fail(LPA_ERR_ARG, "bad arguments");
        hipEvent_t ev[2] = {nullptr, nullptr};
        // per launch (bins, then heavy sets): a family's time is the sum of
        // its launches, wherever they fall in the step
        const size_t nb = e->bins.size(), nl = nb + e->heavy.size();
        std::vector<hipEvent_t> kev(2 * nl, nullptr);
        std::vector<int> kind_of_launch(nl, -1);
        for (size_t i = 0; i < nb; i++) {
            const Bin &b = e->bins[i];
            if (b.kind != LPA_BIN_SKIP && b.kind != LPA_BIN_CMS && b.kind != LPA_BIN_GLOBAL && b.end > b.begin)
                kind_of_launch[i] = b.kind;
        }
        for (size_t j = 0; j < e->heavy.size(); j++)
            if (e->heavy[j]->nh) kind_of_launch[nb + j] = e->heavy[j]->cms ? LPA_BIN_CMS : LPA_BIN_GLOBAL;
        auto destroy = [&] {
            for (auto &x : ev)
                if (x) (void)hipEventDestroy(x);
            for (auto &x : kev)
                if (x) (void)hipEventDestroy(x);
        };
        std::vector<double> kind_sum(LPA_NBIN_KINDS, 0.0);
        const bool had_graph_mode = e->graph_mode;
        try {
            for (auto &x : ev) HIPCHK(hipEventCreate(&x));
            for (auto &x : kev) HIPCHK(hipEventCreate(&x));
            if (method < 0 || method > 2) throw Fail(LPA_ERR_ARG, "method: 0 NOVA, 1 CGLP, 2 CSR");
            if (use_graph && e->host_stream)
                throw Fail(LPA_ERR_STATE, "the stream is host-resident: steps are not captured");
            if (method >= 1 && !e->have_csr) {
                cglp_decode(e);
                e->sync();
            }
            e->graph_mode = use_graph != 0;
            for (int it = 0; it < iters; it++) {
                HIPCHK(hipEventRecord(ev[0], e->stream));
                if (e->graph_mode) enqueue_step(e, method);
                else if (method == 0) e->step(false, kev.data());
                else method_step(e, method);
                HIPCHK(hipEventRecord(ev[1], e->stream));
                HIPCHK(hipMemcpyAsync(e->h_counter, e->counter.p, 8, hipMemcpyDeviceToHost, e->stream));
                e->sync();
                HIPCHK(hipEventElapsedTime(&ms[it], ev[0], ev[1]));
                if (changed) changed[it] = *e->h_counter;
                e->cur ^= 1;
                if (method == 0 && !e->graph_mode && !e->host_stream) {
                    for (size_t l = 0; l < nl; l++) {
                        if (kind_of_launch[l] < 0) continue;
                        float t = 0;
                        HIPCHK(hipEventElapsedTime(&t, kev[2 * l], kev[2 * l + 1]));
                        kind_sum[kind_of_launch[l]] += t;
                    }
                }
            }
        } catch (...) {
            e->graph_mode = had_graph_mode;
            destroy();
            throw;
        }
        e->graph_mode = had_graph_mode;
        for (int k = 0; k < LPA_NBIN_KINDS; k++) e->kind_ms[k] = (float)(kind_sum[k] / iters);
        destroy();
    });
}

int lpa_autotune(lpa_engine *e, int iters, lpa_sched_params *best, float *best_ms) {
    return guard([&] {
        require_ready(e);
        if (iters <= 0 || !best) throw Fail(LPA_ERR_ARG, "bad arguments");
        // the labels are restored after every candidate: tuning never changes results
        const uint64_t tot = e->n * e->K;
        std::vector<uint32_t> saved(tot);
        HIPCHK(hipMemcpy(saved.data(), e->L[e->cur].p, tot * 4, hipMemcpyDeviceToHost));
        const int saved_cur = e->cur;
        auto restore = [&] {
            e->cur = saved_cur;
            HIPCHK(hipMemcpy(e->L[e->cur].p, saved.data(), tot * 4, hipMemcpyHostToDevice));
        };
        const uint32_t wc = wave_cap(e);
        float best_t = 1e30f;
        lpa_sched_params bp = e->sched_used;
        std::vector<float> ms(iters);
        // median step of a candidate layout, after one untimed step (a new
        // layout's first step clears its freshly allocated hub tables)
        auto measure = [&](const lpa_sched_params &q) {
            int rc = lpa_schedule_degree(e, &q);
            if (rc) throw Fail(rc, g_err);
            rc = lpa_time_iterations(e, 0, 1, 0, ms.data(), nullptr);
            if (rc) throw Fail(rc, g_err);
            restore();
            rc = lpa_time_iterations(e, 0, iters, 0, ms.data(), nullptr);
            if (rc) throw Fail(rc, g_err);
            restore();
            std::vector<float> sorted(ms);
            std::sort(sorted.begin(), sorted.end());
            const float med = sorted[sorted.size() / 2];
            if (med < best_t) {
                best_t = med;
                bp = q;
            }
        };
        for (uint32_t lane : {8u, 16u, 32u, 64u}) {
            for (uint32_t wave : {64u, 128u, 256u, 512u, 1024u}) {
                if (wave > wc) continue;
                for (uint32_t quarter : {64u, 128u}) {
                    lpa_sched_params q{lane, wave, 0, 1};
                    q.quarter_max = quarter;
                    measure(q);
                }
            }
        }
        // then the layout options for those thresholds: vertex order inside
        // the bins and the XCD tile mapping
        const lpa_sched_params base = bp;
        for (uint32_t order : {1u, 2u}) {
            for (uint32_t xcd : {1u, 2u}) {
                lpa_sched_params q = base;
                q.bin_order = order;
                q.xcd_tiles = xcd;
                measure(q);
            }
        }
        int rc = lpa_schedule_degree(e, &bp);
        if (rc) throw Fail(rc, g_err);
        *best = bp;
        if (best_ms) *best_ms = best_t;
    });
}

int lpa_last_kernel_times(lpa_engine *e, float *ms_per_kind) {
    return guard([&] {
        if (!e || !ms_per_kind) throw Fail(LPA_ERR_ARG, "null argument");
        std::memcpy(ms_per_kind, e->kind_ms, sizeof e->kind_ms);
    });
}

int lpa_calibration_copy(lpa_engine *e, uint64_t bytes, int reps, float *ms) {
    return guard([&] {
        if (!e || !ms) throw Fail(LPA_ERR_ARG, "null argument");
        if (bytes == 0 || bytes % 16 || reps <= 0) throw Fail(LPA_ERR_RANGE, "bytes must be a positive multiple of 16");
        HIPCHK(hipSetDevice(e->device));
        const uint64_t n = bytes / 16;
        DevBuf<uint4> src, dst;
        src.alloc(n);
        dst.alloc(n);
        HIPCHK(hipMemsetAsync(src.p, 1, bytes, e->stream));
        hipEvent_t a, b;
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        HIPCHK(hipEventRecord(a, e->stream));
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(k_calibration_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, e->stream, src.p,
                               dst.p, n);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(b, e->stream));
        HIPCHK(hipEventSynchronize(b));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, a, b));
        *ms = t / reps;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        e->sync();
    });
}

int lpa_calibration_gather(lpa_engine *e, uint64_t bytes, uint32_t row_words, uint64_t gathers, int reps,
                           float *ms) {
    return guard([&] {
        if (!e || !ms) throw Fail(LPA_ERR_ARG, "null argument");
        if (row_words != 1 && row_words != 2 && row_words != 4 && row_words != 8 && row_words != 16)
            throw Fail(LPA_ERR_RANGE, "row_words must be 1, 2, 4, 8 or 16");
        const uint64_t nrows = bytes / (4ull * row_words);
        if (nrows == 0 || gathers == 0 || reps <= 0) throw Fail(LPA_ERR_RANGE, "empty gather");
        HIPCHK(hipSetDevice(e->device));
        constexpr uint32_t PER_LANE = 64;
        const uint64_t lanes = (gathers + PER_LANE - 1) / PER_LANE;
        const uint64_t blocks = (lanes + 255) / 256;
        if (blocks > 0x7FFFFFFFull) throw Fail(LPA_ERR_RANGE, "too many gathers");
        DevBuf<uint32_t> src, out;
        src.alloc(nrows * row_words);
        out.alloc(blocks * 256);
        HIPCHK(hipMemsetAsync(src.p, 1, nrows * row_words * 4, e->stream));
        auto launch = [&](uint64_t seed) {
            const dim3 g((unsigned)blocks), b(256);
            switch (row_words) {
            case 1: hipLaunchKernelGGL(k_calibration_gather<1>, g, b, 0, e->stream, src.p, nrows, PER_LANE, seed, out.p); break;
            case 2: hipLaunchKernelGGL(k_calibration_gather<2>, g, b, 0, e->stream, src.p, nrows, PER_LANE, seed, out.p); break;
            case 4: hipLaunchKernelGGL(k_calibration_gather<4>, g, b, 0, e->stream, src.p, nrows, PER_LANE, seed, out.p); break;
            case 8: hipLaunchKernelGGL(k_calibration_gather<8>, g, b, 0, e->stream, src.p, nrows, PER_LANE, seed, out.p); break;
            default: hipLaunchKernelGGL(k_calibration_gather<16>, g, b, 0, e->stream, src.p, nrows, PER_LANE, seed, out.p); break;
            }
        };
        launch(12345);   // warm-up
        hipEvent_t a = nullptr, b = nullptr;
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        HIPCHK(hipEventRecord(a, e->stream));
        for (int r = 0; r < reps; r++) launch((uint64_t)r * 0x51ED27ull);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(b, e->stream));
        HIPCHK(hipEventSynchronize(b));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, a, b));
        *ms = t / reps * (float)((double)gathers / (double)(lanes * PER_LANE));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        e->sync();
    });
}

}  // extern "C"
