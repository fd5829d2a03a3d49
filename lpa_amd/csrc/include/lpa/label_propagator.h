// lpa/label_propagator.h -- the reference's propagator API
// (include/label_propagator.h:11-51: LabelPropagator<V,E>::run(niter) ->
// (LP ms, total ms), get_labels()) with the NOVA, NOVA+ and CGLP methods
// (baselines/nova/nova_propogator.cuh:23-403,
//  baselines/novaplus/novaplus_propogator.cuh,
//  baselines/cglp/cglp_propagator.cuh:35-363) running on the gfx950 engine
// through the C ABI of lpa_amd.h.
//
// All three compute synchronous LP with the smallest-label tie-break.  NOVA
// and NOVA+ are the same fused decode+LP step here: the reference's NOVA
// leaves its medium-degree kernel disabled (baselines/nova/kernel.cuh:53),
// which is a defect, not a method.  CGLP decodes to a CSR and runs the LP
// kernels over it, every iteration.
#pragma once
#include <chrono>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "lpa/graph.h"
#include "lpa_amd.h"

namespace lpa_api {

struct LpOptions {
    uint32_t K = 1;                 // label vectors per vertex
    bool semi = false;              // 0xFFFFFFFF = unlabelled, seeds clamped
    uint32_t cms_thresh = 0;        // 0: exact mode everywhere
    uint32_t cms_depth = 4, cms_width = 1024;
    uint64_t cms_seed = 0;
    bool until_converged = false;   // stop early when no label changes
    lpa_sched_params sched{32, 512, 0, 1, 128};
    std::vector<uint8_t> bins;      // explicit per-vertex bins (learned schedule); empty: degree
    std::vector<uint32_t> init;     // initial labels (n*K); empty: L0(v) = v
    bool verbose = true;
    int device = 0;
};

template <typename V, typename E> class LabelPropagator {
  public:
    LabelPropagator() = default;
    explicit LabelPropagator(V n) : labels(new V[n]), n_(n) {}
    virtual ~LabelPropagator() = default;
    virtual std::pair<double, double> run(int niter) = 0;
    std::vector<V> get_labels() {
        auto p = labels.get();
        return std::vector<V>(p, p + n_ * k_);
    }
    std::unique_ptr<V[]> labels;

  protected:
    V n_ = 0;
    uint32_t k_ = 1;
};

namespace detail {

inline void check(int rc, const char *what) {
    if (rc != LPA_OK) throw std::runtime_error(std::string(what) + ": " + lpa_last_error());
}

using clock = std::chrono::steady_clock;
inline double ms_since(clock::time_point t) {
    return std::chrono::duration<double, std::milli>(clock::now() - t).count();
}

}  // namespace detail

// Engine-backed propagator; method 0 = fused (NOVA), 1 = two-kernel (CGLP).
template <typename V, typename E> class EnginePropagator : public LabelPropagator<V, E> {
  public:
    EnginePropagator(int method, V n, uint64_t n_c, const std::vector<V> &words, const std::vector<uint64_t> &offs,
                     const std::vector<V> &vertices, const std::vector<E> &small_offs,
                     const std::vector<V> &small_nbrs, LpOptions opt)
        : LabelPropagator<V, E>(n * opt.K), method_(method), n_c_(n_c), words_(words), offs_(offs),
          vertices_(vertices), small_nbrs_(small_nbrs), opt_(std::move(opt)) {
        this->n_ = n;
        this->k_ = opt_.K;
        small_offs_.assign(small_offs.begin(), small_offs.end());
        if (small_offs_.empty()) small_offs_.push_back(0);
    }
    ~EnginePropagator() override {
        if (e_) lpa_engine_destroy(e_);
    }

    std::pair<double, double> run(int niter) override {
        using namespace detail;
        auto t_all = clock::now();
        preprocess();
        auto t_lp = clock::now();
        for (int i = 0; i < niter; i++) {
            auto t = clock::now();
            uint64_t changed = 0;
            check(method_ == 0 ? lpa_nova_iterate(e_, &changed) : lpa_cglp_iterate(e_, &changed), "iterate");
            if (opt_.verbose) printf("%d: %f (ms), updated %llu\n", i + 1, ms_since(t), (unsigned long long)changed);
            iterations_ = i + 1;
            if (opt_.until_converged && changed == 0) break;
        }
        check(lpa_labels_get(e_, this->labels.get()), "labels");
        const double lp = ms_since(t_lp);
        return {lp, ms_since(t_all)};
    }
    int iterations() const { return iterations_; }

  private:
    void preprocess() {
        using namespace detail;
        check(lpa_engine_create(opt_.device, &e_), "engine");
        const V n = this->n_;
        check(lpa_graph_upload(e_, n, n_c_, words_.data(), words_.size(), offs_.data(), vertices_.data(),
                               small_offs_.data(), small_nbrs_.empty() ? nullptr : small_nbrs_.data()),
              "upload");
        check(lpa_lp_config(e_, opt_.K, opt_.semi ? 1 : 0, opt_.cms_thresh, opt_.cms_depth, opt_.cms_width,
                            opt_.cms_seed),
              "config");
        if (opt_.bins.empty()) check(lpa_schedule_degree(e_, &opt_.sched), "schedule");
        else check(lpa_schedule_bins(e_, opt_.bins.data(), 1), "schedule");
        check(lpa_labels_init(e_, opt_.init.empty() ? nullptr : opt_.init.data()), "labels");
    }

    int method_;
    uint64_t n_c_;
    const std::vector<V> &words_;
    const std::vector<uint64_t> &offs_;
    const std::vector<V> &vertices_;
    std::vector<uint64_t> small_offs_;
    const std::vector<V> &small_nbrs_;
    LpOptions opt_;
    lpa_engine *e_ = nullptr;
    int iterations_ = 0;
};

template <typename V, typename E> class NovaPlusPropagator : public EnginePropagator<V, E> {
  public:
    using CGRGraphT = CGRGraph<V, E>;
    explicit NovaPlusPropagator(std::shared_ptr<CGRGraphT> G, LpOptions opt = {})
        : EnginePropagator<V, E>(0, G->n, G->cgr_offsets.empty() ? 0 : G->cgr_offsets.size() - 1,
                                 G->cgr_neighbors, G->cgr_offsets, G->vertices, G->csr_small_offsets,
                                 G->csr_small_neighbors, std::move(opt)),
          G_(G) {}

  private:
    std::shared_ptr<CGRGraphT> G_;
};

template <typename V, typename E> class NovaPropagator : public NovaPlusPropagator<V, E> {
  public:
    using NovaPlusPropagator<V, E>::NovaPlusPropagator;
};

template <typename V, typename E> class CGLPPropagator : public EnginePropagator<V, E> {
  public:
    using CGRGraphT = CGRGraphPure<V, E>;
    explicit CGLPPropagator(std::shared_ptr<CGRGraphT> G, LpOptions opt = {})
        : EnginePropagator<V, E>(1, G->n, G->cgr_offsets.empty() ? 0 : G->cgr_offsets.size() - 1,
                                 G->cgr_neighbors, G->cgr_offsets, G->vertices, empty_offs_, empty_nbrs_,
                                 std::move(opt)),
          G_(G) {}
    // the mixed graph as well: small vertices are read from their CSR
    explicit CGLPPropagator(std::shared_ptr<CGRGraph<V, E>> G, LpOptions opt = {})
        : EnginePropagator<V, E>(1, G->n, G->cgr_offsets.empty() ? 0 : G->cgr_offsets.size() - 1,
                                 G->cgr_neighbors, G->cgr_offsets, G->vertices, G->csr_small_offsets,
                                 G->csr_small_neighbors, std::move(opt)),
          Gm_(G) {}

  private:
    std::shared_ptr<CGRGraphT> G_;
    std::shared_ptr<CGRGraph<V, E>> Gm_;
    static inline const std::vector<E> empty_offs_{0};
    static inline const std::vector<V> empty_nbrs_{};
};

}  // namespace lpa_api
