// Learned cardinality estimation for join ordering (host side).
//
// The reference estimates a join's cardinality as the largest of its inputs
// after sampling-based filter selectivities (engine/query/QueryPlan.cpp:66-159,
// 719-746; engine/query/QueryGraph.cpp:53-68). Here a small MLP corrects the
// independence estimate of a connected set of relations (log10 residual) from
// features of the set (relation sizes, exact pairwise join selectivities
// measured on the device, shape), together with a confidence; below the
// confidence bar the planner keeps the uncorrected independence estimate.
// Estimates only steer the join order, so results are identical under every
// estimator.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "engine.hpp"
#include "mlp.hpp"
#include "planner.hpp"

namespace cgj {

// x -> 256 -> 256 -> 2 network with its own FP32 training on the GPU
// (forward, backward, Adam) and BF16-input MFMA inference.
class Mlp {
public:
    Mlp(Context& ctx, uint32_t in, uint64_t seed);
    uint32_t in() const { return in_; }
    uint64_t num_params() const;
    void     set_params(const float* p);
    void     get_params(float* p) const;
    // Mean loss over n rows and its gradient w.r.t. every parameter (FP32).
    float    loss_grad(const float* x, const float* y, uint32_t n, float* grads);
    // Adam (L2 weight decay added to the gradient) over `epochs` passes of
    // shuffled batches; returns the last epoch's mean loss.
    float    train(const float* x, const float* y, uint32_t n, uint32_t epochs, uint32_t batch, float lr,
                   float weight_decay, uint64_t seed);
    void     predict(const float* x, uint32_t n, float* out);     // BF16 MFMA, host arrays (n x 2)
    void     predict_f32(const float* x, uint32_t n, float* out); // FP32 forward, host arrays
    // BF16 inference on the stream and heap of `on` (another context of the
    // same device); reads the weights only, so threads may call it at once
    // with contexts of their own after refresh_bf16().
    void     predict_on(Context& on, const float* x, uint32_t n, float* out) const;
    // Converts the first two layers' weights to BF16 after they changed.
    void     refresh_bf16();

private:
    float forward(const float* xb, const float* yb, uint32_t m, bool backward);
    void  ensure_rows(uint32_t m);
    uint64_t off_w1() const { return 0; }
    uint64_t off_b1() const { return uint64_t(in_) * kMlpHidden; }
    uint64_t off_w2() const { return off_b1() + kMlpHidden; }
    uint64_t off_b2() const { return off_w2() + uint64_t(kMlpHidden) * kMlpHidden; }
    uint64_t off_w3() const { return off_b2() + kMlpHidden; }
    uint64_t off_b3() const { return off_w3() + uint64_t(kMlpHidden) * kMlpOut; }

    Context& ctx_;
    uint32_t in_;
    DevBuf   params_, grads_, m_, v_, bf1_, bf2_, loss_;
    DevBuf   a1_, a2_, out_, dout_, d2_, d1_;
    uint32_t rows_   = 0;
    bool     dirty_  = true;
    uint64_t step_   = 0;
};

// Features of a connected relation set of a query (kFeatures floats).
constexpr uint32_t kFeatures = 24;

// Per-query statistics the features use: relation sizes and, per join class,
// the exact 2-way join selectivity of every related relation pair.
struct QueryStats {
    struct Edge {
        uint32_t a, b; // relations, a < b
        uint32_t eq;   // join class
        double   sel;  // |a join b| / (|a| |b|) on this class alone
    };
    std::vector<double> rows;
    std::vector<Edge>   edges;
};

// Featurisation of many sets of one query: the join classes, their members
// and measured edges, and the log sizes, prepared once. No allocation per set.
class SetFeaturizer {
public:
    SetFeaturizer(const Query& q, const QueryStats& st);
    // The edges the estimates use (see active_edges); returns their number (<= 64).
    uint32_t active(uint64_t set, const QueryStats::Edge** out) const;
    double   log10_independence(uint64_t set) const;
    void     featurize(uint64_t set, float* out) const; // kFeatures floats

private:
    struct Cls {
        std::vector<uint32_t>                members; // relations, sorted
        uint64_t                             mask = 0;
        std::vector<const QueryStats::Edge*> edges;
    };
    const QueryStats::Edge* edge(const Cls& c, uint32_t a, uint32_t b) const;
    uint32_t                nrels_ = 0;
    uint64_t                empty_ = 0; // relations without rows
    std::vector<double>     lrows_;
    std::vector<Cls>        classes_;
};

// Edges of `relset` the estimates use: per join class, the members inside the
// set in relation order, consecutive pairs (a spanning chain of the class).
std::vector<const QueryStats::Edge*> active_edges(const Query& q, const QueryStats& st, uint64_t relset);

void featurize(const Query& q, const QueryStats& st, uint64_t relset, float* out);

// log10 of the independence (System R) estimate: product of the relation
// sizes and of the selectivities of the active edges.
double log10_independence(const Query& q, const QueryStats& st, uint64_t relset);

// The learned estimate's base: the independence estimate in log10 units,
// at least 0 (one row). The network predicts log10 |set| minus this base.
double independence_base(const Query& q, const QueryStats& st, uint64_t relset);

// Connected relation sets of the query graph, each once (EnumerateCsg of
// Moerkotte & Neumann, VLDB 2006), singletons included; false when there are
// more than `limit`.
bool connected_subsets(const Query& q, uint32_t limit, std::vector<uint64_t>& out);

enum : uint32_t { EST_MAX = 0, EST_INDEPENDENCE = 1, EST_LEARNED = 2, EST_SAMPLE = 3 };

// Cardinality model of the join ordering.
//   EST_SAMPLE       the reference's planner (the default): the max rule over
//                    sampled relation cardinalities, re-sampled after every
//                    hash-table build (Samples)
//   EST_MAX          the max rule over exact relation sizes
//   EST_INDEPENDENCE relation sizes x exact pairwise selectivities
//   EST_LEARNED      the independence estimate corrected by the MLP's log10
//                    residual for every connected set, batched in one BF16 MFMA
//                    inference; sets whose predicted spread (sigma, log10
//                    units) exceeds `max_sigma` fall back to the sampling
//                    estimate (the max rule over sampled cardinalities)
// Pairwise selectivities come from exact 2-way counts on the device, cached
// per (table, column) pair for the context's lifetime: tables of a shared
// pool are joined by many plans. model() may run on several threads at once,
// each with its own context (counts and inference run on that context).
class Estimator {
public:
    uint32_t             mode      = EST_SAMPLE;
    std::shared_ptr<Mlp> mlp;
    float                max_sigma = 0.5f;
    uint32_t             max_sets  = 1u << 15;

    // `samples`: the query's sampled relations (EST_SAMPLE, and EST_LEARNED's fallback).
    CardFn     model(Context& ctx, const Query& q, const std::vector<const DevTable*>& in, EstimateReport& rep,
                     std::shared_ptr<Samples> samples = nullptr);
    QueryStats stats(Context& ctx, const Query& q, const std::vector<const DevTable*>& in, EstimateReport& rep);
    void       clear_pairs() {
        std::lock_guard<std::mutex> lock(pairs_mu_);
        pair_rows_.clear();
    }
    size_t     cached_pairs() const {
        std::lock_guard<std::mutex> lock(pairs_mu_);
        return pair_rows_.size();
    }

private:
    using PairKey = std::tuple<uint64_t, uint64_t, uint64_t, uint64_t>; // (uid, col) of each side, ordered
    mutable std::mutex          pairs_mu_;
    std::map<PairKey, uint64_t> pair_rows_;
};

} // namespace cgj
