// Batched joins: many independent two-table equi-joins executed by one
// persistent kernel launch over a device work queue, captured in a hipGraph.
//
// Each plan of the batch is one join of two scans (the shape of the contest's
// single-join queries, reference engine/query/QueryPlan.cpp:414-678 run once
// per query). One graph holds: fill kernels that clear the hash tables and
// counters, one persistent launch that takes every job's build tasks from a
// device queue, then one persistent launch per output width that takes the
// probe tasks (the one-probe pipeline over a row range) from their queue. The
// kernel boundary orders every probe after every build; no wave waits on
// another. Per-job tunings come from BatchTuner.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <tuple>
#include <vector>

#include "engine.hpp"

namespace cgj {

class Batch {
public:
    Batch(Context& ctx, const std::vector<cgj_plan>& plans, const std::vector<std::vector<const DevTable*>>& inputs);
    ~Batch();
    Batch(const Batch&)            = delete;
    Batch& operator=(const Batch&) = delete;

    // One graph launch: table fills + the persistent kernel(s). Re-runs with
    // larger page pools if a job's pool overflowed.
    void     run();
    // Searches the tunables per size class (each candidate is one whole-batch
    // run), keeps the fastest per class in the context's tuner and rebuilds.
    void     autotune();
    uint32_t size() const { return uint32_t(jobs_.size()); }
    uint64_t rows(uint32_t j) const;
    void     download(uint32_t j, cgj_result& out);

    double   ms_last   = 0; // device time of the last run (graph launch)
    uint32_t retries   = 0;
    uint64_t tasks     = 0;
    uint32_t waves     = 0; // persistent grid of the widest launch
    uint32_t launches  = 0; // kernel launches in the graph

private:
    struct Job {
        bool                 empty = false;
        std::vector<int32_t> types;
        uint32_t             nout = 0;
        const DevTable*      bt   = nullptr;
        const DevTable*      pt   = nullptr;
        uint32_t             bkey = 0;
        std::vector<uint32_t> battr;
        uint32_t             pkey = 0;
        LeanOutCol           out[kLeanOut]{};
        uint64_t             brows = 0, prows = 0;
        uint32_t             key_bytes = 8;
        BatchTune            tune;
        double               pool_scale = 1.0;
        // layout
        uint32_t cap_log = 0, words = 1, btasks = 0, ptasks = 0, page_chunk = 1;
        uint64_t ht_off = 0, bloom_off = 0, bloom_words = 0, pool_off = 0, pool_pages = 0;
    };

    void layout();
    void capture();
    void launch_once();
    void release_graph();

    Context&           ctx_;
    std::vector<Job>   jobs_;
    DevBuf             ht_, zero_, pool_, jobs_dev_, tasks_dev_;
    std::vector<PipeCounters> counters_;
    std::vector<BatchSync>    sync_;
    uint64_t           sync_off_ = 0, queue_off_ = 0, bloom_base_ = 0;
    std::vector<std::pair<uint32_t, std::pair<uint64_t, uint64_t>>> groups_; // nout -> task range
    hipGraph_t         graph_ = nullptr;
    hipGraphExec_t     exec_  = nullptr;
    hipEvent_t         ev0_ = nullptr, ev1_ = nullptr;
};

} // namespace cgj
