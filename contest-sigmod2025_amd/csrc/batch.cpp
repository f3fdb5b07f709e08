// Batched joins (host side): plan import per job, device layout, work queue,
// graph capture, runs with pool regrowth, and the per-size-class autotuner.
#include "batch.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>

#include "planner.hpp"

namespace cgj {

namespace {

uint32_t log2_floor(uint64_t x) {
    uint32_t l = 0;
    while ((x >> (l + 1)) != 0) ++l;
    return l;
}
uint32_t log2_ceil64(uint64_t x) {
    uint32_t l = 0;
    while ((uint64_t(1) << l) < x) ++l;
    return l;
}
uint64_t round_up(uint64_t x, uint64_t m) { return (x + m - 1) / m * m; }
uint32_t next_pow2_u32(uint64_t x, uint32_t lo, uint32_t hi) {
    uint64_t p = uint64_t(1) << log2_ceil64(std::max<uint64_t>(x, 1));
    return uint32_t(std::min<uint64_t>(std::max<uint64_t>(p, lo), hi));
}

} // namespace

// ---------------------------------------------------------------------------
// tuner
// ---------------------------------------------------------------------------
BatchTuner::Key BatchTuner::key(uint64_t build_rows, uint64_t probe_rows, uint32_t key_bytes) {
    uint64_t b = std::max<uint64_t>(build_rows, 1);
    return Key{log2_floor(b), std::min<uint32_t>(log2_floor(std::max<uint64_t>(probe_rows / b, 1)), 7), key_bytes};
}

BatchTune BatchTuner::pick(uint64_t build_rows, uint64_t probe_rows, uint32_t key_bytes) const {
    auto it = tuned_.find(key(build_rows, probe_rows, key_bytes));
    if (it != tuned_.end()) return it->second;
    BatchTune t;
    // Rules before measurement: a load factor of at most 0.5 keeps probe
    // sequences short (a wave waits for its longest); about four tasks per
    // side keep small jobs cheap to schedule and large ones balanced; no bloom
    // filter (its extra lookup only pays when most probes miss).
    t.max_load        = 0.5;
    t.bloom_bits      = 0;
    t.probe_task_rows = next_pow2_u32(probe_rows / 4, 16384, 262144);
    t.build_task_rows = next_pow2_u32(build_rows / 4, 8192, 65536);
    return t;
}

// ---------------------------------------------------------------------------
// batch
// ---------------------------------------------------------------------------
Batch::Batch(Context& ctx, const std::vector<cgj_plan>& plans, const std::vector<std::vector<const DevTable*>>& inputs)
    : ctx_(ctx) {
    if (plans.size() != inputs.size()) throw std::runtime_error("batch: one input list per plan");
    CGJ_HIP(hipEventCreate(&ev0_));
    CGJ_HIP(hipEventCreate(&ev1_));
    jobs_.resize(plans.size());
    for (size_t j = 0; j < plans.size(); ++j) {
        Job&                  J   = jobs_[j];
        const auto&           ins = inputs[j];
        std::vector<uint64_t> rows;
        std::vector<uint32_t> cols;
        for (const DevTable* t : ins) {
            rows.push_back(t->nrows);
            cols.push_back(uint32_t(t->cols.size()));
        }
        Query q = import_plan(plans[j], rows, cols);
        for (uint32_t a : q.out) J.types.push_back(q.attrs[a].type);
        J.nout = uint32_t(q.out.size());
        if (q.rels.size() != 2)
            throw std::runtime_error("batch: plan " + std::to_string(j) + " is not a join of two scans");
        auto shared = q.shared(1, 2);
        if (shared.size() != 1)
            throw std::runtime_error("batch: plan " + std::to_string(j) + " joins on " + std::to_string(shared.size()) +
                                     " classes (one equi-join key is supported)");
        if (J.nout > uint32_t(kLeanOut))
            throw std::runtime_error("batch: plan " + std::to_string(j) + " outputs more than 4 columns");
        for (int32_t t : J.types)
            if (t == CGJ_VARCHAR) throw std::runtime_error("batch: VARCHAR outputs are not batched");
        const uint32_t key_eq = shared[0];
        const uint32_t b      = q.rels[0].rows <= q.rels[1].rows ? 0 : 1, p = 1 - b;
        J.brows               = q.rels[b].rows;
        J.prows               = q.rels[p].rows;
        J.bt                  = ins[q.rels[b].table];
        J.pt                  = ins[q.rels[p].table];
        J.empty               = J.brows == 0 || J.prows == 0;
        J.bkey                = uint32_t(q.attrs[q.attr_of(b, key_eq)].col);
        J.pkey                = uint32_t(q.attrs[q.attr_of(p, key_eq)].col);
        J.key_bytes           = q.attrs[q.attr_of(b, key_eq)].type == CGJ_INT32 ? 4 : 8;
        std::vector<uint32_t> bclasses;
        for (uint32_t e : q.needed(1ull << b))
            if (e != key_eq) bclasses.push_back(e);
        if (bclasses.size() > size_t(kBatchAttrs))
            throw std::runtime_error("batch: build side carries more than 3 columns");
        for (uint32_t e : bclasses) J.battr.push_back(uint32_t(q.attrs[q.attr_of(b, e)].col));
        for (uint32_t c = 0; c < J.nout; ++c) {
            const QAttr& at = q.attrs[q.out[c]];
            LeanOutCol&  o  = J.out[c];
            o.type          = at.type;
            if (at.eq == key_eq) {
                o.src = SRC_KEY;
            } else if (at.rel == p) {
                o.src = SRC_SCAN;
                o.col = J.pt->cols[at.col].view(0);
            } else {
                o.src  = SRC_HT;
                o.word = 1 + uint32_t(std::find(bclasses.begin(), bclasses.end(), at.eq) - bclasses.begin());
            }
        }
        J.tune = ctx_.batch_tuner.pick(J.brows, J.prows, J.key_bytes);
    }
    layout();
    capture();
}

Batch::~Batch() {
    release_graph();
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
}

void Batch::release_graph() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_  = nullptr;
    graph_ = nullptr;
}

void Batch::layout() {
    const size_t n = jobs_.size();
    // zero arena: counters, sync, one queue counter per output width, blooms
    uint64_t zero_bytes = round_up(n * sizeof(PipeCounters), 256);
    sync_off_           = zero_bytes;
    zero_bytes += round_up(n * sizeof(BatchSync), 256);
    queue_off_ = zero_bytes;
    zero_bytes += 256;
    bloom_base_          = zero_bytes;
    uint64_t ht_words    = 0, bloom_words = 0, pool_bytes = 0;
    for (Job& J : jobs_) {
        if (J.empty) continue;
        const BatchTune& t = J.tune;
        J.words            = 1 + uint32_t(J.battr.size());
        J.cap_log          = std::max<uint32_t>(6, log2_ceil64(uint64_t(std::ceil(double(J.brows) / t.max_load))));
        J.ht_off           = ht_words;
        ht_words += round_up((uint64_t(1) << J.cap_log) * J.words, 32);
        J.bloom_words = t.bloom_bits ? (uint64_t(1) << log2_ceil64(std::max<uint64_t>(1, J.brows * t.bloom_bits / 64))) : 0;
        J.bloom_off   = bloom_words;
        bloom_words += J.bloom_words;
        J.btasks = uint32_t((J.brows + t.build_task_rows - 1) / t.build_task_rows);
        J.ptasks = uint32_t((J.prows + t.probe_task_rows - 1) / t.probe_task_rows);
        double bytes_per_row = 0;
        for (int32_t ty : J.types) bytes_per_row += (ty == CGJ_INT32 ? 4.0 : 8.0) + 0.125;
        const uint64_t data_pages =
            uint64_t(std::ceil(double(J.prows) * J.pool_scale * std::max(bytes_per_row, 0.125) / 8000.0)) + 1;
        J.page_chunk = uint32_t(std::min<uint64_t>(kPageChunk, std::max<uint64_t>(1, data_pages / J.ptasks + 1)));
        J.pool_pages = data_pages + uint64_t(J.ptasks) * (J.nout + J.page_chunk) + 4;
        J.pool_off   = pool_bytes;
        pool_bytes += round_up(J.pool_pages * (kPageSize + 16), 256);
    }
    zero_bytes += bloom_words * sizeof(uint64_t);
    ht_   = DevBuf(&ctx_.heap(), std::max<uint64_t>(ht_words, 2) * sizeof(uint64_t));
    zero_ = DevBuf(&ctx_.heap(), zero_bytes);
    pool_ = DevBuf(&ctx_.heap(), std::max<uint64_t>(pool_bytes, 256));

    uint8_t*      zb       = zero_.get<uint8_t>();
    PipeCounters* counters = reinterpret_cast<PipeCounters*>(zb);
    BatchSync*    sync     = reinterpret_cast<BatchSync*>(zb + sync_off_);
    uint64_t*     blooms   = reinterpret_cast<uint64_t*>(zb + bloom_base_);
    std::vector<BatchJob> host(n);
    std::memset(host.data(), 0, n * sizeof(BatchJob));
    for (size_t j = 0; j < n; ++j) {
        const Job& J = jobs_[j];
        BatchJob&  H = host[j];
        H.sync       = sync + j;
        if (J.empty) continue;
        const uint64_t cap  = uint64_t(1) << J.cap_log;
        uint64_t*      ht   = ht_.get<uint64_t>() + J.ht_off;
        uint64_t*      blm  = J.bloom_words ? blooms + J.bloom_off : nullptr;
        const uint32_t blog = J.bloom_words ? log2_ceil64(J.bloom_words) : 0;
        LeanArgs&      L    = H.probe;
        L.key               = J.pt->cols[J.pkey].view(0);
        L.driver            = DRIVER_PAGES;
        L.morsel_rows       = J.tune.probe_task_rows;
        L.nrows             = J.prows;
        L.nmorsels          = J.ptasks;
        L.ht                = ht;
        L.mask              = cap - 1;
        L.shift             = 64 - J.cap_log;
        L.words             = J.words;
        // 2-word entries (16-byte aligned) are read as aligned pairs; the
        // region is the whole table, so the pair's second slot wraps with it
        L.aligned_pairs     = J.words == 2 ? 1u : 0u;
        L.bloom             = blm;
        L.bloom_log         = blog;
        L.nout              = J.nout;
        for (uint32_t c = 0; c < J.nout; ++c) L.out[c] = J.out[c];
        L.pool       = pool_.get<uint8_t>() + J.pool_off;
        L.pool_pages = J.pool_pages;
        L.dir        = reinterpret_cast<uint32_t*>(L.pool + J.pool_pages * kPageSize);
        L.counters   = counters + j;
        L.page_chunk = J.page_chunk;
        L.abort_pages = ctx_.mem_bytes() / kPageSize;
        BatchBuild& B  = H.build;
        B.key          = J.bt->cols[J.bkey].view(0);
        B.nattr        = uint32_t(J.battr.size());
        for (uint32_t t = 0; t < B.nattr; ++t) B.attr[t] = J.bt->cols[J.battr[t]].view(0);
        B.morsel_rows = J.tune.build_task_rows;
        B.rows        = J.brows;
        B.morsels     = J.btasks;
        B.ht          = ht;
        B.mask        = cap - 1;
        B.shift       = 64 - J.cap_log;
        B.words       = J.words;
        B.bloom       = blm;
        B.bloom_log   = blog;
    }
    jobs_dev_ = DevBuf(&ctx_.heap(), std::max<size_t>(n, 1) * sizeof(BatchJob));
    CGJ_HIP(hipMemcpyAsync(jobs_dev_.get(), host.data(), n * sizeof(BatchJob), hipMemcpyHostToDevice, ctx_.stream()));

    // Work queues: one launch over every build task (largest jobs first),
    // then one launch per output width over the probe tasks.
    std::vector<BatchTask> all;
    groups_.clear();
    std::vector<uint32_t> order;
    for (uint32_t j = 0; j < n; ++j)
        if (!jobs_[j].empty) order.push_back(j);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        return jobs_[x].brows + jobs_[x].prows > jobs_[y].brows + jobs_[y].prows;
    });
    for (uint32_t j : order)
        for (uint32_t t = 0; t < jobs_[j].btasks; ++t) all.push_back(BatchTask{j, 0, t});
    if (!all.empty()) groups_.push_back({kBatchBuildKernel, {0, all.size()}});
    for (uint32_t no = 0; no <= uint32_t(kLeanOut); ++no) {
        const uint64_t first = all.size();
        for (uint32_t j : order)
            if (jobs_[j].nout == no)
                for (uint32_t t = 0; t < jobs_[j].ptasks; ++t) all.push_back(BatchTask{j, 1, t});
        if (all.size() > first) groups_.push_back({no, {first, all.size() - first}});
    }
    tasks      = all.size();
    tasks_dev_ = DevBuf(&ctx_.heap(), std::max<size_t>(all.size(), 1) * sizeof(BatchTask));
    if (!all.empty())
        CGJ_HIP(hipMemcpyAsync(tasks_dev_.get(), all.data(), all.size() * sizeof(BatchTask), hipMemcpyHostToDevice,
                               ctx_.stream()));
    CGJ_HIP(hipStreamSynchronize(ctx_.stream()));
    counters_.assign(n, PipeCounters{});
    sync_.assign(n, BatchSync{});
}

void Batch::capture() {
    release_graph();
    hipStream_t           s = ctx_.stream();
    std::vector<uint32_t> resident;
    for (auto& g : groups_) resident.push_back(batch_resident_waves(g.first)); // queried before capture
    CGJ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    // Clearing with fill kernels, not memset nodes, inside the graph.
    hipError_t e = launch_fill_u64(ht_.get<uint64_t>(), ht_.bytes() / sizeof(uint64_t), kEmpty, s);
    if (e == hipSuccess) e = launch_fill_u64(zero_.get<uint64_t>(), zero_.bytes() / sizeof(uint64_t), 0, s);
    launches = 2;
    waves    = 0;
    auto* queues = reinterpret_cast<unsigned long long*>(zero_.get<uint8_t>() + queue_off_);
    for (size_t g = 0; g < groups_.size() && e == hipSuccess; ++g) {
        const uint32_t no = groups_[g].first;
        const uint64_t first = groups_[g].second.first, cnt = groups_[g].second.second;
        e = launch_batch(jobs_dev_.get<BatchJob>(), tasks_dev_.get<BatchTask>() + first, cnt, queues + g, no,
                         resident[g], s);
        waves = std::max(waves, resident[g]);
        ++launches;
    }
    hipGraph_t  graph = nullptr;
    hipError_t  e2    = hipStreamEndCapture(s, &graph);
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        throw HipError(std::string("batch capture: ") + hipGetErrorString(e));
    }
    CGJ_HIP(e2);
    graph_ = graph;
    CGJ_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
}

void Batch::launch_once() {
    hipStream_t s = ctx_.stream();
    CGJ_HIP(hipEventRecord(ev0_, s));
    CGJ_HIP(hipGraphLaunch(exec_, s));
    CGJ_HIP(hipEventRecord(ev1_, s));
    CGJ_HIP(hipEventSynchronize(ev1_));
    float ms = 0;
    CGJ_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    ms_last = ms;
    const size_t n = jobs_.size();
    if (n) {
        CGJ_HIP(hipMemcpy(counters_.data(), zero_.get(), n * sizeof(PipeCounters), hipMemcpyDeviceToHost));
        CGJ_HIP(hipMemcpy(sync_.data(), zero_.get<uint8_t>() + sync_off_, n * sizeof(BatchSync),
                          hipMemcpyDeviceToHost));
    }
}

void Batch::run() {
    for (int attempt = 0;; ++attempt) {
        launch_once();
        bool grow = false;
        for (size_t j = 0; j < jobs_.size(); ++j) {
            if (sync_[j].dup & 2u)
                throw std::runtime_error("batch: job " + std::to_string(j) + ": hash-table insert found no empty slot");
            if (counters_[j].fault)
                throw std::runtime_error("batch: job " + std::to_string(j) + ": hash-table probe found no empty slot");
            if (counters_[j].overflow && counters_[j].pages >= ctx_.mem_bytes() / kPageSize)
                throw std::runtime_error("batch: job " + std::to_string(j) +
                                         ": result too large, more output pages than device memory holds");
            if (counters_[j].overflow) {
                // the pool was sized for one row per probe row; duplicates made more
                Job& J = jobs_[j];
                J.pool_scale *= std::max(2.0, double(counters_[j].tuples) / double(std::max<uint64_t>(J.prows, 1)) * 1.25);
                grow = true;
            }
        }
        if (!grow) return;
        if (attempt >= 3) throw std::runtime_error("batch: output pools keep overflowing");
        ++retries;
        layout();
        capture();
    }
}

uint64_t Batch::rows(uint32_t j) const {
    if (j >= jobs_.size()) throw std::out_of_range("batch job index");
    return jobs_[j].empty ? 0 : counters_[j].tuples;
}

void Batch::download(uint32_t j, cgj_result& out) {
    if (j >= jobs_.size()) throw std::out_of_range("batch job index");
    const Job& J = jobs_[j];
    PoolView   v;
    v.types = J.types;
    if (!J.empty) {
        const PipeCounters& c = counters_[j];
        v.nrows               = c.tuples;
        v.pool                = pool_.get<uint8_t>() + J.pool_off;
        v.pool_pages          = J.pool_pages;
        v.reserved_pages      = std::min<uint64_t>(c.pages, J.pool_pages);
        for (uint32_t i = 0; i < J.nout; ++i) v.total_pages += c.col_pages[i];
    }
    ctx_.download_pool(v, out);
}

void Batch::autotune() {
    // Coordinate search: each candidate value of one tunable is applied to
    // every job, the batch runs once, and each size class keeps the value with
    // the fewest measured task ticks per input row.
    using Key = BatchTuner::Key;
    auto class_of = [&](const Job& J) { return BatchTuner::key(J.brows, J.prows, J.key_bytes); };
    auto cost_by_class = [&]() {
        std::map<Key, double> cost;
        for (size_t j = 0; j < jobs_.size(); ++j) {
            const Job& J = jobs_[j];
            if (J.empty) continue;
            double ticks = double(sync_[j].build_clocks) + double(sync_[j].probe_clocks);
            cost[class_of(J)] += ticks / double(J.brows + J.prows);
        }
        return cost;
    };
    struct Knob {
        std::vector<double> values;
        void (*apply)(BatchTune&, double);
    };
    const std::vector<Knob> knobs = {
        {{16384, 65536, 262144}, [](BatchTune& t, double v) { t.probe_task_rows = uint32_t(v); }},
        {{8192, 32768, 131072}, [](BatchTune& t, double v) { t.build_task_rows = uint32_t(v); }},
        {{0.5, 0.35}, [](BatchTune& t, double v) { t.max_load = v; }},
        {{0, 16}, [](BatchTune& t, double v) { t.bloom_bits = uint32_t(v); }},
    };
    for (const Knob& k : knobs) {
        std::map<Key, std::pair<double, double>> best; // class -> (cost, value)
        for (double v : k.values) {
            for (Job& J : jobs_) k.apply(J.tune, v);
            layout();
            capture();
            run();
            for (auto& kv : cost_by_class()) {
                auto it = best.find(kv.first);
                if (it == best.end() || kv.second < it->second.first) best[kv.first] = {kv.second, v};
            }
        }
        for (Job& J : jobs_) {
            auto it = best.find(class_of(J));
            if (it != best.end()) k.apply(J.tune, it->second.second);
        }
    }
    for (const Job& J : jobs_)
        if (!J.empty) ctx_.batch_tuner.store(class_of(J), J.tune);
    layout();
    capture();
}

} // namespace cgj
