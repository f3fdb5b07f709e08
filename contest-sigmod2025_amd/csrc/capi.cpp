// C ABI (include/gpujoin.h) and the C++ contest surface (include/contest/plan.h).
#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/contest/plan.h"
#include "../../include/gpujoin.h"
#include "batch.hpp"
#include "engine.hpp"
#include "estimator.hpp"
#include "gen.hpp"
#include "noalloc.hpp"

struct cgj_context {
    std::unique_ptr<cgj::Context> impl;
    std::string                   error;
};
struct cgj_dtable {
    std::unique_ptr<cgj::DevTable> impl;
};
struct cgj_dresult {
    std::unique_ptr<cgj::DevResult> impl;
};
struct cgj_batch {
    std::unique_ptr<cgj::Batch> impl;
};
struct cgj_mlp {
    std::shared_ptr<cgj::Mlp> impl; // shared with a context's learned estimator
};

namespace {

template <class F>
int guarded(cgj_context* ctx, F&& f) {
    try {
        if (ctx) ctx->error.clear();
        f();
        return 0;
    } catch (const std::exception& e) {
        if (ctx) ctx->error = e.what();
        return -1;
    } catch (...) {
        if (ctx) ctx->error = "unknown error";
        return -1;
    }
}

void free_column(cgj_result_column& c) {
    for (uint64_t i = 0; i < c.num_pages; ++i)
        if (c.pages && c.pages[i]) ::operator delete(c.pages[i]);
    delete[] c.pages;
    c.pages     = nullptr;
    c.num_pages = 0;
}

} // namespace

namespace {
std::vector<const cgj::DevTable*> views_of(cgj_dtable* const* inputs, uint32_t n) {
    std::vector<const cgj::DevTable*> v;
    for (uint32_t i = 0; i < n; ++i) {
        if (!inputs[i]) throw std::runtime_error("null input table");
        v.push_back(inputs[i]->impl.get());
    }
    return v;
}
} // namespace

extern "C" {

cgj_context* cgj_context_create(int device) {
    try {
        auto* c = new cgj_context;
        c->impl = std::make_unique<cgj::Context>(device);
        return c;
    } catch (...) {
        return nullptr;
    }
}

void cgj_context_destroy(cgj_context* ctx) { delete ctx; }

const char* cgj_last_error(const cgj_context* ctx) { return ctx ? ctx->error.c_str() : "no context"; }

int cgj_execute(cgj_context* ctx, const cgj_plan* plan, const cgj_host_table* inputs, uint32_t num_inputs,
                cgj_result* out) {
    return guarded(ctx, [&] {
        std::memset(out, 0, sizeof(*out));
        std::vector<std::unique_ptr<cgj::DevTable>> owned;
        std::vector<const cgj::DevTable*>           views;
        // upload only the columns that the plan scans
        std::vector<std::vector<bool>> used(num_inputs);
        for (uint32_t i = 0; i < num_inputs; ++i) used[i].assign(inputs[i].num_columns, false);
        for (uint32_t n = 0; n < plan->num_nodes; ++n) {
            const cgj_node& nd = plan->nodes[n];
            if (nd.kind != CGJ_NODE_SCAN || nd.base_table_id >= num_inputs) continue;
            for (uint32_t a = 0; a < nd.num_output_attrs; ++a)
                if (nd.output_attrs[a].index < used[nd.base_table_id].size())
                    used[nd.base_table_id][nd.output_attrs[a].index] = true;
        }
        using clock  = std::chrono::steady_clock;
        auto ms      = [](clock::time_point a, clock::time_point b) {
            return std::chrono::duration<double, std::milli>(b - a).count();
        };
        const auto t0 = clock::now();
        for (uint32_t i = 0; i < num_inputs; ++i) {
            auto dt = ctx->impl->upload(inputs[i], &used[i]);
            views.push_back(dt.get());
            owned.push_back(std::move(dt));
        }
        const auto t1  = clock::now();
        auto       res = ctx->impl->execute(*plan, views);
        const auto t2  = clock::now();
        ctx->impl->download(*res, *out);
        const auto t3 = clock::now();
        ctx->impl->stats.ms_host_upload   = ms(t0, t1);
        ctx->impl->stats.ms_host_download = ms(t2, t3);
    });
}

void cgj_result_release(cgj_result* res) {
    if (!res || !res->columns) return;
    for (uint32_t c = 0; c < res->num_columns; ++c) free_column(res->columns[c]);
    delete[] res->columns;
    res->columns     = nullptr;
    res->num_columns = 0;
}

cgj_dtable* cgj_dtable_upload(cgj_context* ctx, const cgj_host_table* table) {
    cgj_dtable* out = nullptr;
    guarded(ctx, [&] {
        auto* t = new cgj_dtable;
        t->impl = ctx->impl->upload(*table);
        out     = t;
    });
    return out;
}

void cgj_dtable_free(cgj_context*, cgj_dtable* t) { delete t; }

uint64_t cgj_dtable_num_rows(const cgj_dtable* t) { return t ? t->impl->nrows : 0; }

int cgj_dtable_download(cgj_context* ctx, const cgj_dtable* t, uint32_t column, cgj_result_column* out) {
    return guarded(ctx, [&] {
        std::memset(out, 0, sizeof(*out));
        if (!t || column >= t->impl->cols.size()) throw std::out_of_range("no such column");
        try {
            ctx->impl->download_column(t->impl->cols[column], *out);
        } catch (...) {
            free_column(*out);
            throw;
        }
    });
}

int cgj_dtable_prepare(cgj_context* ctx, cgj_dtable* t) {
    return guarded(ctx, [&] {
        if (!t) throw std::runtime_error("null table");
        std::vector<cgj::DevColumn*> cols;
        for (auto& c : t->impl->cols)
            if (c.npages) cols.push_back(&c); // placeholders of unscanned columns have none
        ctx->impl->prepare_columns(cols);
    });
}

uint64_t cgj_dtable_num_pages(const cgj_dtable* t, uint32_t column) {
    if (!t || column >= t->impl->cols.size()) return 0;
    return t->impl->cols[column].npages;
}

int cgj_execute_many(cgj_context* ctx, const cgj_plan* plans, uint32_t num_plans, cgj_dtable* const* inputs,
                     const uint32_t* num_inputs, uint32_t workers, cgj_dresult** out) {
    return guarded(ctx, [&] {
        for (uint32_t i = 0; i < num_plans; ++i) out[i] = nullptr;
        std::vector<const cgj_plan*>                  ps;
        std::vector<std::vector<const cgj::DevTable*>> views(num_plans);
        uint64_t                                      k = 0;
        for (uint32_t i = 0; i < num_plans; ++i) {
            ps.push_back(&plans[i]);
            for (uint32_t j = 0; j < num_inputs[i]; ++j, ++k) {
                if (!inputs[k]) throw std::runtime_error("null input table");
                views[i].push_back(inputs[k]->impl.get());
            }
        }
        auto res = ctx->impl->execute_many(ps, views, workers);
        for (uint32_t i = 0; i < num_plans; ++i) {
            out[i]       = new cgj_dresult;
            out[i]->impl = std::move(res[i]);
        }
    });
}

int cgj_execute_device(cgj_context* ctx, const cgj_plan* plan, cgj_dtable* const* inputs, uint32_t num_inputs,
                       cgj_dresult** out) {
    return guarded(ctx, [&] {
        *out = nullptr;
        std::vector<const cgj::DevTable*> views;
        for (uint32_t i = 0; i < num_inputs; ++i) views.push_back(inputs[i]->impl.get());
        auto* r = new cgj_dresult;
        try {
            r->impl = ctx->impl->execute(*plan, views);
        } catch (...) {
            delete r;
            throw;
        }
        *out = r;
    });
}

cgj_batch* cgj_batch_create(cgj_context* ctx, const cgj_plan* plans, uint32_t num_plans, cgj_dtable* const* inputs,
                            const uint32_t* num_inputs) {
    cgj_batch* b = nullptr;
    guarded(ctx, [&] {
        std::vector<cgj_plan>                           ps(plans, plans + num_plans);
        std::vector<std::vector<const cgj::DevTable*>> ins(num_plans);
        uint64_t                                        k = 0;
        for (uint32_t i = 0; i < num_plans; ++i)
            for (uint32_t t = 0; t < num_inputs[i]; ++t) ins[i].push_back(inputs[k++]->impl.get());
        auto nb  = std::make_unique<cgj_batch>();
        nb->impl = std::make_unique<cgj::Batch>(*ctx->impl, ps, ins);
        b        = nb.release();
    });
    return b;
}

int cgj_batch_run(cgj_context* ctx, cgj_batch* b) {
    return guarded(ctx, [&] { b->impl->run(); });
}

int cgj_batch_autotune(cgj_context* ctx, cgj_batch* b) {
    return guarded(ctx, [&] { b->impl->autotune(); });
}

int cgj_batch_info(const cgj_batch* b, cgj_batch_stats* out) {
    if (!b || !out) return -1;
    out->ms_last  = b->impl->ms_last;
    out->tasks    = b->impl->tasks;
    out->waves    = b->impl->waves;
    out->launches = b->impl->launches;
    out->retries  = b->impl->retries;
    out->jobs     = b->impl->size();
    return 0;
}

uint64_t cgj_batch_num_rows(const cgj_batch* b, uint32_t plan) {
    if (!b || plan >= b->impl->size()) return 0;
    return b->impl->rows(plan);
}

int cgj_batch_download(cgj_context* ctx, cgj_batch* b, uint32_t plan, cgj_result* out) {
    return guarded(ctx, [&] {
        std::memset(out, 0, sizeof(*out));
        b->impl->download(plan, *out);
    });
}

void cgj_batch_free(cgj_context*, cgj_batch* b) { delete b; }

cgj_mlp* cgj_mlp_create(cgj_context* ctx, uint32_t in_features, uint64_t seed) {
    cgj_mlp* m = nullptr;
    guarded(ctx, [&] {
        auto nm  = std::make_unique<cgj_mlp>();
        nm->impl = std::make_shared<cgj::Mlp>(*ctx->impl, in_features, seed);
        m        = nm.release();
    });
    return m;
}

uint64_t cgj_mlp_num_params(const cgj_mlp* m) { return m ? m->impl->num_params() : 0; }

int cgj_mlp_set_params(cgj_context* ctx, cgj_mlp* m, const float* params) {
    return guarded(ctx, [&] { m->impl->set_params(params); });
}

int cgj_mlp_get_params(cgj_context* ctx, const cgj_mlp* m, float* params) {
    return guarded(ctx, [&] { m->impl->get_params(params); });
}

int cgj_mlp_loss_grad(cgj_context* ctx, cgj_mlp* m, const float* x, const float* y, uint32_t n, float* loss,
                      float* grads) {
    return guarded(ctx, [&] { *loss = m->impl->loss_grad(x, y, n, grads); });
}

int cgj_mlp_train(cgj_context* ctx, cgj_mlp* m, const float* x, const float* y, uint32_t n, uint32_t epochs,
                  uint32_t batch, float lr, float weight_decay, uint64_t seed, float* final_loss) {
    return guarded(ctx, [&] {
        float l = m->impl->train(x, y, n, epochs, batch, lr, weight_decay, seed);
        if (final_loss) *final_loss = l;
    });
}

int cgj_mlp_predict(cgj_context* ctx, cgj_mlp* m, const float* x, uint32_t n, float* out) {
    return guarded(ctx, [&] { m->impl->predict(x, n, out); });
}

int cgj_mlp_predict_f32(cgj_context* ctx, cgj_mlp* m, const float* x, uint32_t n, float* out) {
    return guarded(ctx, [&] { m->impl->predict_f32(x, n, out); });
}

void cgj_mlp_free(cgj_context*, cgj_mlp* m) { delete m; }

size_t cgj_stats_size(void) { return sizeof(cgj_stats); }

int cgj_set_option(cgj_context* ctx, const char* name, double value) {
    if (!ctx || !name) return -1;
    return ctx->impl->set_option(name, value) ? 0 : -1;
}

int cgj_count_device(cgj_context* ctx, const cgj_plan* plan, cgj_dtable* const* inputs, uint32_t num_inputs,
                     uint64_t* rows) {
    return guarded(ctx, [&] {
        std::vector<const cgj::DevTable*> views;
        for (uint32_t i = 0; i < num_inputs; ++i) views.push_back(inputs[i]->impl.get());
        *rows = ctx->impl->count(*plan, views);
    });
}

int cgj_sample_cardinalities(cgj_context* ctx, const cgj_plan* plan, cgj_dtable* const* inputs, uint32_t num_inputs,
                             double* cards, uint32_t cap, uint32_t* num_rels) {
    return guarded(ctx, [&] {
        const auto views = views_of(inputs, num_inputs);
        cgj::PlannedQuery pq;
        std::vector<uint64_t> rows;
        std::vector<uint32_t> cols;
        for (auto* t : views) {
            rows.push_back(t->nrows);
            cols.push_back(uint32_t(t->cols.size()));
        }
        const cgj::Query q = cgj::import_plan(*plan, rows, cols);
        *num_rels          = uint32_t(q.rels.size());
        bool empty         = false;
        for (const auto& r : q.rels) empty = empty || r.rows == 0;
        std::shared_ptr<cgj::Samples> s;
        if (!empty) s = ctx->impl->take_samples(q, views);
        for (uint32_t r = 0; r < q.rels.size() && r < cap; ++r) cards[r] = s ? s->card(r) : 0.0;
    });
}

int cgj_count_device_noalloc(cgj_context* ctx, const cgj_plan* plan, cgj_dtable* const* inputs, uint32_t num_inputs,
                             uint64_t* rows) {
    return guarded(ctx, [&] { *rows = cgj::count_join_noalloc(*ctx->impl, *plan, views_of(inputs, num_inputs)); });
}

uint32_t cgj_num_features(void) { return cgj::kFeatures; }

int cgj_set_estimator(cgj_context* ctx, uint32_t mode, cgj_mlp* mlp, float max_sigma) {
    return guarded(ctx, [&] {
        if (mode > cgj::EST_SAMPLE) throw std::runtime_error("unknown estimator mode");
        if (mode == cgj::EST_LEARNED && (!mlp || mlp->impl->in() != cgj::kFeatures))
            throw std::runtime_error("the learned estimator needs a network with cgj_num_features() inputs");
        auto& e     = ctx->impl->estimator();
        e.mode      = mode;
        e.mlp       = mode == cgj::EST_LEARNED ? mlp->impl : nullptr;
        e.max_sigma = max_sigma;
    });
}

int cgj_clear_statistics(cgj_context* ctx) {
    return guarded(ctx, [&] { ctx->impl->estimator().clear_pairs(); });
}

int64_t cgj_plan_features(cgj_context* ctx, const cgj_plan* plan, cgj_dtable* const* inputs, uint32_t num_inputs,
                          uint64_t* sets, float* features, uint64_t cap) {
    int64_t n  = -1;
    int     rc = guarded(ctx, [&] {
        auto                  views = views_of(inputs, num_inputs);
        std::vector<uint64_t> rows;
        std::vector<uint32_t> cols;
        for (auto* t : views) {
            rows.push_back(t->nrows);
            cols.push_back(uint32_t(t->cols.size()));
        }
        cgj::Query            q = cgj::import_plan(*plan, rows, cols);
        cgj::EstimateReport   rep;
        cgj::QueryStats       st = ctx->impl->estimator().stats(*ctx->impl, q, views, rep);
        std::vector<uint64_t> all;
        if (!cgj::connected_subsets(q, 1u << 20, all)) throw std::runtime_error("too many connected relation sets");
        uint64_t k = 0;
        for (uint64_t set : all) {
            if (!(set & (set - 1))) continue;
            if (k < cap) {
                sets[k] = set;
                cgj::featurize(q, st, set, features + k * cgj::kFeatures);
            }
            ++k;
        }
        n = int64_t(k);
    });
    return rc == 0 ? n : -1;
}

int cgj_count_subset(cgj_context* ctx, const cgj_plan* plan, cgj_dtable* const* inputs, uint32_t num_inputs,
                     uint64_t relset, uint64_t* rows) {
    return guarded(ctx, [&] { *rows = ctx->impl->count_subset(*plan, views_of(inputs, num_inputs), relset); });
}

int cgj_plan_joins(cgj_context* ctx, const cgj_plan* plan, cgj_dtable* const* inputs, uint32_t num_inputs,
                   uint64_t* build_sets, uint64_t* probe_sets, double* cards, uint32_t cap, uint32_t* num_joins) {
    return guarded(ctx, [&] {
        auto joins = ctx->impl->plan_tree(*plan, views_of(inputs, num_inputs));
        *num_joins = uint32_t(joins.size());
        for (uint32_t i = 0; i < joins.size() && i < cap; ++i) {
            build_sets[i] = joins[i].build;
            probe_sets[i] = joins[i].probe;
            cards[i]      = joins[i].card;
        }
    });
}

double cgj_copy_probe(cgj_context* ctx, uint64_t bytes, uint32_t reps) {
    double ms = -1.0;
    guarded(ctx, [&] { ms = ctx->impl->copy_probe(bytes, reps); });
    return ms;
}

uint64_t cgj_dresult_num_rows(const cgj_dresult* r) { return r ? r->impl->nrows : 0; }
uint64_t cgj_dresult_num_pages(const cgj_dresult* r) { return r ? r->impl->total_pages : 0; }

int cgj_dresult_download(cgj_context* ctx, const cgj_dresult* r, cgj_result* out) {
    return guarded(ctx, [&] {
        std::memset(out, 0, sizeof(*out));
        ctx->impl->download(*r->impl, *out);
    });
}

void cgj_dresult_free(cgj_context*, cgj_dresult* r) { delete r; }

int cgj_last_stats(const cgj_context* ctx, cgj_stats* out) {
    if (!ctx || !out) return -1;
    *out = ctx->impl->stats;
    return 0;
}

int cgj_scan_decode(cgj_context* ctx, const cgj_dtable* t, uint32_t column, uint64_t* slots) {
    return guarded(ctx, [&] {
        if (column >= t->impl->cols.size()) throw std::runtime_error("column out of range");
        const cgj::DevColumn& c = t->impl->cols[column];
        if (c.nrows && !c.npages) throw std::runtime_error("column was not uploaded");
        cgj::DevBuf out(&ctx->impl->heap(), std::max<uint64_t>(c.nrows, 1) * sizeof(uint64_t));
        CGJ_HIP(cgj::launch_decode(c.view(0), c.nrows, out.get<uint64_t>(), ctx->impl->stream()));
        CGJ_HIP(hipMemcpyAsync(slots, out.get(), c.nrows * sizeof(uint64_t), hipMemcpyDeviceToHost,
                               ctx->impl->stream()));
        CGJ_HIP(hipStreamSynchronize(ctx->impl->stream()));
    });
}

int cgj_encode_column(int32_t type, const void* values, const char* chars, const uint8_t* valid, uint64_t n,
                      cgj_result_column* out) {
    try {
        std::memset(out, 0, sizeof(*out));
        Column col(static_cast<DataType>(type));
        switch (type) {
        case CGJ_INT32: {
            ColumnInserter<int32_t> ins(col);
            auto*                   v = static_cast<const int32_t*>(values);
            for (uint64_t i = 0; i < n; ++i) (valid && !valid[i]) ? ins.insert_null() : ins.insert(v[i]);
            ins.finalize();
            break;
        }
        case CGJ_INT64: {
            ColumnInserter<int64_t> ins(col);
            auto*                   v = static_cast<const int64_t*>(values);
            for (uint64_t i = 0; i < n; ++i) (valid && !valid[i]) ? ins.insert_null() : ins.insert(v[i]);
            ins.finalize();
            break;
        }
        case CGJ_FP64: {
            ColumnInserter<double> ins(col);
            auto*                  v = static_cast<const double*>(values);
            for (uint64_t i = 0; i < n; ++i) (valid && !valid[i]) ? ins.insert_null() : ins.insert(v[i]);
            ins.finalize();
            break;
        }
        case CGJ_VARCHAR: {
            ColumnInserter<std::string> ins(col);
            auto*                       off = static_cast<const uint64_t*>(values);
            for (uint64_t i = 0; i < n; ++i) {
                if (valid && !valid[i])
                    ins.insert_null();
                else
                    ins.insert(std::string_view(chars + off[i], off[i + 1] - off[i]));
            }
            ins.finalize();
            break;
        }
        default: return -1;
        }
        out->type      = type;
        out->num_pages = col.pages.size();
        out->pages     = new void*[col.pages.size() ? col.pages.size() : 1];
        for (size_t i = 0; i < col.pages.size(); ++i) out->pages[i] = col.pages[i];
        col.pages.clear(); // ownership moves to `out`
        return 0;
    } catch (...) {
        return -1;
    }
}

void cgj_column_release(cgj_result_column* col) {
    if (col) free_column(*col);
}

cgj_dtable* cgj_dtable_generate(cgj_context* ctx, uint64_t seed, uint64_t num_rows, const cgj_gen_column* columns,
                                uint32_t num_columns) {
    cgj_dtable* out = nullptr;
    guarded(ctx, [&] {
        auto* t = new cgj_dtable;
        t->impl = ctx->impl->generate(seed, num_rows, columns, num_columns);
        out     = t;
    });
    return out;
}

int cgj_host_generate(uint64_t seed, uint64_t first_row, uint64_t num_rows, const cgj_gen_column* g,
                      cgj_result_column* out) {
    try {
        std::memset(out, 0, sizeof(*out));
        if (g->type == CGJ_VARCHAR) {
            if (g->kind != CGJ_GEN_ASCII) return -1;
            // ASCII strings drawn from the value stream: length uniform in
            // [lo, hi]; a `miss_frac` share of rows carries a long string.
            std::vector<uint8_t>  valid(num_rows);
            std::vector<uint64_t> offs(num_rows + 1, 0);
            std::string           chars;
            uint64_t              base = cgj::stream_base(seed, g->stream);
            for (uint64_t i = 0; i < num_rows; ++i) {
                uint64_t r = first_row + i;
                valid[i]   = !cgj::gen_is_null(seed, g->null_stream, g->null_frac, r);
                uint64_t x = cgj::draw(base, r);
                uint64_t len;
                if (cgj::unit(cgj::sm_mix(x ^ 0x5851f42d4c957f2dull)) < g->miss_frac)
                    len = 8186 + x % (16000 - 8186);
                else
                    len = g->lo + (g->hi >= g->lo ? x % (g->hi - g->lo + 1) : 0);
                if (valid[i]) {
                    uint64_t z = x;
                    for (uint64_t k = 0; k < len; ++k) {
                        if ((k & 7) == 0) z = cgj::sm_mix(z + cgj::kGolden);
                        chars.push_back(char('a' + (z >> (8 * (k & 7))) % 26));
                    }
                }
                offs[i + 1] = chars.size();
            }
            return cgj_encode_column(CGJ_VARCHAR, offs.data(), chars.data(), valid.data(), num_rows, out);
        }
        std::vector<uint8_t>  valid(num_rows);
        std::vector<uint64_t> vals(num_rows);
        for (uint64_t i = 0; i < num_rows; ++i) {
            uint64_t r = first_row + i;
            valid[i]   = !cgj::gen_is_null(seed, g->null_stream, g->null_frac, r);
            vals[i]    = cgj::gen_value(g->type, g->kind, seed, g->stream, g->domain ? g->domain : 1, g->zipf_s,
                                        g->miss_frac, g->lo, g->hi, r);
        }
        if (g->type == CGJ_INT32) {
            std::vector<int32_t> v32(num_rows);
            for (uint64_t i = 0; i < num_rows; ++i) v32[i] = int32_t(uint32_t(vals[i]));
            return cgj_encode_column(g->type, v32.data(), nullptr, valid.data(), num_rows, out);
        }
        return cgj_encode_column(g->type, vals.data(), nullptr, valid.data(), num_rows, out);
    } catch (...) {
        return -1;
    }
}

} // extern "C"

// ---------------------------------------------------------------------------
// contest surface: Contest::build_context / execute / destroy_context
// (reference engine/Execute.cpp:28-49, include/plan.h:337-344)
// ---------------------------------------------------------------------------
namespace Contest {

void* build_context() {
    cgj_context* ctx = cgj_context_create(0);
    if (!ctx) throw std::runtime_error("cannot create the GPU context");
    return ctx;
}

void destroy_context(void* context) { cgj_context_destroy(static_cast<cgj_context*>(context)); }

ColumnarTable execute(const Plan& plan, void* context) {
    auto* ctx = static_cast<cgj_context*>(context);
    std::unique_ptr<cgj_context, void (*)(cgj_context*)> own(nullptr, cgj_context_destroy);
    if (!ctx) {
        ctx = cgj_context_create(0);
        if (!ctx) throw std::runtime_error("cannot create the GPU context");
        own.reset(ctx);
    }
    // POD view of the plan
    std::vector<std::vector<cgj_attr>> attrs(plan.nodes.size());
    std::vector<cgj_node>              nodes(plan.nodes.size());
    for (size_t i = 0; i < plan.nodes.size(); ++i) {
        const PlanNode& pn = plan.nodes[i];
        for (auto& [idx, dt] : pn.output_attrs) attrs[i].push_back(cgj_attr{idx, int32_t(dt)});
        cgj_node& n        = nodes[i];
        n                  = cgj_node{};
        n.output_attrs     = attrs[i].data();
        n.num_output_attrs = uint32_t(attrs[i].size());
        if (auto* s = std::get_if<ScanNode>(&pn.data)) {
            n.kind          = CGJ_NODE_SCAN;
            n.base_table_id = s->base_table_id;
        } else {
            const JoinNode& j = std::get<JoinNode>(pn.data);
            n.kind            = CGJ_NODE_JOIN;
            n.build_left      = j.build_left ? 1 : 0;
            n.left            = j.left;
            n.right           = j.right;
            n.left_attr       = j.left_attr;
            n.right_attr      = j.right_attr;
        }
    }
    std::vector<std::vector<cgj_host_column>> cols(plan.inputs.size());
    std::vector<cgj_host_table>               tables(plan.inputs.size());
    for (size_t t = 0; t < plan.inputs.size(); ++t) {
        for (const Column& c : plan.inputs[t].columns)
            cols[t].push_back(cgj_host_column{reinterpret_cast<const void* const*>(c.pages.data()), c.pages.size(),
                                              int32_t(c.type)});
        tables[t] = cgj_host_table{plan.inputs[t].num_rows, cols[t].data(), uint32_t(cols[t].size())};
    }
    cgj_plan   p{nodes.data(), uint32_t(nodes.size()), plan.root};
    cgj_result r{};
    if (cgj_execute(ctx, &p, tables.data(), uint32_t(tables.size()), &r) != 0)
        throw std::runtime_error(std::string("execute: ") + cgj_last_error(ctx));
    ColumnarTable out;
    out.num_rows = r.num_rows;
    for (uint32_t c = 0; c < r.num_columns; ++c) {
        Column col(static_cast<DataType>(r.columns[c].type));
        for (uint64_t i = 0; i < r.columns[c].num_pages; ++i) {
            col.pages.push_back(static_cast<Page*>(r.columns[c].pages[i]));
            r.columns[c].pages[i] = nullptr;
        }
        out.columns.push_back(std::move(col));
    }
    cgj_result_release(&r);
    return out;
}

} // namespace Contest
