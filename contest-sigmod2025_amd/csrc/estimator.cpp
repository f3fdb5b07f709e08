// Learned cardinality estimation: the MLP (training and inference on the GPU)
// and the features of a connected relation set.
#include "estimator.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <stdexcept>
#include <unordered_map>

namespace cgj {

// ---------------------------------------------------------------------------
// MLP
// ---------------------------------------------------------------------------
Mlp::Mlp(Context& ctx, uint32_t in, uint64_t seed) : ctx_(ctx), in_(in) {
    if (in == 0 || in > kMlpMaxIn) throw std::runtime_error("mlp: 1..64 input features");
    const uint64_t     n = num_params();
    std::vector<float> p(n, 0.f);
    std::mt19937_64    rng(seed);
    // He-uniform initialisation for the ReLU layers, Glorot-uniform for the output
    auto fill = [&](uint64_t off, uint64_t count, double bound) {
        std::uniform_real_distribution<double> u(-bound, bound);
        for (uint64_t i = 0; i < count; ++i) p[off + i] = float(u(rng));
    };
    fill(off_w1(), uint64_t(in_) * kMlpHidden, std::sqrt(6.0 / in_));
    fill(off_w2(), uint64_t(kMlpHidden) * kMlpHidden, std::sqrt(6.0 / kMlpHidden));
    fill(off_w3(), uint64_t(kMlpHidden) * kMlpOut, std::sqrt(6.0 / (kMlpHidden + kMlpOut)));
    params_ = DevBuf(&ctx_.heap(), n * sizeof(float));
    grads_  = DevBuf(&ctx_.heap(), n * sizeof(float));
    m_      = DevBuf(&ctx_.heap(), n * sizeof(float));
    v_      = DevBuf(&ctx_.heap(), n * sizeof(float));
    loss_   = DevBuf(&ctx_.heap(), sizeof(float));
    const uint32_t in_pad = (in_ + 7) / 8 * 8;
    bf1_    = DevBuf(&ctx_.heap(), uint64_t(in_pad) * kMlpHidden * sizeof(uint16_t));
    bf2_    = DevBuf(&ctx_.heap(), uint64_t(kMlpHidden) * kMlpHidden * sizeof(uint16_t));
    set_params(p.data());
}

uint64_t Mlp::num_params() const { return off_b3() + kMlpOut; }

// Every copy and fill goes through the context's stream (non-blocking with
// respect to the null stream), followed by a synchronisation where the host
// reads or releases the memory.
void Mlp::set_params(const float* p) {
    const uint64_t n = num_params();
    hipStream_t    s = ctx_.stream();
    CGJ_HIP(hipMemcpyAsync(params_.get(), p, n * sizeof(float), hipMemcpyHostToDevice, s));
    CGJ_HIP(hipMemsetAsync(m_.get(), 0, n * sizeof(float), s));
    CGJ_HIP(hipMemsetAsync(v_.get(), 0, n * sizeof(float), s));
    CGJ_HIP(hipStreamSynchronize(s));
    step_  = 0;
    dirty_ = true;
}

void Mlp::get_params(float* p) const {
    hipStream_t s = ctx_.stream();
    CGJ_HIP(hipMemcpyAsync(p, params_.get(), num_params() * sizeof(float), hipMemcpyDeviceToHost, s));
    CGJ_HIP(hipStreamSynchronize(s));
}

void Mlp::ensure_rows(uint32_t m) {
    if (m <= rows_) return;
    rows_ = m;
    a1_   = DevBuf(&ctx_.heap(), uint64_t(m) * kMlpHidden * sizeof(float));
    a2_   = DevBuf(&ctx_.heap(), uint64_t(m) * kMlpHidden * sizeof(float));
    d1_   = DevBuf(&ctx_.heap(), uint64_t(m) * kMlpHidden * sizeof(float));
    d2_   = DevBuf(&ctx_.heap(), uint64_t(m) * kMlpHidden * sizeof(float));
    out_  = DevBuf(&ctx_.heap(), uint64_t(m) * kMlpOut * sizeof(float));
    dout_ = DevBuf(&ctx_.heap(), uint64_t(m) * kMlpOut * sizeof(float));
}

// Forward over m rows of device inputs; with `backward`, the loss gradient into
// grads_. Returns the summed loss when targets are given.
float Mlp::forward(const float* xb, const float* yb, uint32_t m, bool backward) {
    hipStream_t  s  = ctx_.stream();
    const float* P  = params_.get<float>();
    float*       G  = grads_.get<float>();
    const uint32_t H = kMlpHidden, O = kMlpOut;
    ensure_rows(m);
    float* A1 = a1_.get<float>();
    float* A2 = a2_.get<float>();
    float* Y  = out_.get<float>();
    CGJ_HIP(launch_gemm(m, H, in_, xb, in_, false, P + off_w1(), H, false, A1, H, P + off_b1(), true, s));
    CGJ_HIP(launch_gemm(m, H, H, A1, H, false, P + off_w2(), H, false, A2, H, P + off_b2(), true, s));
    CGJ_HIP(launch_gemm(m, O, H, A2, H, false, P + off_w3(), O, false, Y, O, P + off_b3(), false, s));
    if (!yb) return 0.f;
    float* dY = dout_.get<float>();
    CGJ_HIP(hipMemsetAsync(loss_.get(), 0, sizeof(float), s));
    CGJ_HIP(launch_nll(Y, yb, m, dY, loss_.get<float>(), s));
    if (backward) {
        float* D2 = d2_.get<float>();
        float* D1 = d1_.get<float>();
        CGJ_HIP(launch_gemm(H, O, m, A2, H, true, dY, O, false, G + off_w3(), O, nullptr, false, s));
        CGJ_HIP(launch_colsum(dY, m, O, G + off_b3(), s));
        CGJ_HIP(launch_gemm(m, H, O, dY, O, false, P + off_w3(), O, true, D2, H, nullptr, false, s));
        CGJ_HIP(launch_relu_grad(D2, A2, uint64_t(m) * H, s));
        CGJ_HIP(launch_gemm(H, H, m, A1, H, true, D2, H, false, G + off_w2(), H, nullptr, false, s));
        CGJ_HIP(launch_colsum(D2, m, H, G + off_b2(), s));
        CGJ_HIP(launch_gemm(m, H, H, D2, H, false, P + off_w2(), H, true, D1, H, nullptr, false, s));
        CGJ_HIP(launch_relu_grad(D1, A1, uint64_t(m) * H, s));
        CGJ_HIP(launch_gemm(in_, H, m, xb, in_, true, D1, H, false, G + off_w1(), H, nullptr, false, s));
        CGJ_HIP(launch_colsum(D1, m, H, G + off_b1(), s));
    }
    float loss = 0;
    CGJ_HIP(hipMemcpyAsync(&loss, loss_.get(), sizeof(float), hipMemcpyDeviceToHost, s));
    CGJ_HIP(hipStreamSynchronize(s));
    return loss;
}

float Mlp::loss_grad(const float* x, const float* y, uint32_t n, float* grads) {
    if (n == 0) throw std::runtime_error("mlp: no rows");
    hipStream_t s = ctx_.stream();
    DevBuf X(&ctx_.heap(), uint64_t(n) * in_ * sizeof(float)), Y(&ctx_.heap(), uint64_t(n) * sizeof(float));
    CGJ_HIP(hipMemcpyAsync(X.get(), x, uint64_t(n) * in_ * sizeof(float), hipMemcpyHostToDevice, s));
    CGJ_HIP(hipMemcpyAsync(Y.get(), y, uint64_t(n) * sizeof(float), hipMemcpyHostToDevice, s));
    const float sum = forward(X.get<float>(), Y.get<float>(), n, true);
    CGJ_HIP(hipMemcpyAsync(grads, grads_.get(), num_params() * sizeof(float), hipMemcpyDeviceToHost, s));
    CGJ_HIP(hipStreamSynchronize(s));
    return sum / float(n);
}

float Mlp::train(const float* x, const float* y, uint32_t n, uint32_t epochs, uint32_t batch, float lr,
                 float weight_decay, uint64_t seed) {
    if (n == 0 || batch == 0) throw std::runtime_error("mlp: no rows");
    hipStream_t s = ctx_.stream();
    DevBuf X(&ctx_.heap(), uint64_t(n) * in_ * sizeof(float)), Y(&ctx_.heap(), uint64_t(n) * sizeof(float));
    DevBuf Xb(&ctx_.heap(), uint64_t(batch) * in_ * sizeof(float)), Yb(&ctx_.heap(), uint64_t(batch) * sizeof(float));
    DevBuf idx(&ctx_.heap(), uint64_t(n) * sizeof(uint32_t));
    CGJ_HIP(hipMemcpyAsync(X.get(), x, uint64_t(n) * in_ * sizeof(float), hipMemcpyHostToDevice, s));
    CGJ_HIP(hipMemcpyAsync(Y.get(), y, uint64_t(n) * sizeof(float), hipMemcpyHostToDevice, s));
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::mt19937_64 rng(seed);
    const float     b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double          epoch_loss = 0;
    for (uint32_t e = 0; e < epochs; ++e) {
        std::shuffle(perm.begin(), perm.end(), rng);
        CGJ_HIP(hipMemcpyAsync(idx.get(), perm.data(), uint64_t(n) * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        CGJ_HIP(hipStreamSynchronize(s)); // perm is rewritten by the next shuffle
        epoch_loss = 0;
        for (uint32_t off = 0; off < n; off += batch) {
            const uint32_t m = std::min(batch, n - off);
            CGJ_HIP(launch_gather(X.get<float>(), Y.get<float>(), idx.get<uint32_t>() + off, m, in_, Xb.get<float>(),
                                  Yb.get<float>(), s));
            epoch_loss += forward(Xb.get<float>(), Yb.get<float>(), m, true);
            ++step_;
            const float bc1 = 1.f - std::pow(b1, float(step_)), bc2 = 1.f - std::pow(b2, float(step_));
            CGJ_HIP(launch_adam(params_.get<float>(), grads_.get<float>(), m_.get<float>(), v_.get<float>(),
                                num_params(), lr, b1, b2, eps, bc1, bc2, weight_decay, s));
        }
    }
    CGJ_HIP(hipStreamSynchronize(s));
    dirty_ = true;
    return float(epoch_loss / n);
}

void Mlp::refresh_bf16() {
    if (!dirty_) return;
    hipStream_t    s      = ctx_.stream();
    const uint32_t in_pad = (in_ + 7) / 8 * 8;
    CGJ_HIP(hipMemsetAsync(bf1_.get(), 0, uint64_t(in_pad) * kMlpHidden * sizeof(uint16_t), s));
    CGJ_HIP(launch_to_bf16(params_.get<float>() + off_w1(), bf1_.get<uint16_t>(), uint64_t(in_) * kMlpHidden, s));
    CGJ_HIP(launch_to_bf16(params_.get<float>() + off_w2(), bf2_.get<uint16_t>(), uint64_t(kMlpHidden) * kMlpHidden, s));
    dirty_ = false;
}

void Mlp::predict(const float* x, uint32_t n, float* out) {
    if (n == 0) return;
    hipStream_t s = ctx_.stream();
    refresh_bf16();
    DevBuf X(&ctx_.heap(), uint64_t(n) * in_ * sizeof(float)), Y(&ctx_.heap(), uint64_t(n) * kMlpOut * sizeof(float));
    CGJ_HIP(hipMemcpyAsync(X.get(), x, uint64_t(n) * in_ * sizeof(float), hipMemcpyHostToDevice, s));
    const float* P = params_.get<float>();
    CGJ_HIP(launch_mlp_infer(X.get<float>(), n, in_, bf1_.get<uint16_t>(), P + off_b1(), bf2_.get<uint16_t>(),
                             P + off_b2(), P + off_w3(), P + off_b3(), Y.get<float>(), s));
    CGJ_HIP(hipMemcpyAsync(out, Y.get(), uint64_t(n) * kMlpOut * sizeof(float), hipMemcpyDeviceToHost, s));
    CGJ_HIP(hipStreamSynchronize(s));
}

void Mlp::predict_f32(const float* x, uint32_t n, float* out) {
    if (n == 0) return;
    hipStream_t s = ctx_.stream();
    DevBuf      X(&ctx_.heap(), uint64_t(n) * in_ * sizeof(float));
    CGJ_HIP(hipMemcpyAsync(X.get(), x, uint64_t(n) * in_ * sizeof(float), hipMemcpyHostToDevice, s));
    forward(X.get<float>(), nullptr, n, false);
    CGJ_HIP(hipMemcpyAsync(out, out_.get(), uint64_t(n) * kMlpOut * sizeof(float), hipMemcpyDeviceToHost, s));
    CGJ_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
std::vector<const QueryStats::Edge*> active_edges(const Query& q, const QueryStats& st, uint64_t relset) {
    std::vector<const QueryStats::Edge*> out;
    for (uint32_t e = 0; e < q.eq_attrs.size(); ++e) {
        if (!q.is_join_eq(e)) continue;
        std::vector<uint32_t> members;
        for (uint32_t a : q.eq_attrs[e]) {
            const uint32_t r = q.attrs[a].rel;
            if ((relset >> r) & 1) members.push_back(r);
        }
        std::sort(members.begin(), members.end());
        for (size_t i = 1; i < members.size(); ++i)
            for (const auto& ed : st.edges)
                if (ed.eq == e && ed.a == members[i - 1] && ed.b == members[i]) {
                    out.push_back(&ed);
                    break;
                }
    }
    return out;
}

static double log10_rows(double r) { return std::log10(std::max(r, 1.0)); }
static double log10_sel(double s) { return std::log10(std::max(s, 1e-30)); }

double log10_independence(const Query& q, const QueryStats& st, uint64_t relset) {
    double l = 0;
    for (uint32_t r = 0; r < q.rels.size(); ++r)
        if ((relset >> r) & 1) {
            if (st.rows[r] <= 0) return -30.0; // an empty input empties the join
            l += log10_rows(st.rows[r]);
        }
    for (const auto* e : active_edges(q, st, relset)) l += log10_sel(e->sel);
    return l;
}

double independence_base(const Query& q, const QueryStats& st, uint64_t relset) {
    return std::max(log10_independence(q, st, relset), 0.0);
}

void featurize(const Query& q, const QueryStats& st, uint64_t relset, float* f) {
    std::fill(f, f + kFeatures, 0.f);
    std::vector<uint32_t> rels;
    for (uint32_t r = 0; r < q.rels.size(); ++r)
        if ((relset >> r) & 1) rels.push_back(r);
    std::vector<double> lr;
    for (uint32_t r : rels) lr.push_back(log10_rows(st.rows[r]));
    std::vector<double>   ls, fan_big, fan_small;
    std::vector<uint32_t> degree(q.rels.size(), 0);
    for (const auto* ed : active_edges(q, st, relset)) {
        const double l  = log10_sel(ed->sel);
        const double la = log10_rows(st.rows[ed->a]), lb = log10_rows(st.rows[ed->b]);
        ls.push_back(l);
        fan_big.push_back(l + std::max(la, lb));   // log matches per row of the smaller side
        fan_small.push_back(l + std::min(la, lb)); // ... of the larger side
        ++degree[ed->a];
        ++degree[ed->b];
    }
    auto sum = [](const std::vector<double>& v) { return std::accumulate(v.begin(), v.end(), 0.0); };
    const double n = double(rels.size()), e = double(ls.size());
    f[0]           = float(n / 12.0);
    f[1]           = float(e / 12.0);
    if (!lr.empty()) {
        f[2] = float(sum(lr) / 32.0);
        f[3] = float(*std::max_element(lr.begin(), lr.end()) / 8.0);
        f[4] = float(*std::min_element(lr.begin(), lr.end()) / 8.0);
        f[5] = float(sum(lr) / n / 8.0);
    }
    if (!ls.empty()) {
        f[6] = float(sum(ls) / 32.0);
        f[7] = float(*std::max_element(ls.begin(), ls.end()) / 8.0);
        f[8] = float(*std::min_element(ls.begin(), ls.end()) / 8.0);
        f[9] = float(sum(ls) / e / 8.0);
    }
    f[10] = float(std::max(log10_independence(q, st, relset), -8.0) / 8.0); // independence estimate
    f[11] = lr.empty() ? 0.f : float(*std::max_element(lr.begin(), lr.end()) / 8.0); // max-of-inputs rule
    uint32_t maxdeg = 0, leaves = 0;
    for (uint32_t r : rels) {
        maxdeg = std::max(maxdeg, degree[r]);
        leaves += degree[r] == 1 ? 1 : 0;
    }
    f[12] = maxdeg >= 3 ? 1.f : 0.f;
    f[13] = float(maxdeg / 12.0);
    f[14] = float(leaves / 12.0);
    std::vector<double> srt = lr;
    std::sort(srt.rbegin(), srt.rend());
    for (uint32_t i = 0; i < 3 && i < srt.size(); ++i) f[15 + i] = float(srt[i] / 8.0);
    std::vector<double> sls = ls;
    std::sort(sls.begin(), sls.end());
    for (uint32_t i = 0; i < 3 && i < sls.size(); ++i) f[18 + i] = float(sls[i] / 8.0);
    f[21] = float(sum(fan_big) / 16.0);
    f[22] = float(sum(fan_small) / 16.0);
    f[23] = 1.f;
}

// ---------------------------------------------------------------------------
// connected sets
// ---------------------------------------------------------------------------
namespace {
struct CsgEnum {
    const std::vector<uint64_t>& adj;
    uint32_t                     limit;
    std::vector<uint64_t>&       out;
    bool                         ok = true;

    uint64_t neighbours(uint64_t s) const {
        uint64_t n = 0;
        for (uint64_t t = s; t; t &= t - 1) n |= adj[__builtin_ctzll(t)];
        return n & ~s;
    }
    // Emits every connected superset of `s` grown through neighbours outside
    // `x`; each set is emitted once (the exclusion set grows with the frontier).
    void grow(uint64_t s, uint64_t x) {
        const uint64_t nb = neighbours(s) & ~x;
        if (!nb || !ok) return;
        for (uint64_t sub = nb; sub; sub = (sub - 1) & nb) {
            if (out.size() >= limit) {
                ok = false;
                return;
            }
            out.push_back(s | sub);
        }
        for (uint64_t sub = nb; sub; sub = (sub - 1) & nb) grow(s | sub, x | nb);
    }
};
} // namespace

bool connected_subsets(const Query& q, uint32_t limit, std::vector<uint64_t>& out) {
    out.clear();
    const auto     adj = q.adjacency();
    const uint32_t n   = uint32_t(q.rels.size());
    CsgEnum        en{adj, limit, out};
    for (uint32_t v = n; v-- > 0;) {
        if (out.size() >= limit) return false;
        out.push_back(1ull << v);
        // relations below v are excluded: sets whose lowest relation is v
        en.grow(1ull << v, (1ull << v) | ((1ull << v) - 1));
        if (!en.ok) return false;
    }
    return true;
}

// ---------------------------------------------------------------------------
// estimator
// ---------------------------------------------------------------------------
QueryStats Estimator::stats(Context& ctx, const Query& q, const std::vector<const DevTable*>& in,
                            EstimateReport& rep) {
    QueryStats st;
    for (const auto& r : q.rels) st.rows.push_back(double(r.rows));
    for (uint32_t e = 0; e < q.eq_attrs.size(); ++e) {
        if (!q.is_join_eq(e)) continue;
        const auto& members = q.eq_attrs[e];
        for (size_t i = 0; i < members.size(); ++i)
            for (size_t j = 0; j < members.size(); ++j) {
                const QAttr& x = q.attrs[members[i]];
                const QAttr& y = q.attrs[members[j]];
                if (x.rel >= y.rel) continue;
                const DevTable* tx = in[q.rels[x.rel].table];
                const DevTable* ty = in[q.rels[y.rel].table];
                PairKey         k  = std::make_tuple(tx->uid, x.col, ty->uid, y.col);
                if (std::make_pair(tx->uid, x.col) > std::make_pair(ty->uid, y.col))
                    k = std::make_tuple(ty->uid, y.col, tx->uid, x.col);
                auto it = pair_rows_.find(k);
                if (it == pair_rows_.end()) {
                    uint64_t rows = 0;
                    if (tx->nrows && ty->nrows) {
                        // two scans joined on the pair's columns
                        Query pq;
                        pq.rels.resize(2);
                        pq.rels[0].table = 0;
                        pq.rels[0].rows  = tx->nrows;
                        pq.rels[1].table = 1;
                        pq.rels[1].rows  = ty->nrows;
                        pq.attrs         = {QAttr{0, x.col, x.type, 0}, QAttr{1, y.col, y.type, 0}};
                        pq.eq_attrs      = {{0, 1}};
                        pq.rels[0].attrs = {0};
                        pq.rels[0].eqs   = {0};
                        pq.rels[1].attrs = {1};
                        pq.rels[1].eqs   = {0};
                        const auto t0    = std::chrono::steady_clock::now();
                        rows             = ctx.count_query(pq, {tx, ty});
                        rep.ms_stats +=
                            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                        ++rep.stat_joins;
                    }
                    it = pair_rows_.emplace(k, rows).first;
                }
                const double denom = double(q.rels[x.rel].rows) * double(q.rels[y.rel].rows);
                st.edges.push_back(QueryStats::Edge{x.rel, y.rel, e, denom > 0 ? double(it->second) / denom : 0.0});
            }
    }
    return st;
}

CardFn Estimator::model(Context& ctx, const Query& q, const std::vector<const DevTable*>& in, EstimateReport& rep) {
    std::vector<double> base;
    for (const auto& r : q.rels) base.push_back(double(r.rows));
    if (mode == EST_MAX || q.rels.size() < 2) return max_card_model(q, base);
    auto st = std::make_shared<QueryStats>(stats(ctx, q, in, rep));
    if (mode == EST_INDEPENDENCE) {
        return [st, &q](uint64_t set) { return std::pow(10.0, log10_independence(q, *st, set)); };
    }
    if (!mlp || mlp->in() != kFeatures) throw std::runtime_error("learned estimator without a network of matching width");
    const auto            t0 = std::chrono::steady_clock::now();
    std::vector<uint64_t> sets;
    auto                  est = std::make_shared<std::unordered_map<uint64_t, double>>();
    if (connected_subsets(q, max_sets, sets)) {
        std::vector<uint64_t> multi;
        for (uint64_t s : sets)
            if (s & (s - 1)) multi.push_back(s);
        rep.sets = uint32_t(multi.size());
        std::vector<float> x(multi.size() * kFeatures), y(multi.size() * kMlpOut);
        for (size_t i = 0; i < multi.size(); ++i) featurize(q, *st, multi[i], &x[i * kFeatures]);
        mlp->predict(x.data(), uint32_t(multi.size()), y.data());
        // The network predicts a correction (mu) to the independence estimate
        // in log10 units, and its spread sigma = exp(s).
        const float max_s = std::log(std::max(max_sigma, 1e-6f));
        for (size_t i = 0; i < multi.size(); ++i) {
            const double base_l = independence_base(q, *st, multi[i]);
            double       l      = base_l + y[2 * i];
            if (!(y[2 * i + 1] <= max_s) || !std::isfinite(l)) {
                l = base_l;
                ++rep.fallbacks;
            }
            (*est)[multi[i]] = std::pow(10.0, std::min(l, 300.0));
        }
    }
    rep.ms_model += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return [st, est, base, &q](uint64_t set) {
        if (!(set & (set - 1))) return base[__builtin_ctzll(set)];
        auto it = est->find(set);
        if (it != est->end()) return it->second;
        return std::pow(10.0, log10_independence(q, *st, set)); // more sets than the cap
    };
}

} // namespace cgj
