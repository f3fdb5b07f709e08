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
    CGJ_HIP(hipStreamSynchronize(s)); // other contexts' streams read the BF16 weights next
    dirty_ = false;
}

void Mlp::predict(const float* x, uint32_t n, float* out) {
    refresh_bf16();
    predict_on(ctx_, x, n, out);
}

void Mlp::predict_on(Context& on, const float* x, uint32_t n, float* out) const {
    if (n == 0) return;
    if (dirty_) throw std::runtime_error("network weights changed since their BF16 conversion");
    // One device block for features and outputs, and the copies through the
    // context's pinned buffer (direct transfers, no staging of pageable memory).
    hipStream_t    s   = on.stream();
    const size_t   xb  = size_t(n) * in_ * sizeof(float);
    const size_t   yb  = size_t(n) * kMlpOut * sizeof(float);
    const size_t   yo  = (xb + 255) / 256 * 256;
    DevBuf         dev(&on.heap(), yo + yb);
    uint8_t*       pin = on.small_pinned(yo + yb);
    std::memcpy(pin, x, xb);
    CGJ_HIP(hipMemcpyAsync(dev.get(), pin, xb, hipMemcpyHostToDevice, s));
    const float* P = params_.get<float>();
    float*       Y = reinterpret_cast<float*>(dev.get<uint8_t>() + yo);
    CGJ_HIP(launch_mlp_infer(dev.get<float>(), n, in_, bf1_.get<uint16_t>(), P + off_b1(), bf2_.get<uint16_t>(),
                             P + off_b2(), P + off_w3(), P + off_b3(), Y, s));
    CGJ_HIP(hipMemcpyAsync(pin + yo, Y, yb, hipMemcpyDeviceToHost, s));
    CGJ_HIP(hipStreamSynchronize(s));
    std::memcpy(out, pin + yo, yb);
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
// Per-query structures for featurising many sets: the join classes with their
// member relations (sorted) and their measured edges, the log sizes.
SetFeaturizer::SetFeaturizer(const Query& q, const QueryStats& st) {
    nrels_ = uint32_t(q.rels.size());
    for (uint32_t r = 0; r < nrels_; ++r) lrows_.push_back(std::log10(std::max(st.rows[r], 1.0)));
    for (uint32_t r = 0; r < nrels_; ++r) empty_ |= st.rows[r] <= 0 ? (1ull << r) : 0;
    for (uint32_t e = 0; e < q.eq_attrs.size(); ++e) {
        if (!q.is_join_eq(e)) continue;
        Cls c;
        for (uint32_t a : q.eq_attrs[e]) c.members.push_back(q.attrs[a].rel);
        std::sort(c.members.begin(), c.members.end());
        for (uint32_t r : c.members) c.mask |= 1ull << r;
        for (const auto& ed : st.edges)
            if (ed.eq == e) c.edges.push_back(&ed);
        classes_.push_back(std::move(c));
    }
}

const QueryStats::Edge* SetFeaturizer::edge(const Cls& c, uint32_t a, uint32_t b) const {
    for (const auto* ed : c.edges)
        if (ed->a == a && ed->b == b) return ed;
    return nullptr;
}

uint32_t SetFeaturizer::active(uint64_t set, const QueryStats::Edge** out) const {
    uint32_t n = 0;
    for (const Cls& c : classes_) {
        const uint64_t m = c.mask & set;
        if (!(m & (m - 1))) continue; // fewer than two members inside the set
        if (c.members.size() == 2) {
            if (!c.edges.empty()) out[n++] = c.edges[0];
            continue;
        }
        // consecutive members inside the set, in relation order
        int32_t prev = -1;
        for (uint32_t r : c.members) {
            if (!((set >> r) & 1)) continue;
            if (prev >= 0)
                if (const auto* ed = edge(c, uint32_t(prev), r)) out[n++] = ed;
            prev = int32_t(r);
        }
    }
    return n;
}

double SetFeaturizer::log10_independence(uint64_t set) const {
    if (set & empty_) return -30.0; // an empty input empties the join
    const QueryStats::Edge* ed[64];
    double                  l = 0;
    for (uint64_t t = set; t; t &= t - 1) l += lrows_[__builtin_ctzll(t)];
    const uint32_t n = active(set, ed);
    for (uint32_t i = 0; i < n; ++i) l += std::log10(std::max(ed[i]->sel, 1e-30));
    return l;
}

void SetFeaturizer::featurize(uint64_t set, float* f) const {
    std::fill(f, f + kFeatures, 0.f);
    const QueryStats::Edge* ed[64];
    const uint32_t          ne = active(set, ed);
    double                  lr[64], ls[64];
    uint32_t                nr = 0;
    double                  sum_lr = 0, max_lr = -1e300, min_lr = 1e300;
    for (uint64_t t = set; t; t &= t - 1) {
        const double v = lrows_[__builtin_ctzll(t)];
        lr[nr++]       = v;
        sum_lr += v;
        max_lr = std::max(max_lr, v);
        min_lr = std::min(min_lr, v);
    }
    uint8_t degree[64] = {0};
    double  sum_ls = 0, max_ls = -1e300, min_ls = 1e300, fan_big = 0, fan_small = 0;
    for (uint32_t i = 0; i < ne; ++i) {
        const double l  = std::log10(std::max(ed[i]->sel, 1e-30));
        const double la = lrows_[ed[i]->a], lb = lrows_[ed[i]->b];
        ls[i]           = l;
        sum_ls += l;
        max_ls = std::max(max_ls, l);
        min_ls = std::min(min_ls, l);
        fan_big += l + std::max(la, lb);   // log matches per row of the smaller side
        fan_small += l + std::min(la, lb); // ... of the larger side
        ++degree[ed[i]->a];
        ++degree[ed[i]->b];
    }
    const double n = double(nr), e = double(ne);
    f[0]           = float(n / 12.0);
    f[1]           = float(e / 12.0);
    if (nr) {
        f[2] = float(sum_lr / 32.0);
        f[3] = float(max_lr / 8.0);
        f[4] = float(min_lr / 8.0);
        f[5] = float(sum_lr / n / 8.0);
    }
    if (ne) {
        f[6] = float(sum_ls / 32.0);
        f[7] = float(max_ls / 8.0);
        f[8] = float(min_ls / 8.0);
        f[9] = float(sum_ls / e / 8.0);
    }
    const double indep = (set & empty_) ? -30.0 : sum_lr + sum_ls;
    f[10]              = float(std::max(indep, -8.0) / 8.0); // independence estimate
    f[11]              = nr ? float(max_lr / 8.0) : 0.f;     // max-of-inputs rule
    uint32_t maxdeg = 0, leaves = 0;
    for (uint64_t t = set; t; t &= t - 1) {
        const uint32_t d = degree[__builtin_ctzll(t)];
        maxdeg           = std::max(maxdeg, d);
        leaves += d == 1 ? 1 : 0;
    }
    f[12] = maxdeg >= 3 ? 1.f : 0.f;
    f[13] = float(maxdeg / 12.0);
    f[14] = float(leaves / 12.0);
    std::sort(lr, lr + nr, [](double x, double y) { return x > y; });
    for (uint32_t i = 0; i < 3 && i < nr; ++i) f[15 + i] = float(lr[i] / 8.0);
    std::sort(ls, ls + ne);
    for (uint32_t i = 0; i < 3 && i < ne; ++i) f[18 + i] = float(ls[i] / 8.0);
    f[21] = float(fan_big / 16.0);
    f[22] = float(fan_small / 16.0);
    f[23] = 1.f;
}

std::vector<const QueryStats::Edge*> active_edges(const Query& q, const QueryStats& st, uint64_t relset) {
    SetFeaturizer           fz(q, st);
    const QueryStats::Edge* ed[64];
    const uint32_t          n = fz.active(relset, ed);
    return std::vector<const QueryStats::Edge*>(ed, ed + n);
}

double log10_independence(const Query& q, const QueryStats& st, uint64_t relset) {
    return SetFeaturizer(q, st).log10_independence(relset);
}

double independence_base(const Query& q, const QueryStats& st, uint64_t relset) {
    return std::max(log10_independence(q, st, relset), 0.0);
}

void featurize(const Query& q, const QueryStats& st, uint64_t relset, float* f) {
    SetFeaturizer(q, st).featurize(relset, f);
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
                uint64_t rows   = 0;
                bool     cached = false;
                {
                    std::lock_guard<std::mutex> lock(pairs_mu_);
                    auto                        it = pair_rows_.find(k);
                    if (it != pair_rows_.end()) {
                        rows   = it->second;
                        cached = true;
                    }
                }
                if (!cached) {
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
                    // counted outside the lock: two threads may count one pair, with equal results
                    std::lock_guard<std::mutex> lock(pairs_mu_);
                    pair_rows_.emplace(k, rows);
                }
                const double denom = double(q.rels[x.rel].rows) * double(q.rels[y.rel].rows);
                st.edges.push_back(QueryStats::Edge{x.rel, y.rel, e, denom > 0 ? double(rows) / denom : 0.0});
            }
    }
    return st;
}

CardFn Estimator::model(Context& ctx, const Query& q, const std::vector<const DevTable*>& in, EstimateReport& rep,
                        std::shared_ptr<Samples> samples) {
    std::vector<double> base;
    for (const auto& r : q.rels) base.push_back(double(r.rows));
    // the sampling estimate of a set: the largest sampled cardinality in it,
    // read when asked (re-filtering after builds updates the samples)
    auto sampled = [samples, base](uint64_t set) {
        double c = 0;
        for (uint64_t t = set; t; t &= t - 1) {
            const uint32_t r = uint32_t(__builtin_ctzll(t));
            c                = std::max(c, samples ? samples->card(r) : base[r]);
        }
        return c;
    };
    if (mode == EST_SAMPLE) return sampled;
    if (mode == EST_MAX || q.rels.size() < 2) return max_card_model(q, base);
    auto st = std::make_shared<QueryStats>(stats(ctx, q, in, rep));
    if (mode == EST_INDEPENDENCE) {
        auto fz = std::make_shared<SetFeaturizer>(q, *st);
        return [st, fz](uint64_t set) { return std::pow(10.0, fz->log10_independence(set)); };
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
        const SetFeaturizer fz(q, *st);
        for (size_t i = 0; i < multi.size(); ++i) fz.featurize(multi[i], &x[i * kFeatures]);
        mlp->refresh_bf16(); // a no-op unless the weights changed (never while threads plan)
        mlp->predict_on(ctx, x.data(), uint32_t(multi.size()), y.data());
        // The network predicts a correction (mu) to the independence estimate
        // in log10 units, and its spread sigma = exp(s).
        const float max_s = std::log(std::max(max_sigma, 1e-6f));
        for (size_t i = 0; i < multi.size(); ++i) {
            const double base_l = std::max(fz.log10_independence(multi[i]), 0.0);
            double       l      = base_l + y[2 * i];
            if (!(y[2 * i + 1] <= max_s) || !std::isfinite(l)) {
                (*est)[multi[i]] = -1.0; // low confidence: the sampling estimate
                ++rep.fallbacks;
                continue;
            }
            (*est)[multi[i]] = std::pow(10.0, std::min(l, 300.0));
        }
    }
    rep.ms_model += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    auto fz = std::make_shared<SetFeaturizer>(q, *st);
    return [st, est, base, fz, sampled](uint64_t set) {
        if (!(set & (set - 1))) return base[__builtin_ctzll(set)];
        auto it = est->find(set);
        if (it != est->end()) return it->second >= 0 ? it->second : sampled(set);
        return std::pow(10.0, fz->log10_independence(set)); // more sets than the cap
    };
}

} // namespace cgj
