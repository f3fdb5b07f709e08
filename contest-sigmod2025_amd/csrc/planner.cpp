// Plan import, join ordering and join tuning (host).
#include "planner.hpp"

#include <algorithm>
#include <map>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_map>

#include "engine.hpp"

namespace cgj {

std::set<uint32_t> Query::produced(uint64_t relset) const {
    std::set<uint32_t> s;
    for (uint32_t r = 0; r < rels.size(); ++r)
        if ((relset >> r) & 1) s.insert(rels[r].eqs.begin(), rels[r].eqs.end());
    return s;
}

std::vector<uint32_t> Query::needed(uint64_t relset) const {
    std::set<uint32_t> inside  = produced(relset);
    std::set<uint32_t> outside = produced(all_rels() & ~relset);
    for (uint32_t a : out) outside.insert(attrs[a].eq);
    std::vector<uint32_t> r;
    for (uint32_t e : inside)
        if (outside.count(e)) r.push_back(e);
    return r;
}

std::vector<uint32_t> Query::shared(uint64_t a, uint64_t b) const {
    std::set<uint32_t>    pa = produced(a), pb = produced(b);
    std::vector<uint32_t> r;
    for (uint32_t e : pa)
        if (pb.count(e)) r.push_back(e);
    return r;
}

uint32_t Query::attr_of(uint32_t rel, uint32_t eq) const {
    for (uint32_t a : rels[rel].attrs)
        if (attrs[a].eq == eq) return a;
    throw std::runtime_error("relation does not produce class");
}

std::vector<uint64_t> Query::adjacency() const {
    std::vector<uint64_t> adj(rels.size(), 0);
    for (uint32_t e = 0; e < eq_attrs.size(); ++e) {
        uint64_t members = 0;
        for (uint32_t a : eq_attrs[e]) members |= 1ull << attrs[a].rel;
        for (uint32_t r = 0; r < rels.size(); ++r)
            if ((members >> r) & 1) adj[r] |= members & ~(1ull << r);
    }
    return adj;
}

namespace {
struct UnionFind {
    std::vector<uint32_t> p;
    uint32_t add() {
        p.push_back(uint32_t(p.size()));
        return uint32_t(p.size() - 1);
    }
    uint32_t find(uint32_t x) {
        while (p[x] != x) x = p[x] = p[p[x]];
        return x;
    }
    void merge(uint32_t a, uint32_t b) {
        a = find(a);
        b = find(b);
        if (a != b) p[std::max(a, b)] = std::min(a, b);
    }
};
} // namespace

Query import_plan(const cgj_plan& plan, const std::vector<uint64_t>& table_rows,
                  const std::vector<uint32_t>& table_cols) {
    Query     q;
    UnionFind uf;
    // returns the node's output attribute list
    std::function<std::vector<uint32_t>(uint64_t, int)> rec = [&](uint64_t id, int depth) {
        if (id >= plan.num_nodes) throw std::runtime_error("plan node id out of range");
        if (depth > 256) throw std::runtime_error("plan is too deep");
        const cgj_node&       n = plan.nodes[id];
        std::vector<uint32_t> out;
        if (n.kind == CGJ_NODE_SCAN) {
            if (n.base_table_id >= table_rows.size()) throw std::runtime_error("scan of a missing input table");
            if (q.rels.size() >= 64) throw std::runtime_error("more than 64 scans in one plan");
            uint32_t rel = uint32_t(q.rels.size());
            QRel     r;
            r.table = n.base_table_id;
            r.rows  = table_rows[n.base_table_id];
            for (uint32_t i = 0; i < n.num_output_attrs; ++i) {
                uint64_t col = n.output_attrs[i].index;
                if (r.rows > 0 && col >= table_cols[n.base_table_id])
                    throw std::runtime_error("scan of a missing column");
                uint32_t a = uint32_t(q.attrs.size());
                q.attrs.push_back(QAttr{rel, col, n.output_attrs[i].type, 0});
                uf.add();
                r.attrs.push_back(a);
                out.push_back(a);
            }
            q.rels.push_back(std::move(r));
            return out;
        }
        if (n.kind != CGJ_NODE_JOIN) throw std::runtime_error("unknown plan node kind");
        std::vector<uint32_t> l = rec(n.left, depth + 1);
        std::vector<uint32_t> r = rec(n.right, depth + 1);
        if (n.left_attr >= l.size() || n.right_attr >= r.size()) throw std::runtime_error("join attribute out of range");
        uf.merge(l[n.left_attr], r[n.right_attr]);
        for (uint32_t i = 0; i < n.num_output_attrs; ++i) {
            uint64_t idx = n.output_attrs[i].index;
            if (idx < l.size())
                out.push_back(l[idx]);
            else if (idx - l.size() < r.size())
                out.push_back(r[idx - l.size()]);
            else
                throw std::runtime_error("output attribute out of range");
        }
        return out;
    };
    q.out = rec(plan.root, 0);

    std::unordered_map<uint32_t, uint32_t> cls;
    for (uint32_t a = 0; a < q.attrs.size(); ++a) {
        uint32_t root = uf.find(a);
        auto     it   = cls.find(root);
        if (it == cls.end()) {
            it = cls.emplace(root, uint32_t(q.eq_attrs.size())).first;
            q.eq_attrs.emplace_back();
        }
        q.attrs[a].eq = it->second;
        q.eq_attrs[it->second].push_back(a);
    }
    for (auto& r : q.rels) {
        for (uint32_t a : r.attrs) {
            uint32_t e = q.attrs[a].eq;
            if (std::find(r.eqs.begin(), r.eqs.end(), e) != r.eqs.end())
                throw std::runtime_error("a relation joins two of its own columns");
            r.eqs.push_back(e);
        }
        std::sort(r.eqs.begin(), r.eqs.end());
    }
    for (uint32_t e = 0; e < q.eq_attrs.size(); ++e) {
        if (!q.is_join_eq(e)) continue;
        for (uint32_t a : q.eq_attrs[e])
            if (q.attrs[a].type == CGJ_VARCHAR) throw std::runtime_error("VARCHAR join keys are not supported");
    }
    return q;
}

bool QFilter::pass(uint64_t v) const {
    if (v == ~0ull) return false; // NULL
    switch (kind) {
    case FLT_EQ: return v == a;
    case FLT_EQ2: return v == a || v == b;
    case FLT_RANGE: return v >= a && v <= b;
    case FLT_TINY: return tiny[tiny_slot(v, tiny_hash, tiny_bits)] == v;
    default: throw std::runtime_error("restriction kind not evaluated on the host");
    }
}

bool small_restriction(std::vector<uint64_t> keys, uint32_t attr, QFilter& out) {
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    if (keys.empty() || keys.size() > 32) return false;
    out      = QFilter{};
    out.attr = attr;
    if (keys.size() == 1) {
        out.kind = FLT_EQ;
        out.a    = keys[0];
        return true;
    }
    if (keys.size() == 2) {
        out.kind = FLT_EQ2;
        out.a    = keys[0];
        out.b    = keys[1];
        return true;
    }
    if (keys.back() - keys.front() == keys.size() - 1) { // a contiguous run
        out.kind = FLT_RANGE;
        out.a    = keys.front();
        out.b    = keys.back();
        return true;
    }
    if (keys.size() > 16) return false;
    static const std::pair<uint32_t, uint32_t> kTry[] = {{3, TINY_MOD},  {4, TINY_MOD},   {5, TINY_MOD},  {4, TINY_FIBO},
                                                         {5, TINY_FIBO}, {4, TINY_FIBO2}, {5, TINY_FIBO2}};
    for (auto [bits, hash] : kTry) {
        std::vector<uint64_t> slots(size_t(1) << bits, ~0ull);
        bool                  ok = true;
        for (uint64_t k : keys) {
            uint64_t& sl = slots[tiny_slot(k, hash, bits)];
            if (sl != ~0ull) {
                ok = false;
                break;
            }
            sl = k;
        }
        if (!ok) continue;
        out.kind      = FLT_TINY;
        out.tiny_bits = bits;
        out.tiny_hash = hash;
        out.tiny      = std::move(slots);
        return true;
    }
    return false;
}

SingletonReduction eliminate_singletons(const Query& q, const std::vector<std::vector<uint64_t>>& values) {
    SingletonReduction red;
    const uint32_t     n = uint32_t(q.rels.size());
    for (uint32_t a : q.out) red.types.push_back(q.attrs[a].type);
    const auto adj       = q.adjacency();
    auto       connected = [&](uint64_t set) {
        if (!set) return false;
        uint64_t seen = set & (~set + 1), frontier = seen;
        while (frontier) {
            uint64_t next = 0;
            for (uint64_t t = frontier; t; t &= t - 1) next |= adj[__builtin_ctzll(t)];
            frontier = next & set & ~seen;
            seen |= frontier;
        }
        return seen == set;
    };
    // remove one-row relations in order while the rest stays connected
    uint64_t keep = q.all_rels();
    for (uint32_t r = 0; r < n; ++r) {
        if (q.rels[r].rows != 1 || r >= values.size() || values[r].size() != q.rels[r].attrs.size()) continue;
        if (__builtin_popcountll(keep) <= 1) break;
        const uint64_t rest = keep & ~(1ull << r);
        if (connected(rest)) keep = rest;
    }
    red.removed.assign(n, -1);
    if (keep == q.all_rels()) {
        red.q = q;
        for (uint32_t r = 0; r < n; ++r) red.removed[r] = int32_t(r);
        for (uint32_t i = 0; i < q.out.size(); ++i) red.out_pos.push_back(i);
        return red;
    }
    // the constants of the removed relations' classes
    std::map<uint32_t, uint64_t> cst;
    for (uint32_t r = 0; r < n; ++r) {
        if ((keep >> r) & 1) continue;
        ++red.eliminated;
        for (size_t i = 0; i < q.rels[r].attrs.size(); ++i) {
            const uint32_t e = q.attrs[q.rels[r].attrs[i]].eq;
            const uint64_t v = values[r][i];
            if (q.is_join_eq(e) && v == ~0ull) red.empty = true; // NULL never joins
            auto it = cst.find(e);
            if (it != cst.end() && it->second != v) red.empty = true;
            cst[e] = v;
        }
    }
    // the remaining relations, their attributes and classes
    Query&                s = red.q;
    std::vector<int32_t>  eqmap(q.eq_attrs.size(), -1), attrmap(q.attrs.size(), -1);
    for (uint32_t r = 0; r < n; ++r)
        if ((keep >> r) & 1) {
            red.removed[r] = int32_t(s.rels.size());
            QRel nr;
            nr.table = q.rels[r].table;
            nr.rows  = q.rels[r].rows;
            s.rels.push_back(nr);
        }
    for (uint32_t a = 0; a < q.attrs.size(); ++a) {
        const QAttr& at = q.attrs[a];
        if (red.removed[at.rel] < 0) continue;
        if (eqmap[at.eq] < 0) {
            eqmap[at.eq] = int32_t(s.eq_attrs.size());
            s.eq_attrs.emplace_back();
        }
        attrmap[a] = int32_t(s.attrs.size());
        s.attrs.push_back(QAttr{uint32_t(red.removed[at.rel]), at.col, at.type, uint32_t(eqmap[at.eq])});
        s.eq_attrs[size_t(eqmap[at.eq])].push_back(uint32_t(attrmap[a]));
        s.rels[size_t(red.removed[at.rel])].attrs.push_back(uint32_t(attrmap[a]));
        s.rels[size_t(red.removed[at.rel])].eqs.push_back(uint32_t(eqmap[at.eq]));
    }
    for (auto& r : s.rels) {
        std::sort(r.eqs.begin(), r.eqs.end());
        r.eqs.erase(std::unique(r.eqs.begin(), r.eqs.end()), r.eqs.end());
    }
    // outputs: from a remaining member of the class, else a constant column
    for (uint32_t i = 0; i < q.out.size(); ++i) {
        const uint32_t a = q.out[i];
        int32_t        m = attrmap[a];
        for (size_t k = 0; m < 0 && k < q.eq_attrs[q.attrs[a].eq].size(); ++k)
            m = attrmap[q.eq_attrs[q.attrs[a].eq][k]];
        if (m >= 0) {
            s.out.push_back(uint32_t(m));
            red.out_pos.push_back(i);
            continue;
        }
        ConstOut c;
        c.pos   = i;
        c.type  = q.attrs[a].type;
        c.rel   = q.attrs[a].rel;
        c.slot  = uint32_t(std::find(q.rels[c.rel].attrs.begin(), q.rels[c.rel].attrs.end(), a) -
                          q.rels[c.rel].attrs.begin());
        c.value = values[c.rel][c.slot];
        red.consts.push_back(c);
    }
    // restrictions: the existing ones of remaining relations, and each class
    // constant on every remaining member
    for (const QFilter& f : q.filters)
        if (attrmap[f.attr] >= 0) {
            QFilter g = f;
            g.attr    = uint32_t(attrmap[f.attr]);
            s.filters.push_back(g);
        }
    for (uint32_t a = 0; a < q.attrs.size(); ++a) {
        auto it = cst.find(q.attrs[a].eq);
        if (attrmap[a] < 0 || it == cst.end()) continue;
        QFilter f;
        f.attr = uint32_t(attrmap[a]);
        f.kind = FLT_EQ;
        f.a    = it->second;
        if (it->second == ~0ull) red.empty = true; // a remaining member equal to NULL: no row
        s.filters.push_back(f);
    }
    return red;
}

Query restrict_query(const Query& q, uint64_t relset) {
    const uint32_t n = uint32_t(q.rels.size());
    if (relset == 0 || (n < 64 && (relset >> n) != 0)) throw std::runtime_error("relation set out of range");
    // connected?
    const auto adj  = q.adjacency();
    uint64_t   seen = relset & (~relset + 1), frontier = seen;
    while (frontier) {
        uint64_t next = 0;
        for (uint64_t t = frontier; t; t &= t - 1) next |= adj[__builtin_ctzll(t)];
        frontier = next & relset & ~seen;
        seen |= frontier;
    }
    if (seen != relset) throw std::runtime_error("relation set is not connected by join predicates");
    Query                 s;
    std::vector<int32_t>  relmap(n, -1), eqmap(q.eq_attrs.size(), -1);
    for (uint32_t r = 0; r < n; ++r)
        if ((relset >> r) & 1) {
            relmap[r] = int32_t(s.rels.size());
            QRel nr;
            nr.table = q.rels[r].table;
            nr.rows  = q.rels[r].rows;
            s.rels.push_back(nr);
        }
    for (uint32_t a = 0; a < q.attrs.size(); ++a) {
        const QAttr& at = q.attrs[a];
        if (relmap[at.rel] < 0) continue;
        if (eqmap[at.eq] < 0) {
            eqmap[at.eq] = int32_t(s.eq_attrs.size());
            s.eq_attrs.emplace_back();
        }
        const uint32_t na = uint32_t(s.attrs.size());
        s.attrs.push_back(QAttr{uint32_t(relmap[at.rel]), at.col, at.type, uint32_t(eqmap[at.eq])});
        s.eq_attrs[eqmap[at.eq]].push_back(na);
        s.rels[relmap[at.rel]].attrs.push_back(na);
        s.rels[relmap[at.rel]].eqs.push_back(uint32_t(eqmap[at.eq]));
    }
    for (auto& r : s.rels) std::sort(r.eqs.begin(), r.eqs.end());
    return s;
}

CardFn max_card_model(const Query& q, const std::vector<double>& base_card) {
    (void)q;
    return [base_card](uint64_t set) {
        double c = 0;
        for (uint32_t r = 0; r < base_card.size(); ++r)
            if ((set >> r) & 1) c = std::max(c, base_card[r]);
        return c;
    };
}

namespace {

inline uint64_t low_mask(uint32_t v) { return v >= 63 ? ~0ull : ((1ull << (v + 1)) - 1); } // bits 0..v

struct Dp {
    const Query&                   q;
    const std::vector<PlanVertex>& vx;
    const CardFn&                  card; // over vertex masks
    std::vector<uint64_t>          adj;  // vertex adjacency
    PlanTree                       tree;
    // Per vertex set: the cheapest plan's cost and build side (the probe side
    // is the rest), and the set's cardinality. A dense array indexed by the
    // set up to kDenseVertices vertices, a hash map above. The tree is built
    // from the root's entry once the search is done.
    struct Entry {
        double   cost  = std::numeric_limits<double>::infinity();
        double   card  = -1.0;
        uint64_t build = 0;
    };
    static constexpr uint32_t                 kDenseVertices = 18;
    bool                                      dense;
    std::vector<Entry>                        dense_;
    std::unordered_map<uint64_t, Entry>       sparse_;
    std::vector<std::pair<uint64_t, uint64_t>> pairs;
    size_t                                    budget   = 4u << 20;
    bool                                      overflow = false;

    Dp(const Query& q_, const std::vector<PlanVertex>& v, const CardFn& c)
        : q(q_), vx(v), card(c), adj(v.size(), 0), dense(v.size() <= kDenseVertices) {
        for (size_t i = 0; i < vx.size(); ++i)
            for (size_t j = i + 1; j < vx.size(); ++j)
                if (!q.shared(vx[i].set, vx[j].set).empty()) {
                    adj[i] |= 1ull << j;
                    adj[j] |= 1ull << i;
                }
        if (dense) {
            dense_.resize(size_t(1) << vx.size());
            // neighbourhood of every vertex set, from the set without its lowest vertex
            nb_.resize(size_t(1) << vx.size());
            for (uint64_t t = 1; t < nb_.size(); ++t) nb_[t] = nb_[t & (t - 1)] | adj[__builtin_ctzll(t)];
        }
    }
    std::vector<uint64_t> nb_; // dense only: union of the members' adjacency
    Entry& entry(uint64_t s) { return dense ? dense_[s] : sparse_[s]; }
    const Entry* find(uint64_t s) const {
        if (dense) return &dense_[s];
        auto it = sparse_.find(s);
        return it == sparse_.end() ? nullptr : &it->second;
    }

    uint64_t base(uint64_t vmask) const {
        uint64_t s = 0;
        for (uint64_t t = vmask; t; t &= t - 1) s |= vx[__builtin_ctzll(t)].set;
        return s;
    }
    uint64_t neigh(uint64_t s) const {
        if (dense) return nb_[s] & ~s;
        uint64_t n = 0;
        for (uint64_t t = s; t; t &= t - 1) n |= adj[__builtin_ctzll(t)];
        return n & ~s;
    }
    double card_of(uint64_t s) {
        Entry& e = entry(s);
        if (e.card < 0) e.card = std::max(card(s), 0.0);
        return e.card;
    }
    // A built table (a single vertex with a key) may only be a build side,
    // joined on its key class.
    bool can_join(uint64_t b, uint64_t p) const {
        if (!(p & (p - 1)) && vx[__builtin_ctzll(p)].key >= 0) return false;
        if (!(b & (b - 1)) && vx[__builtin_ctzll(b)].key >= 0) {
            const auto sh = q.shared(base(b), base(p));
            return std::find(sh.begin(), sh.end(), uint32_t(vx[__builtin_ctzll(b)].key)) != sh.end();
        }
        return true;
    }
    void emit_pair(uint64_t a, uint64_t b) {
        if (pairs.size() >= budget) {
            overflow = true;
            return;
        }
        pairs.emplace_back(a, b);
    }
    void cmp_rec(uint64_t s1, uint64_t s2, uint64_t x) {
        if (overflow) return;
        uint64_t n = neigh(s2) & ~x;
        if (!n) return;
        for (uint64_t s = n; s; s = (s - 1) & n) emit_pair(s1, s2 | s);
        for (uint64_t s = n; s; s = (s - 1) & n) cmp_rec(s1, s2 | s, x | n);
    }
    void emit_csg(uint64_t s1) {
        if (overflow) return;
        uint32_t mn = __builtin_ctzll(s1);
        uint64_t x  = s1 | low_mask(mn);
        uint64_t n  = neigh(s1) & ~x;
        for (int v = 63; v >= 0; --v) {
            if (!((n >> v) & 1)) continue;
            uint64_t s2 = 1ull << v;
            emit_pair(s1, s2);
            cmp_rec(s1, s2, x | (low_mask(v) & n));
        }
    }
    void csg_rec(uint64_t s, uint64_t x) {
        if (overflow) return;
        uint64_t n = neigh(s) & ~x;
        if (!n) return;
        for (uint64_t t = n; t; t = (t - 1) & n) emit_csg(s | t);
        for (uint64_t t = n; t; t = (t - 1) & n) csg_rec(s | t, x | n);
    }

    // Cost C(T) = C(B) + C(P) + |T| + 2|B| of building b and probing it with p.
    void consider(uint64_t b, uint64_t p) {
        if (!can_join(b, p)) return;
        const Entry* eb = find(b);
        const Entry* ep = find(p);
        if (!eb || !ep || eb->cost == std::numeric_limits<double>::infinity() ||
            ep->cost == std::numeric_limits<double>::infinity())
            return;
        const double cb = eb->cost, cp = ep->cost, nb = eb->card;
        const uint64_t t = b | p;
        const double   c = cb + cp + card_of(t) + 2.0 * nb;
        Entry&         e = entry(t);
        if (e.cost <= c) return;
        e.cost  = c;
        e.build = b;
    }

    void leaves() {
        for (uint32_t v = 0; v < vx.size(); ++v) {
            card_of(1ull << v);
            entry(1ull << v).cost = 0;
        }
    }

    // Greedy operator ordering: repeatedly join the two connected components
    // with the smallest joint cardinality (the fallback for very large graphs).
    void greedy() {
        std::vector<uint64_t> comps;
        for (uint32_t v = 0; v < vx.size(); ++v) comps.push_back(1ull << v);
        while (comps.size() > 1) {
            double bc = 1e300;
            size_t bi = 0, bj = 0;
            for (size_t i = 0; i < comps.size(); ++i)
                for (size_t j = i + 1; j < comps.size(); ++j) {
                    if (!(neigh(comps[i]) & comps[j])) continue;
                    if (!can_join(comps[i], comps[j]) && !can_join(comps[j], comps[i])) continue;
                    double c = card_of(comps[i] | comps[j]);
                    if (c < bc) {
                        bc = c;
                        bi = i;
                        bj = j;
                    }
                }
            if (bc == 1e300) throw std::runtime_error("join graph is not connected");
            uint64_t   a = comps[bi], b = comps[bj];
            const bool ab = can_join(a, b), ba = can_join(b, a);
            if (ab && (!ba || card_of(a) <= card_of(b)))
                consider(a, b);
            else
                consider(b, a);
            comps[bi] = a | b;
            comps.erase(comps.begin() + long(bj));
        }
    }

    void run() {
        leaves();
        uint32_t n = uint32_t(vx.size());
        for (int v = int(n) - 1; v >= 0; --v) {
            uint64_t s = 1ull << v;
            emit_csg(s);
            csg_rec(s, low_mask(uint32_t(v)));
        }
        if (overflow) {
            greedy();
            return;
        }
        // smaller sets first, so both halves of a pair are planned before it
        // (a stable counting sort by the pair's vertex count)
        std::vector<size_t> start(n + 2, 0);
        for (auto& pr : pairs) ++start[size_t(__builtin_popcountll(pr.first | pr.second)) + 1];
        for (uint32_t k = 1; k <= n + 1; ++k) start[k] += start[k - 1];
        std::vector<std::pair<uint64_t, uint64_t>> sorted(pairs.size());
        for (auto& pr : pairs) sorted[start[size_t(__builtin_popcountll(pr.first | pr.second))]++] = pr;
        pairs.clear();
        pairs.shrink_to_fit();
        for (auto& pr : sorted) {
            consider(pr.first, pr.second);
            consider(pr.second, pr.first);
        }
    }

    // The plan of vertex set s as tree nodes (children first); -1 if none.
    int emit_tree(uint64_t s) {
        const Entry* e = find(s);
        if (!e || e->cost == std::numeric_limits<double>::infinity()) return -1;
        TreeNode n;
        n.card = e->card;
        n.cost = e->cost;
        if (!(s & (s - 1))) {
            const PlanVertex& v = vx[__builtin_ctzll(s)];
            n.rel   = v.rel;
            n.group = v.group;
            n.set   = v.set;
        } else {
            const uint64_t b = e->build;
            n.build          = emit_tree(b);
            n.probe          = emit_tree(s ^ b);
            if (n.build < 0 || n.probe < 0) return -1;
            n.set = tree.nodes[size_t(n.build)].set | tree.nodes[size_t(n.probe)].set;
            // emit_tree may have grown the node vector: the entry is re-read
            const Entry* e2 = find(s);
            n.card          = e2->card;
            n.cost          = e2->cost;
        }
        tree.nodes.push_back(n);
        return int(tree.nodes.size() - 1);
    }
};

} // namespace

PlanTree plan_joins_vertices(const Query& q, const std::vector<PlanVertex>& vertices, const CardFn& card) {
    if (vertices.empty()) throw std::runtime_error("plan has no scans");
    if (vertices.size() > 64) throw std::runtime_error("more than 64 join-graph vertices");
    Dp dp(q, vertices, card);
    dp.run();
    const uint64_t all  = vertices.size() >= 64 ? ~0ull : ((1ull << vertices.size()) - 1);
    const int      root = dp.emit_tree(all);
    if (root < 0) throw std::runtime_error("join graph is not connected");
    dp.tree.root = root;
    return std::move(dp.tree);
}

PlanTree plan_joins(const Query& q, const CardFn& card) {
    std::vector<PlanVertex> v(q.rels.size());
    for (uint32_t r = 0; r < q.rels.size(); ++r) {
        v[r].set = 1ull << r;
        v[r].rel = int(r);
    }
    return plan_joins_vertices(q, v, card); // vertex masks are relation masks here
}

JoinTuning autotune_join(uint64_t build_rows, uint64_t probe_rows, uint32_t key_bytes) {
    (void)key_bytes;
    JoinTuning t;
    // Capacity: a wave waits for its longest probe sequence, so the load factor
    // bounds the probe more than the table's size does. At a load factor of at
    // most 0.5 (capacity = next power of two above 2x the keys) a miss scans ~2
    // slots instead of ~8 at 0.75; on the 50M-key INT64 join this halved the
    // probe kernel (20 ms -> 12 ms) for 1 ms more build. Small tables are cache
    // resident and sparser still.
    t.max_load = build_rows < (1u << 20) ? 0.4 : 0.5;
    // Bloom filter (one 64-bit block per key, 4 bits set): worth its extra
    // lookup when the probe side is much larger than the build side, which in
    // foreign-key joins signals pre-filtered (mostly missing) build keys, and
    // when the filter stays cache resident (<= 16M keys at 16 bits per key).
    t.bloom_bits  = (probe_rows >= 8 * build_rows && build_rows <= (16u << 20)) ? 16 : 0;
    t.partitioned = true;
    return t;
}

} // namespace cgj
