// Plan import and join ordering (host).
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "../../include/gpujoin.h"

namespace cgj {

struct JoinTuning;

// A scanned column of a relation.
struct QAttr {
    uint32_t rel;
    uint64_t col;
    int32_t  type;
    uint32_t eq;
};

// One scan node of the plan (the same base table may be scanned twice).
struct QRel {
    uint64_t              table;
    uint64_t              rows = 0;
    std::vector<uint32_t> attrs;
    std::vector<uint32_t> eqs; // classes produced (sorted, unique)
};

// A restriction of one scanned attribute with values known on the host
// (reference engine/query/Restriction.hpp; the kinds of
// engine/storage/RestrictionLogic.cpp:40-447): FLT_EQ (v == a), FLT_EQ2
// (v == a or b), FLT_RANGE (a <= v <= b), FLT_TINY (v in a direct-mapped table
// of 2^tiny_bits slots). NULL never passes. The FLT_* and TINY_* codes are
// those of the device restrictions (kernels.hpp).
struct QFilter {
    uint32_t              attr = 0;
    uint32_t              kind = 0;
    uint64_t              a = 0, b = 0;
    uint32_t              tiny_bits = 0, tiny_hash = 0;
    std::vector<uint64_t> tiny; // 2^tiny_bits slots, ~0 = empty
    bool                  pass(uint64_t v) const;
};

// The restriction that keeps exactly the given keys (at most 32 distinct),
// chosen as the reference's setupRestriction does (RestrictionLogic.cpp:397-447):
// one key -> FLT_EQ, two -> FLT_EQ2, a contiguous run -> FLT_RANGE, up to 16
// -> the first collision-free tiny table of 8/16/32 slots by value (then 16/32
// slots by Fibonacci hash, then by folded Fibonacci hash). False when none of
// these applies (more than 32 keys, or no tiny table fits).
bool small_restriction(std::vector<uint64_t> keys, uint32_t attr, QFilter& out);

struct Query {
    std::vector<QRel>                  rels;
    std::vector<QAttr>                 attrs;
    std::vector<std::vector<uint32_t>> eq_attrs; // class -> attributes
    std::vector<uint32_t>              out;      // root output attributes
    std::vector<QFilter>               filters;  // restrictions of scanned attributes

    bool     is_join_eq(uint32_t eq) const { return eq_attrs[eq].size() > 1; }
    uint64_t all_rels() const { return rels.size() >= 64 ? ~0ull : ((1ull << rels.size()) - 1); }
    // classes produced by a set of relations
    std::set<uint32_t> produced(uint64_t relset) const;
    // classes of `relset` needed outside it: joined with another relation or in the output
    std::vector<uint32_t> needed(uint64_t relset) const;
    // classes shared by two disjoint relation sets (their join predicates)
    std::vector<uint32_t> shared(uint64_t a, uint64_t b) const;
    // attribute of relation `rel` in class `eq`
    uint32_t attr_of(uint32_t rel, uint32_t eq) const;
    // relations adjacent to each relation through join classes
    std::vector<uint64_t> adjacency() const;
};

// Builds the query from a contest plan (reference engine/query/PlanImport.cpp:24-89).
// `table_rows[i]` / `table_cols[i]` describe input table i.
Query import_plan(const cgj_plan& plan, const std::vector<uint64_t>& table_rows,
                  const std::vector<uint32_t>& table_cols);

// The sub-query of a connected relation set: its relations (in order), their
// scanned columns and the join classes among them; no output columns.
Query restrict_query(const Query& q, uint64_t relset);

// Join tree: leaves are relations; an inner node joins `build` (hashed) with
// `probe` (streamed through the pipeline).
struct TreeNode {
    int      rel   = -1; // leaf: a base relation
    int      group = -1; // leaf: a hash table built earlier (index into the caller's vertex list)
    int      build = -1;
    int      probe = -1;
    uint64_t set   = 0;  // base relations covered
    double   card  = 0;
    double   cost  = 0;
};
struct PlanTree {
    std::vector<TreeNode> nodes;
    int                   root = -1;
    bool                  leaf(int n) const { return nodes[n].rel >= 0 || nodes[n].group >= 0; }
};

// A vertex of the join graph being planned: a base relation, or a hash table
// built by an earlier pipeline over a set of relations. A built table may only
// be the build side of a join, on its key class (reference
// engine/query/QueryGraph.cpp canJoin: a hash table cannot be probed as a
// stream, and it is keyed on one class).
struct PlanVertex {
    uint64_t set   = 0;  // base relations covered
    int      rel   = -1; // base relation, or -1
    int      group = -1; // built table id, or -1
    int32_t  key   = -1; // built table: its key class
};

// Cardinality of a connected relation set.
using CardFn = std::function<double(uint64_t relset)>;

// Reference heuristic (engine/query/QueryGraph.cpp:53-68): the cardinality of a
// join is the largest cardinality among its relations.
CardFn max_card_model(const Query& q, const std::vector<double>& base_card);

// Dynamic programming over connected sub-graph / complement pairs (DPccp,
// Moerkotte & Neumann, VLDB 2006) with cost(T) = cost(build) + cost(probe) +
// card(T) + 2 card(build): every hash-table tuple is written and read once
// more than a streamed one.
PlanTree plan_joins(const Query& q, const CardFn& card);
// The same DP over arbitrary vertices; `card` receives vertex masks (bit i =
// vertices[i]). Tree node sets are base-relation masks.
PlanTree plan_joins_vertices(const Query& q, const std::vector<PlanVertex>& vertices, const CardFn& card);

// What the cardinality model did while a query was planned.
struct EstimateReport {
    uint32_t sets       = 0; // connected sets estimated (two or more relations)
    uint32_t fallbacks  = 0; // of them, estimated by the independence rule (low confidence)
    uint32_t stat_joins = 0; // 2-way counts run for the statistics catalog
    double   ms_stats   = 0; // time in those counts
    double   ms_model   = 0; // featurisation + inference
};

// Row samples of each relation's join columns, the reference's sampling
// estimate (engine/query/QueryPlan.cpp:719-746 computeSamples, :66-72
// recomputeCardinality): up to 64 rows per relation; a relation's estimated
// cardinality is rows x (sample rows passing every restriction so far) /
// (sample rows), or rows / (2 x sample rows) when none passes. The
// restrictions are NOT NULL on every join column, then membership in each
// hash table built on one of the relation's join classes.
struct Samples {
    std::vector<uint64_t>                                  rows;    // relation rows
    std::vector<uint32_t>                                  size;    // sample rows (<= 64)
    std::vector<std::map<uint32_t, std::vector<uint64_t>>> keys;    // join class -> sample values (slots)
    std::vector<uint64_t>                                  matches; // bit i: sample row i passes so far
    // per relation: the sample rows found in each hash table probed by the
    // re-filter (the table's selectivity on this relation, for probe order)
    std::vector<std::map<const void*, uint64_t>>           pass;
    double card(uint32_t r) const {
        if (size[r] == 0) return 0;
        const int m = __builtin_popcountll(matches[r]);
        return double(rows[r]) * (m ? double(m) / size[r] : 1.0 / (2.0 * size[r]));
    }
};

// An output column whose every row holds one value: an attribute of a
// relation removed as a singleton (one row) whose class no remaining
// relation produces. VARCHAR values carry their bytes.
struct ConstOut {
    uint32_t    pos   = 0; // position among the query's output columns
    uint32_t    rel   = 0; // the removed relation and its attribute (index in its attrs) that holds it
    uint32_t    slot  = 0;
    int32_t     type  = 0;
    uint64_t    value = ~0ull; // the slot value (~0 = NULL)
    std::string str;           // VARCHAR: the string
};

// Singleton elimination (reference engine/query/QueryPlan.cpp:680-717): every
// relation of one row is removed while more than one relation remains and the
// others stay connected; its values become constants of their classes. A
// class with a constant restricts every remaining member to that value
// (FLT_EQ); an output attribute whose class keeps no member becomes a constant
// column. `values[r][i]` = the slot of attribute q.rels[r].attrs[i] for each
// singleton r. The result is empty when a joined value is NULL or two
// constants of one class differ.
struct SingletonReduction {
    Query                 q;          // the remaining relations (outputs without the constants)
    std::vector<ConstOut> consts;
    std::vector<uint32_t> out_pos;    // column j of q.out -> position among the original outputs
    std::vector<int32_t>  types;      // the original output types
    std::vector<int32_t>  removed;    // original relation -> -1 (removed) or its index in q
    bool                  empty = false;
    uint32_t              eliminated = 0;
};
SingletonReduction eliminate_singletons(const Query& q, const std::vector<std::vector<uint64_t>>& values);

// A query and the join tree chosen for it.
struct PlannedQuery {
    Query                    q;
    std::vector<ConstOut>    consts;     // constant output columns (singleton elimination)
    std::vector<uint32_t>    out_pos;    // q.out column -> output position (empty: identity)
    std::vector<int32_t>     types;      // output types (all columns)
    bool                     empty = false; // the result is known to be empty
    uint32_t                 singletons = 0;
    PlanTree                 tree;
    EstimateReport           rep;
    double                   ms_plan = 0; // estimation + join ordering, statistics counts excluded
    uint32_t                 mode    = 0; // estimator mode that planned it
    CardFn                   card;        // that model's cardinality of base-relation sets
    std::shared_ptr<Samples> samples;     // sampled relations (modes that use them)
};

// Per-join tuning chosen from (build rows, probe rows, key width).
JoinTuning autotune_join(uint64_t build_rows, uint64_t probe_rows, uint32_t key_bytes);

} // namespace cgj
