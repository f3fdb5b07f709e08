// Host-side checks of the planner's restriction and singleton logic, linked
// against the library (no GPU is touched).
//
// small_restriction must pick the kind the reference's setupRestriction
// picks (engine/storage/RestrictionLogic.cpp:397-447) and, whatever the kind,
// pass exactly the given keys (checked by brute force over the keys, their
// neighbours and other values). eliminate_singletons must follow
// engine/query/QueryPlan.cpp:680-717 on hand-built queries.
#include <cstdio>
#include <set>
#include <vector>

#include "../../contest-sigmod2025_amd/csrc/kernels.hpp"
#include "../../contest-sigmod2025_amd/csrc/planner.hpp"

using namespace cgj;

static int fails = 0;
#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);         \
            ++fails;                                                         \
        }                                                                    \
    } while (0)

static bool exact(const QFilter& f, const std::vector<uint64_t>& keys) {
    std::set<uint64_t> in(keys.begin(), keys.end());
    std::vector<uint64_t> probe = {0, 1, 2, 31, 32, 33, 0xffffffffull, 0xfffffffeull, ~0ull - 1, 1ull << 40};
    for (uint64_t k : keys) {
        probe.push_back(k);
        probe.push_back(k + 1);
        probe.push_back(k - 1);
        probe.push_back(k + 8);
        probe.push_back(k + 16);
        probe.push_back(k + 32);
    }
    for (uint64_t v = 0; v < 300; ++v) probe.push_back(v * 7919);
    for (uint64_t v : probe)
        if (v != ~0ull && f.pass(v) != (in.count(v) > 0)) return false;
    return !f.pass(~0ull); // NULL never passes
}

static void kinds() {
    struct Case {
        std::vector<uint64_t> keys;
        bool                  ok;
        uint32_t              kind;
    };
    std::vector<uint64_t> seventeen, thirtythree;
    for (uint64_t i = 0; i < 17; ++i) seventeen.push_back(3 * i);
    for (uint64_t i = 0; i < 33; ++i) thirtythree.push_back(i);
    const std::vector<Case> cases = {
        {{5}, true, FLT_EQ},
        {{7, 7, 7}, true, FLT_EQ},                 // duplicates collapse
        {{7, 3}, true, FLT_EQ2},
        {{10, 11, 12, 13}, true, FLT_RANGE},       // Between
        {{2, 0, 1}, true, FLT_RANGE},              // from 0: Lt
        {{0xfffffffdull, 0xfffffffeull, 0xffffffffull}, true, FLT_RANGE}, // to the u32 maximum: Gt
        {{1, 3, 5}, true, FLT_TINY},
        {{1, 9, 17}, true, FLT_TINY},              // collide mod 8 and mod 16
        {{0, 32, 64}, true, FLT_TINY},             // collide in every modular table
        {seventeen, false, 0},                     // over 16 keys, not contiguous
        {thirtythree, false, 0},                   // over 32 keys
        {{}, false, 0},
    };
    for (const Case& c : cases) {
        QFilter   f;
        const bool ok = small_restriction(c.keys, 7, f);
        CHECK(ok == c.ok);
        if (!ok) continue;
        CHECK(f.kind == c.kind);
        CHECK(f.attr == 7);
        CHECK(exact(f, c.keys));
    }
    // the first tiny table that fits: by value into 8 slots, then 16, 32
    QFilter f;
    CHECK(small_restriction({1, 3, 5}, 0, f) && f.tiny_hash == TINY_MOD && f.tiny_bits == 3);
    CHECK(small_restriction({1, 9, 17}, 0, f) && f.tiny_hash == TINY_MOD && f.tiny_bits == 5);
    CHECK(small_restriction({0, 32, 64}, 0, f) && f.tiny_hash != TINY_MOD);
    CHECK(small_restriction({1, 2, 4}, 0, f) && f.kind == FLT_TINY); // {1,2} contiguous but not all three
    CHECK(small_restriction(thirtythree, 0, f) == false);
    std::vector<uint64_t> run;
    for (uint64_t i = 100; i < 132; ++i) run.push_back(i); // 32 contiguous keys: a range
    CHECK(small_restriction(run, 0, f) && f.kind == FLT_RANGE && f.a == 100 && f.b == 131);
}

// Relations r: attribute list (column, class); classes with several members join.
static Query make_query(const std::vector<uint64_t>& rows, const std::vector<std::vector<uint32_t>>& classes_of,
                        uint32_t nclasses, const std::vector<std::pair<uint32_t, uint32_t>>& out) {
    Query q;
    q.eq_attrs.resize(nclasses);
    for (uint32_t r = 0; r < rows.size(); ++r) {
        QRel rel;
        rel.table = r;
        rel.rows  = rows[r];
        for (uint32_t i = 0; i < classes_of[r].size(); ++i) {
            const uint32_t a = uint32_t(q.attrs.size());
            q.attrs.push_back(QAttr{r, i, T_INT32, classes_of[r][i]});
            q.eq_attrs[classes_of[r][i]].push_back(a);
            rel.attrs.push_back(a);
            rel.eqs.push_back(classes_of[r][i]);
        }
        std::sort(rel.eqs.begin(), rel.eqs.end());
        q.rels.push_back(rel);
    }
    for (auto [r, i] : out) q.out.push_back(q.rels[r].attrs[i]);
    return q;
}

static void singletons() {
    // R0(x, y) - S1(x, z) - R2(y): S1 is a leaf on x, z only an output
    {
        Query q = make_query({100, 1, 50}, {{0, 1}, {0, 2}, {1}}, 3, {{1, 1}, {0, 0}, {2, 0}});
        auto  red = eliminate_singletons(q, {{}, {42, 7}, {}});
        CHECK(red.eliminated == 1 && !red.empty);
        CHECK(red.q.rels.size() == 2 && red.removed[1] == -1);
        CHECK(red.consts.size() == 1 && red.consts[0].pos == 0 && red.consts[0].value == 7);
        CHECK((red.out_pos == std::vector<uint32_t>{1, 2}));
        CHECK(red.q.filters.size() == 1 && red.q.filters[0].kind == FLT_EQ && red.q.filters[0].a == 42);
        CHECK(red.q.attrs[red.q.filters[0].attr].rel == 0); // on R0's x
        CHECK(red.q.out.size() == 2 && red.q.attrs[red.q.out[0]].rel == 0 && red.q.attrs[red.q.out[1]].rel == 1);
        CHECK(red.types.size() == 3);
    }
    // R0(x) - S1(x, y) - R2(y): removing S1 disconnects R0 from R2: kept
    {
        Query q   = make_query({100, 1, 50}, {{0}, {0, 1}, {1}}, 2, {{0, 0}});
        auto  red = eliminate_singletons(q, {{}, {4, 5}, {}});
        CHECK(red.eliminated == 0 && red.q.rels.size() == 3 && red.q.filters.empty());
    }
    // a NULL joined value empties the result
    {
        Query q   = make_query({100, 1}, {{0}, {0}}, 1, {{0, 0}});
        auto  red = eliminate_singletons(q, {{}, {~0ull}});
        CHECK(red.eliminated == 1 && red.empty);
    }
    // two singletons fixing one class to different values: empty
    {
        Query q   = make_query({100, 1, 1}, {{0}, {0}, {0}}, 1, {{0, 0}});
        auto  red = eliminate_singletons(q, {{}, {3}, {4}});
        CHECK(red.eliminated == 2 && red.empty);
        auto same = eliminate_singletons(q, {{}, {3}, {3}});
        CHECK(same.eliminated == 2 && !same.empty && same.q.rels.size() == 1 && same.q.filters.size() == 1);
    }
    // every relation a singleton: removed in order, the last one stays; the
    // removed one's output-only class becomes a constant column
    {
        Query q   = make_query({1, 1}, {{0, 1}, {0}}, 2, {{0, 1}, {1, 0}});
        auto  red = eliminate_singletons(q, {{8, 9}, {8}});
        CHECK(red.eliminated == 1 && red.q.rels.size() == 1 && red.removed[0] == -1 && !red.empty);
        CHECK(red.consts.size() == 1 && red.consts[0].pos == 0 && red.consts[0].value == 9);
        CHECK((red.out_pos == std::vector<uint32_t>{1}));
        CHECK(red.q.filters.size() == 1 && red.q.filters[0].a == 8);
    }
}

int main() {
    kinds();
    singletons();
    if (fails) {
        std::printf("%d check(s) failed\n", fails);
        return 1;
    }
    std::printf("planner checks passed\n");
    return 0;
}
