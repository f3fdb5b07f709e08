// CPU oracle for the join-pipeline hot path. TEST INFRASTRUCTURE ONLY: the
// tests, the smoke check and the benchmark's CPU-baseline leg load this
// library as a checker; the product library never links or calls it.
//
// It restates, in plain single-threaded C++, what the reference computes for a
// contest Plan:
//   * page reading (reference engine/query/DataSource.hpp:14-79 accessors and
//     engine/op/TableScan.cpp:135-324 reader): rows per page, NULL bitmap,
//     values at +4 (INT32) / +8 (INT64, FP64), VARCHAR end offsets at +4 with
//     the characters after them, long strings over 0xffff/0xfffe pages;
//   * values travel as 64-bit slots with the NULL sentinel ~0
//     (reference engine/query/RuntimeValue.hpp:10): INT32 is zero-extended,
//     FP64 keeps its bits, so an INT64 of -1 or an FP64 with all bits set
//     reads back as NULL exactly as in the reference;
//   * joins are inner equi-joins on 64-bit slots where NULL keys never match
//     (every join attribute gets a NotNull restriction, reference
//     engine/query/QueryPlan.cpp:749-753) and every duplicate contributes
//     (reference engine/pipeline/JoinPipeline.hpp:112-131);
//   * the result is written with the reference's output-page bit budget
//     (reference engine/op/TableTarget.cpp:111-391): fixed-width columns spend
//     1 bit per row plus the value width per non-NULL row out of
//     8*(8192-4[-4]) bits; VARCHAR spends 1 bit per NULL and 17+8*len bits per
//     short string; long strings are copied page by page.
// The tree is evaluated in the order the Plan gives (build side per
// build_left); the reference re-plans, which changes no result row.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../include/gpujoin.h"

namespace {

constexpr uint64_t kNull = ~0ull;
constexpr size_t   kPage = CGJ_PAGE_SIZE;

struct StringPool {
    std::vector<std::string> strings;
    uint64_t add(std::string s) {
        strings.push_back(std::move(s));
        return strings.size() - 1;
    }
};

uint16_t rd16(const uint8_t* p) {
    uint16_t v;
    std::memcpy(&v, p, 2);
    return v;
}

// Decodes one column into slots (reference DataSource::Page accessors).
void decode_column(const cgj_host_column& col, uint64_t num_rows, StringPool& pool,
                   std::vector<uint64_t>& out) {
    out.clear();
    out.reserve(num_rows);
    for (uint64_t pi = 0; pi < col.num_pages; ++pi) {
        const uint8_t* page = static_cast<const uint8_t*>(col.pages[pi]);
        uint16_t       nr   = rd16(page);
        uint16_t       nv   = rd16(page + 2);
        if (col.type == CGJ_VARCHAR && nr >= 0xfffe) {
            if (nr == 0xfffe) continue; // continuation, consumed by its start page
            std::string s;
            uint64_t    pj = pi;
            do {
                const uint8_t* q   = static_cast<const uint8_t*>(col.pages[pj]);
                uint16_t       len = rd16(q + 2);
                s.append(reinterpret_cast<const char*>(q + 4), len);
                ++pj;
            } while (pj < col.num_pages && rd16(static_cast<const uint8_t*>(col.pages[pj])) == 0xfffe);
            out.push_back(pool.add(std::move(s)));
            continue;
        }
        const uint8_t* bitmap   = page + kPage - (nr + 7u) / 8u;
        bool           no_nulls = nv == nr;
        uint32_t       k        = 0; // non-null index
        for (uint32_t r = 0; r < nr; ++r) {
            bool valid = no_nulls || ((bitmap[r / 8] >> (r % 8)) & 1u);
            if (!valid) {
                out.push_back(kNull);
                continue;
            }
            switch (col.type) {
            case CGJ_INT32: {
                uint32_t v;
                std::memcpy(&v, page + 4 + 4 * k, 4);
                out.push_back(v);
                break;
            }
            case CGJ_INT64:
            case CGJ_FP64: {
                uint64_t v;
                std::memcpy(&v, page + 8 + 8 * k, 8);
                out.push_back(v);
                break;
            }
            case CGJ_VARCHAR: {
                const uint8_t* offs  = page + 4;
                uint16_t       end   = rd16(offs + 2 * k);
                uint16_t       start = k ? rd16(offs + 2 * (k - 1)) : 0;
                const char*    chars = reinterpret_cast<const char*>(page + 4 + 2 * nv);
                out.push_back(pool.add(std::string(chars + start, end - start)));
                break;
            }
            default: throw std::runtime_error("unknown column type");
            }
            ++k;
        }
    }
    if (out.size() != num_rows) throw std::runtime_error("page rows do not sum to num_rows");
}

struct Rel {
    std::vector<int32_t>               types;
    std::vector<std::vector<uint64_t>> cols; // column-major
    uint64_t                           rows = 0;
};

struct Evaluator {
    const cgj_plan*       plan;
    const cgj_host_table* inputs;
    uint32_t              num_inputs;
    StringPool            pool;

    // Join key of slot `v` of type `t` as hashable bytes (strings compare by content).
    std::string key_of(uint64_t v, int32_t t) const {
        if (t == CGJ_VARCHAR) return "s" + pool.strings[v];
        std::string k(9, '\0');
        k[0] = 'v';
        std::memcpy(&k[1], &v, 8);
        return k;
    }

    Rel eval(uint64_t node_id) {
        const cgj_node& n = plan->nodes[node_id];
        Rel             out;
        if (n.kind == CGJ_NODE_SCAN) {
            if (n.base_table_id >= num_inputs) throw std::runtime_error("bad base table id");
            const cgj_host_table& t = inputs[n.base_table_id];
            out.rows                = t.num_rows;
            for (uint32_t a = 0; a < n.num_output_attrs; ++a) {
                uint64_t ci = n.output_attrs[a].index;
                out.types.push_back(n.output_attrs[a].type);
                std::vector<uint64_t> slots;
                if (t.num_rows == 0) {
                    out.cols.push_back({});
                    continue;
                }
                if (ci >= t.num_columns) throw std::runtime_error("bad column index");
                decode_column(t.columns[ci], t.num_rows, pool, slots);
                out.cols.push_back(std::move(slots));
            }
            return out;
        }
        Rel      left  = eval(n.left);
        Rel      right = eval(n.right);
        Rel&     build = n.build_left ? left : right;
        Rel&     probe = n.build_left ? right : left;
        uint64_t bkey  = n.build_left ? n.left_attr : n.right_attr;
        uint64_t pkey  = n.build_left ? n.right_attr : n.left_attr;
        size_t   nl    = left.cols.size();

        std::unordered_multimap<std::string, uint64_t> ht;
        ht.reserve(build.rows * 2 + 1);
        for (uint64_t r = 0; r < build.rows; ++r) {
            uint64_t v = build.cols[bkey][r];
            if (v == kNull) continue;
            ht.emplace(key_of(v, build.types[bkey]), r);
        }
        for (uint32_t a = 0; a < n.num_output_attrs; ++a) {
            uint64_t idx = n.output_attrs[a].index;
            out.types.push_back(idx < nl ? left.types[idx] : right.types[idx - nl]);
            out.cols.emplace_back();
        }
        for (uint64_t r = 0; r < probe.rows; ++r) {
            uint64_t v = probe.cols[pkey][r];
            if (v == kNull) continue;
            auto range = ht.equal_range(key_of(v, probe.types[pkey]));
            for (auto it = range.first; it != range.second; ++it) {
                uint64_t lr = n.build_left ? it->second : r;
                uint64_t rr = n.build_left ? r : it->second;
                for (uint32_t a = 0; a < n.num_output_attrs; ++a) {
                    uint64_t idx = n.output_attrs[a].index;
                    out.cols[a].push_back(idx < nl ? left.cols[idx][lr] : right.cols[idx - nl][rr]);
                }
                ++out.rows;
            }
        }
        return out;
    }
};

// Output page writer with the reference's bit budget (TableTarget.cpp:111-391).
struct PageWriter {
    int32_t            type;
    std::vector<void*> pages;
    uint8_t*           cur       = nullptr;
    int64_t            available = 0;
    uint32_t           rows      = 0;
    uint32_t           nonnull   = 0;
    std::vector<uint8_t> bits;
    std::string        chars;

    explicit PageWriter(int32_t t) : type(t), bits(kPage, 0) { open(); }

    int64_t budget() const { return 8 * int64_t(kPage - 4 - ((type == CGJ_INT64 || type == CGJ_FP64) ? 4 : 0)); }

    void open() {
        cur = static_cast<uint8_t*>(::operator new(kPage));
        std::memset(cur, 0, kPage);
        pages.push_back(cur);
        available = budget();
        rows = nonnull = 0;
        chars.clear();
    }
    void close() {
        uint16_t nr = uint16_t(rows), nv = uint16_t(nonnull);
        std::memcpy(cur, &nr, 2);
        std::memcpy(cur + 2, &nv, 2);
        if (type == CGJ_VARCHAR) std::memcpy(cur + 4 + 2 * nonnull, chars.data(), chars.size());
        size_t nb = (rows + 7u) / 8u;
        std::memcpy(cur + kPage - nb, bits.data(), nb);
        std::fill(bits.begin(), bits.begin() + nb, 0);
    }
    void push_bit(bool v) {
        if (v) bits[rows / 8] |= uint8_t(1u << (rows % 8));
        ++rows;
    }
    void step_fixed(uint64_t v) {
        int64_t need = 1 + (v == kNull ? 0 : (type == CGJ_INT32 ? 32 : 64));
        available -= need;
        if (available < 0) {
            close();
            open();
            available -= need;
        }
        if (v != kNull) {
            if (type == CGJ_INT32) {
                uint32_t x = uint32_t(v);
                std::memcpy(cur + 4 + 4 * nonnull, &x, 4);
            } else {
                std::memcpy(cur + 8 + 8 * nonnull, &v, 8);
            }
            ++nonnull;
        }
        push_bit(v != kNull);
    }
    void step_string(uint64_t v, const StringPool& pool) {
        if (v == kNull) {
            available -= 1;
            if (available < 0) {
                close();
                open();
                available -= 1;
            }
            push_bit(false);
            return;
        }
        const std::string& s = pool.strings[v];
        if (s.size() > kPage - 7) {
            // long string: close the current page and copy the string's pages
            close();
            size_t done = 0;
            bool   first = true;
            // reuse the freshly opened page slot as the first long-string page
            open();
            while (done < s.size()) {
                uint16_t tag  = first ? 0xffff : 0xfffe;
                uint16_t part = uint16_t(std::min(s.size() - done, kPage - 4));
                std::memcpy(cur, &tag, 2);
                std::memcpy(cur + 2, &part, 2);
                std::memcpy(cur + 4, s.data() + done, part);
                done += part;
                first = false;
                open();
            }
            return;
        }
        int64_t need = 1 + 16 + 8 * int64_t(s.size());
        available -= need;
        if (available < 0) {
            close();
            open();
            available -= need;
        }
        chars += s;
        uint16_t end = uint16_t(chars.size());
        std::memcpy(cur + 4 + 2 * nonnull, &end, 2);
        ++nonnull;
        push_bit(true);
    }
    // Ends the column; drops a trailing page that holds no rows.
    void finish() {
        if (rows > 0) {
            close();
        } else {
            ::operator delete(cur);
            pages.pop_back();
        }
    }
};

// ---- full-size checks: threaded decode and order-independent fingerprints ----

// Runs f(t) for t in [0, n) on n threads; rethrows the first failure.
template <class F>
void on_threads(uint32_t n, F&& f) {
    std::vector<std::thread>        th;
    std::vector<std::exception_ptr> err(n);
    for (uint32_t t = 0; t < n; ++t)
        th.emplace_back([&, t] {
            try {
                f(t);
            } catch (...) {
                err[t] = std::current_exception();
            }
        });
    for (auto& x : th) x.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

// Decodes a fixed-width column into slots with `threads` threads: page row
// counts first (their prefix gives each thread's first output row), then each
// thread decodes a contiguous range of pages as decode_column does.
std::vector<uint64_t> decode_fixed_parallel(const cgj_host_column& col, uint64_t num_rows, uint32_t threads) {
    if (col.type == CGJ_VARCHAR) throw std::runtime_error("fixed-width columns only");
    const uint64_t        np = col.num_pages;
    std::vector<uint64_t> start(np + 1, 0);
    for (uint64_t p = 0; p < np; ++p) start[p + 1] = start[p] + rd16(static_cast<const uint8_t*>(col.pages[p]));
    if (start[np] != num_rows) throw std::runtime_error("page rows do not sum to num_rows");
    std::vector<uint64_t> out(num_rows);
    threads = std::max<uint32_t>(1, std::min<uint64_t>(threads, np ? np : 1));
    on_threads(threads, [&](uint32_t t) {
        for (uint64_t p = np * t / threads; p < np * (t + 1) / threads; ++p) {
            const uint8_t* page = static_cast<const uint8_t*>(col.pages[p]);
            const uint16_t nr = rd16(page), nv = rd16(page + 2);
            const uint8_t* bm = page + kPage - (nr + 7u) / 8u;
            uint64_t*      o  = out.data() + start[p];
            uint32_t       k  = 0;
            for (uint32_t r = 0; r < nr; ++r) {
                if (nv != nr && !((bm[r / 8] >> (r % 8)) & 1u)) {
                    o[r] = kNull;
                    continue;
                }
                if (col.type == CGJ_INT32) {
                    uint32_t v;
                    std::memcpy(&v, page + 4 + 4 * k, 4);
                    o[r] = v;
                } else {
                    std::memcpy(&o[r], page + 8 + 8 * size_t(k), 8);
                }
                ++k;
            }
        }
    });
    return out;
}

// SplitMix64 finaliser (Steele et al. 2014).
inline uint64_t smix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Row fingerprint: depends on every slot and its column position; a table's
// fingerprint is the sum over its rows (order independent, duplicates count).
inline uint64_t row_hash(const uint64_t* slots, uint32_t n) {
    uint64_t h = 0x243f6a8885a308d3ull;
    for (uint32_t c = 0; c < n; ++c) h = smix(h ^ (slots[c] + 0x9e3779b97f4a7c15ull * (c + 1)));
    return h;
}

// Open-addressing multimap from a non-NULL key to build rows, split into
// 2^pbits partitions by the key hash so that threads build disjoint parts.
struct KeyRows {
    uint32_t                           pbits = 0;
    std::vector<std::vector<uint64_t>> keys, rows; // per partition: slots (kNull = empty)
    std::vector<uint64_t>              masks;

    uint32_t part(uint64_t h) const { return pbits ? uint32_t(h >> (64 - pbits)) : 0u; }

    void build(const std::vector<uint64_t>& bk, uint32_t threads) {
        pbits = 0;
        while ((1u << pbits) < threads) ++pbits;
        const uint32_t np = 1u << pbits;
        keys.assign(np, {});
        rows.assign(np, {});
        masks.assign(np, 0);
        std::vector<uint64_t> count(np, 0);
        for (uint64_t v : bk)
            if (v != kNull) ++count[part(smix(v))];
        on_threads(np, [&](uint32_t p) {
            uint64_t cap = 16;
            while (cap < 2 * count[p]) cap <<= 1;
            keys[p].assign(cap, kNull);
            rows[p].assign(cap, 0);
            masks[p] = cap - 1;
            for (uint64_t r = 0; r < bk.size(); ++r) {
                const uint64_t v = bk[r];
                if (v == kNull) continue;
                const uint64_t h = smix(v);
                if (part(h) != p) continue;
                uint64_t s = h & masks[p];
                while (keys[p][s] != kNull) s = (s + 1) & masks[p];
                keys[p][s] = v;
                rows[p][s] = r;
            }
        });
    }
    // Calls f(build row) for every build row holding key v.
    template <class F>
    void each(uint64_t v, F&& f) const {
        const uint64_t h = smix(v);
        const uint32_t p = part(h);
        for (uint64_t s = h & masks[p]; keys[p][s] != kNull; s = (s + 1) & masks[p])
            if (keys[p][s] == v) f(rows[p][s]);
    }
};

void set_error(char* err, size_t len, const char* msg) {
    if (err && len) std::snprintf(err, len, "%s", msg);
}

} // namespace

extern "C" {

// Evaluates `plan` over `inputs` on the CPU; result pages in `out`.
// Returns 0 on success, -1 on error (message in err).
int oracle_execute(const cgj_plan* plan, const cgj_host_table* inputs, uint32_t num_inputs,
                   cgj_result* out, char* err, size_t errlen) {
    try {
        Evaluator ev{plan, inputs, num_inputs, {}};
        Rel       res = ev.eval(plan->root);
        out->num_rows    = res.rows;
        out->num_columns = uint32_t(res.cols.size());
        out->columns     = new cgj_result_column[res.cols.size()];
        for (size_t c = 0; c < res.cols.size(); ++c) {
            PageWriter w(res.types[c]);
            if (res.rows == 0) {
                w.finish();
            } else {
                for (uint64_t r = 0; r < res.rows; ++r) {
                    if (res.types[c] == CGJ_VARCHAR)
                        w.step_string(res.cols[c][r], ev.pool);
                    else
                        w.step_fixed(res.cols[c][r]);
                }
                w.finish();
            }
            out->columns[c].type      = res.types[c];
            out->columns[c].num_pages = w.pages.size();
            out->columns[c].pages     = new void*[w.pages.size() ? w.pages.size() : 1];
            for (size_t i = 0; i < w.pages.size(); ++i) out->columns[c].pages[i] = w.pages[i];
        }
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

void oracle_result_release(cgj_result* res) {
    if (!res || !res->columns) return;
    for (uint32_t c = 0; c < res->num_columns; ++c) {
        for (uint64_t i = 0; i < res->columns[c].num_pages; ++i) ::operator delete(res->columns[c].pages[i]);
        delete[] res->columns[c].pages;
    }
    delete[] res->columns;
    res->columns     = nullptr;
    res->num_columns = 0;
}

// Decodes a fixed-width column into 64-bit slots (NULL = ~0).
int oracle_decode_fixed(const cgj_host_column* col, uint64_t num_rows, uint64_t* slots, char* err,
                        size_t errlen) {
    try {
        if (col->type == CGJ_VARCHAR) throw std::runtime_error("varchar column");
        StringPool            pool;
        std::vector<uint64_t> out;
        decode_column(*col, num_rows, pool, out);
        std::memcpy(slots, out.data(), out.size() * 8);
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

// Row count and fingerprint (sum of row_hash over rows) of a table of
// fixed-width columns, decoded with `threads` threads. The order of rows and
// the packing of pages do not change it.
int oracle_table_fingerprint(const cgj_host_table* t, uint32_t threads, uint64_t* count, uint64_t* fp, char* err,
                             size_t errlen) {
    try {
        std::vector<std::vector<uint64_t>> cols;
        for (uint32_t c = 0; c < t->num_columns; ++c)
            cols.push_back(decode_fixed_parallel(t->columns[c], t->num_rows, threads));
        const uint32_t        n = t->num_columns, nt = std::max<uint32_t>(threads, 1);
        std::vector<uint64_t> part(nt, 0);
        on_threads(nt, [&](uint32_t k) {
            uint64_t acc = 0, row[64];
            for (uint64_t r = t->num_rows * k / nt; r < t->num_rows * (k + 1) / nt; ++r) {
                for (uint32_t c = 0; c < n; ++c) row[c] = cols[c][r];
                acc += row_hash(row, n);
            }
            part[k] = acc;
        });
        uint64_t total = 0;
        for (uint64_t v : part) total += v;
        *count = t->num_rows;
        *fp    = total;
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

// Expected row count and fingerprint of the inner equi-join of two tables on
// build column `bkey` = probe column `pkey` (NULL never matches, duplicates
// multiply), projected on `nout` columns (side 0 = build, 1 = probe, and the
// column of that table), by a partitioned open-addressing hash join on the
// CPU with `threads` threads. Fixed-width columns only.
int oracle_join_fingerprint(const cgj_host_table* build, uint32_t bkey, const cgj_host_table* probe, uint32_t pkey,
                            const uint32_t* out_side, const uint32_t* out_col, uint32_t nout, uint32_t threads,
                            uint64_t* count, uint64_t* fp, char* err, size_t errlen) {
    try {
        if (nout > 64) throw std::runtime_error("too many output columns");
        const uint32_t nt = std::max<uint32_t>(threads, 1);
        // decode every column that is read (key columns and outputs)
        std::vector<std::vector<uint64_t>> bc(build->num_columns), pc(probe->num_columns);
        auto need = [&](uint32_t side, uint32_t col) {
            const cgj_host_table* t = side ? probe : build;
            auto&                 v = side ? pc : bc;
            if (col >= t->num_columns) throw std::runtime_error("bad column index");
            if (v[col].empty() && t->num_rows) v[col] = decode_fixed_parallel(t->columns[col], t->num_rows, nt);
        };
        need(0, bkey);
        need(1, pkey);
        for (uint32_t o = 0; o < nout; ++o) need(out_side[o], out_col[o]);
        KeyRows kr;
        kr.build(bc[bkey], nt);
        std::vector<uint64_t> cnt(nt, 0), acc(nt, 0);
        const uint64_t        np = probe->num_rows;
        if (np)
            on_threads(nt, [&](uint32_t k) {
                uint64_t row[64], c = 0, a = 0;
                for (uint64_t r = np * k / nt; r < np * (k + 1) / nt; ++r) {
                    const uint64_t v = pc[pkey][r];
                    if (v == kNull) continue;
                    kr.each(v, [&](uint64_t b) {
                        for (uint32_t o = 0; o < nout; ++o) row[o] = out_side[o] ? pc[out_col[o]][r] : bc[out_col[o]][b];
                        a += row_hash(row, nout);
                        ++c;
                    });
                }
                cnt[k] = c;
                acc[k] = a;
            });
        uint64_t c = 0, a = 0;
        for (uint32_t k = 0; k < nt; ++k) {
            c += cnt[k];
            a += acc[k];
        }
        *count = c;
        *fp    = a;
        return 0;
    } catch (const std::exception& e) {
        set_error(err, errlen, e.what());
        return -1;
    }
}

// Joins two in-memory key arrays (NULL = ~0 never matches) and returns the
// number of matching pairs: the count form of a single equi-join, used as a
// size-independent check at full benchmark size.
uint64_t oracle_join_count(const uint64_t* build, uint64_t nb, const uint64_t* probe, uint64_t np) {
    std::unordered_map<uint64_t, uint64_t> mult;
    mult.reserve(nb * 2 + 1);
    for (uint64_t i = 0; i < nb; ++i)
        if (build[i] != kNull) ++mult[build[i]];
    uint64_t total = 0;
    for (uint64_t i = 0; i < np; ++i) {
        if (probe[i] == kNull) continue;
        auto it = mult.find(probe[i]);
        if (it != mult.end()) total += it->second;
    }
    return total;
}

} // extern "C"
