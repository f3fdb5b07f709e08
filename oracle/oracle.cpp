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
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
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
