// A contest-harness stand-in: reads plans and input rows as text on stdin,
// builds each Plan with the contest's ColumnInserter, runs it through the C++
// contest surface (Contest::build_context / execute / destroy_context, the
// calls the reference's tests make: tests/unit_tests.cpp:19-664), decodes the
// result pages itself and prints the rows. The result ColumnarTable is
// destroyed by its own destructor, which releases every page with `delete`.
//
// Input, one item per line (values: NULL, decimal integers, FP64 as 16 hex
// digits of its bits, VARCHAR as hex bytes, "-" for the empty string):
//   case <name>
//   table <nrows> <type>...            (INT32 INT64 FP64 VARCHAR)
//   row <value>...                     (nrows lines after each table line)
//   scan <table> <index>:<type>...
//   join <build_left> <left> <right> <left_attr> <right_attr> <index>:<type>...
//   root <node>
//   run
// Output per case: "result <rows> <columns>", then one line per row.
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/contest/plan.h"

namespace {

DataType parse_type(const std::string& s) {
    if (s == "INT32") return DataType::INT32;
    if (s == "INT64") return DataType::INT64;
    if (s == "FP64") return DataType::FP64;
    if (s == "VARCHAR") return DataType::VARCHAR;
    throw std::runtime_error("unknown type " + s);
}

std::string unhex(const std::string& h) {
    if (h == "-") return std::string();
    std::string out;
    for (size_t i = 0; i + 1 < h.size(); i += 2) out.push_back(char(std::stoi(h.substr(i, 2), nullptr, 16)));
    return out;
}

std::string hex(const char* p, size_t n) {
    if (n == 0) return "-";
    static const char* d = "0123456789abcdef";
    std::string        s;
    for (size_t i = 0; i < n; ++i) {
        s.push_back(d[uint8_t(p[i]) >> 4]);
        s.push_back(d[uint8_t(p[i]) & 15]);
    }
    return s;
}

std::vector<std::tuple<size_t, DataType>> parse_attrs(std::istringstream& in) {
    std::vector<std::tuple<size_t, DataType>> out;
    std::string                               tok;
    while (in >> tok) {
        auto c = tok.find(':');
        out.emplace_back(size_t(std::stoull(tok.substr(0, c))), parse_type(tok.substr(c + 1)));
    }
    return out;
}

// Decodes one result column into printable values, row by row, following the
// page format (u16 rows, u16 non-NULL values, values, LSB-first bitmap at the
// page end; VARCHAR end offsets, long-string pages 0xffff / 0xfffe).
std::vector<std::string> decode_column(const Column& col) {
    std::vector<std::string> out;
    for (size_t pi = 0; pi < col.pages.size(); ++pi) {
        const auto*    page = reinterpret_cast<const uint8_t*>(col.pages[pi]->data);
        uint16_t       nr, nv;
        std::memcpy(&nr, page, 2);
        std::memcpy(&nv, page + 2, 2);
        if (col.type == DataType::VARCHAR && nr == 0xffff) {
            std::string s(reinterpret_cast<const char*>(page + 4), nv);
            while (pi + 1 < col.pages.size()) {
                const auto* q = reinterpret_cast<const uint8_t*>(col.pages[pi + 1]->data);
                uint16_t    t, n;
                std::memcpy(&t, q, 2);
                if (t != 0xfffe) break;
                std::memcpy(&n, q + 2, 2);
                s.append(reinterpret_cast<const char*>(q + 4), n);
                ++pi;
            }
            out.push_back(hex(s.data(), s.size()));
            continue;
        }
        const uint8_t* bm = page + PAGE_SIZE - (nr + 7) / 8;
        uint32_t       k  = 0;
        for (uint32_t r = 0; r < nr; ++r) {
            if (!((bm[r / 8] >> (r % 8)) & 1)) {
                out.push_back("NULL");
                continue;
            }
            char buf[32];
            switch (col.type) {
            case DataType::INT32: {
                int32_t v;
                std::memcpy(&v, page + 4 + 4 * k, 4);
                std::snprintf(buf, sizeof(buf), "%d", v);
                out.push_back(buf);
                break;
            }
            case DataType::INT64: {
                int64_t v;
                std::memcpy(&v, page + 8 + 8 * size_t(k), 8);
                std::snprintf(buf, sizeof(buf), "%" PRId64, v);
                out.push_back(buf);
                break;
            }
            case DataType::FP64: {
                uint64_t v;
                std::memcpy(&v, page + 8 + 8 * size_t(k), 8);
                std::snprintf(buf, sizeof(buf), "%016" PRIx64, v);
                out.push_back(buf);
                break;
            }
            case DataType::VARCHAR: {
                uint16_t end, begin = 0;
                std::memcpy(&end, page + 4 + 2 * k, 2);
                if (k) std::memcpy(&begin, page + 4 + 2 * (k - 1), 2);
                out.push_back(hex(reinterpret_cast<const char*>(page + 4 + 2 * nv + begin), end - begin));
                break;
            }
            }
            ++k;
        }
    }
    return out;
}

void add_table(Plan& plan, size_t nrows, const std::vector<DataType>& types, const std::vector<std::vector<std::string>>& rows) {
    ColumnarTable t;
    t.num_rows = nrows;
    for (size_t c = 0; c < types.size(); ++c) {
        Column col(types[c]);
        switch (types[c]) {
        case DataType::INT32: {
            ColumnInserter<int32_t> ins(col);
            for (auto& r : rows) r[c] == "NULL" ? ins.insert_null() : ins.insert(int32_t(std::stol(r[c])));
            ins.finalize();
            break;
        }
        case DataType::INT64: {
            ColumnInserter<int64_t> ins(col);
            for (auto& r : rows) r[c] == "NULL" ? ins.insert_null() : ins.insert(int64_t(std::stoll(r[c])));
            ins.finalize();
            break;
        }
        case DataType::FP64: {
            ColumnInserter<double> ins(col);
            for (auto& r : rows) {
                if (r[c] == "NULL") {
                    ins.insert_null();
                } else {
                    uint64_t bits = std::stoull(r[c], nullptr, 16);
                    double   d;
                    std::memcpy(&d, &bits, 8);
                    ins.insert(d);
                }
            }
            ins.finalize();
            break;
        }
        case DataType::VARCHAR: {
            ColumnInserter<std::string> ins(col);
            for (auto& r : rows) {
                if (r[c] == "NULL") {
                    ins.insert_null();
                } else {
                    std::string s = unhex(r[c]);
                    ins.insert(s);
                }
            }
            ins.finalize();
            break;
        }
        }
        t.columns.push_back(std::move(col));
    }
    plan.new_input(std::move(t));
}

} // namespace

// Prints a table's rows as decoded by decode_column (the --echo-inputs check).
void print_table(const char* tag, const ColumnarTable& t) {
    std::vector<std::vector<std::string>> cols;
    for (const Column& c : t.columns) {
        cols.push_back(decode_column(c));
        if (cols.back().size() != t.num_rows) throw std::runtime_error("column rows differ from num_rows");
    }
    std::printf("%s %zu %zu\n", tag, t.num_rows, t.columns.size());
    for (size_t r = 0; r < t.num_rows; ++r) {
        std::string s;
        for (size_t c = 0; c < cols.size(); ++c) s += (c ? " " : "") + cols[c][r];
        std::printf("%s\n", s.c_str());
    }
}

int main(int argc, char** argv) {
    // --echo-inputs: no GPU; each `run` prints the plan's input tables as
    // written by ColumnInserter and decoded back
    const bool echo = argc > 1 && std::strcmp(argv[1], "--echo-inputs") == 0;
    try {
        void*       context = echo ? nullptr : Contest::build_context();
        std::string line, name;
        Plan        plan;
        while (std::getline(std::cin, line)) {
            std::istringstream in(line);
            std::string        cmd;
            in >> cmd;
            if (cmd == "case") {
                plan = Plan();
                std::getline(in, name);
            } else if (cmd == "table") {
                size_t                n;
                std::string           ty;
                std::vector<DataType> types;
                in >> n;
                while (in >> ty) types.push_back(parse_type(ty));
                std::vector<std::vector<std::string>> rows(n);
                for (size_t r = 0; r < n; ++r) {
                    std::getline(std::cin, line);
                    std::istringstream rin(line);
                    std::string        tag, v;
                    rin >> tag;
                    while (rin >> v) rows[r].push_back(v);
                    if (tag != "row" || rows[r].size() != types.size()) throw std::runtime_error("bad row line");
                }
                add_table(plan, n, types, rows);
            } else if (cmd == "scan") {
                size_t t;
                in >> t;
                plan.new_scan_node(t, parse_attrs(in));
            } else if (cmd == "join") {
                int    bl;
                size_t l, r, la, ra;
                in >> bl >> l >> r >> la >> ra;
                plan.new_join_node(bl != 0, l, r, la, ra, parse_attrs(in));
            } else if (cmd == "root") {
                in >> plan.root;
            } else if (cmd == "run") {
                if (echo) {
                    for (const ColumnarTable& t : plan.inputs) print_table("input", t);
                } else {
                    ColumnarTable res = Contest::execute(plan, context);
                    print_table("result", res);
                    // `res` releases its pages with `delete` here
                }
                std::fflush(stdout);
            }
        }
        if (context) Contest::destroy_context(context);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "contest_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
