// SIGMOD 2025 contest API surface, restated for this library.
//
// The struct layouts below must be ABI-identical to the contest header
// (reference include/plan.h:27-149) so that a harness compiled against the
// contest's plan.h can link this library in place of src/execute.cpp. The
// page writer `ColumnInserter<T>` follows the contest's page format rules
// (reference include/plan.h:151-335, README.md:131-147): it is used by tests
// and benchmarks to build inputs exactly as the contest loader does.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <string>
#include <string_view>
#include <tuple>
#include <variant>
#include <vector>

#include "attribute.h"

enum class NodeType {
    HashJoin,
    Scan,
};

struct ScanNode {
    size_t base_table_id;
};

struct JoinNode {
    bool   build_left;
    size_t left;
    size_t right;
    size_t left_attr;
    size_t right_attr;
};

struct PlanNode {
    std::variant<ScanNode, JoinNode>          data;
    std::vector<std::tuple<size_t, DataType>> output_attrs;

    PlanNode(std::variant<ScanNode, JoinNode> d, std::vector<std::tuple<size_t, DataType>> attrs)
        : data(std::move(d)), output_attrs(std::move(attrs)) {}
};

constexpr size_t PAGE_SIZE = 8192;

struct alignas(8) Page {
    std::byte data[PAGE_SIZE];
};

// A column owns its pages; they are released with plain `delete`.
struct Column {
    DataType           type;
    std::vector<Page*> pages;

    explicit Column(DataType t) : type(t), pages() {}
    Column(Column&& o) noexcept : type(o.type), pages(std::move(o.pages)) { o.pages.clear(); }
    Column& operator=(Column&& o) noexcept {
        if (this != &o) {
            release();
            type  = o.type;
            pages = std::move(o.pages);
            o.pages.clear();
        }
        return *this;
    }
    Column(const Column&)            = delete;
    Column& operator=(const Column&) = delete;
    ~Column() { release(); }

    Page* new_page() {
        Page* p = new Page;
        pages.push_back(p);
        return p;
    }

private:
    void release() {
        for (Page* p : pages) delete p;
        pages.clear();
    }
};

struct ColumnarTable {
    size_t              num_rows{0};
    std::vector<Column> columns;
};

struct Plan {
    std::vector<PlanNode>      nodes;
    std::vector<ColumnarTable> inputs;
    size_t                     root;

    size_t new_join_node(bool build_left, size_t left, size_t right, size_t left_attr,
                         size_t right_attr, std::vector<std::tuple<size_t, DataType>> output_attrs) {
        JoinNode j{build_left, left, right, left_attr, right_attr};
        nodes.emplace_back(j, std::move(output_attrs));
        return nodes.size() - 1;
    }

    size_t new_scan_node(size_t base_table_id, std::vector<std::tuple<size_t, DataType>> output_attrs) {
        nodes.emplace_back(ScanNode{base_table_id}, std::move(output_attrs));
        return nodes.size() - 1;
    }

    size_t new_input(ColumnarTable input) {
        inputs.emplace_back(std::move(input));
        return inputs.size() - 1;
    }
};

// Page writer for fixed-width columns. Page layout: u16 row count, u16 non-null
// count, the non-null values packed from the first `sizeof(T)`-aligned offset
// (4 for INT32, 8 for INT64/FP64), and an LSB-first validity bitmap in the last
// ceil(rows/8) bytes of the page. A page is closed as soon as the next value
// (or NULL) would no longer fit next to the bitmap.
template <class T>
struct ColumnInserter {
    static constexpr size_t kFirst = sizeof(T) < 4 ? 4 : sizeof(T);

    Column&              column;
    size_t               page_idx  = 0;
    uint16_t             rows      = 0;
    size_t               value_end = kFirst;
    std::vector<uint8_t> bits;

    explicit ColumnInserter(Column& c) : column(c), bits(PAGE_SIZE, 0) {}

    std::byte* current() {
        if (page_idx == column.pages.size()) column.new_page();
        return column.pages[page_idx]->data;
    }

    void close_page() {
        std::byte* p        = current();
        uint16_t   non_null = static_cast<uint16_t>((value_end - kFirst) / sizeof(T));
        std::memcpy(p, &rows, 2);
        std::memcpy(p + 2, &non_null, 2);
        size_t nbytes = (rows + 7u) / 8u;
        std::memcpy(p + PAGE_SIZE - nbytes, bits.data(), nbytes);
        std::fill(bits.begin(), bits.begin() + nbytes, 0);
        ++page_idx;
        rows      = 0;
        value_end = kFirst;
    }

    void insert(T v) {
        if (value_end + 4 + rows / 8 + 1 > PAGE_SIZE) close_page();
        std::byte* p = current();
        std::memcpy(p + value_end, &v, sizeof(T));
        value_end += sizeof(T);
        bits[rows / 8] |= static_cast<uint8_t>(1u << (rows % 8));
        ++rows;
    }

    void insert_null() {
        if (value_end + rows / 8 + 1 > PAGE_SIZE) close_page();
        current();
        bits[rows / 8] &= static_cast<uint8_t>(~(1u << (rows % 8)));
        ++rows;
    }

    void finalize() {
        if (rows != 0) close_page();
    }
};

// Page writer for VARCHAR columns. Short-string page: u16 rows, u16 non-null
// count, u16 end offsets of the non-null strings at +4, the characters right
// after the offsets, bitmap at the page end. Strings longer than
// PAGE_SIZE - 7 bytes go to dedicated long-string pages: the first carries
// row count 0xffff, continuations 0xfffe, each with a u16 byte count at +2 and
// the bytes from +4.
template <>
struct ColumnInserter<std::string> {
    Column&              column;
    size_t               page_idx   = 0;
    uint16_t             rows       = 0;
    uint16_t             chars      = 0;
    size_t               offset_end = 4;
    std::vector<char>    text;
    std::vector<uint8_t> bits;

    explicit ColumnInserter(Column& c) : column(c), text(PAGE_SIZE), bits(PAGE_SIZE, 0) {}

    std::byte* current() {
        if (page_idx == column.pages.size()) column.new_page();
        return column.pages[page_idx]->data;
    }

    void close_page() {
        std::byte* p        = current();
        uint16_t   non_null = static_cast<uint16_t>((offset_end - 4) / 2);
        std::memcpy(p, &rows, 2);
        std::memcpy(p + 2, &non_null, 2);
        std::memcpy(p + offset_end, text.data(), chars);
        size_t nbytes = (rows + 7u) / 8u;
        std::memcpy(p + PAGE_SIZE - nbytes, bits.data(), nbytes);
        std::fill(bits.begin(), bits.begin() + nbytes, 0);
        ++page_idx;
        rows       = 0;
        chars      = 0;
        offset_end = 4;
    }

    void insert_long(std::string_view s) {
        size_t done = 0;
        bool   first = true;
        while (done < s.size()) {
            std::byte* p    = current();
            uint16_t   tag  = first ? 0xffff : 0xfffe;
            uint16_t   part = static_cast<uint16_t>(std::min(s.size() - done, PAGE_SIZE - 4));
            std::memcpy(p, &tag, 2);
            std::memcpy(p + 2, &part, 2);
            std::memcpy(p + 4, s.data() + done, part);
            done += part;
            first = false;
            ++page_idx;
        }
    }

    void insert(std::string_view s) {
        if (s.size() > PAGE_SIZE - 7) {
            if (rows > 0) close_page();
            insert_long(s);
            return;
        }
        if (offset_end + 2 + chars + s.size() + rows / 8 + 1 > PAGE_SIZE) close_page();
        std::memcpy(text.data() + chars, s.data(), s.size());
        chars = static_cast<uint16_t>(chars + s.size());
        std::byte* p = current();
        std::memcpy(p + offset_end, &chars, 2);
        offset_end += 2;
        bits[rows / 8] |= static_cast<uint8_t>(1u << (rows % 8));
        ++rows;
    }

    void insert_null() {
        if (offset_end + chars + rows / 8 + 1 > PAGE_SIZE) close_page();
        current();
        bits[rows / 8] &= static_cast<uint8_t>(~(1u << (rows % 8)));
        ++rows;
    }

    void finalize() {
        if (rows != 0) close_page();
    }
};

namespace Contest {

void* build_context();
void  destroy_context(void*);

ColumnarTable execute(const Plan& plan, void* context);

} // namespace Contest
