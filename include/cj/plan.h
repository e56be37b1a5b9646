// cj/plan.h -- the contest API (Plan / ColumnarTable / 8 KiB Page and
// Contest::execute), re-declared for the MI355X engine.
//
// Layout and source compatible with the reference's declarations
// (include/data_model/plan.h:10-385, include/foundation/attribute.h:7-12;
// the public contest specification in README.md:24-104), so code written
// against the reference compiles and links against libcalamarijoins_amd.so
// unchanged.  Contest::execute, build_context and destroy_context are
// implemented by the engine (calamarijoins_amd/csrc/contest_api.cpp) on top
// of the C ABI in cj_api.h.
//
// Ownership follows the reference: Column deletes its pages unless it holds
// a MappedMemory reference; execute() returns columns whose pages belong to
// the context and stay valid until the next execute() or destroy_context()
// on that context.
#pragma once

#include <sys/mman.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <string>
#include <string_view>
#include <tuple>
#include <variant>
#include <vector>

enum class DataType {
    INT32,
    INT64,
    FP64,
    VARCHAR,
};

// A reference-counted mapping that backs the pages of some columns
// (reference plan.h:10-44).  The destructor unmaps the region.
class MappedMemory {
  public:
    void *addr;
    size_t length;
    size_t refs;
    MappedMemory(void *a, size_t len) : addr(a), length(len), refs(0) {}
    MappedMemory(const MappedMemory &) = delete;
    MappedMemory &operator=(const MappedMemory &) = delete;
    MappedMemory(MappedMemory &&o) noexcept : addr(o.addr), length(o.length), refs(o.refs) {
        o.addr = nullptr;
        o.length = 0;
        o.refs = 0;
    }
    MappedMemory &operator=(MappedMemory &&o) noexcept {
        if (this != &o) {
            addr = o.addr;
            length = o.length;
            refs = o.refs;
            o.addr = nullptr;
            o.length = 0;
            o.refs = 0;
        }
        return *this;
    }
    ~MappedMemory() {
        if (addr) munmap(addr, length);
    }
};

enum class NodeType {
    HashJoin,
    Scan,
};

struct ScanNode {
    size_t base_table_id;
};

struct JoinNode {
    bool build_left;  // a hint: the engine builds on the side with fewer rows
    size_t left;
    size_t right;
    size_t left_attr;
    size_t right_attr;
};

struct PlanNode {
    std::variant<ScanNode, JoinNode> data;
    std::vector<std::tuple<size_t, DataType>> output_attrs;

    PlanNode(std::variant<ScanNode, JoinNode> d, std::vector<std::tuple<size_t, DataType>> attrs)
        : data(std::move(d)), output_attrs(std::move(attrs)) {}
};

constexpr size_t PAGE_SIZE = 8192;

struct alignas(8) Page {
    std::byte data[PAGE_SIZE];
};

struct Column {
    DataType type;
    std::vector<Page *> pages;
    MappedMemory *mapped_memory;

    explicit Column(DataType t) : type(t), pages(), mapped_memory(nullptr) {}
    Column(Column &&o) noexcept : type(o.type), pages(std::move(o.pages)), mapped_memory(o.mapped_memory) {
        o.pages.clear();
        o.mapped_memory = nullptr;
    }
    Column &operator=(Column &&o) noexcept {
        if (this != &o) {
            release();
            type = o.type;
            pages = std::move(o.pages);
            o.pages.clear();
            mapped_memory = o.mapped_memory;
            o.mapped_memory = nullptr;
        }
        return *this;
    }
    Column(const Column &) = delete;
    Column &operator=(const Column &) = delete;
    ~Column() { release(); }

    Page *new_page() {
        pages.push_back(new Page);
        return pages.back();
    }
    void assign_mapped_memory(MappedMemory *m) {
        mapped_memory = m;
        ++mapped_memory->refs;
    }

  private:
    void release() {
        if (mapped_memory) {
            if (--mapped_memory->refs == 0) delete mapped_memory;
            mapped_memory = nullptr;
            pages.clear();
            return;
        }
        for (Page *p : pages) delete p;
        pages.clear();
    }
};

struct ColumnarTable {
    size_t num_rows{0};
    std::vector<Column> columns;
};

struct Plan {
    std::vector<PlanNode> nodes;
    std::vector<ColumnarTable> inputs;
    size_t root;

    size_t new_join_node(bool build_left, size_t left, size_t right, size_t left_attr, size_t right_attr,
                         std::vector<std::tuple<size_t, DataType>> output_attrs) {
        nodes.emplace_back(JoinNode{build_left, left, right, left_attr, right_attr}, std::move(output_attrs));
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

// Page writer for building input columns (the API's ColumnInserter,
// reference plan.h:187-366): fixed-width values are appended until the
// next value plus the bitmap would overflow the page.
template <class T>
struct ColumnInserter {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "INT32, INT64 or FP64");
    Column &column;
    size_t page_idx = 0;
    uint16_t rows = 0;
    uint16_t values = 0;
    std::vector<uint8_t> bitmap;

    static constexpr size_t begin() { return sizeof(T) < 4 ? 4 : sizeof(T); }
    explicit ColumnInserter(Column &c) : column(c), bitmap(PAGE_SIZE, 0) {}

    std::byte *page() {
        if (page_idx == column.pages.size()) column.new_page();
        return column.pages[page_idx]->data;
    }
    void flush() {
        std::byte *p = page();
        std::memcpy(p, &rows, 2);
        std::memcpy(p + 2, &values, 2);
        size_t nb = (rows + 7) / 8;
        std::memcpy(p + PAGE_SIZE - nb, bitmap.data(), nb);
        std::fill(bitmap.begin(), bitmap.end(), 0);
        ++page_idx;
        rows = values = 0;
    }
    void insert(T v) {
        if (begin() + (size_t)(values + 1) * sizeof(T) + rows / 8 + 1 > PAGE_SIZE) flush();
        std::memcpy(page() + begin() + (size_t)values * sizeof(T), &v, sizeof(T));
        bitmap[rows / 8] |= (uint8_t)(1u << (rows % 8));
        ++values;
        ++rows;
    }
    void insert_null() {
        if (begin() + (size_t)values * sizeof(T) + rows / 8 + 1 > PAGE_SIZE) flush();
        ++rows;
    }
    void finalize() {
        if (rows) flush();
    }
};

template <>
struct ColumnInserter<std::string> {
    Column &column;
    size_t page_idx = 0;
    uint16_t rows = 0;
    std::vector<uint16_t> offsets;
    std::vector<char> chars;
    std::vector<uint8_t> bitmap;

    explicit ColumnInserter(Column &c) : column(c), bitmap(PAGE_SIZE, 0) {}

    std::byte *page() {
        if (page_idx == column.pages.size()) column.new_page();
        return column.pages[page_idx]->data;
    }
    void flush() {
        std::byte *p = page();
        uint16_t nv = (uint16_t)offsets.size();
        std::memcpy(p, &rows, 2);
        std::memcpy(p + 2, &nv, 2);
        std::memcpy(p + 4, offsets.data(), offsets.size() * 2);
        std::memcpy(p + 4 + offsets.size() * 2, chars.data(), chars.size());
        size_t nb = (rows + 7) / 8;
        std::memcpy(p + PAGE_SIZE - nb, bitmap.data(), nb);
        std::fill(bitmap.begin(), bitmap.end(), 0);
        offsets.clear();
        chars.clear();
        ++page_idx;
        rows = 0;
    }
    void insert_long(std::string_view v) {
        bool first = true;
        for (size_t off = 0; off < v.size();) {
            std::byte *p = page();
            uint16_t tag = first ? 0xffff : 0xfffe;
            uint16_t n = (uint16_t)std::min(v.size() - off, PAGE_SIZE - 4);
            std::memcpy(p, &tag, 2);
            std::memcpy(p + 2, &n, 2);
            std::memcpy(p + 4, v.data() + off, n);
            off += n;
            first = false;
            ++page_idx;
        }
    }
    void insert(std::string_view v) {
        if (v.size() > PAGE_SIZE - 7) {
            if (rows) flush();
            insert_long(v);
            return;
        }
        if (4 + (offsets.size() + 1) * 2 + chars.size() + v.size() + rows / 8 + 1 > PAGE_SIZE) flush();
        chars.insert(chars.end(), v.begin(), v.end());
        offsets.push_back((uint16_t)chars.size());
        bitmap[rows / 8] |= (uint8_t)(1u << (rows % 8));
        ++rows;
    }
    void insert_null() {
        if (4 + offsets.size() * 2 + chars.size() + rows / 8 + 1 > PAGE_SIZE) flush();
        ++rows;
    }
    void finalize() {
        if (rows) flush();
    }
};

namespace Contest {

// Per-phase wall-clock breakdown of one execute() (reference plan.h:370-378).
struct TimingStats {
    int64_t hashtable_build_ms = 0;
    int64_t hash_join_probe_ms = 0;
    int64_t nested_loop_join_ms = 0;
    int64_t materialize_ms = 0;
    int64_t setup_ms = 0;
    int64_t total_execution_ms = 0;
    int64_t intermediate_ms = 0;
};

// Creates the engine context (GPU 0, streams, memory pool).
void *build_context();
void destroy_context(void *context);
// Executes `plan` on the GPU.  `context` may be null (a process-wide
// default context is used).  Throws std::runtime_error on rejected plans
// and std::bad_alloc when device memory runs out.
ColumnarTable execute(const Plan &plan, void *context, TimingStats *stats_out = nullptr);

}  // namespace Contest
