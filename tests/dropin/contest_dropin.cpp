// contest_dropin.cpp -- a caller of the contest API exactly as the
// reference's callers use it (tests/unit_tests.cpp:28-30,
// tests/read_sql.cpp:1380): Plan objects with Column-owned pages,
// Contest::build_context / execute / destroy_context, TimingStats.
//
// Reads a spec on stdin, one directive per line:
//   case <name>                       start a plan
//   input <path.tbl>                  add an input table (pages copied into
//                                     Column-owned `new Page`s)
//   scan <table> <k> (<idx> <type>)*  ScanNode
//   join <build_left> <left> <right> <left_attr> <right_attr> <k> (<idx> <type>)*
//   root <node>
//   run <out.tbl>                     execute; write the result to out.tbl
// and prints one line per run:
//   result <name> rows <n> cols <k> pages <p0,p1,..> empty_ok <0|1>
//          stable <0|1> timing <build probe nl mat setup total inter>
// where `stable` says the result pages kept their bytes until the next
// execute() (checksummed after the call and again just before the next
// one) and `empty_ok` that an empty result has columns without pages and
// without a mapping.  Exit status 0 unless a call threw.
#include "cj/plan.h"
#include "cj_api.h"

#include <cstdio>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

ColumnarTable load_input(const std::string &path) {
    cj_tbl *h = nullptr;
    cj_table_desc d{};
    if (cj_tbl_open(path.c_str(), &h, &d) != CJ_OK) throw std::runtime_error(cj_last_error());
    ColumnarTable t;
    t.num_rows = d.num_rows;
    for (uint32_t c = 0; c < d.num_columns; ++c) {
        t.columns.emplace_back((DataType)d.columns[c].type);
        Column &col = t.columns.back();
        for (uint32_t p = 0; p < d.columns[c].num_pages; ++p) {
            Page *pg = col.new_page();
            std::memcpy(pg->data, d.columns[c].pages[p], PAGE_SIZE);
        }
    }
    cj_tbl_close(h);
    return t;
}

std::vector<std::tuple<size_t, DataType>> read_attrs(std::istringstream &in) {
    size_t k;
    in >> k;
    std::vector<std::tuple<size_t, DataType>> a;
    for (size_t i = 0; i < k; ++i) {
        size_t idx;
        int t;
        in >> idx >> t;
        a.emplace_back(idx, (DataType)t);
    }
    return a;
}

uint64_t checksum(const ColumnarTable &t) {
    uint64_t h = 1469598103934665603ull;
    for (const Column &c : t.columns)
        for (const Page *p : c.pages) {
            const unsigned char *b = reinterpret_cast<const unsigned char *>(p->data);
            for (size_t i = 0; i < PAGE_SIZE; i += 8) {
                uint64_t w;
                std::memcpy(&w, b + i, 8);
                h = (h ^ w) * 1099511628211ull;
            }
        }
    return h;
}

void write_result(const ColumnarTable &t, const std::string &path) {
    std::vector<std::vector<const void *>> ptrs(t.columns.size());
    std::vector<cj_column_desc> cols(t.columns.size());
    for (size_t i = 0; i < t.columns.size(); ++i) {
        for (const Page *p : t.columns[i].pages) ptrs[i].push_back(p);
        cols[i] = cj_column_desc{(int32_t)t.columns[i].type, (uint32_t)ptrs[i].size(), ptrs[i].data()};
    }
    cj_table_desc d{t.num_rows, (uint32_t)cols.size(), cols.data()};
    if (cj_tbl_write(path.c_str(), &d) != CJ_OK) throw std::runtime_error(cj_last_error());
}

}  // namespace

int main() {
    void *ctx = Contest::build_context();
    int status = 0;
    Plan plan;
    std::string name;
    ColumnarTable last;  // the previous result: valid until the next execute()
    uint64_t last_sum = 0;
    bool have_last = false;
    std::string line;
    std::vector<std::string> pending;  // printed once the next execute() confirms stability
    auto flush_pending = [&](bool stable) {
        for (auto &p : pending) std::cout << p << " stable " << (stable ? 1 : 0) << "\n";
        pending.clear();
        std::cout.flush();
    };
    try {
        while (std::getline(std::cin, line)) {
            std::istringstream in(line);
            std::string op;
            in >> op;
            if (op == "case") {
                in >> name;
                plan = Plan();
            } else if (op == "input") {
                std::string path;
                in >> path;
                plan.new_input(load_input(path));
            } else if (op == "scan") {
                size_t table;
                in >> table;
                plan.new_scan_node(table, read_attrs(in));
            } else if (op == "join") {
                int bl;
                size_t l, r, la, ra;
                in >> bl >> l >> r >> la >> ra;
                plan.new_join_node(bl != 0, l, r, la, ra, read_attrs(in));
            } else if (op == "root") {
                in >> plan.root;
            } else if (op == "run") {
                std::string out;
                in >> out;
                // the previous result's pages must still hold their bytes
                if (have_last) flush_pending(checksum(last) == last_sum);
                Contest::TimingStats ts;
                ColumnarTable res = Contest::execute(plan, ctx, &ts);
                last = ColumnarTable();  // the earlier result object goes away after the call
                bool empty_ok = true;
                std::ostringstream pages;
                for (size_t i = 0; i < res.columns.size(); ++i) {
                    pages << (i ? "," : "") << res.columns[i].pages.size();
                    if (res.num_rows == 0 && (!res.columns[i].pages.empty() || res.columns[i].mapped_memory))
                        empty_ok = false;
                }
                write_result(res, out);
                std::ostringstream r;
                r << "result " << name << " rows " << res.num_rows << " cols " << res.columns.size() << " pages "
                  << (res.columns.empty() ? "-" : pages.str()) << " empty_ok " << (empty_ok ? 1 : 0) << " timing "
                  << ts.hashtable_build_ms << " " << ts.hash_join_probe_ms << " " << ts.nested_loop_join_ms << " "
                  << ts.materialize_ms << " " << ts.setup_ms << " " << ts.total_execution_ms << " "
                  << ts.intermediate_ms;
                pending.push_back(r.str());
                last = std::move(res);
                last_sum = checksum(last);
                have_last = true;
            }
        }
        if (have_last) flush_pending(checksum(last) == last_sum);
    } catch (const std::exception &e) {
        flush_pending(false);
        std::cout << "error " << name << " " << e.what() << "\n";
        status = 1;
    }
    last = ColumnarTable();
    Contest::destroy_context(ctx);
    return status;
}
