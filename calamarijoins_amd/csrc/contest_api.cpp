// contest_api.cpp -- Contest::build_context / destroy_context / execute
// (reference include/data_model/plan.h:380-385, src/execute.cpp:311-329)
// implemented on the C ABI of this library.
#include "../../include/cj/plan.h"
#include "../../include/cj_api.h"

#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>

namespace {

struct ContestContext {
    cj_ctx *ctx = nullptr;
    cj_result last{};  // pages handed out by the previous execute()
    bool has_last = false;
    std::mutex mu;
};

// Pages of results point into context-owned memory; columns hold this
// never-released mapping so ~Column leaves them alone (the reference uses
// the same device, materialize.h:45-53).
MappedMemory *borrowed_pages() {
    static MappedMemory *m = [] {
        auto *p = new MappedMemory(nullptr, 0);
        p->refs = 1u << 30;
        return p;
    }();
    return m;
}

void check(int rc) {
    if (rc == CJ_OK) return;
    std::string msg = cj_last_error();
    if (rc == CJ_ERR_OOM) throw std::bad_alloc();
    throw std::runtime_error("calamarijoins_amd: " + msg);
}

ContestContext *default_context() {
    static std::unique_ptr<ContestContext> d;
    static std::once_flag once;
    std::call_once(once, [] {
        d = std::make_unique<ContestContext>();
        cj_ctx_options o{};
        o.cache_decoded = 1;
        check(cj_ctx_create(&o, &d->ctx));
    });
    return d.get();
}

cj_attr to_attr(const std::tuple<size_t, DataType> &t) {
    return cj_attr{(uint64_t)std::get<0>(t), (int32_t)std::get<1>(t)};
}

}  // namespace

namespace Contest {

void *build_context() {
    auto *c = new ContestContext();
    cj_ctx_options o{};
    o.cache_decoded = 1;
    int rc = cj_ctx_create(&o, &c->ctx);
    if (rc != CJ_OK) {
        delete c;
        check(rc);
    }
    return c;
}

void destroy_context(void *context) {
    auto *c = static_cast<ContestContext *>(context);
    if (!c) return;
    if (c->has_last) cj_result_release(c->ctx, &c->last);
    cj_ctx_destroy(c->ctx);
    delete c;
}

ColumnarTable execute(const Plan &plan, void *context, TimingStats *stats_out) {
    auto *c = context ? static_cast<ContestContext *>(context) : default_context();
    std::lock_guard<std::mutex> g(c->mu);
    if (c->has_last) {  // outputs are valid until the next execute()
        cj_result_release(c->ctx, &c->last);
        c->has_last = false;
    }
    std::vector<std::vector<cj_column_desc>> cols(plan.inputs.size());
    std::vector<cj_table_desc> tables(plan.inputs.size());
    for (size_t i = 0; i < plan.inputs.size(); ++i) {
        const ColumnarTable &t = plan.inputs[i];
        for (const Column &col : t.columns)
            cols[i].push_back(cj_column_desc{(int32_t)col.type, (uint32_t)col.pages.size(),
                                             reinterpret_cast<const void *const *>(col.pages.data())});
        tables[i] = cj_table_desc{(uint64_t)t.num_rows, (uint32_t)cols[i].size(), cols[i].data()};
    }
    std::vector<std::vector<cj_attr>> attrs(plan.nodes.size());
    std::vector<cj_node_desc> nodes(plan.nodes.size());
    for (size_t i = 0; i < plan.nodes.size(); ++i) {
        const PlanNode &n = plan.nodes[i];
        for (const auto &a : n.output_attrs) attrs[i].push_back(to_attr(a));
        cj_node_desc d{};
        if (const auto *s = std::get_if<ScanNode>(&n.data)) {
            d.kind = CJ_NODE_SCAN;
            d.base_table_id = s->base_table_id;
        } else {
            const auto &j = std::get<JoinNode>(n.data);
            d.kind = CJ_NODE_JOIN;
            d.build_left = j.build_left;
            d.left = j.left;
            d.right = j.right;
            d.left_attr = j.left_attr;
            d.right_attr = j.right_attr;
        }
        d.num_output_attrs = (uint32_t)attrs[i].size();
        d.output_attrs = attrs[i].data();
        nodes[i] = d;
    }
    cj_plan_desc pd{(uint32_t)nodes.size(), nodes.data(), (uint32_t)tables.size(), tables.data(), (uint64_t)plan.root};
    cj_result res{};
    cj_timing tm{};
    check(cj_execute(c->ctx, &pd, 0, &res, &tm));
    ColumnarTable out;
    out.num_rows = res.num_rows;
    bool any_pages = false;
    for (uint32_t i = 0; i < res.num_columns; ++i) {
        Column col((DataType)res.columns[i].type);
        if (res.num_rows) {
            for (uint32_t p = 0; p < res.columns[i].num_pages; ++p)
                col.pages.push_back(reinterpret_cast<Page *>(res.columns[i].pages[p]));
            col.assign_mapped_memory(borrowed_pages());
            any_pages = true;
        }
        out.columns.push_back(std::move(col));
    }
    if (any_pages) {
        c->last = res;
        c->has_last = true;
    } else {
        cj_result_release(c->ctx, &res);
    }
    if (stats_out) {
        stats_out->hashtable_build_ms = tm.hashtable_build_ms;
        stats_out->hash_join_probe_ms = tm.hash_join_probe_ms;
        stats_out->nested_loop_join_ms = tm.nested_loop_join_ms;
        stats_out->materialize_ms = tm.materialize_ms;
        stats_out->setup_ms = tm.setup_ms;
        stats_out->total_execution_ms = tm.total_execution_ms;
        stats_out->intermediate_ms = tm.intermediate_ms;
    }
    return out;
}

}  // namespace Contest
