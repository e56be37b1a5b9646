// batch_host.cpp -- host side of the batched radix join (batch.hpp): plan
// analysis, the per-template device workspace, hipGraph capture of the
// first phase sequence, one read-back of the per-plan result sizes, the
// write and pack phase and the results; plus the autotuner that picks the
// hash load factor and the persistent grid size per batch template class.
//
// A plan joins in the batch when it is one join of two base-table scans of
// resident tables, with INT32 keys, fixed-width output columns (at most 8,
// at most 4 distinct non-key columns per side) and a build side of 8 to 2^22
// rows (smaller builds take the nested loop, which the per-plan executor
// runs); every other plan runs through Context::execute.  The build side is
// the side with fewer rows, ties left (join_setup.h:108-136), as in the
// per-plan executor.  An output column that is a join key column of either
// side is the matched key itself.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>

#include "batch.hpp"
#include "engine.hpp"

namespace cj {

namespace {

struct PlanInfo {
    bool eligible = false;
    DevColumn *bk = nullptr, *pk = nullptr;
    std::vector<DevColumn *> bcar, pcar;  // carried columns per side
    struct Out {
        uint8_t src;  // 0 key, 1 build carried, 2 probe carried
        uint8_t j;
        int32_t type;
    };
    std::vector<Out> outs;
};

PlanInfo analyse(Context &ctx, const cj_plan_desc &p) {
    PlanInfo pi;
    if (p.root >= p.num_nodes) return pi;
    const cj_node_desc &j = p.nodes[p.root];
    if (j.kind != CJ_NODE_JOIN || j.left >= p.num_nodes || j.right >= p.num_nodes) return pi;
    const cj_node_desc &L = p.nodes[j.left], &R = p.nodes[j.right];
    if (L.kind != CJ_NODE_SCAN || R.kind != CJ_NODE_SCAN) return pi;
    if (L.base_table_id >= p.num_inputs || R.base_table_id >= p.num_inputs) return pi;
    const cj_table_desc &lt = p.inputs[L.base_table_id], &rt = p.inputs[R.base_table_id];
    auto col_of = [&](const cj_node_desc &s, const cj_table_desc &t, uint64_t a) -> DevColumn * {
        if (a >= s.num_output_attrs) return nullptr;
        const uint64_t c = s.output_attrs[a].index;
        if (c >= t.num_columns) return nullptr;
        if (t.columns[c].type != s.output_attrs[a].type) return nullptr;
        DevColumn *d = ctx.find_resident(t.columns[c]);
        return d && d->loaded ? d : nullptr;
    };
    const bool build_left = lt.num_rows <= rt.num_rows;
    const cj_node_desc &B = build_left ? L : R, &P = build_left ? R : L;
    const cj_table_desc &bt = build_left ? lt : rt, &ptt = build_left ? rt : lt;
    if (bt.num_rows < 8) return pi;  // nested loop (execute.cpp:236-238)
    if (bt.num_rows > kBatchMaxBuildRows) return pi;
    if (ptt.num_rows >= (1ull << 31)) return pi;  // plan-local probe tuples and result rows are 32-bit
    const uint64_t ba = build_left ? j.left_attr : j.right_attr, pa = build_left ? j.right_attr : j.left_attr;
    pi.bk = col_of(B, bt, ba);
    pi.pk = col_of(P, ptt, pa);
    if (!pi.bk || !pi.pk || pi.bk->type != CJ_INT32 || pi.pk->type != CJ_INT32) return pi;
    if (j.num_output_attrs > kBatchMaxOut) return pi;
    const uint64_t lw = L.num_output_attrs;
    for (uint32_t a = 0; a < j.num_output_attrs; ++a) {
        const uint64_t ci = j.output_attrs[a].index;
        const int32_t ty = j.output_attrs[a].type;
        if (ty == CJ_VARCHAR || ci >= lw + R.num_output_attrs) return pi;
        const bool from_left = ci < lw;
        const cj_node_desc &S = from_left ? L : R;
        const cj_table_desc &st = from_left ? lt : rt;
        DevColumn *c = col_of(S, st, from_left ? ci : ci - lw);
        if (!c || c->type != ty) return pi;
        const bool from_build = from_left == build_left;
        if (c == (from_build ? pi.bk : pi.pk)) {
            pi.outs.push_back({0, 0, ty});
            continue;
        }
        auto &car = from_build ? pi.bcar : pi.pcar;
        auto it = std::find(car.begin(), car.end(), c);
        if (it == car.end()) {
            if (car.size() == kBatchMaxCarry) return pi;
            car.push_back(c);
            it = car.end() - 1;
        }
        pi.outs.push_back({(uint8_t)(from_build ? 1 : 2), (uint8_t)(it - car.begin()), ty});
    }
    pi.eligible = true;
    return pi;
}

uint32_t ceil_log2(uint64_t x) {
    uint32_t l = 0;
    while ((1ull << l) < x) ++l;
    return l;
}

uint32_t width_of(int32_t type) { return type == CJ_INT32 ? 4u : 8u; }

double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

}  // namespace

// Device workspace of one batch template: sized by its unit counts, reused
// (and its phase-1 graph replayed) while a batch of the same shape repeats.
struct BatchTemplate {
    std::vector<void *> blocks;
    BatchArgs args{};
    BPlan *d_plans = nullptr;
    uint32_t *d_starts = nullptr;  // 3 x (n + 1)
    BDecode *d_jobs = nullptr;
    uint32_t *d_job_start = nullptr;
    uint64_t *scan_tmp = nullptr;
    size_t scan_elems = 0;
    uint64_t *totals = nullptr;  // [n]
    uint32_t *valid_all = nullptr;
    uint64_t valid_words = 0;
    std::vector<void *> dec_values;
    std::vector<uint32_t *> dec_valid;
    std::vector<BPlan> hplans;  // device pointers of the per-plan arrays
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t njobs = 0, ndpages = 0;
    BatchGrids grids{};
    void release(DevicePool &pool) {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        for (void *b : blocks) pool.free(b);
        blocks.clear();
    }
};

static double run_batch(Context &ctx, const std::vector<uint32_t> &batched, const std::vector<PlanInfo> &info,
                        const BatchConfig &cfg, uint32_t flags, cj_result *out, cj_batch_stats *bs);

void Context::execute_batch(const cj_plan_desc *plans, uint32_t n, uint32_t flags, cj_result *out,
                            cj_batch_stats *stats) {
    auto t_start = std::chrono::steady_clock::now();
    for (uint32_t i = 0; i < n; ++i) std::memset(&out[i], 0, sizeof out[i]);
    cj_batch_stats st_local{};
    cj_batch_stats &bs = stats ? *stats : st_local;
    std::memset(&bs, 0, sizeof bs);
    std::vector<PlanInfo> info(n);
    std::vector<uint32_t> batched, single;
    for (uint32_t i = 0; i < n; ++i) {
        info[i] = analyse(*this, plans[i]);
        (info[i].eligible ? batched : single).push_back(i);
    }
    // plans outside the batch shape run one at a time
    for (uint32_t i : single) execute(plans[i], flags, &out[i], nullptr);
    bs.plans_batched = (uint32_t)batched.size();
    bs.plans_single = (uint32_t)single.size();
    if (batched.empty()) {
        bs.total_ms = ms_since(t_start);
        return;
    }
    const uint32_t nb_plans = (uint32_t)batched.size();
    uint64_t sum_probe = 0, sum_build = 0;
    for (uint32_t i : batched) {
        sum_probe += info[i].pk->num_rows;
        sum_build += info[i].bk->num_rows;
    }
    // configuration: the tuned one for this template class, else the default
    const uint64_t cls = ((uint64_t)ceil_log2(nb_plans + 1) << 16) | ((uint64_t)ceil_log2(sum_probe + 1) << 8) |
                         ceil_log2(sum_build + 1);
    BatchConfig cfg{2, 11, 16384, 1, 4096};  // the best of the 256-join workload's sweep
    bool tuned = false;
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = batch_tuning.find(cls);
        if (batch_pinned) {
            cfg = batch_pin;  // cj_batch_config_set
            tuned = true;
        } else if (it != batch_tuning.end()) {
            cfg = it->second;
            tuned = true;
        }
    }
    if (!tuned && (flags & CJ_BATCH_AUTOTUNE)) {
        // load factor x radix bits (slots per partition) x join tile x join
        // scheduling x grid (each kernel's resident grid, or 4096 workgroups)
        std::vector<BatchConfig> cands;
        for (uint32_t lf : {2u, 4u})
            for (uint32_t lp : {11u, 12u})
                for (uint32_t chunk : {4096u, 16384u})
                    for (uint32_t jq : {0u, 1u})
                        for (uint32_t grid : {0u, 4096u}) cands.push_back({lf, lp, chunk, jq, grid});
        double best = 1e30;
        std::vector<uint64_t> rows0;  // every configuration must give the same per-plan row counts
        for (const BatchConfig &c : cands) {
            // the second run replays the captured graph with the pool warm
            double ms = 0;
            for (int rep = 0; rep < 2; ++rep) {
                std::vector<cj_result> tmp(n);
                ms = run_batch(*this, batched, info, c, flags | CJ_EXEC_OUTPUT_DEVICE, tmp.data(), nullptr);
                std::vector<uint64_t> rows;
                for (uint32_t i : batched) rows.push_back(tmp[i].num_rows);
                for (uint32_t i : batched) release(&tmp[i]);
                if (rows0.empty()) rows0 = rows;
                else if (rows != rows0) throw DeviceError("batch configurations disagree on result sizes");
            }
            if (ms < best) {
                best = ms;
                cfg = c;
            }
        }
        std::lock_guard<std::mutex> g(mu);
        batch_tuning[cls] = cfg;
        bs.tuned = 1;
    }
    bs.device_ms = run_batch(*this, batched, info, cfg, flags, out, &bs);
    bs.load_factor = cfg.load_factor;
    bs.log_part = cfg.log_part;
    bs.chunk = cfg.chunk;
    bs.join_queue = cfg.join_queue;
    bs.total_ms = ms_since(t_start);
}

static double run_batch(Context &ctx, const std::vector<uint32_t> &batched, const std::vector<PlanInfo> &info,
                        const BatchConfig &cfg, uint32_t flags, cj_result *out, cj_batch_stats *bs) {
    DevicePool &pool = ctx.pool;
    PinnedPool &pinned = ctx.pinned;
    const uint32_t n = (uint32_t)batched.size();
    hipStream_t s = ctx.acquire_stream();
    struct Rel_ {
        Context &c;
        hipStream_t s;
        ~Rel_() { c.release_stream(s); }
    } rel_guard{ctx, s};
    CJ_HIP(hipSetDevice(ctx.device));
    // the kernels' LDS arrays hold 2^kBatchMaxLogPart slots, scanned kWG at a time
    if (cfg.log_part < 8 || cfg.log_part > kBatchMaxLogPart || cfg.chunk == 0 || cfg.chunk > kBatchMaxChunk ||
        cfg.grid > 65536)
        throw DeviceError("batch configuration out of range");
    // ---- shape of the batch
    // prefixes: build pages, probe pages, partitions; then build and probe
    // pages of the plans scattered through LDS staging / directly
    std::vector<uint32_t> starts(7 * (n + 1), 0);
    uint32_t *bps = starts.data(), *pps = bps + (n + 1), *pts = pps + (n + 1);
    uint32_t *bss = pts + (n + 1), *bsd = bss + (n + 1), *pss = bsd + (n + 1), *psd = pss + (n + 1);
    uint32_t staged_parts = 1, direct_parts = 1;
    std::vector<uint32_t> log_slots(n);
    uint64_t dir_total = 0, sum_nb = 0, sum_np = 0;
    uint32_t max_parts = 1, max_tuple_bytes = 4;
    std::vector<DevColumn *> dec_cols;  // carried columns decoded into dense arrays first
    std::map<DevColumn *, uint32_t> dec_index;
    // an INT32 column without NULLs whose pages hold the same row count is
    // read in place by the partition pass (no decode)
    auto in_place = [](const DevColumn *c) {
        return c->type == CJ_INT32 && c->dense_rows > 0 && c->num_rows < (1ull << 32);
    };
    // the source of carried column c of a side: its pages, or its decoded arrays
    auto set_src = [&](BSide &sd, size_t c, DevColumn *col, const BatchTemplate &t) {
        if (in_place(col)) {
            sd.src[c] = col->pages, sd.src_valid[c] = nullptr, sd.src_rows[c] = col->dense_rows;
        } else {
            const uint32_t j = dec_index.at(col);
            sd.src[c] = static_cast<const uint8_t *>(t.dec_values[j]), sd.src_valid[c] = t.dec_valid[j];
            sd.src_rows[c] = 0;
        }
    };
    uint64_t key = 1469598103934665603ull;
    auto mixk = [&](uint64_t v) { key = (key ^ v) * 1099511628211ull; };
    for (uint32_t k = 0; k < n; ++k) {
        const PlanInfo &pi = info[batched[k]];
        const uint64_t want = std::max<uint64_t>(1ull << cfg.log_part, (uint64_t)cfg.load_factor * pi.bk->num_rows);
        log_slots[k] = std::min<uint32_t>(ceil_log2(want), 30);
        const uint32_t parts = 1u << (log_slots[k] - cfg.log_part);
        max_parts = std::max(max_parts, parts);
        dir_total += 1ull << log_slots[k];
        sum_nb += pi.bk->num_rows;
        sum_np += pi.pk->num_rows;
        bps[k + 1] = bps[k] + pi.bk->num_pages;
        pps[k + 1] = pps[k] + pi.pk->num_pages;
        const bool staged = parts <= kStagedMaxParts;
        (staged ? staged_parts : direct_parts) = std::max(staged ? staged_parts : direct_parts, parts);
        bss[k + 1] = bss[k] + (staged ? pi.bk->num_pages : 0);
        bsd[k + 1] = bsd[k] + (staged ? 0 : pi.bk->num_pages);
        pss[k + 1] = pss[k] + (staged ? pi.pk->num_pages : 0);
        psd[k + 1] = psd[k] + (staged ? 0 : pi.pk->num_pages);
        pts[k + 1] = pts[k] + parts;
        uint32_t tb = 4;
        for (DevColumn *c : pi.bcar) tb += width_of(c->type) + (c->type == CJ_INT32 ? 0 : 1);
        max_tuple_bytes = std::max(max_tuple_bytes, tb);
        for (const auto *side : {&pi.bcar, &pi.pcar})
            for (DevColumn *c : *side)
                if (in_place(c)) {
                    mixk(c->dense_rows), mixk(c->num_pages);
                } else if (!dec_index.count(c)) {
                    dec_index[c] = (uint32_t)dec_cols.size();
                    dec_cols.push_back(c);
                }
        mixk(pi.bk->num_rows), mixk(pi.pk->num_rows), mixk(pi.bk->num_pages), mixk(pi.pk->num_pages);
        mixk(log_slots[k]), mixk(pi.bcar.size() | pi.pcar.size() << 8);
        for (DevColumn *c : pi.bcar) mixk(c->type);
        for (DevColumn *c : pi.pcar) mixk(c->type);
    }
    const uint32_t nbp = bps[n], npp = pps[n], nparts = pts[n];
    const uint64_t max_units64 = nparts + sum_np / cfg.chunk + 1;
    if (max_units64 > 0xffffffffull) throw UnsupportedError("batch too large");
    const uint32_t max_units = (uint32_t)max_units64;
    // A join workgroup stages a build partition of at most lds_cap tuples in
    // LDS (slot ends as u16, keys, carried values; a larger partition is
    // probed in place): twice the largest mean partition of the batch, so a
    // workgroup's LDS stays small and several workgroups share a CU (the join
    // passes are latency-bound).
    uint64_t max_mean = 1;
    for (uint32_t k = 0; k < n; ++k) {
        const uint64_t parts = 1ull << (log_slots[k] - cfg.log_part);
        max_mean = std::max<uint64_t>(max_mean, (info[batched[k]].bk->num_rows + parts - 1) / parts);
    }
    const uint32_t dyn_budget = kBatchLdsBytes - (2u << cfg.log_part) - 4 * 2048 - 512;  // slot ends, match window
    const uint32_t cap_budget = std::min<uint32_t>((dyn_budget / max_tuple_bytes) & ~15u, 65535u & ~15u);
    const uint32_t lds_cap =
        (uint32_t)std::min<uint64_t>(cap_budget, (std::max<uint64_t>(256, 2 * max_mean) + 15) & ~15ull);
    // a partition staged in LDS is also one the build orders by key
    if (lds_cap > kBatchSortCap) throw InvalidError("batch: join LDS capacity above the build's sort capacity");
    std::vector<BDecode> jobs(dec_cols.size());
    std::vector<uint32_t> job_start(dec_cols.size() + 1, 0);
    uint64_t valid_words = 0, val_bytes = 0;
    for (size_t j = 0; j < dec_cols.size(); ++j) {
        job_start[j + 1] = job_start[j] + dec_cols[j]->num_pages;
        valid_words += dec_cols[j]->num_rows / 32 + 2;
        val_bytes += ((dec_cols[j]->num_rows + 2) * width_of(dec_cols[j]->type) + 15) & ~15ull;
        // the page count is captured as the decode unit count of the graph:
        // equal row counts with another NULL density give other page counts
        mixk(dec_cols[j]->num_rows), mixk(dec_cols[j]->type), mixk(dec_cols[j]->num_pages);
    }
    mixk(n), mixk(cfg.load_factor), mixk(cfg.log_part), mixk(cfg.chunk), mixk(cfg.join_queue), mixk(cfg.grid), mixk(dec_cols.size());
    // ---- the template workspace (reused while the batch shape repeats)
    std::shared_ptr<BatchTemplate> tpl;
    {
        std::lock_guard<std::mutex> g(ctx.mu);
        auto it = ctx.batch_templates.find(key);
        if (it != ctx.batch_templates.end()) {
            tpl = it->second;
            ctx.batch_templates.erase(it);  // in use: not shared with a concurrent batch
        }
    }
    // a template taken from (or not yet in) the context's cache goes back to
    // it at the end; if anything throws before that, its blocks and graph
    // are released here
    struct TplGuard {
        std::shared_ptr<BatchTemplate> &t;
        DevicePool &pool;
        bool armed = true;
        ~TplGuard() {
            if (armed && t) t->release(pool);
        }
    } tpl_guard{tpl, pool};
    bool fresh = false;
    if (!tpl) {
        fresh = true;
        tpl = std::make_shared<BatchTemplate>();
        auto blk = [&](size_t bytes) {
            void *p = pool.alloc(bytes + 64);
            tpl->blocks.push_back(p);
            return static_cast<uint8_t *>(p);
        };
        BatchArgs &a = tpl->args;
        tpl->d_plans = reinterpret_cast<BPlan *>(blk(n * sizeof(BPlan)));
        tpl->d_starts = reinterpret_cast<uint32_t *>(blk(starts.size() * 4));
        a.bcount = reinterpret_cast<uint64_t *>(blk((nparts + 1) * 8));
        a.bcur = reinterpret_cast<uint64_t *>(blk((nparts + 1) * 8));
        a.boff = reinterpret_cast<uint64_t *>(blk((nparts + 1) * 8));
        a.bstage_keys = reinterpret_cast<int32_t *>(blk((sum_nb + 1) * 4));
        a.bkeys = reinterpret_cast<int32_t *>(blk((sum_nb + 1) * 4));
        a.dir = reinterpret_cast<uint32_t *>(blk(dir_total * 4));
        a.pcount = reinterpret_cast<uint64_t *>(blk((nparts + 1) * 8));
        a.pcur = reinterpret_cast<uint64_t *>(blk((nparts + 1) * 8));
        a.poff = reinterpret_cast<uint64_t *>(blk((nparts + 1) * 8));
        a.pkeys = reinterpret_cast<int32_t *>(blk((sum_np + 1) * 4));
        a.ucount = reinterpret_cast<uint64_t *>(blk((nparts + 1) * 8));
        a.uoff = reinterpret_cast<uint64_t *>(blk((nparts + 1) * 8));
        a.mcount = reinterpret_cast<uint64_t *>(blk(((uint64_t)max_units + 1) * 8));
        a.moff = reinterpret_cast<uint64_t *>(blk(((uint64_t)max_units + 1) * 8));
        tpl->totals = reinterpret_cast<uint64_t *>(blk(n * 8));
        tpl->scan_elems = scan_tmp_elems(std::max<uint64_t>(std::max<uint64_t>(nparts, max_units), 1));
        tpl->scan_tmp = reinterpret_cast<uint64_t *>(blk(tpl->scan_elems * 8));
        tpl->d_jobs = reinterpret_cast<BDecode *>(blk(std::max<size_t>(jobs.size(), 1) * sizeof(BDecode)));
        tpl->d_job_start = reinterpret_cast<uint32_t *>(blk(job_start.size() * 4));
        tpl->valid_all = reinterpret_cast<uint32_t *>(blk(std::max<uint64_t>(valid_words, 1) * 4));
        tpl->valid_words = valid_words;
        uint8_t *vals = blk(std::max<uint64_t>(val_bytes, 16));
        uint64_t vo = 0, wo = 0;
        for (size_t j = 0; j < dec_cols.size(); ++j) {
            tpl->dec_values.push_back(vals + vo);
            tpl->dec_valid.push_back(tpl->valid_all + wo);
            vo += ((dec_cols[j]->num_rows + 2) * width_of(dec_cols[j]->type) + 15) & ~15ull;
            wo += dec_cols[j]->num_rows / 32 + 2;
        }
        // per-plan carried arrays (plan-local tuple index, sized by the side's rows)
        tpl->hplans.resize(n);
        uint64_t dir_base = 0;
        for (uint32_t k = 0; k < n; ++k) {
            const PlanInfo &pi = info[batched[k]];
            BPlan &b = tpl->hplans[k];
            std::memset(&b, 0, sizeof b);
            b.log_slots = log_slots[k];
            b.dir_base = dir_base;
            dir_base += 1ull << log_slots[k];
            auto side = [&](BSide &sd, const DevColumn *kc, const std::vector<DevColumn *> &car, bool build) {
                sd.k_pages = kc->pages;
                sd.k_rs = kc->row_start;
                sd.ncar = (uint32_t)car.size();
                const uint64_t rows = kc->num_rows + 1;
                for (size_t c = 0; c < car.size(); ++c) {
                    const uint32_t w = width_of(car[c]->type);
                    sd.car_w[c] = w;
                    set_src(sd, c, car[c], *tpl);
                    sd.car[c] = blk(rows * w);
                    if (w == 8) sd.car_vld[c] = blk(rows);
                    if (build) {
                        b.bslot_car[c] = blk(rows * w);
                        if (w == 8) b.bslot_vld[c] = blk(rows);
                    }
                }
            };
            side(b.b, pi.bk, pi.bcar, true);
            side(b.p, pi.pk, pi.pcar, false);
        }
        a.plans = tpl->d_plans;
        a.nplans = n;
        a.bpage_start = tpl->d_starts;
        a.ppage_start = tpl->d_starts + (n + 1);
        a.part_start = tpl->d_starts + 2 * (n + 1);
        a.bpage_start_s = tpl->d_starts + 3 * (n + 1);
        a.bpage_start_d = tpl->d_starts + 4 * (n + 1);
        a.ppage_start_s = tpl->d_starts + 5 * (n + 1);
        a.ppage_start_d = tpl->d_starts + 6 * (n + 1);
        a.max_parts = max_parts;
        a.staged_parts = staged_parts;
        a.direct_parts = direct_parts;
        a.nparts = nparts;
        a.max_units = max_units;
        a.lds_cap = lds_cap;
        a.tuple_bytes = max_tuple_bytes;
        a.log_part = cfg.log_part;
        a.chunk = cfg.chunk;
        a.join_queue = cfg.join_queue;
        a.queue = reinterpret_cast<uint32_t *>(blk(kBatchQueues * 4));
        tpl->njobs = (uint32_t)jobs.size();
        tpl->ndpages = job_start.back();
        tpl->grids = batch_resident_grids(a, ctx.device);
        if (cfg.grid)
            tpl->grids =
                BatchGrids{cfg.grid, cfg.grid, cfg.grid, cfg.grid, cfg.grid, cfg.grid, cfg.grid, cfg.grid, cfg.grid, cfg.grid};
    }
    // a template whose captured decode units disagree with this batch's page
    // counts would decode past (or short of) the pages: never replay it
    if (tpl->ndpages != job_start.back() || tpl->njobs != (uint32_t)jobs.size())
        throw DeviceError("batch template does not match the batch's carried columns");
    for (size_t j = 0; j < dec_cols.size(); ++j)
        jobs[j] = BDecode{dec_cols[j]->pages, dec_cols[j]->row_start, (int32_t)width_of(dec_cols[j]->type),
                          tpl->dec_values[j], tpl->dec_valid[j]};
    // key pages of a reused template may belong to other tables of the same shape
    for (uint32_t k = 0; k < n; ++k) {
        const PlanInfo &pi = info[batched[k]];
        BPlan &b = tpl->hplans[k];
        b.b.k_pages = pi.bk->pages, b.b.k_rs = pi.bk->row_start;
        b.p.k_pages = pi.pk->pages, b.p.k_rs = pi.pk->row_start;
        for (size_t c = 0; c < pi.bcar.size(); ++c) set_src(b.b, c, pi.bcar[c], *tpl);
        for (size_t c = 0; c < pi.pcar.size(); ++c) set_src(b.p, c, pi.pcar[c], *tpl);
    }
    // ---- upload the descriptors (pinned staging; the copies finish before the graph runs)
    const size_t up_bytes =
        n * sizeof(BPlan) + starts.size() * 4 + jobs.size() * sizeof(BDecode) + job_start.size() * 4 + 64;
    uint8_t *stage = static_cast<uint8_t *>(pinned.alloc(up_bytes));
    struct PinGuard {
        PinnedPool &p;
        void *b;
        ~PinGuard() {
            if (b) p.free(b);
        }
    } pg{pinned, stage};
    size_t o = 0;
    auto up = [&](void *dst, const void *src, size_t bytes) {
        if (!bytes) return;
        std::memcpy(stage + o, src, bytes);
        CJ_HIP(hipMemcpyAsync(dst, stage + o, bytes, hipMemcpyHostToDevice, s));
        o += bytes;
    };
    up(tpl->d_plans, tpl->hplans.data(), n * sizeof(BPlan));
    up(tpl->d_starts, starts.data(), starts.size() * 4);
    up(tpl->d_jobs, jobs.data(), jobs.size() * sizeof(BDecode));
    up(tpl->d_job_start, job_start.data(), job_start.size() * 4);
    // ---- phase 1: decode, partition both sides, build, units, match counts (captured once)
    BatchTemplate &T = *tpl;
    BatchArgs &A = T.args;
    const BatchGrids &G = T.grids;
    auto phase1 = [&](hipStream_t ps) {
        fill_u64(A.bcount, 0, nparts + 1, ps);
        fill_u64(A.pcount, 0, nparts + 1, ps);
        fill_u64(A.mcount, 0, (uint64_t)max_units + 1, ps);
        fill_u32(T.valid_all, 0, T.valid_words, ps);
        fill_u32(A.queue, 0, kBatchQueues, ps);  // the join work queues (a fill kernel, not a memset)
        batch_decode(T.d_jobs, T.d_job_start, T.njobs, T.ndpages, G.decode, ps);
        batch_part_count(A, false, nbp, G.part_count, ps);
        scan_exclusive_u64(A.bcount, const_cast<uint64_t *>(A.boff), nparts, ps, T.scan_tmp, T.scan_elems);
        batch_copy_u64(A.boff, A.bcur, nparts + 1, ps);
        batch_part_scatter(A, false, A.bpage_start_s, bss[n], true, G.part_scatter, ps);
        batch_part_scatter(A, false, A.bpage_start_d, bsd[n], false, G.part_scatter_direct, ps);
        batch_part_count(A, true, npp, G.part_count, ps);
        scan_exclusive_u64(A.pcount, const_cast<uint64_t *>(A.poff), nparts, ps, T.scan_tmp, T.scan_elems);
        batch_copy_u64(A.poff, A.pcur, nparts + 1, ps);
        batch_part_scatter(A, true, A.ppage_start_s, pss[n], true, G.part_scatter, ps);
        batch_part_scatter(A, true, A.ppage_start_d, psd[n], false, G.part_scatter_direct, ps);
        batch_build(A, G.build, G.build_big, ps);
        batch_units(A, ps);
        scan_exclusive_u64(A.ucount, const_cast<uint64_t *>(A.uoff), nparts, ps, T.scan_tmp, T.scan_elems);
        batch_join_count(A, G.join_count, ps);
        scan_exclusive_u64(A.mcount, const_cast<uint64_t *>(A.moff), max_units, ps, T.scan_tmp, T.scan_elems);
        batch_totals(A, T.totals, ps);
        CJ_HIP(hipGetLastError());
    };
    // device time = the phase-1 graph + the phase-2 launches (host work
    // between them, the capture included, is not counted)
    hipEvent_t ev[4];
    for (hipEvent_t &e : ev) CJ_HIP(hipEventCreate(&e));
    struct EvGuard {
        hipEvent_t *e;
        ~EvGuard() {
            for (int i = 0; i < 4; ++i) hipEventDestroy(e[i]);
        }
    } eg{ev};
    if (fresh || !T.exec) {
        CJ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            phase1(s);
        } catch (...) {
            hipGraph_t g = nullptr;
            hipStreamEndCapture(s, &g);
            if (g) hipGraphDestroy(g);
            throw;
        }
        CJ_HIP(hipStreamEndCapture(s, &T.graph));
        CJ_HIP(hipGraphInstantiate(&T.exec, T.graph, nullptr, nullptr, 0));
    }
    CJ_HIP(hipEventRecord(ev[0], s));
    CJ_HIP(hipGraphLaunch(T.exec, s));
    CJ_HIP(hipEventRecord(ev[1], s));
    // ---- one read-back: the result size of every plan
    std::vector<uint64_t> totals(n);
    {
        uint64_t *h = static_cast<uint64_t *>(pinned.alloc(n * 8));  // a pageable copy is staged
        PinGuard pt{pinned, h};
        CJ_HIP(hipMemcpyAsync(h, T.totals, n * 8, hipMemcpyDeviceToHost, s));
        CJ_HIP(hipStreamSynchronize(s));
        std::copy(h, h + n, totals.begin());
    }
    uint64_t out_rows = 0;
    for (uint64_t t : totals) out_rows += t;
    // ---- phase 2: result pages (headers and bitmaps), values into their slots, pages with NULLs packed
    std::vector<void *> tmp_blocks;
    struct TmpGuard {
        DevicePool &p;
        std::vector<void *> &v;
        ~TmpGuard() {
            for (void *b : v) p.free(b);
        }
    } tg{pool, tmp_blocks};
    auto tblk = [&](size_t bytes) {
        void *p = pool.alloc(bytes + 64);
        tmp_blocks.push_back(p);
        return static_cast<uint8_t *>(p);
    };
    std::vector<BOut> bouts(n);
    std::vector<BPageCol> pcols;
    std::vector<uint32_t> page_start(1, 0);
    std::vector<std::vector<std::pair<uint8_t *, uint32_t>>> out_blocks(n);
    try {
        for (uint32_t k = 0; k < n; ++k) {
            const PlanInfo &pi = info[batched[k]];
            BOut &bo = bouts[k];
            std::memset(&bo, 0, sizeof bo);
            bo.nout = (uint32_t)pi.outs.size();
            if (totals[k] >= (1ull << 32)) throw UnsupportedError("a batched plan's result exceeds 2^32 rows");
            bo.rows = (uint32_t)totals[k];
            for (size_t c = 0; c < pi.outs.size(); ++c) {
                const auto &oc = pi.outs[c];
                const uint32_t w = width_of(oc.type);
                bo.src[c] = oc.src, bo.j[c] = oc.j, bo.w[c] = (uint8_t)w;
                const uint32_t R = w == 4 ? kRowsI32 : kRowsI64;
                const uint64_t np = (totals[k] + R - 1) / R;
                if (page_start.back() + np > 0xffffffffull) throw UnsupportedError("batch result too large");
                uint8_t *blk = np ? static_cast<uint8_t *>(pool.alloc((size_t)np * kPage + 16)) : nullptr;
                out_blocks[k].push_back({blk, (uint32_t)np});
                bo.out[c] = blk;
                bo.pg0[c] = page_start.back();
                if (!np) continue;
                pcols.push_back(BPageCol{blk, totals[k], (int32_t)w});
                page_start.push_back(page_start.back() + (uint32_t)np);
            }
        }
        const uint32_t npages = page_start.back();
        BOut *d_outs = reinterpret_cast<BOut *>(tblk(n * sizeof(BOut)));
        BPageCol *d_cols = reinterpret_cast<BPageCol *>(tblk(std::max<size_t>(pcols.size(), 1) * sizeof(BPageCol)));
        uint32_t *d_ps = reinterpret_cast<uint32_t *>(tblk(page_start.size() * 4));
        uint32_t *d_nulls = reinterpret_cast<uint32_t *>(tblk(std::max<size_t>(npages, 1) * 4));
        const size_t b2 = n * sizeof(BOut) + pcols.size() * sizeof(BPageCol) + page_start.size() * 4 + 64;
        uint8_t *stage2 = static_cast<uint8_t *>(pinned.alloc(b2));
        PinGuard pg2{pinned, stage2};
        size_t o2 = 0;
        auto up2 = [&](void *dst, const void *src, size_t bytes) {
            if (!bytes) return;
            std::memcpy(stage2 + o2, src, bytes);
            CJ_HIP(hipMemcpyAsync(dst, stage2 + o2, bytes, hipMemcpyHostToDevice, s));
            o2 += bytes;
        };
        up2(d_outs, bouts.data(), n * sizeof(BOut));
        up2(d_cols, pcols.data(), pcols.size() * sizeof(BPageCol));
        up2(d_ps, page_start.data(), page_start.size() * 4);
        CJ_HIP(hipEventRecord(ev[2], s));
        batch_page_init(d_cols, d_ps, (uint32_t)pcols.size(), npages, d_nulls, G.page_init, s);
        if (out_rows) batch_join_write(A, d_outs, d_nulls, G.join_write, s);
        batch_page_fix(d_cols, d_ps, (uint32_t)pcols.size(), npages, d_nulls, G.page_fix, s);
        CJ_HIP(hipEventRecord(ev[3], s));
        CJ_HIP(hipStreamSynchronize(s));
    } catch (...) {
        hipStreamSynchronize(s);
        for (auto &v : out_blocks)
            for (auto &b : v)
                if (b.first) pool.free(b.first);
        throw;
    }
    float ms1 = 0, ms2 = 0;
    CJ_HIP(hipEventElapsedTime(&ms1, ev[0], ev[1]));
    CJ_HIP(hipEventElapsedTime(&ms2, ev[2], ev[3]));
    const double dev_ms = (double)ms1 + ms2;
    // ---- results (the output blocks now belong to them)
    for (uint32_t k = 0; k < n; ++k) {
        const PlanInfo &pi = info[batched[k]];
        auto *impl = new ResultImpl();
        impl->cols.resize(pi.outs.size());
        impl->host_page_ptrs.resize(pi.outs.size());
        for (size_t c = 0; c < pi.outs.size(); ++c) {
            impl->cols[c].type = pi.outs[c].type;
            impl->cols[c].num_pages = out_blocks[k][c].second;
            impl->cols[c].device_pages = out_blocks[k][c].first;
            impl->dev_pages.push_back(out_blocks[k][c].first);
        }
        cj_result &r = out[batched[k]];
        r.num_rows = totals[k];
        r.num_columns = (uint32_t)impl->cols.size();
        r.columns = impl->cols.data();
        r.opaque = impl;
        if (!(flags & CJ_EXEC_OUTPUT_DEVICE)) {
            ctx.to_host(&r);
            for (size_t c = 0; c < impl->cols.size(); ++c) {
                if (impl->dev_pages[c]) pool.free(impl->dev_pages[c]);
                impl->dev_pages[c] = nullptr;
                impl->cols[c].device_pages = nullptr;
            }
        }
    }
    if (bs) {
        bs->grid = T.grids.join_write;
        bs->output_rows = out_rows;
        bs->graph_replayed = fresh ? 0 : 1;
    }
    {
        std::lock_guard<std::mutex> g(ctx.mu);
        auto &slot = ctx.batch_templates[key];
        if (slot) slot->release(pool);
        slot = tpl;
        tpl_guard.armed = false;
        // keep the workspace of a few recent shapes only
        while (ctx.batch_templates.size() > 4) {
            auto it = ctx.batch_templates.begin();
            if (it->first == key) ++it;
            it->second->release(pool);
            ctx.batch_templates.erase(it);
        }
    }
    return dev_ms;
}

void Context::release_batch_templates() {
    std::lock_guard<std::mutex> g(mu);
    for (auto &kv : batch_templates) kv.second->release(pool);
    batch_templates.clear();
}

}  // namespace cj
