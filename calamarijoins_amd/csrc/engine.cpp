// engine.cpp -- context, memory pools, resident-table registry and the plan
// executor of the MI355X join engine.
//
// The executor restates the reference's depth-first recursion
// (src/execute.cpp:84-300): every JoinNode resolves its children, builds on
// the side with fewer rows (ties go left, join_setup.h:108-136; the plan's
// build_left hint is not consulted), uses the nested loop when the build
// side has fewer than 8 rows (execute.cpp:236-238) and the radix-partitioned
// unchained hash table otherwise, and materializes pages at the root only.
// Intermediates are late-materialized: instead of the reference's 4-byte
// value_t columns (intermediate.h:40-172) a child result is a set of row-id
// lists, one per base scan that contributes columns, and values are read
// from the base pages only at the root.  The value_t semantics that leak
// into results are kept: keys read through an intermediate treat NULL and
// INT32_MIN alike (hash_join.h:120-127, nested_loop.h:85), a hash build over
// an intermediate inserts them as INT32_MIN (hashtable.h:357), and INT32
// output values equal to INT32_MIN come out NULL (page_builders.h:123).
#include "engine.hpp"

#include <atomic>
#include "fused.hpp"
#include "pagemat.hpp"
#include "filter.hpp"
#include "noalloc.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <set>
#include <thread>

namespace cj {

// ========================================================== device pool
static size_t pool_round(size_t n) {
    if (n <= (1u << 20)) return (n + 4095) & ~size_t(4095);
    const size_t g = size_t(2) << 20;
    return (n + g - 1) & ~(g - 1);
}

DevicePool::~DevicePool() {
    for (auto &kv : sizes_) hipFree(kv.first);
}

void DevicePool::trim() {
    for (auto &kv : free_) {
        sizes_.erase(kv.second);
        reserved_ -= kv.first;
        hipFree(kv.second);
    }
    free_.clear();
}

PoolMeter &pool_meter() {
    static thread_local PoolMeter m;
    return m;
}

static void meter_take(int64_t sz) {
    PoolMeter &m = pool_meter();
    if (!m.on) return;
    m.cur += sz;
    m.peak = std::max(m.peak, m.cur);
}

void *DevicePool::alloc(size_t bytes) {
    size_t sz = pool_round(bytes ? bytes : 1);
    std::lock_guard<std::mutex> g(mu_);
    auto it = free_.lower_bound(sz);
    if (it != free_.end() && it->first <= sz + sz / 2 + (size_t(4) << 20)) {
        void *p = it->second;
        in_use_ += it->first;
        peak_ = std::max(peak_, in_use_);
        meter_take((int64_t)it->first);
        free_.erase(it);
        return p;
    }
    if (budget_ && reserved_ + sz > budget_) trim();
    if (budget_ && reserved_ + sz > budget_)
        throw OomError("device budget exhausted: need " + std::to_string(sz) + " B, reserved " +
                       std::to_string(reserved_) + " of " + std::to_string(budget_));
    void *p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        trim();
        if (hipMalloc(&p, sz) != hipSuccess) {
            (void)hipGetLastError();
            throw OomError("hipMalloc of " + std::to_string(sz) + " bytes failed");
        }
    }
    sizes_[p] = sz;
    reserved_ += sz;
    in_use_ += sz;
    peak_ = std::max(peak_, in_use_);
    meter_take((int64_t)sz);
    return p;
}

void DevicePool::free(void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu_);
    auto it = sizes_.find(p);
    if (it == sizes_.end()) return;
    in_use_ -= it->second;
    meter_take(-(int64_t)it->second);
    free_.emplace(it->second, p);
}

PinnedPool::~PinnedPool() {
    for (auto &kv : sizes_) hipHostFree(kv.first);
}
void *PinnedPool::alloc(size_t bytes) {
    size_t sz = pool_round(bytes ? bytes : 1);
    std::lock_guard<std::mutex> g(mu_);
    auto it = free_.lower_bound(sz);
    if (it != free_.end() && it->first <= 2 * sz + (size_t(4) << 20)) {
        void *p = it->second;
        free_.erase(it);
        return p;
    }
    void *p = nullptr;
    if (hipHostMalloc(&p, sz, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        throw OomError("hipHostMalloc of " + std::to_string(sz) + " bytes failed");
    }
    sizes_[p] = sz;
    return p;
}
void PinnedPool::free(void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu_);
    auto it = sizes_.find(p);
    if (it != sizes_.end()) free_.emplace(it->second, p);
}

// ========================================================== kernel timer
void KernelTimer::begin(const char *name, uint64_t bytes) {
    if (!on) return;
    Rec r;
    r.name = name;
    r.bytes = bytes;
    CJ_HIP(hipEventCreate(&r.a));
    CJ_HIP(hipEventCreate(&r.b));
    CJ_HIP(hipEventRecord(r.a, s));
    recs.push_back(r);
}
void KernelTimer::end() {
    if (!on || recs.empty()) return;
    CJ_HIP(hipEventRecord(recs.back().b, s));
}
void KernelTimer::collect(std::map<std::string, KernelAgg> &acc) {
    for (auto &r : recs) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            auto &e = acc[r.name];
            e.launches += 1;
            e.ms += ms;
            e.bytes += r.bytes;
        }
        hipEventDestroy(r.a);
        hipEventDestroy(r.b);
    }
    recs.clear();
}

struct TimedScope {
    KernelTimer &t;
    TimedScope(KernelTimer &tt, const char *n, uint64_t bytes) : t(tt) { t.begin(n, bytes); }
    ~TimedScope() { t.end(); }
};

// ========================================================== phase clock
uint64_t next_column_uid() {
    static std::atomic<uint64_t> next{1};
    return next.fetch_add(1);
}

hipEvent_t Context::take_event() {
    {
        std::lock_guard<std::mutex> g(emu);
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
    }
    hipEvent_t e;
    CJ_HIP(hipEventCreate(&e));
    return e;
}
void Context::give_event(hipEvent_t e) {
    std::lock_guard<std::mutex> g(emu);
    event_pool.push_back(e);
}

void PhaseClock::mark() {
    hipEvent_t e = ctx->take_event();
    marks.emplace_back(e, stack.empty() ? -1 : stack.back());
    CJ_HIP(hipEventRecord(e, s));
}
void PhaseClock::push(int phase) {
    if (!on) return;
    stack.push_back(phase);
    mark();
}
void PhaseClock::pop() {
    if (!on || stack.empty()) return;
    stack.pop_back();
    mark();
}
void PhaseClock::collect(double *ms) {
    for (size_t i = 0; i + 1 < marks.size(); ++i) {
        if (marks[i].second < 0) continue;
        float t = 0;
        if (hipEventElapsedTime(&t, marks[i].first, marks[i + 1].first) == hipSuccess) ms[marks[i].second] += t;
    }
}
PhaseClock::~PhaseClock() {
    for (auto &m : marks) ctx->give_event(m.first);
}

struct PhaseScope {
    PhaseClock &c;
    PhaseScope(PhaseClock &cc, int phase) : c(cc) { c.push(phase); }
    ~PhaseScope() { c.pop(); }
};

// Algorithmic (compulsory) bytes of each launch, for the per-kernel
// roofline: what the kernel must read and write at minimum.
static uint64_t decode_bytes(const DevColumn &c) {
    uint64_t w = c.type == CJ_INT32 ? 4 : 8;
    return (uint64_t)c.num_pages * kPage + c.num_rows * w + c.num_rows / 8;
}
static uint64_t key_bytes(const KeySource &k) {
    if (k.mode == KEY_COLUMNAR) return (uint64_t)k.num_pages * kPage;
    return k.rows * 8 + k.rows / 8;  // row id + key value (+ validity)
}
static uint64_t fixed_mat_bytes(uint64_t rows, int32_t type) {
    uint64_t w = type == CJ_INT32 ? 4 : 8;
    return rows * (4 + w) + rows / 8 + rows * w + rows / 8;  // row ids + values read, page values + bitmap written
}

// ============================================================== context
Context::Context(const cj_ctx_options &o) : opts(o), pool(0) {
    device = o.device;
    CJ_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    CJ_HIP(hipGetDeviceProperties(&prop, device));
    dev_name = prop.name;
    if (dev_name.empty()) dev_name = prop.gcnArchName;
    else dev_name += std::string(" (") + prop.gcnArchName + ")";
    num_cus = prop.multiProcessorCount;
    max_lds = prop.sharedMemPerBlock;
    noalloc_init(device);  // this GPU's scratch of cj_execute_noalloc (held until the process ends)
    size_t free_b = 0, total_b = 0;
    CJ_HIP(hipMemGetInfo(&free_b, &total_b));
    pool.set_budget(o.device_budget ? o.device_budget : (uint64_t)(free_b * 0.92));
    int ns = o.num_streams > 0 ? o.num_streams : 4;
    streams.resize(ns);
    for (auto &st : streams) CJ_HIP(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
}

Context::~Context() {
    hipSetDevice(device);
    hipDeviceSynchronize();
    for (hipEvent_t e : event_pool) hipEventDestroy(e);
    release_batch_templates();
    for (auto &kv : tables)
        for (auto &c : kv.second->cols) free_column(*c);
    tables.clear();
    for (auto &st : streams) hipStreamDestroy(st.s);
}

hipStream_t Context::acquire_stream() {
    std::unique_lock<std::mutex> g(smu);
    for (;;) {
        for (auto &st : streams)
            if (!st.busy) {
                st.busy = true;
                return st.s;
            }
        scv.wait(g);
    }
}
void Context::release_stream(hipStream_t s) {
    {
        std::lock_guard<std::mutex> g(smu);
        for (auto &st : streams)
            if (st.s == s) st.busy = false;
    }
    scv.notify_one();
}

void Context::free_column(DevColumn &c) {
    pool.free(c.pages);
    pool.free(c.row_start);
    pool.free(c.values);
    pool.free(c.valid);
    pool.free(c.byte_prefix);  // byte_start lives in the row_start block
    c.pages = nullptr;
    c.row_start = nullptr;
    c.values = nullptr;
    c.valid = nullptr;
    c.byte_start = nullptr;
    c.byte_prefix = nullptr;
    c.loaded = false;
}

static uint32_t host_rows_of(const void *page) {
    uint16_t n;
    std::memcpy(&n, page, 2);
    return n == 0xffff ? 1u : (n == 0xfffe ? 0u : n);
}

// Structural check of one page header against the format (reference
// README.md:65-104) so that no kernel indexes outside its page: value /
// offset area plus bitmap must fit 8 KiB; long-string pages only in VARCHAR
// columns and 0xfffe only after 0xffff / 0xfffe.
static void validate_page(const uint8_t *p, int32_t type, uint32_t idx, uint16_t prev_rows) {
    uint16_t n, nv;
    std::memcpy(&n, p, 2);
    std::memcpy(&nv, p + 2, 2);
    if (n == 0xffff || n == 0xfffe) {
        if (type != CJ_VARCHAR) throw InvalidError("long-string page in a fixed-width column");
        if (nv > kPage - 4) throw InvalidError("long-string page holds more than 8188 chars");
        if (n == 0xfffe && (idx == 0 || (prev_rows != 0xffff && prev_rows != 0xfffe)))
            throw InvalidError("0xfffe page does not continue a long string");
        return;
    }
    if (nv > n) throw InvalidError("page has more values than rows");
    size_t bm = (n + 7u) / 8u;
    size_t need = type == CJ_VARCHAR ? 4 + 2 * (size_t)nv : (type == CJ_INT32 ? 4 + 4 * (size_t)nv : 8 + 8 * (size_t)nv);
    if (need + bm > kPage) throw InvalidError("page values and bitmap overflow 8 KiB");
    if (nv < n) {  // a page with NULLs: its values are the rows whose bit is set, so the two counts agree
        const uint8_t *b = p + kPage - bm;
        uint32_t ones = 0;
        for (size_t i = 0; i + 1 < bm; ++i) ones += (uint32_t)__builtin_popcount(b[i]);
        const uint32_t tail = n & 7u;
        ones += (uint32_t)__builtin_popcount(tail ? b[bm - 1] & ((1u << tail) - 1u) : b[bm - 1]);
        if (ones != nv) throw InvalidError("page bitmap marks a different number of values than its header");
    }
}

// Upload a column's pages (gathered through pinned staging, 32 MiB chunks,
// double-buffered) and its row index.
void Context::upload_column(DevColumn &c, const cj_column_desc &d, hipStream_t s,
                            const std::function<void(uint32_t p0, uint32_t np, uint8_t *dst)> &fill) {
    c.type = d.type;
    c.num_pages = d.num_pages;
    if (!fill)
        for (uint32_t i = 0; i < d.num_pages; ++i)
            if (!d.pages[i]) throw InvalidError("null page pointer");
    // host index: row_start (first row of each page) and byte_start (first
    // bitmap byte of each page in the page index), both length num_pages + 1,
    // built from the page headers as the pages pass through the staging buffer
    std::vector<uint64_t> rs(2 * (size_t)d.num_pages + 2);
    uint64_t *bs = rs.data() + d.num_pages + 1;
    uint64_t run = 0, bytes = 0;
    uint16_t prev = 0;
    // INT32 pages without NULLs, all of one row count but a shorter last
    // one, are read in place instead of decoded (dense_rows)
    bool dense = d.type == CJ_INT32 && d.num_pages > 0, dense_short = false;
    uint32_t dense_rows = 0;
    c.pages = static_cast<uint8_t *>(pool.alloc((size_t)d.num_pages * kPage + 16));
    c.row_start = dalloc<uint64_t>(2 * (size_t)d.num_pages + 2);
    c.byte_start = nullptr;  // lives inside the row_start block
    const uint32_t chunk = 4096;
    uint8_t *stage[2] = {nullptr, nullptr};
    hipEvent_t ev[2];
    size_t stage_bytes = (size_t)std::min<uint32_t>(chunk, std::max<uint32_t>(d.num_pages, 1)) * kPage;
    stage[0] = static_cast<uint8_t *>(pinned.alloc(std::max(stage_bytes, rs.size() * 8)));
    stage[1] = static_cast<uint8_t *>(pinned.alloc(stage_bytes));
    CJ_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    CJ_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    bool pending[2] = {false, false};
    struct Cleanup {
        Context &ctx;
        hipStream_t s;
        hipEvent_t *ev;
        uint8_t **stage;
        ~Cleanup() {
            hipStreamSynchronize(s);
            hipEventDestroy(ev[0]);
            hipEventDestroy(ev[1]);
            ctx.pinned.free(stage[0]);
            ctx.pinned.free(stage[1]);
        }
    } cleanup{*this, s, ev, stage};
    int nthreads = opts.host_threads > 0 ? opts.host_threads : (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    for (uint32_t p0 = 0, k = 0; p0 < d.num_pages; p0 += chunk, k ^= 1) {
        uint32_t np = std::min(chunk, d.num_pages - p0);
        if (pending[k]) CJ_HIP(hipEventSynchronize(ev[k]));
        uint8_t *dst = stage[k];
        if (fill) {
            fill(p0, np, dst);
        } else {
            auto copy_range = [&](uint32_t a, uint32_t b) {
                for (uint32_t i = a; i < b; ++i) std::memcpy(dst + (size_t)i * kPage, d.pages[p0 + i], kPage);
            };
            if (np >= 512 && nthreads > 1) {
                std::vector<std::thread> th;
                uint32_t per = (np + nthreads - 1) / nthreads;
                for (int t = 0; t < nthreads; ++t) {
                    uint32_t a = t * per, b = std::min(np, a + per);
                    if (a < b) th.emplace_back(copy_range, a, b);
                }
                for (auto &x : th) x.join();
            } else {
                copy_range(0, np);
            }
        }
        for (uint32_t i = 0; i < np; ++i) {
            const uint8_t *p = dst + (size_t)i * kPage;
            validate_page(p, d.type, p0 + i, prev);
            std::memcpy(&prev, p, 2);
            uint16_t nvals;
            std::memcpy(&nvals, p + 2, 2);
            if (prev != nvals || prev == 0 || prev >= 0xfffe || dense_short) dense = false;
            if (dense_rows == 0) dense_rows = prev;
            else if (prev > dense_rows) dense = false;
            else if (prev < dense_rows) dense_short = true;
            rs[p0 + i] = run;
            bs[p0 + i] = bytes;
            run += host_rows_of(p);
            if (prev < 0xfffe) {
                bytes += (prev + 7u) / 8u;
                c.max_page_rows = std::max<uint32_t>(c.max_page_rows, prev);
            }
        }
        CJ_HIP(hipMemcpyAsync(c.pages + (size_t)p0 * kPage, dst, (size_t)np * kPage, hipMemcpyHostToDevice, s));
        CJ_HIP(hipEventRecord(ev[k], s));
        pending[k] = true;
    }
    if (run > 0xffffffffull) throw UnsupportedError("column has more than 2^32 rows");
    rs[d.num_pages] = run;
    bs[d.num_pages] = bytes;
    c.num_rows = run;
    c.bitmap_bytes = bytes;
    c.dense_rows = dense ? dense_rows : 0;
    CJ_HIP(hipStreamSynchronize(s));
    std::memcpy(stage[0], rs.data(), rs.size() * 8);
    CJ_HIP(hipMemcpyAsync(c.row_start, stage[0], rs.size() * 8, hipMemcpyHostToDevice, s));
    CJ_HIP(hipStreamSynchronize(s));
    c.byte_start = c.row_start + d.num_pages + 1;
    c.loaded = true;
}

void Context::ensure_decoded(DevColumn &c, hipStream_t s) {
    // caller holds c.mu when c is shared
    if (c.values) return;
    size_t w = c.type == CJ_INT32 ? 4 : 8;
    c.values = pool.alloc(c.num_rows * w + 16);
    size_t words = c.num_rows / 32 + 2;
    c.valid = static_cast<uint32_t *>(pool.alloc(words * 4));
    fill_u32(c.valid, 0, words, s);
    if (c.type == CJ_VARCHAR) decode_varchar(c.pages, c.num_pages, c.row_start, (uint64_t *)c.values, c.valid, s);
    else decode_fixed(c.pages, c.num_pages, c.row_start, (int)w, c.values, c.valid, s);
}

uint64_t Context::register_table(const cj_table_desc &t, const PageFill &fill, std::shared_ptr<void> keepalive) {
    hipStream_t s = acquire_stream();
    auto tab = std::make_unique<DevTable>();
    tab->keepalive = std::move(keepalive);
    try {
        tab->resident = true;
        tab->num_rows = t.num_rows;
        for (uint32_t i = 0; i < t.num_columns; ++i) {
            auto c = std::make_unique<DevColumn>();
            tab->cols.push_back(std::move(c));
            if (fill)
                upload_column(*tab->cols.back(), t.columns[i], s,
                              [&](uint32_t p0, uint32_t np, uint8_t *dst) { fill(i, p0, np, dst); });
            else
                upload_column(*tab->cols.back(), t.columns[i], s);
            tab->host_pages.emplace_back(t.columns[i].pages, t.columns[i].pages + t.columns[i].num_pages);
        }
    } catch (...) {
        for (auto &c : tab->cols) free_column(*c);
        release_stream(s);
        throw;
    }
    release_stream(s);
    std::lock_guard<std::mutex> g(mu);
    uint64_t id = next_id++;
    tab->id = id;
    for (uint32_t i = 0; i < t.num_columns; ++i)
        if (t.columns[i].num_pages)
            by_first_page[{t.columns[i].pages[0], t.columns[i].num_pages}] = {tab.get(), i};
    tables[id] = std::move(tab);
    return id;
}

void Context::unregister_table(uint64_t id) {
    std::lock_guard<std::mutex> g(mu);
    auto it = tables.find(id);
    if (it == tables.end()) throw InvalidError("unknown table id");
    for (auto jt = by_first_page.begin(); jt != by_first_page.end();)
        jt = jt->second.first == it->second.get() ? by_first_page.erase(jt) : std::next(jt);
    CJ_HIP(hipDeviceSynchronize());
    for (auto &c : it->second->cols) free_column(*c);
    tables.erase(it);
}

DevColumn *Context::find_resident(const cj_column_desc &c) {
    if (!c.num_pages) return nullptr;
    std::lock_guard<std::mutex> g(mu);
    auto it = by_first_page.find({c.pages[0], c.num_pages});
    if (it == by_first_page.end()) return nullptr;
    DevTable *t = it->second.first;
    const auto &hp = t->host_pages[it->second.second];
    if (hp.back() != c.pages[c.num_pages - 1]) return nullptr;
    DevColumn *col = t->cols[it->second.second].get();
    return col->type == c.type ? col : nullptr;
}

void Context::kernel_stats(std::vector<cj_kernel_stat> &out) {
    std::lock_guard<std::mutex> g(mu);
    for (auto &kv : kstats) {
        cj_kernel_stat s{};
        std::snprintf(s.name, sizeof s.name, "%s", kv.first.c_str());
        s.launches = kv.second.launches;
        s.total_ms = kv.second.ms;
        s.bytes = kv.second.bytes;
        out.push_back(s);
    }
}
void Context::kernel_stats_reset() {
    std::lock_guard<std::mutex> g(mu);
    kstats.clear();
}

// ============================================================== executor
namespace {

struct Slot {
    uint32_t table;          // plan input index
    const uint32_t *rids;    // row ids into the base table (nullptr = identity)
    // non-null: rids holds slots of a direct map of row ids, row id = map[slot]
    // (resolved only where a row id is read: keys, output columns)
    const uint32_t *map = nullptr;
};
struct ColRef {
    uint32_t slot;
    uint32_t col;            // base column index
    int32_t type;
};
struct Rel {
    bool columnar = false;   // a ScanNode (single identity slot)
    uint64_t rows = 0;
    std::vector<Slot> slots;
    std::vector<ColRef> cols;
};

struct ColView {
    DevColumn *col = nullptr;
    const void *values = nullptr;
    const uint32_t *valid = nullptr;
};

class Exec {
  public:
    Exec(Context &c, const cj_plan_desc &p, hipStream_t s) : ctx(c), plan(p), st(s) {
        timer.on = c.opts.kernel_timing != 0;
        timer.s = s;
        clock.s = s;
        clock.ctx = &c;
        cols.resize(p.num_inputs);
        for (uint32_t i = 0; i < p.num_inputs; ++i) cols[i].resize(p.inputs[i].num_columns, nullptr);
    }
    ~Exec() {
        for (void *p : temps) ctx.pool.free(p);  // (none in no-allocation mode)
        for (auto &c : own_cols) ctx.free_column(*c);
        if (rb) ctx.pinned.free(rb);
        if (rb_event) (void)hipEventDestroy(rb_event);
    }

    Context &ctx;
    const cj_plan_desc &plan;
    hipStream_t st;
    KernelTimer timer;
    PhaseClock clock;  // on when the caller asked for timing
    std::vector<void *> temps;
    std::vector<std::unique_ptr<DevColumn>> own_cols;
    std::vector<std::vector<DevColumn *>> cols;
    std::map<DevColumn *, ColView> views;
    std::map<DevColumn *, const uint16_t *> page_indexes;
    std::set<DevColumn *> referenced;
    uint64_t scratch_elems = 0;
    uint64_t *scratch = nullptr;
    double h2d_ms = 0;

    // No-allocation mode (cj_execute_noalloc): temporaries are carved from
    // the GPU's fixed scratch and result pages go to the caller's buffers.
    uint8_t *arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    void *const *user_out = nullptr;
    const uint32_t *user_cap = nullptr;
    uint32_t user_n = 0;
    bool noalloc() const { return arena != nullptr; }

    template <class T>
    T *tmp(size_t n) {
        if (arena) {
            const size_t bytes = (n * sizeof(T) + 16 + 255) & ~size_t(255);
            if (arena_used + bytes > arena_cap)
                throw OomError("no-allocation scratch exhausted: the plan needs more than " +
                               std::to_string(arena_cap) + " bytes of temporaries (" +
                               std::to_string(arena_used + bytes) + " so far)");
            T *p = reinterpret_cast<T *>(arena + arena_used);
            arena_used += bytes;
            return p;
        }
        T *p = ctx.dalloc<T>(n);
        temps.push_back(p);
        return p;
    }
    // Room for `need` pages of result column `attr`: a pool block, or the
    // caller's buffer in no-allocation mode.
    uint8_t *out_pages(uint32_t attr, uint64_t need) {
        if (!noalloc()) return static_cast<uint8_t *>(ctx.pool.alloc((size_t)need * kPage + 16));
        if (attr >= user_n) throw InvalidError("cj_execute_noalloc: fewer output buffers than result columns");
        if (need > user_cap[attr])
            throw OomError("cj_execute_noalloc: result column " + std::to_string(attr) + " needs room for " +
                           std::to_string(need) + " pages, its buffer holds " + std::to_string(user_cap[attr]));
        return static_cast<uint8_t *>(user_out[attr]);
    }

    DevColumn *column(uint32_t table, uint32_t c) {
        if (cols[table][c]) return cols[table][c];
        const cj_column_desc &d = plan.inputs[table].columns[c];
        DevColumn *dc = ctx.find_resident(d);
        if (!dc && noalloc())
            throw UnsupportedError("cj_execute_noalloc needs every input registered as a resident table");
        if (!dc) {
            auto t0 = std::chrono::steady_clock::now();
            own_cols.push_back(std::make_unique<DevColumn>());  // owned first: freed if the upload throws
            dc = own_cols.back().get();
            ctx.upload_column(*dc, d, st);
            h2d_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        // every column of a table must describe the same rows: row ids of one
        // column index all the others
        for (DevColumn *o : cols[table])
            if (o && o->num_rows != dc->num_rows)
                throw InvalidError("columns of input table " + std::to_string(table) + " hold different row counts");
        cols[table][c] = dc;
        referenced.insert(dc);
        return dc;
    }

    // Dense decode of a base column, cached in the context for resident
    // tables when cache_decoded is set, otherwise private to this call.
    ColView view(DevColumn *c) {
        auto it = views.find(c);
        if (it != views.end()) return it->second;
        PhaseScope ps(clock, CJ_PHASE_SETUP);
        ColView v;
        v.col = c;
        bool shared = false;
        for (auto &o : own_cols) shared |= o.get() == c;
        shared = !shared;
        bool cached = false;
        if (shared && noalloc()) {  // a decode cached earlier serves; none is made here
            std::lock_guard<std::mutex> g(c->mu);
            cached = c->values != nullptr;
        }
        if (cached || (shared && ctx.opts.cache_decoded && !noalloc())) {
            std::lock_guard<std::mutex> g(c->mu);
            if (!c->values) {
                TimedScope ts(timer, c->type == CJ_VARCHAR ? "decode_varchar" : "decode_fixed", decode_bytes(*c));
                ctx.ensure_decoded(*c, st);
                CJ_HIP(hipStreamSynchronize(st));
            }
            v.values = c->values;
            v.valid = c->valid;
        } else if (!shared) {
            TimedScope ts(timer, c->type == CJ_VARCHAR ? "decode_varchar" : "decode_fixed", decode_bytes(*c));
            ctx.ensure_decoded(*c, st);
            v.values = c->values;
            v.valid = c->valid;
        } else {
            size_t w = c->type == CJ_INT32 ? 4 : 8;
            void *vals = tmp<uint8_t>(c->num_rows * w + 16);
            size_t words = c->num_rows / 32 + 2;
            uint32_t *valid = tmp<uint32_t>(words);
            fill_u32(valid, 0, words, st);
            TimedScope ts(timer, c->type == CJ_VARCHAR ? "decode_varchar" : "decode_fixed", decode_bytes(*c));
            if (c->type == CJ_VARCHAR) decode_varchar(c->pages, c->num_pages, c->row_start, (uint64_t *)vals, valid, st);
            else decode_fixed(c->pages, c->num_pages, c->row_start, (int)w, vals, valid, st);
            v.values = vals;
            v.valid = valid;
        }
        views[c] = v;
        return v;
    }

    uint64_t *scan_scratch(uint64_t n) {
        uint64_t need = scan_tmp_elems(n);
        if (need > scratch_elems) {
            scratch = tmp<uint64_t>(need);
            scratch_elems = need;
        }
        return scratch;
    }
    void scan(const uint64_t *in, uint64_t *out, uint64_t n) {
        TimedScope ts(timer, "scan", 16 * n);
        uint64_t *sc = scan_scratch(n);
        scan_exclusive_u64(in, out, n, st, sc, scratch_elems);
    }
    // Pinned host words for counter read-backs (a copy into pageable memory
    // is staged and takes tens of microseconds).
    static constexpr size_t kReadbackWords = 256;
    uint64_t *rb = nullptr;
    hipEvent_t rb_event = nullptr;  // a read-back's copy is done (the stream may go on)
    uint64_t *readback() {
        if (noalloc()) {  // pageable words (the copy is staged by the runtime)
            thread_local uint64_t words[kReadbackWords];
            return words;
        }
        if (!rb) rb = static_cast<uint64_t *>(ctx.pinned.alloc(kReadbackWords * 8));
        return rb;
    }
    // First output page of each key page for 4- and 8-byte build columns
    // (exclusive scans of ceil(matches / page capacity)), for counts `cnt`.
    const uint64_t *po_cnt = nullptr;
    uint64_t *po4 = nullptr, *po8 = nullptr;
    void page_offsets(const uint64_t *cnt, uint32_t kp) {
        uint64_t *p4 = tmp<uint64_t>(kp + 1), *p8 = tmp<uint64_t>(kp + 1);
        po4 = tmp<uint64_t>(kp + 1);
        po8 = tmp<uint64_t>(kp + 1);
        key_page_out_pages(cnt, kp, p4, p8, st);
        scan(p4, po4, kp);
        scan(p8, po8, kp);
        po_cnt = cnt;
    }
    uint64_t read_u64(const uint64_t *dev) {
        uint64_t *h = readback();
        CJ_HIP(hipMemcpyAsync(h, dev, 8, hipMemcpyDeviceToHost, st));
        CJ_HIP(hipStreamSynchronize(st));
        return h[0];
    }
    // Several device counters with one read-back.
    std::vector<uint64_t> read_u64s(const std::vector<const uint64_t *> &ptrs) {
        std::vector<uint64_t> v(ptrs.size());
        uint64_t *h = ptrs.size() <= kReadbackWords ? readback() : v.data();
        for (size_t i = 0; i < ptrs.size(); i += 16) {
            U64List l{};
            l.n = (uint32_t)std::min<size_t>(16, ptrs.size() - i);
            for (uint32_t k = 0; k < l.n; ++k) l.p[k] = ptrs[i + k];
            uint64_t *d = tmp<uint64_t>(16);
            gather_u64s(l, d, st);
            CJ_HIP(hipMemcpyAsync(h + i, d, 8 * l.n, hipMemcpyDeviceToHost, st));
        }
        CJ_HIP(hipStreamSynchronize(st));
        if (h != v.data()) std::copy(h, h + ptrs.size(), v.begin());
        return v;
    }

    Rel scan_rel(const cj_node_desc &n) {
        if (n.base_table_id >= plan.num_inputs) throw InvalidError("scan base_table_id out of range");
        const cj_table_desc &t = plan.inputs[n.base_table_id];
        Rel r;
        r.columnar = true;
        r.rows = t.num_rows;
        r.slots.push_back(Slot{(uint32_t)n.base_table_id, nullptr});
        for (uint32_t a = 0; a < n.num_output_attrs; ++a) {
            const cj_attr &at = n.output_attrs[a];
            if (at.index >= t.num_columns) throw InvalidError("scan output attr out of range");
            if (t.columns[at.index].type != at.type) throw UnsupportedError("scan output attr type mismatch");
            r.cols.push_back(ColRef{0, (uint32_t)at.index, at.type});
        }
        return r;
    }

    Rel resolve(uint64_t idx, int depth) {
        if (idx >= plan.num_nodes) throw InvalidError("node index out of range");
        if (depth > 256) throw InvalidError("plan too deep (cycle?)");
        const cj_node_desc &n = plan.nodes[idx];
        if (n.kind == CJ_NODE_SCAN) return scan_rel(n);
        if (n.kind != CJ_NODE_JOIN) throw InvalidError("unknown node kind");
        return join(n, false, depth, nullptr);
    }

    KeySource key_source(const Rel &r, uint32_t attr) {
        const ColRef &cr = r.cols[attr];
        DevColumn *c = column(r.slots[cr.slot].table, cr.col);
        KeySource k{};
        if (r.columnar) {
            k.mode = KEY_COLUMNAR;
            k.pages = c->pages;
            k.num_pages = c->num_pages;
            k.row_start = c->row_start;
        } else {
            k.mode = KEY_INTER;
            if (c->dense_rows) {  // read in place: no decode
                k.dense_pages = c->pages;
                k.dense_rows = c->dense_rows;
            } else {
                ColView v = view(c);
                k.values = static_cast<const int32_t *>(v.values);
                k.valid_bits = v.valid;
            }
            k.rids = resolved_rids(r.slots[cr.slot], r.rows);
            k.rows = r.rows;
        }
        return k;
    }

    // Hash build: partition pass (histogram, scan, scatter) then one
    // workgroup per partition.
    HashTableDev build_hash(const KeySource &k, uint64_t rows, uint64_t cap_rows) {
        PhaseScope ps(clock, CJ_PHASE_HASHTABLE_BUILD);
        int log_slots = 11;  // directory >= 2048 slots (hashtable.h:255-264)
        while ((1ull << log_slots) < rows) ++log_slots;
        if (log_slots > 32) throw UnsupportedError("hash build side above 2^32 rows");
        // partitions of 2^11 .. 2^14 slots, about 256 of them for parallelism
        // and at most 4096 (the partition pass's LDS histogram): beyond 2^26
        // slots a partition holds more slots than one LDS pass counts, and
        // its workgroup counts them 2^14 (128 KiB of LDS) at a time
        int log_part = std::min(std::max(log_slots - 8, 11), 14);
        if (log_part > log_slots) log_part = log_slots;
        if (log_slots - log_part > 12) log_part = log_slots - 12;
        int log_sub = std::min(log_part, 14);
        while (log_sub > 10 && (8ull << log_sub) + 256 > ctx.max_lds) --log_sub;
        HashGeom g{log_slots, log_part, log_sub, 1u << (log_slots - log_part)};
        uint32_t P = g.num_partitions;
        uint64_t *pc = tmp<uint64_t>(P + 1);
        uint64_t *poff = tmp<uint64_t>(P + 1);
        uint64_t *cur = tmp<uint64_t>(P + 1);
        fill_u64(pc, 0, P + 1, st);
        {
            TimedScope ts(timer, "part_count", key_bytes(k));
            build_partition_hist(k, g, pc, st);
        }
        scan(pc, poff, P);
        CJ_HIP(hipMemcpyAsync(cur, poff, (P + 1) * 8, hipMemcpyDeviceToDevice, st));
        uint64_t cap = std::max<uint64_t>(std::max(rows, cap_rows), 1);
        int32_t *tk = tmp<int32_t>(cap);
        uint32_t *tr = tmp<uint32_t>(cap);
        {
            TimedScope ts(timer, "part_scatter", key_bytes(k) + 8 * cap);
            build_partition_scatter(k, g, cur, tk, tr, st);
        }
        uint64_t *dir = tmp<uint64_t>(1ull << log_slots);
        int32_t *keys = tmp<int32_t>(cap);
        uint32_t *rids = tmp<uint32_t>(cap);
        {
            TimedScope ts(timer, "build_partitions", 16 * cap + (8ull << log_slots));
            build_partitions(g, poff, tk, tr, dir, keys, rids, st);
        }
        return HashTableDev{dir, keys, rids, log_slots, rows};
    }

    Rel join(const cj_node_desc &n, bool is_root, int depth, cj_result *out);
    bool fused_ok(const cj_node_desc &n, const Rel &B, const Rel &P, uint32_t bw, const std::vector<uint32_t> &remap);
    void fused_root(const cj_node_desc &n, const Rel &B, const Rel &P, uint32_t bw, const std::vector<uint32_t> &remap,
                    const ProbeArgs &a, const uint64_t *off, uint64_t total);
    void aligned_root(const cj_node_desc &n, const Rel &B, const Rel &P, uint32_t bw,
                      const std::vector<uint32_t> &remap, const ProbeArgs &a, const uint64_t *cnt,
                      const uint32_t *match_bits, uint64_t total);
    bool try_direct(const cj_node_desc &n, const Rel &B, uint32_t battr, uint32_t bw,
                    const std::vector<uint32_t> &remap, DirectTable *dt, bool rids);
    const uint16_t *page_index(DevColumn *c);
    uint32_t flags = 0;
    std::vector<std::pair<uint32_t, uint32_t>> direct_field;  // per output attr: record offset, flag bit
    // Row ids of a slot of `rows` rows (map slots read through the map).
    const uint32_t *resolved_rids(const Slot &s, uint64_t rows) {
        if (!s.map || !rows) return s.rids;
        uint32_t *g = tmp<uint32_t>(rows);
        TimedScope ts(timer, "gather_rids", 12 * rows);
        gather_u32(s.map, s.rids, g, rows, st);
        return g;
    }
    void add_gather(GatherSet &g, const uint32_t *old, uint32_t *dst, const uint32_t *idx, uint64_t total);
    void flush_gather(GatherSet &g, const uint32_t *idx, uint64_t total);
    void materialize_root(const cj_node_desc &n, const Rel &B, const Rel &P, uint32_t bw,
                          const std::vector<uint32_t> &remap, uint64_t total, const uint32_t *ob, const uint32_t *op,
                          cj_result *out);
    // pages of base column c gathered by row ids (rows of them); int32_quirk:
    // INT32_MIN values are written as NULL (result page builders only)
    // (attr: the result column the pages belong to)
    std::pair<uint8_t *, uint32_t> materialize_column(DevColumn *c, int32_t type, const uint32_t *rids, uint64_t rows,
                                                      bool int32_quirk, uint32_t attr);
    // filter of input table 0: the rows passing the program, every column
    void filter_root(const std::vector<FNode> &prog, const std::vector<uint32_t> &used,
                     const std::vector<uint8_t> &lits);

    // output placement
    bool device_output = false;
    uint64_t bytes_out = 0;
    std::vector<std::pair<uint8_t *, std::pair<uint32_t, int32_t>>> out_cols;  // pages, (num_pages, type)
    uint64_t out_rows = 0;
    std::vector<int32_t> out_types;
};

Rel Exec::join(const cj_node_desc &n, bool is_root, int depth, cj_result *out) {
    Rel L = resolve(n.left, depth + 1);
    Rel R = resolve(n.right, depth + 1);
    if (n.left_attr >= L.cols.size() || n.right_attr >= R.cols.size()) throw InvalidError("join attr out of range");
    if (L.cols[n.left_attr].type != CJ_INT32 || R.cols[n.right_attr].type != CJ_INT32)
        throw UnsupportedError("join keys must be INT32");
    const bool build_left = L.rows <= R.rows;
    Rel &B = build_left ? L : R;
    Rel &P = build_left ? R : L;
    const uint32_t battr = (uint32_t)(build_left ? n.left_attr : n.right_attr);
    const uint32_t pattr = (uint32_t)(build_left ? n.right_attr : n.left_attr);
    const uint32_t lw = (uint32_t)L.cols.size(), bw = (uint32_t)B.cols.size();
    std::vector<uint32_t> remap(n.num_output_attrs);
    bool need_b = false, need_p = false;
    for (uint32_t a = 0; a < n.num_output_attrs; ++a) {
        uint64_t ci = n.output_attrs[a].index;
        if (ci >= (uint64_t)lw + R.cols.size()) throw InvalidError("join output attr out of range");
        remap[a] = build_left ? (uint32_t)ci : (ci < lw ? bw + (uint32_t)ci : (uint32_t)(ci - lw));
        (remap[a] < bw ? need_b : need_p) = true;
    }

    KeySource bk = key_source(B, battr);
    KeySource pk = key_source(P, pattr);
    const bool use_nl = B.rows < 8;
    HashTableDev ht{};
    const uint32_t *lazy_map = nullptr;  // the build side's row ids, read through map slots
    NLKeys *nl = nullptr;
    uint64_t total = 0;
    uint32_t *ob = nullptr, *op = nullptr;
    const bool empty = B.rows == 0 || P.rows == 0;
    if (!empty) {
        const bool fused = is_root && fused_ok(n, B, P, bw, remap);
        ProbeArgs fa{};
        fa.mode = kModeHash;
        if (use_nl) {
            PhaseScope ps(clock, CJ_PHASE_NESTED_LOOP_JOIN);
            nl = tmp<NLKeys>(1);
            TimedScope ts(timer, "nl_collect", 0);
            nl_collect(bk, nl, st);
            fa.mode = kModeNL;
        } else if (B.columnar && ((fused && try_direct(n, B, battr, bw, remap, &fa.dt, false)) ||
                                  (!fused && try_direct(n, B, battr, bw, remap, &fa.dt, true)))) {
            // a direct map: with the projected build payloads for the fused
            // root writers, or with the build row ids for pair lists
            fa.mode = kModeDirect;
            if (!fused && !is_root) {
                // an intermediate keeps the map slots; row ids are read
                // through the map only for the rows that reach a key or an
                // output column
                lazy_map = reinterpret_cast<const uint32_t *>(fa.dt.records);
                fa.dt.rid_offset = -1;
            }
        } else {
            // tuple capacity: the rows the key pages actually hold (the
            // declared num_rows only sizes the directory)
            uint64_t cap = B.rows;
            if (B.columnar) cap = column(B.slots[0].table, B.cols[battr].col)->num_rows;
            ht = build_hash(bk, B.rows, cap);
        }
        fa.key = pk;
        fa.ht = ht;
        fa.nl = nl;
        // counting pass .. pair lists: the probe phase (the root writers
        // below open the materialize phase)
        PhaseScope probe_phase(clock, use_nl ? CJ_PHASE_NESTED_LOOP_JOIN : CJ_PHASE_HASH_JOIN_PROBE);
        uint32_t units = key_source_units(pk);
        uint64_t *cnt = tmp<uint64_t>(units + 1);
        uint64_t *off = tmp<uint64_t>(units + 1);
        // page-aligned root output (pagemat.hpp): every probe row matches at
        // most once, so the count pass also records which rows matched
        // (with a direct map every probe row matches at most once: the match
        // bits also let the pair writer skip a second lookup)
        const bool bits = fa.mode == kModeDirect && pk.mode == KEY_COLUMNAR && !(flags & CJ_EXEC_NO_ALIGNED);
        // a row-id list probing a direct map: the count pass records match
        // bits so the pair writer reads the keys of the matched rows only
        const bool inter_bits = fa.mode == kModeDirect && pk.mode == KEY_INTER && !fused &&
                                !(flags & CJ_EXEC_NO_ALIGNED);
        // (not in no-allocation mode: the page-aligned writers' page bound is
        // only known after the call, so caller buffers could not be sized)
        const bool aligned = fused && bits && !noalloc();
        uint32_t *match_bits = nullptr;
        const uint64_t probe_rows = pk.mode == KEY_COLUMNAR ? column(P.slots[0].table, P.cols[pattr].col)->num_rows
                                                            : pk.mode == KEY_INTER ? pk.rows : 0;
        if (bits || inter_bits) {
            const uint64_t words = probe_rows / 32 + 2;
            match_bits = tmp<uint32_t>(words);
            // the row-id list's count pass stores every word its rows cover
            if (bits) fill_u32(match_bits, 0, words, st);
        }
        {
            TimedScope ts(timer, use_nl ? "nl_probe_count" : "probe_count", key_bytes(pk));
            if (bits) probe_count_match(fa, cnt, match_bits, column(P.slots[0].table, P.cols[pattr].col)->dense_rows, st);
            else if (inter_bits) inter_count_match(fa, cnt, match_bits, st);
            else probe_count_columnar(fa, cnt, st);
        }
        if (aligned && units <= kMultiScanMax) {
            // the page-aligned writers' output page offsets depend only on the
            // counts: one launch scans the counts and the build columns' page
            // counts (pages of kCap4 / kCap8 rows per key page)
            po4 = tmp<uint64_t>(units + 1);
            po8 = tmp<uint64_t>(units + 1);
            const ScanJob jobs[3] = {{cnt, off, units, 0}, {cnt, po4, units, kCap4}, {cnt, po8, units, kCap8}};
            {
                TimedScope ts(timer, "scan3", 32ull * units);
                multi_scan_u64(jobs, 3, tmp<uint64_t>(multi_scan_blocks(jobs, 3) + 1), st);
            }
            po_cnt = cnt;
            // both key columns are base-table columns and every probe row
            // matches at most once: the count is a property of the two
            // columns, read back once and kept on the probe key column
            DevColumn *pkc = column(P.slots[0].table, P.cols[pattr].col);
            const uint64_t buid = column(B.slots[0].table, B.cols[battr].col)->uid;
            bool known = false;
            {
                std::lock_guard<std::mutex> g(pkc->mu);
                for (auto &mt : pkc->match_totals)
                    if (mt.first == buid) total = mt.second, known = true;
            }
            // (trusted only where it leads to the page-aligned writers, whose
            // writes are bounded by their blocks; other strategies read the
            // count back)
            if (known && !(total && 2 * total >= probe_rows)) known = false;
            if (!known) {
                total = read_u64(off + units);
                std::lock_guard<std::mutex> g(pkc->mu);
                bool have = false;
                for (auto &mt : pkc->match_totals) have |= mt.first == buid;
                if (!have) pkc->match_totals.emplace_back(buid, total);
            }
        } else if (aligned) {
            scan(cnt, off, units);
            uint64_t *h = readback();
            CJ_HIP(hipMemcpyAsync(h, off + units, 8, hipMemcpyDeviceToHost, st));
            CJ_HIP(hipStreamSynchronize(st));
            page_offsets(cnt, units);
            total = h[0];
        } else {
            scan(cnt, off, units);
            total = read_u64(off + units);
        }
        if (total > 0xffffffffull) throw UnsupportedError("join result exceeds 2^32 rows");
        // dense matching (at least half the probe rows) keeps the filtered
        // pages at least half full on average; sparser results take the
        // block writers, which fill every page
        if (total && aligned && 2 * total >= probe_rows) {
            aligned_root(n, B, P, bw, remap, fa, cnt, match_bits, total);
            return Rel{};
        }
        if (total && fused) {
            fused_root(n, B, P, bw, remap, fa, off, total);
            return Rel{};
        }
        if (total) {
            if (need_b) ob = tmp<uint32_t>(total);
            if (need_p) op = tmp<uint32_t>(total);
            TimedScope ts(timer, use_nl ? "nl_probe_write" : "probe_write", key_bytes(pk) + total * (4ull * (need_b + need_p)));
            if (bits) direct_pairs(fa, match_bits, off, ob, op, st);
            else if (inter_bits) inter_pairs(fa, match_bits, off, ob, op, st);
            else probe_write_columnar(fa, off, ob, op, st);
        }
    }

    if (is_root) {
        materialize_root(n, B, P, bw, remap, total, ob, op, out);
        return Rel{};
    }
    // Non-root: late-materialized intermediate (row-id list per slot).
    PhaseScope inter_phase(clock, CJ_PHASE_INTERMEDIATE);
    Rel res;
    res.columnar = false;
    res.rows = total;
    std::map<std::pair<int, uint32_t>, uint32_t> slot_map;
    GatherSet gs[2] = {};  // the lists each side's index list gathers
    for (uint32_t a = 0; a < n.num_output_attrs; ++a) {
        bool fb = remap[a] < bw;
        const Rel &S = fb ? B : P;
        uint32_t local = fb ? remap[a] : remap[a] - bw;
        const ColRef &cr = S.cols[local];
        auto key = std::make_pair(fb ? 0 : 1, cr.slot);
        auto it = slot_map.find(key);
        uint32_t ns;
        if (it == slot_map.end()) {
            const uint32_t *idx = fb ? ob : op;
            const uint32_t *old = S.slots[cr.slot].rids;
            const uint32_t *nr = idx;
            if (old && total) {
                uint32_t *g = tmp<uint32_t>(total);
                add_gather(gs[fb ? 0 : 1], old, g, idx, total);
                nr = g;
            }
            // the build side of a lazy direct map: ob holds map slots
            const uint32_t *map = old ? S.slots[cr.slot].map : (fb ? lazy_map : nullptr);
            ns = (uint32_t)res.slots.size();
            res.slots.push_back(Slot{S.slots[cr.slot].table, total ? nr : nullptr, total ? map : nullptr});
            if (!total) {  // an empty list still needs a non-null marker
                res.slots.back().rids = tmp<uint32_t>(1);
            }
            slot_map[key] = ns;
        } else {
            ns = it->second;
        }
        res.cols.push_back(ColRef{ns, cr.col, cr.type});
    }
    flush_gather(gs[0], ob, total);
    flush_gather(gs[1], op, total);
    return res;
}

// Queue dst = old[idx] into the side's gather set (launched when full).
void Exec::add_gather(GatherSet &g, const uint32_t *old, uint32_t *dst, const uint32_t *idx, uint64_t total) {
    if (g.count == kGatherSet) flush_gather(g, idx, total);
    g.src[g.count] = old;
    g.dst[g.count] = dst;
    ++g.count;
}

void Exec::flush_gather(GatherSet &g, const uint32_t *idx, uint64_t total) {
    if (!g.count) return;
    TimedScope ts(timer, "gather_rids", (4 + 8 * (uint64_t)g.count) * total);
    gather_u32_set(g, idx, total, st);
    g.count = 0;
}

// The fused probe + materialization applies to a root join whose probe side
// is a base-table scan (payload pages read in place) and whose output columns
// are all fixed-width.
bool Exec::fused_ok(const cj_node_desc &n, const Rel &B, const Rel &P, uint32_t bw,
                    const std::vector<uint32_t> &remap) {
    if (flags & CJ_EXEC_NO_FAST_PATH) return false;
    if (!P.columnar || n.num_output_attrs == 0 || n.num_output_attrs > kMatMaxCols) return false;
    for (uint32_t a = 0; a < n.num_output_attrs; ++a) {
        int32_t t = n.output_attrs[a].type;
        if (t == CJ_VARCHAR) return false;
        const Rel &S = remap[a] < bw ? B : P;
        const ColRef &cr = S.cols[remap[a] < bw ? remap[a] : remap[a] - bw];
        if (cr.type != t) return false;  // the general path reports the mismatch
    }
    return true;
}

// Page index of a base column (see fused.hip): cached with resident tables
// when decoded columns are cached, otherwise private to this call.
const uint16_t *Exec::page_index(DevColumn *c) {
    if (c->byte_prefix) return c->byte_prefix;
    PhaseScope ps(clock, CJ_PHASE_SETUP);
    bool own = false;
    for (auto &o : own_cols) own |= o.get() == c;
    TimedScope ts(timer, "page_index", c->bitmap_bytes * 3);
    if (!noalloc() && (own || ctx.opts.cache_decoded)) {
        std::lock_guard<std::mutex> g(c->mu);
        if (!c->byte_prefix) {
            uint16_t *bp = static_cast<uint16_t *>(ctx.pool.alloc(c->bitmap_bytes * 2 + 16));
            page_byte_prefix(c->pages, c->num_pages, c->byte_start, bp, st);
            if (!own) CJ_HIP(hipStreamSynchronize(st));
            c->byte_prefix = bp;
        }
        return c->byte_prefix;
    }
    auto it = page_indexes.find(c);
    if (it != page_indexes.end()) return it->second;
    uint16_t *bp = tmp<uint16_t>(c->bitmap_bytes + 8);
    page_byte_prefix(c->pages, c->num_pages, c->byte_start, bp, st);
    page_indexes[c] = bp;
    return bp;
}

// Direct-mapped build side (direct.hip) for a columnar build whose non-NULL
// keys are unique and span at most ~2x its rows.  Returns false (and the
// caller builds the hash table) when the keys are too sparse or repeat.
bool Exec::try_direct(const cj_node_desc &n, const Rel &B, uint32_t battr, uint32_t bw,
                      const std::vector<uint32_t> &remap, DirectTable *dt, bool rids) {
    if (flags & CJ_EXEC_NO_DIRECT) return false;
    DevColumn *kc = column(B.slots[0].table, B.cols[battr].col);
    if (kc->num_rows == 0) return false;
    PhaseScope ps(clock, CJ_PHASE_HASHTABLE_BUILD);
    // record: 8-byte fields, then 4-byte fields, then the u32 of NULL flags
    DirectBuildArgs a{};
    std::vector<uint32_t> build_attrs;
    if (!rids)
        for (uint32_t i = 0; i < n.num_output_attrs; ++i)
            if (remap[i] < bw) build_attrs.push_back(i);
    if (build_attrs.size() > kDirectMaxFields) return false;
    ColView kv;
    if (!kc->dense_rows) kv = view(kc);  // else the keys are read in place
    direct_field.assign(n.num_output_attrs, {0u, 0u});
    uint32_t off = 0;
    for (int pass = 0; pass < 2; ++pass)
        for (uint32_t f = 0; f < build_attrs.size(); ++f) {
            uint32_t i = build_attrs[f];
            int32_t type = n.output_attrs[i].type;
            uint32_t w = type == CJ_INT32 ? 4 : 8;
            if ((pass == 0) != (w == 8)) continue;
            const ColRef &cr = B.cols[remap[i]];
            DevColumn *fc = column(B.slots[0].table, cr.col);
            if (type == CJ_INT32 && fc->dense_rows) {  // read in place
                a.fields[f] = DirectField{4, off, nullptr, nullptr, fc->pages, fc->dense_rows};
            } else {
                ColView v = view(fc);
                a.fields[f] = DirectField{(int32_t)w, off, v.values, v.valid, nullptr, 0u};
            }
            direct_field[i] = {off, f};
            off += w;
        }
    a.nfields = (uint32_t)build_attrs.size();
    a.rid_offset = -1;
    if (rids) {
        a.rid_offset = (int32_t)off;
        off += 4;
    }
    a.flags_offset = off;
    a.stride = (off + 4 + 15) & ~15u;  // whole 16-byte chunks: one vector load each
    if (rids) {  // records are the build row ids alone
        a.flags_offset = 0;
        a.stride = 4;
    }
    a.keys = static_cast<const int32_t *>(kv.values);
    a.key_valid = kv.valid;
    a.key_pages = kc->dense_rows ? kc->pages : nullptr;
    a.key_rows = kc->dense_rows;
    a.rows = kc->num_rows;
    // the table is sized for the widest range the direct map accepts, so the
    // key range, the build and the duplicate check need one read-back
    a.max_range = 2 * kc->num_rows + 1024;
    // a map of row ids is 16 bytes a slot with nothing to fill but the keyed
    // rows: sparser keys (filtered primary keys) still take it
    if (rids) a.max_range = std::min<uint64_t>(16 * kc->num_rows + 1024, uint64_t(1) << 25);
    // statistics recorded by an earlier build over this key column (a
    // registered column's keys do not change): the key range, and whether
    // the keys are unique and fill it, are known before anything runs
    uint32_t hk[4];
    bool known, counted = false;
    {
        std::lock_guard<std::mutex> g(kc->mu);
        known = kc->dkey_known;
        counted = kc->dkey_counted;
        if (known) hk[0] = (uint32_t)kc->dkey_lo, hk[1] = (uint32_t)kc->dkey_hi, hk[2] = kc->dkey_keyed, hk[3] = kc->dkey_present;
    }
    if (known) {
        const int32_t lo = (int32_t)hk[0], hi = (int32_t)hk[1];
        if (lo > hi) return false;
        const uint64_t range = (uint64_t)((int64_t)hi - (int64_t)lo) + 1;
        if (range > a.max_range) return false;
        if (counted) {
            if (hk[2] != hk[3]) return false;
            a.known = 1;
            a.full = hk[3] == range;
            a.lo = lo;
            a.hi = hi;
        } else {
            known = false;  // the earlier build's map was too small to count: this one counts
        }
    }
    uint32_t *st4 = tmp<uint32_t>(4);  // min, max, keyed rows, present slots (written by kernels)
    int32_t *mm = reinterpret_cast<int32_t *>(st4);
    if (!known) {
        TimedScope ts(timer, "key_minmax", kc->num_rows * 4 + kc->num_rows / 8);
        key_minmax(a.keys, a.key_valid, a.key_pages, a.key_rows, kc->num_rows, mm, tmp<int32_t>(2 * kMinmaxParts), st);
    }
    a.minmax = mm;
    a.counters = st4 + 2;
    a.partials = tmp<uint32_t>(2 * kDirectParts);
    const uint64_t words = a.max_range / 32 + 4;
    a.presence = tmp<uint32_t>(words);
    a.records = tmp<uint8_t>(a.max_range * a.stride + 16);  // + the row-id map's last 16-byte clear
    {
        // zeroed records + one record per row + the bitmap pass over the records
        TimedScope ts(timer, "direct_build",
                      rids ? kc->num_rows * 8ull : kc->num_rows * (4ull + a.stride + a.stride) + kc->num_rows / 8);
        direct_build(a, st);
    }
    if (!known) {
        uint32_t *rb = reinterpret_cast<uint32_t *>(readback());
        CJ_HIP(hipMemcpyAsync(rb, st4, 16, hipMemcpyDeviceToHost, st));
        CJ_HIP(hipStreamSynchronize(st));
        std::copy(rb, rb + 4, hk);
        const int32_t lo = (int32_t)hk[0], hi = (int32_t)hk[1];
        std::lock_guard<std::mutex> g(kc->mu);
        kc->dkey_lo = lo, kc->dkey_hi = hi, kc->dkey_keyed = hk[2], kc->dkey_present = hk[3];
        kc->dkey_known = true;
        // the counters count only when the map held the key range
        kc->dkey_counted = lo <= hi && (uint64_t)((int64_t)hi - (int64_t)lo) + 1 <= a.max_range;
    }
    const uint32_t *h = hk;
    const int32_t lo = (int32_t)h[0], hi = (int32_t)h[1];
    if (lo > hi) return false;  // no non-NULL key
    const uint64_t range = (uint64_t)((int64_t)hi - (int64_t)lo) + 1;
    if (range > a.max_range) return false;
    if (h[2] != h[3]) return false;  // duplicate keys
    // keys that fill their range (primary keys 0..n-1) match every in-range
    // probe key, so the probes skip the bitmap
    *dt = DirectTable{lo, range, a.presence, a.records, a.stride, a.flags_offset, h[3] == range ? 1u : 0u,
                      a.rid_offset};
    return true;
}

// Columns one page-writer workgroup handles: one column per workgroup keeps
// registers low and puts more independent workgroups in flight; the ids of a
// block are then read once per column, mostly from L2.
static uint32_t mat_cols_per_wg(bool from_build, uint32_t width, uint32_t ncols) {
    (void)from_build;
    (void)width;
    (void)ncols;
    return 1;
}

void Exec::fused_root(const cj_node_desc &n, const Rel &B, const Rel &P, uint32_t bw,
                      const std::vector<uint32_t> &remap, const ProbeArgs &a, const uint64_t *off, uint64_t total) {
    PhaseScope ps(clock, CJ_PHASE_MATERIALIZE);
    const bool direct = a.mode == kModeDirect;
    bool need_b = false, need_p = false;
    for (uint32_t i = 0; i < n.num_output_attrs; ++i) (remap[i] < bw ? need_b : need_p) = true;
    uint32_t *ob = need_b ? tmp<uint32_t>(total) : nullptr;
    uint32_t *op = need_p ? tmp<uint32_t>(total) : nullptr;
    {
        TimedScope ts(timer, "probe_write", key_bytes(a.key) + 4 * total * ((uint64_t)need_b + need_p));
        probe_write_columnar(a, off, ob, op, st);
    }
    out_rows = total;
    out_types.resize(n.num_output_attrs);
    out_cols.resize(n.num_output_attrs);
    // page writers per (side, width); each workgroup writes cols_per_wg
    // columns of one block of result rows
    const uint64_t blocks = (total + kBlockRows - 1) / kBlockRows;
    // row ids of a build-side slot read through its map, resolved once for
    // all the columns that come from it
    std::map<uint32_t, const uint32_t *> resolved;
    auto rids_of = [&](uint32_t slot) {
        auto it = resolved.find(slot);
        if (it != resolved.end()) return it->second;
        return resolved[slot] = resolved_rids(B.slots[slot], B.rows);
    };
    for (int side = 0; side < 2; ++side)
        for (uint32_t width = 4; width <= 8; width += 4) {
            MatArgs m{};
            m.from_build = side == 0;
            m.direct = m.from_build && direct;
            m.dt = a.dt;
            m.ids = m.from_build ? ob : op;
            m.rows = total;
            uint64_t bytes = 0;
            for (uint32_t i = 0; i < n.num_output_attrs; ++i) {
                const int32_t type = n.output_attrs[i].type;
                const bool fb = remap[i] < bw;
                if (fb != (bool)m.from_build || (type == CJ_INT32 ? 4u : 8u) != width) continue;
                const Rel &S = fb ? B : P;
                const ColRef &cr = S.cols[fb ? remap[i] : remap[i] - bw];
                DevColumn *c = column(S.slots[cr.slot].table, cr.col);
                MatCol &mc = m.cols[m.ncols++];
                mc.width = (int32_t)width;
                mc.is_int32 = type == CJ_INT32;
                bytes += 4 * total;  // ids
                if (m.direct) {
                    mc.rec_offset = direct_field[i].first;
                    mc.rec_bit = direct_field[i].second;
                    bytes += total * (uint64_t)(width + 4);  // field + NULL flags
                } else if (fb) {
                    ColView v = view(c);
                    mc.values = v.values;
                    mc.valid_bits = v.valid;
                    mc.rids = rids_of(cr.slot);  // fb: S is B
                    bytes += total * (uint64_t)width + total / 8;
                } else {
                    mc.pages = c->pages;
                    mc.num_pages = c->num_pages;
                    mc.row_start = c->row_start;
                    mc.byte_start = c->byte_start;
                    mc.byte_prefix = page_index(c);
                    uint32_t *wf = tmp<uint32_t>(blocks + 1);
                    {
                        TimedScope ts(timer, "window_first", blocks * 12);
                        window_first(op, total, kBlockRows, c->row_start, c->num_pages, wf, st);
                    }
                    mc.win_first = wf;
                    bytes += (uint64_t)c->num_pages * kPage;
                }
                // INT32: page k per block; 8-byte: pages 2k, 2k + 1 (the last may be unused)
                const uint64_t np = width == 4 ? blocks : (total + kRowsPage8 - 1) / kRowsPage8;
                mc.out = out_pages(i, width == 4 ? blocks : 2 * blocks);
                bytes += total * (uint64_t)width + total / 8;
                out_types[i] = type;
                out_cols[i] = {mc.out, {(uint32_t)np, type}};
            }
            if (!m.ncols) continue;
            m.cols_per_wg = mat_cols_per_wg(m.from_build != 0, width, m.ncols);
            static const char *names[2][2] = {{"mat_build_w4", "mat_build_w8"}, {"mat_probe_w4", "mat_probe_w8"}};
            TimedScope ts(timer, names[side][width / 8], bytes);
            materialize_side(m, st);
        }
}

// Page-aligned root output (pagemat.hip): build-side pages per key page of
// the probe side, probe-side pages per payload page.  Page counts are read
// back once, together, after the writers have been queued.
void Exec::aligned_root(const cj_node_desc &n, const Rel &B, const Rel &P, uint32_t bw,
                        const std::vector<uint32_t> &remap, const ProbeArgs &a, const uint64_t *cnt,
                        const uint32_t *match_bits, uint64_t total) {
    PhaseScope ps(clock, CJ_PHASE_MATERIALIZE);
    out_rows = total;
    out_types.assign(n.num_output_attrs, 0);
    out_cols.assign(n.num_output_attrs, {nullptr, {0u, 0}});
    const uint32_t kp = a.key.num_pages;
    std::vector<const uint64_t *> count_at;  // device page count per output column
    std::vector<uint32_t> count_col;
    std::vector<uint64_t> bounds;  // pages of each build-side column's block
    // ---- build side
    BuildPagesArgs ba{};
    ba.key = a.key;
    ba.dt = a.dt;
    ba.match_bits = match_bits;
    // algorithmic bytes: key pages read once + output values and bitmaps (the
    // direct-map records are a build structure and are not counted)
    uint64_t bbytes = key_bytes(a.key);
    for (uint32_t i = 0; i < n.num_output_attrs; ++i) {
        if (remap[i] >= bw) continue;
        if (!ba.p4_off) {
            if (po_cnt != cnt) page_offsets(cnt, kp);
            ba.p4_off = po4;
            ba.p8_off = po8;
        }
        const int32_t type = n.output_attrs[i].type;
        const uint32_t width = type == CJ_INT32 ? 4 : 8;
        // at most one partial page per key page beyond the full ones
        const uint64_t bound = total / (width == 4 ? kCap4 : kCap8) + kp + 1;
        BuildPageCol &c = ba.cols[ba.ncols++];
        c.width = (int32_t)width;
        c.is_int32 = type == CJ_INT32;
        c.rec_offset = direct_field[i].first;
        c.rec_bit = direct_field[i].second;
        c.out = out_pages(i, bound);
        c.cap_pages = bound;
        bounds.push_back(bound);
        bbytes += total * (uint64_t)width + total / 8;
        out_types[i] = type;
        out_cols[i] = {c.out, {0u, type}};
        count_at.push_back((width == 4 ? ba.p4_off : ba.p8_off) + kp);
        count_col.push_back(i);
    }
    if (ba.ncols) {
        TimedScope ts(timer, "build_pages", bbytes);
        build_pages(ba, st);
    }
    // ---- probe side: payload page q writes output page q (no scan of the
    // pages with a match first); pages without a matched row are counted and
    // closed up afterwards (rare: this strategy needs half the rows to match)
    const size_t nbuild = count_at.size();
    std::vector<DevColumn *> probe_src;
    uint64_t *empties = nullptr;
    uint32_t nprobe = 0;
    for (uint32_t i = 0; i < n.num_output_attrs; ++i) nprobe += remap[i] >= bw;
    if (nprobe) {
        empties = tmp<uint64_t>(nprobe);
        fill_u64(empties, 0, nprobe, st);
    }
    for (uint32_t i = 0; i < n.num_output_attrs; ++i) {
        if (remap[i] < bw) continue;
        const int32_t type = n.output_attrs[i].type;
        const ColRef &cr = P.cols[remap[i] - bw];
        DevColumn *c = column(P.slots[cr.slot].table, cr.col);
        FilterArgs f{};
        f.pages = c->pages;
        f.num_pages = c->num_pages;
        f.row_start = c->row_start;
        f.match_bits = match_bits;
        f.pos = nullptr;
        f.empty_pages = empties + probe_src.size();
        f.width = type == CJ_INT32 ? 4 : 8;
        f.is_int32 = type == CJ_INT32;
        f.max_page_rows = c->max_page_rows;
        f.out = out_pages(i, c->num_pages);
        out_types[i] = type;
        out_cols[i] = {f.out, {0u, type}};
        count_at.push_back(f.empty_pages);
        count_col.push_back(i);
        probe_src.push_back(c);
        TimedScope ts(timer, f.width == 4 ? "probe_filter_w4" : "probe_filter_w8",
                      (uint64_t)c->num_pages * kPage + total * (uint64_t)f.width + total / 8);
        probe_filter(f, st);
    }
    const std::vector<uint64_t> np = read_u64s(count_at);
    for (size_t k = 0; k < nbuild; ++k) {
        if (np[k] > bounds[k])
            throw DeviceError("a join result outgrew its page block: a registered table's pages changed");
        out_cols[count_col[k]].second.first = (uint32_t)np[k];
    }
    for (size_t k = nbuild; k < np.size(); ++k) {
        DevColumn *c = probe_src[k - nbuild];
        const uint32_t i = count_col[k];
        out_cols[i].second.first = c->num_pages - (uint32_t)np[k];
        if (!np[k]) continue;
        // close the gaps: the pages with a matched row, in order, into a new block
        uint64_t *flag = tmp<uint64_t>(c->num_pages + 1), *pos = tmp<uint64_t>(c->num_pages + 1);
        {
            TimedScope ts(timer, "page_match_flags", (uint64_t)c->num_pages * 16 + c->num_rows / 8);
            page_match_flags(c->pages, c->num_pages, c->row_start, match_bits, flag, st);
        }
        scan(flag, pos, c->num_pages);
        uint8_t *dst = static_cast<uint8_t *>(ctx.pool.alloc((size_t)c->num_pages * kPage + 16));
        {
            TimedScope ts(timer, "compact_pages", 2ull * c->num_pages * kPage);
            compact_pages(out_cols[i].first, c->num_pages, pos, dst, st);
        }
        CJ_HIP(hipStreamSynchronize(st));
        ctx.pool.free(out_cols[i].first);
        out_cols[i].first = dst;
    }
}

void Exec::materialize_root(const cj_node_desc &n, const Rel &B, const Rel &P, uint32_t bw,
                            const std::vector<uint32_t> &remap, uint64_t total, const uint32_t *ob,
                            const uint32_t *op, cj_result *) {
    PhaseScope ps(clock, CJ_PHASE_MATERIALIZE);
    out_rows = total;
    out_types.clear();
    out_cols.clear();
    std::map<std::pair<int, uint32_t>, const uint32_t *> final_rids;
    std::vector<const uint32_t *> attr_rids(n.num_output_attrs, nullptr);
    GatherSet gs[2] = {};  // the lists each side's index list gathers
    for (uint32_t a = 0; a < n.num_output_attrs; ++a) {
        const int32_t type = n.output_attrs[a].type;
        out_types.push_back(type);
        bool fb = remap[a] < bw;
        const Rel &S = fb ? B : P;
        uint32_t local = fb ? remap[a] : remap[a] - bw;
        const ColRef &cr = S.cols[local];
        if (cr.type != type) throw UnsupportedError("output attr type does not match its source column");
        if (total == 0) {  // create_empty_result (materialize.h:61-69)
            out_cols.push_back({nullptr, {0u, type}});
            continue;
        }
        auto key = std::make_pair(fb ? 0 : 1, cr.slot);
        const uint32_t *rids;
        auto it = final_rids.find(key);
        if (it != final_rids.end()) {
            rids = it->second;
        } else {
            const uint32_t *idx = fb ? ob : op;
            const uint32_t *old = S.slots[cr.slot].rids;
            if (old) {
                uint32_t *g = tmp<uint32_t>(total);
                add_gather(gs[fb ? 0 : 1], old, g, idx, total);
                rids = g;
            } else {
                rids = idx;
            }
            final_rids[key] = rids;
        }
        attr_rids[a] = rids;
        out_cols.push_back({nullptr, {0u, type}});
    }
    if (total == 0) return;
    flush_gather(gs[0], ob, total);
    flush_gather(gs[1], op, total);
    std::map<std::pair<int, uint32_t>, const uint32_t *> resolved;  // map slots -> row ids
    for (uint32_t a = 0; a < n.num_output_attrs; ++a) {
        const bool fb = remap[a] < bw;
        const Rel &S = fb ? B : P;
        const ColRef &cr = S.cols[fb ? remap[a] : remap[a] - bw];
        const int32_t type = n.output_attrs[a].type;
        const uint32_t *rids = attr_rids[a];
        if (const uint32_t *map = S.slots[cr.slot].map) {
            auto key = std::make_pair(fb ? 0 : 1, cr.slot);
            auto it = resolved.find(key);
            if (it == resolved.end()) {
                uint32_t *g = tmp<uint32_t>(total);
                TimedScope ts(timer, "gather_rids", 12 * total);
                gather_u32(map, rids, g, total, st);
                it = resolved.emplace(key, g).first;
            }
            rids = it->second;
        }
        auto [pages, np] = materialize_column(column(S.slots[cr.slot].table, cr.col), type, rids, total, true, a);
        out_cols[a] = {pages, {np, type}};
    }
}

std::pair<uint8_t *, uint32_t> Exec::materialize_column(DevColumn *c, int32_t type, const uint32_t *rids,
                                                        uint64_t total, bool int32_quirk, uint32_t attr) {
    const bool in_place = type == CJ_INT32 && c->type == CJ_INT32 && c->dense_rows;
    ColView v;
    if (!in_place) v = view(c);
    uint8_t *pages;
    uint32_t np;
    if (type == CJ_VARCHAR) {
        const uint32_t chunk = varchar_chunk_rows(total);
        uint64_t chunks = (total + chunk - 1) / chunk;
        uint64_t *cp = tmp<uint64_t>(chunks + 1);
        uint64_t *co = tmp<uint64_t>(chunks + 1);
        VarcharSource vs{c->pages, static_cast<const uint64_t *>(v.values), v.valid, rids};
        {
            TimedScope ts(timer, "varchar_plan", 12 * total);
            varchar_plan(vs, total, chunk, cp, st);
        }
        scan(cp, co, chunks);
        uint64_t npg = read_u64(co + chunks);
        np = (uint32_t)npg;
        pages = out_pages(attr, np);
        TimedScope ts(timer, "varchar_write", (uint64_t)np * kPage);
        varchar_write(vs, total, chunk, co, pages, st);
    } else {
        uint32_t R = type == CJ_INT32 ? kRowsI32 : kRowsI64;
        np = (uint32_t)((total + R - 1) / R);
        pages = out_pages(attr, np);
        FixedSource fs{type == CJ_INT32 ? 4 : 8, type == CJ_INT32 && int32_quirk, v.values, v.valid, rids,
                       in_place ? c->pages : nullptr, in_place ? c->dense_rows : 0u};
        TimedScope ts(timer, type == CJ_INT32 ? "materialize_i32" : "materialize_i64", fixed_mat_bytes(total, type));
        materialize_fixed(fs, total, R, pages, st);
    }
    return {pages, np};
}

void Exec::filter_root(const std::vector<FNode> &prog, const std::vector<uint32_t> &used,
                       const std::vector<uint8_t> &lits) {
    const cj_table_desc &t = plan.inputs[0];
    uint64_t rows = 0;
    if (t.num_columns) rows = column(0, 0)->num_rows;
    for (uint32_t c = 1; c < t.num_columns; ++c) column(0, c);  // same row count checked there
    // the referenced columns, decoded
    std::vector<FCol> fc(used.size());
    for (size_t i = 0; i < used.size(); ++i) {
        DevColumn *c = column(0, used[i]);
        ColView v = view(c);
        fc[i] = FCol{c->type, v.values, v.valid, c->pages,
                     c->type == CJ_VARCHAR ? static_cast<const uint64_t *>(v.values) : nullptr};
    }
    FNode *d_prog = tmp<FNode>(prog.size());
    FCol *d_cols = tmp<FCol>(std::max<size_t>(fc.size(), 1));
    uint8_t *d_lits = tmp<uint8_t>(std::max<size_t>(lits.size(), 8));
    CJ_HIP(hipMemcpyAsync(d_prog, prog.data(), prog.size() * sizeof(FNode), hipMemcpyHostToDevice, st));
    if (!fc.empty()) CJ_HIP(hipMemcpyAsync(d_cols, fc.data(), fc.size() * sizeof(FCol), hipMemcpyHostToDevice, st));
    if (!lits.empty()) CJ_HIP(hipMemcpyAsync(d_lits, lits.data(), lits.size(), hipMemcpyHostToDevice, st));
    uint64_t total = 0;
    uint32_t *rids = nullptr;
    if (rows) {
        uint32_t *sel = tmp<uint32_t>(filter_sel_words(rows));
        {
            uint64_t bytes = 0;
            for (uint32_t u : used) bytes += (uint64_t)column(0, u)->num_pages * kPage;
            TimedScope ts(timer, "filter_eval", bytes);
            filter_eval(d_prog, (uint32_t)prog.size(), d_cols, d_lits, rows, sel, st);
        }
        const uint64_t chunks = (rows + kSelChunk - 1) / kSelChunk;
        uint64_t *cnt = tmp<uint64_t>(chunks + 1);
        uint64_t *off = tmp<uint64_t>(chunks + 1);
        sel_count(sel, rows, cnt, st);
        scan(cnt, off, chunks);
        total = read_u64(off + chunks);
        rids = tmp<uint32_t>(total + 1);
        TimedScope ts(timer, "filter_rids", rows / 8 + 4 * total);
        sel_rids(sel, rows, off, rids, st);
    }
    out_rows = total;
    PhaseScope ps(clock, CJ_PHASE_MATERIALIZE);
    // a filtered base table keeps its values: no INT32_MIN quirk
    // (from_inner_to_column, build_table.cpp:322-360)
    for (uint32_t c = 0; c < t.num_columns; ++c) {
        const int32_t type = t.columns[c].type;
        out_types.push_back(type);
        if (total == 0) {
            out_cols.push_back({nullptr, {0u, type}});
            continue;
        }
        auto [pages, np] = materialize_column(column(0, c), type, rids, total, false, c);
        out_cols.push_back({pages, {np, type}});
    }
}

}  // namespace

// Algorithmic output bytes of a result column, from its page headers:
// non-NULL value bytes (VARCHAR: chars + 2 per value) plus bitmap bytes.
static uint64_t host_payload_bytes(const uint8_t *pages, uint32_t np, int32_t type) {
    uint64_t b = 0;
    for (uint32_t i = 0; i < np; ++i) {
        const uint8_t *p = pages + (size_t)i * kPage;
        uint16_t n, nv;
        std::memcpy(&n, p, 2);
        std::memcpy(&nv, p + 2, 2);
        if (type == CJ_VARCHAR && (n == 0xffff || n == 0xfffe)) { b += nv; continue; }
        b += (n + 7) / 8;
        if (type == CJ_VARCHAR) {
            uint16_t last = 0;
            if (nv) std::memcpy(&last, p + 4 + 2 * (nv - 1), 2);
            b += last + 2ull * nv;
        } else {
            b += (uint64_t)nv * (type == CJ_INT32 ? 4 : 8);
        }
    }
    return b;
}

// Runs one execution on a stream of the context: `prepare` brings the inputs
// in (not device time), `body` fills the Exec's output columns between the
// two timing events; then the result is assembled (device pages, or host
// pages through pinned memory) and the timing filled.
static void run_exec(Context &ctx, const cj_plan_desc &plan, uint32_t flags, cj_result *out, cj_timing *timing,
                     const std::function<void(Exec &)> &prepare, const std::function<void(Exec &)> &body) {
    auto t_start = std::chrono::steady_clock::now();
    DevicePool &pool = ctx.pool;
    PinnedPool &pinned = ctx.pinned;
    hipStream_t s = ctx.acquire_stream();
    struct Rel_ {
        Context &c;
        hipStream_t s;
        ~Rel_() { c.release_stream(s); }
    } rel_guard{ctx, s};
    CJ_HIP(hipSetDevice(ctx.device));
    Exec ex(ctx, plan, s);
    ex.device_output = (flags & CJ_EXEC_OUTPUT_DEVICE) != 0;
    ex.flags = flags;
    ex.clock.on = timing != nullptr;
    hipEvent_t e0, e1, e2;
    CJ_HIP(hipEventCreate(&e0));
    CJ_HIP(hipEventCreate(&e1));
    CJ_HIP(hipEventCreate(&e2));
    auto *impl = new ResultImpl();
    try {
        prepare(ex);
        CJ_HIP(hipEventRecord(e0, s));
        body(ex);
        CJ_HIP(hipEventRecord(e1, s));
        out->num_rows = ex.out_rows;
        out->num_columns = (uint32_t)ex.out_cols.size();
        impl->cols.resize(ex.out_cols.size());
        impl->host_page_ptrs.resize(ex.out_cols.size());
        uint64_t total_pages = 0;
        for (size_t i = 0; i < ex.out_cols.size(); ++i) {
            impl->cols[i].type = ex.out_cols[i].second.second;
            impl->cols[i].num_pages = ex.out_cols[i].second.first;
            impl->cols[i].device_pages = ex.out_cols[i].first;
            impl->dev_pages.push_back(ex.out_cols[i].first);
            total_pages += ex.out_cols[i].second.first;
        }
        double d2h = 0;
        if (!ex.device_output && total_pages) {
            impl->host_block = pinned.alloc(total_pages * kPage);
            uint64_t at = 0;
            CJ_HIP(hipEventRecord(e1, s));
            ex.clock.push(CJ_PHASE_MATERIALIZE);  // the pages' copy to host memory
            for (size_t i = 0; i < impl->cols.size(); ++i) {
                auto &c = impl->cols[i];
                uint8_t *h = static_cast<uint8_t *>(impl->host_block) + at * kPage;
                if (c.num_pages)
                    CJ_HIP(hipMemcpyAsync(h, c.device_pages, (size_t)c.num_pages * kPage, hipMemcpyDeviceToHost, s));
                impl->host_page_ptrs[i].resize(c.num_pages);
                for (uint32_t q = 0; q < c.num_pages; ++q) impl->host_page_ptrs[i][q] = h + (size_t)q * kPage;
                at += c.num_pages;
            }
            CJ_HIP(hipEventRecord(e2, s));
            ex.clock.pop();
        }
        CJ_HIP(hipStreamSynchronize(s));
        if (!ex.device_output) {
            for (size_t i = 0; i < impl->cols.size(); ++i) {
                impl->cols[i].pages = impl->host_page_ptrs[i].data();
                pool.free(impl->dev_pages[i]);
                impl->dev_pages[i] = nullptr;
                impl->cols[i].device_pages = nullptr;
            }
            if (total_pages) {
                float ms = 0;
                hipEventElapsedTime(&ms, e1, e2);
                d2h = ms;
            }
        }
        out->columns = impl->cols.data();
        out->opaque = impl;
        if (timing) {
            float dev_ms = 0;
            hipEventElapsedTime(&dev_ms, e0, e1);
            std::memset(timing, 0, sizeof *timing);
            timing->device_ms = dev_ms;
            timing->h2d_ms = ex.h2d_ms;
            timing->d2h_ms = d2h;
            ex.clock.collect(timing->phase_ms);
            timing->phase_ms[CJ_PHASE_SETUP] += ex.h2d_ms;  // host clock: uploads before the first kernel
            int64_t *whole[CJ_NUM_PHASES] = {&timing->hashtable_build_ms, &timing->hash_join_probe_ms,
                                             &timing->nested_loop_join_ms, &timing->materialize_ms,
                                             &timing->setup_ms, &timing->intermediate_ms};
            for (int k = 0; k < CJ_NUM_PHASES; ++k) *whole[k] = (int64_t)timing->phase_ms[k];
            uint64_t in_bytes = 0;
            for (DevColumn *c : ex.referenced) in_bytes += (uint64_t)c->num_pages * kPage;
            uint64_t out_bytes = 0;
            if (ex.device_output) {
                uint64_t *acc = ex.tmp<uint64_t>(1);
                fill_u64(acc, 0, 1, s);
                for (auto &c : impl->cols)
                    if (c.num_pages)
                        page_payload_bytes(static_cast<const uint8_t *>(c.device_pages), c.num_pages, c.type, acc, s);
                out_bytes = ex.read_u64(acc);
            } else {
                for (size_t i = 0; i < impl->cols.size(); ++i) {
                    auto &c = impl->cols[i];
                    if (c.num_pages)
                        out_bytes += host_payload_bytes(static_cast<const uint8_t *>(impl->host_page_ptrs[i][0]),
                                                        c.num_pages, c.type);
                }
            }
            timing->bytes_alg = in_bytes + out_bytes;
        }
    } catch (...) {
        hipStreamSynchronize(s);
        for (void *p : impl->dev_pages) pool.free(p);
        for (auto &c : ex.out_cols)
            if (std::find(impl->dev_pages.begin(), impl->dev_pages.end(), c.first) == impl->dev_pages.end())
                pool.free(c.first);
        pinned.free(impl->host_block);
        delete impl;
        std::memset(out, 0, sizeof *out);
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        hipEventDestroy(e2);
        throw;
    }
    if (ex.timer.on) {
        std::lock_guard<std::mutex> g(ctx.mu);
        ex.timer.collect(ctx.kstats);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipEventDestroy(e2);
    if (timing) {
        auto ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
        timing->total_execution_ms = (int64_t)ms;
    }
}

// FNV-1a over a plan's structure and its inputs' pages: equal for repeats of
// one plan over the same (resident) tables.
static uint64_t plan_signature(const cj_plan_desc &p) {
    uint64_t h = 0xcbf29ce484222325ull;
    auto mix = [&](uint64_t v) {
        for (int i = 0; i < 8; ++i) {
            h ^= (v >> (8 * i)) & 0xff;
            h *= 0x100000001b3ull;
        }
    };
    mix(p.num_nodes), mix(p.root), mix(p.num_inputs);
    for (uint32_t i = 0; i < p.num_nodes; ++i) {
        const cj_node_desc &n = p.nodes[i];
        mix((uint64_t)n.kind), mix(n.left), mix(n.right), mix(n.left_attr), mix(n.right_attr);
        mix(n.build_left), mix(n.base_table_id), mix(n.num_output_attrs);
        for (uint32_t a = 0; a < n.num_output_attrs; ++a) mix(n.output_attrs[a].index), mix((uint64_t)n.output_attrs[a].type);
    }
    for (uint32_t i = 0; i < p.num_inputs; ++i) {
        const cj_table_desc &t = p.inputs[i];
        mix(t.num_rows), mix(t.num_columns);
        for (uint32_t c = 0; c < t.num_columns; ++c) {
            const cj_column_desc &col = t.columns[c];
            mix((uint64_t)col.type), mix(col.num_pages);
            mix(col.num_pages ? (uint64_t)(uintptr_t)col.pages[0] : 0);
            mix(col.num_pages ? (uint64_t)(uintptr_t)col.pages[col.num_pages - 1] : 0);
        }
    }
    return h;
}

uint64_t Context::learned_footprint(const cj_plan_desc &plan) {
    const uint64_t sig = plan_signature(plan);
    std::lock_guard<std::mutex> g(learned_mu);
    auto it = learned.find(sig);
    return it == learned.end() ? 0 : it->second;
}

void Context::execute(const cj_plan_desc &plan, uint32_t flags, cj_result *out, cj_timing *timing) {
    std::memset(out, 0, sizeof *out);
    if (plan.root >= plan.num_nodes) throw InvalidError("plan root out of range");
    if (plan.nodes[plan.root].kind != CJ_NODE_JOIN) throw UnsupportedError("plan root must be a join node");
    for (uint32_t i = 0; i < plan.num_inputs; ++i)
        for (uint32_t c = 0; c < plan.inputs[i].num_columns; ++c)
            if (plan.inputs[i].columns[c].type < CJ_INT32 || plan.inputs[i].columns[c].type > CJ_VARCHAR)
                throw InvalidError("bad column type");
    // meter this thread's pool blocks while the plan runs (its
    // temporaries and result pages) and record the peak by plan signature
    PoolMeter &meter = pool_meter();
    struct MeterScope {
        PoolMeter &m;
        explicit MeterScope(PoolMeter &mm) : m(mm) { m = PoolMeter{true, 0, 0}; }
        ~MeterScope() { m.on = false; }
    } ms{meter};
    run_exec(
        *this, plan, flags, out, timing,
        [&](Exec &ex) {
            // inputs first (uploads of non-resident columns are not device time)
            for (uint32_t i = 0; i < plan.num_nodes; ++i) {
                const cj_node_desc &nd = plan.nodes[i];
                if (nd.kind != CJ_NODE_SCAN || nd.base_table_id >= plan.num_inputs) continue;
                for (uint32_t a = 0; a < nd.num_output_attrs; ++a)
                    if (nd.output_attrs[a].index < plan.inputs[nd.base_table_id].num_columns)
                        ex.column((uint32_t)nd.base_table_id, (uint32_t)nd.output_attrs[a].index);
            }
        },
        [&](Exec &ex) { ex.join(plan.nodes[plan.root], true, 0, out); });
    const uint64_t peak = (uint64_t)std::max<int64_t>(meter.peak, 0);
    const uint64_t sig = plan_signature(plan);
    std::lock_guard<std::mutex> g(learned_mu);
    uint64_t &l = learned[sig];
    l = std::max(l, peak);
}

void Context::execute_noalloc(const cj_plan_desc &plan, uint32_t flags, void *const *out_pages,
                              const uint32_t *out_cap_pages, uint32_t num_out, uint64_t *out_rows,
                              uint32_t *out_num_pages) {
    if (plan.root >= plan.num_nodes) throw InvalidError("plan root out of range");
    if (plan.nodes[plan.root].kind != CJ_NODE_JOIN) throw UnsupportedError("plan root must be a join node");
    for (uint32_t i = 0; i < plan.num_inputs; ++i)
        for (uint32_t c = 0; c < plan.inputs[i].num_columns; ++c)
            if (plan.inputs[i].columns[c].type < CJ_INT32 || plan.inputs[i].columns[c].type > CJ_VARCHAR)
                throw InvalidError("bad column type");
    for (uint32_t i = 0; i < num_out; ++i)
        if (out_cap_pages[i] && !out_pages[i]) throw InvalidError("null output buffer");
    std::lock_guard<std::mutex> one_at_a_time(noalloc_mutex(device));
    hipStream_t s = acquire_stream();
    struct Rel_ {
        Context &c;
        hipStream_t s;
        ~Rel_() {
            hipStreamSynchronize(s);  // no kernel may still use the scratch when the next call starts
            c.release_stream(s);
        }
    } rel_guard{*this, s};
    Exec ex(*this, plan, s);
    ex.timer.on = false;
    ex.device_output = true;
    ex.flags = flags;
    ex.arena = noalloc_scratch(device);
    ex.arena_cap = kNoAllocScratchBytes;
    ex.user_out = out_pages;
    ex.user_cap = out_cap_pages;
    ex.user_n = num_out;
    for (uint32_t i = 0; i < plan.num_nodes; ++i) {  // every scanned column must be resident
        const cj_node_desc &nd = plan.nodes[i];
        if (nd.kind != CJ_NODE_SCAN || nd.base_table_id >= plan.num_inputs) continue;
        for (uint32_t a = 0; a < nd.num_output_attrs; ++a)
            if (nd.output_attrs[a].index < plan.inputs[nd.base_table_id].num_columns)
                ex.column((uint32_t)nd.base_table_id, (uint32_t)nd.output_attrs[a].index);
    }
    ex.join(plan.nodes[plan.root], true, 0, nullptr);
    CJ_HIP(hipStreamSynchronize(s));
    if (ex.out_cols.size() > num_out) throw InvalidError("cj_execute_noalloc: fewer output buffers than result columns");
    *out_rows = ex.out_rows;
    for (uint32_t i = 0; i < num_out; ++i) out_num_pages[i] = i < ex.out_cols.size() ? ex.out_cols[i].second.first : 0;
}

void Context::filter(const cj_table_desc &t, const cj_filter_node *nodes, uint32_t n, uint32_t flags, cj_result *out,
                     cj_timing *timing) {
    std::memset(out, 0, sizeof *out);
    if (n == 0 || n > kFilterMaxNodes) throw InvalidError("a filter program has 1 to 64 nodes");
    if (t.num_columns && !t.columns) throw InvalidError("null column array");
    for (uint32_t c = 0; c < t.num_columns; ++c)
        if (t.columns[c].type < CJ_INT32 || t.columns[c].type > CJ_VARCHAR) throw InvalidError("bad column type");
    std::vector<FNode> prog(n);
    std::vector<uint32_t> used;
    std::map<uint32_t, uint32_t> slot;
    std::vector<uint8_t> lits;
    uint32_t depth = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const cj_filter_node &x = nodes[i];
        FNode &f = prog[i];
        std::memset(&f, 0, sizeof f);
        f.kind = x.kind;
        if (x.kind == CJ_F_CMP) {
            if (x.column >= t.num_columns) throw InvalidError("filter column out of range");
            const int32_t type = t.columns[x.column].type;
            if (x.op < CJ_OP_EQ || x.op > CJ_OP_IS_NOT_NULL) throw InvalidError("bad filter operator");
            const bool nulltest = x.op == CJ_OP_IS_NULL || x.op == CJ_OP_IS_NOT_NULL;
            const bool like = x.op == CJ_OP_LIKE || x.op == CJ_OP_NOT_LIKE;
            if (!nulltest) {
                // the literal types the reference's column comparisons read (statement.cpp:47-164)
                const int32_t want = type == CJ_VARCHAR ? CJ_LIT_STR : (type == CJ_FP64 ? CJ_LIT_FLOAT : CJ_LIT_INT);
                if (x.lit_kind != want) throw InvalidError("filter literal type does not match its column");
                if (like && type != CJ_VARCHAR) throw InvalidError("LIKE needs a VARCHAR column");
                if (x.lit_kind == CJ_LIT_STR) {
                    if (x.slen && !x.sval) throw InvalidError("null string literal");
                    if (like && std::memchr(x.sval, '\n', x.slen))
                        throw UnsupportedError("LIKE pattern with a newline");
                    f.soff = (uint32_t)lits.size();
                    f.slen = x.slen;
                    lits.insert(lits.end(), x.sval, x.sval + x.slen);
                }
            }
            f.op = x.op;
            f.lit_kind = x.lit_kind;
            f.ival = x.ival;
            f.fval = x.fval;
            auto it = slot.find(x.column);
            if (it == slot.end()) {
                it = slot.emplace(x.column, (uint32_t)used.size()).first;
                used.push_back(x.column);
            }
            f.col = it->second;
            if (++depth > kFilterMaxDepth) throw InvalidError("filter program nests deeper than 32");
        } else if (x.kind == CJ_F_NOT) {
            if (depth < 1) throw InvalidError("NOT without an operand");
        } else if (x.kind == CJ_F_AND || x.kind == CJ_F_OR) {
            if (depth < 2) throw InvalidError("AND / OR without two operands");
            --depth;
        } else {
            throw InvalidError("bad filter node kind");
        }
    }
    if (depth != 1) throw InvalidError("a filter program leaves exactly one result");
    const cj_plan_desc plan{0, nullptr, 1, &t, 0};
    run_exec(
        *this, plan, flags, out, timing,
        [&](Exec &ex) {
            for (uint32_t c = 0; c < t.num_columns; ++c) ex.column(0, c);
        },
        [&](Exec &ex) { ex.filter_root(prog, used, lits); });
}

void Context::release(cj_result *res) {
    if (!res || !res->opaque) return;
    auto *impl = static_cast<ResultImpl *>(res->opaque);
    for (void *p : impl->dev_pages) pool.free(p);
    pinned.free(impl->host_block);
    delete impl;
    std::memset(res, 0, sizeof *res);
}

void Context::to_host(cj_result *res) {
    if (!res || !res->opaque) throw InvalidError("empty result");
    auto *impl = static_cast<ResultImpl *>(res->opaque);
    if (impl->host_block) return;
    uint64_t total_pages = 0;
    for (auto &c : impl->cols) total_pages += c.num_pages;
    if (!total_pages) return;
    impl->host_block = pinned.alloc(total_pages * kPage);
    uint64_t at = 0;
    for (size_t i = 0; i < impl->cols.size(); ++i) {
        auto &c = impl->cols[i];
        uint8_t *h = static_cast<uint8_t *>(impl->host_block) + at * kPage;
        if (c.num_pages) CJ_HIP(hipMemcpy(h, c.device_pages, (size_t)c.num_pages * kPage, hipMemcpyDeviceToHost));
        impl->host_page_ptrs[i].resize(c.num_pages);
        for (uint32_t q = 0; q < c.num_pages; ++q) impl->host_page_ptrs[i][q] = h + (size_t)q * kPage;
        c.pages = impl->host_page_ptrs[i].data();
        at += c.num_pages;
    }
}

}  // namespace cj
