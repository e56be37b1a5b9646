// engine.hpp -- host side of the MI355X join engine: context, device memory
// pool, HBM-resident table registry and the plan executor.
#pragma once

#include "../../include/cj_api.h"
#include "cj_internal.hpp"

#include <atomic>
#include <condition_variable>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace cj {

// Device bytes one thread's allocations hold while it meters them (the
// blocks it takes from the pool, less the ones it returns): the footprint of
// one execution, which runs on one thread.
struct PoolMeter {
    bool on = false;
    int64_t cur = 0, peak = 0;
};
PoolMeter &pool_meter();  // this thread's

// Caching device allocator: blocks are rounded to 2 MiB (small ones to
// 4 KiB) and recycled best-fit, so steady-state execution calls no hipMalloc.
class DevicePool {
  public:
    explicit DevicePool(uint64_t budget) : budget_(budget) {}
    void set_budget(uint64_t b) { budget_ = b; }
    ~DevicePool();
    void *alloc(size_t bytes);
    void free(void *p);
    void trim();  // return cached blocks to the driver
    uint64_t in_use() const { return in_use_; }
    uint64_t reserved() const { return reserved_; }
    uint64_t peak_in_use() const { return peak_; }
    void reset_peak() {
        std::lock_guard<std::mutex> g(mu_);
        peak_ = in_use_;
    }
    uint64_t budget() const { return budget_; }

  private:
    std::mutex mu_;
    std::multimap<size_t, void *> free_;
    std::unordered_map<void *, size_t> sizes_;
    uint64_t budget_;
    uint64_t reserved_ = 0;  // bytes obtained from hipMalloc
    uint64_t in_use_ = 0;
    uint64_t peak_ = 0;
};

// Cached pinned host buffers (result pages, upload staging).
class PinnedPool {
  public:
    ~PinnedPool();
    void *alloc(size_t bytes);
    void free(void *p);

  private:
    std::mutex mu_;
    std::multimap<size_t, void *> free_;
    std::unordered_map<void *, size_t> sizes_;
};

// A base-table column in HBM.
uint64_t next_column_uid();

struct DevColumn {
    int32_t type = 0;
    uint32_t num_pages = 0;
    uint64_t num_rows = 0;      // rows described by the pages
    uint8_t *pages = nullptr;   // num_pages * 8 KiB, contiguous
    uint64_t *row_start = nullptr;  // [num_pages + 1] first row of each page
    void *values = nullptr;     // decoded: i32 / u64 per row, VARCHAR refs
    uint32_t *valid = nullptr;  // decoded validity bitmap
    uint64_t *byte_start = nullptr;   // [num_pages + 1] first bitmap byte of each page
    uint16_t *byte_prefix = nullptr;  // page index: non-NULL rows before each bitmap byte
    uint64_t bitmap_bytes = 0;        // sum over pages of ceil(rows / 8)
    uint32_t max_page_rows = 0;       // largest n_rows of a (non long-string) page
    uint32_t dense_rows = 0;          // INT32 without NULLs, this many rows on every page but
                                      // a shorter last one: rows are read in place (0: no)
    bool loaded = false;        // pages + row_start present
    // key statistics of a direct-map build over this column, recorded by the
    // first such build and reused by later ones (a registered column's
    // values do not change), so they skip the key range pass and the
    // build's read-back: the non-NULL keys' min and max, and, when that
    // build's map held the range (dkey_counted), the keyed rows and the
    // slots they filled
    bool dkey_known = false, dkey_counted = false;
    int32_t dkey_lo = 0, dkey_hi = 0;
    uint32_t dkey_keyed = 0, dkey_present = 0;
    // as a probe key column: the match count of a page-aligned join against a
    // direct map of each build key column (by that column's uid), for the
    // same reason
    std::vector<std::pair<uint64_t, uint64_t>> match_totals;
    const uint64_t uid = next_column_uid();  // unique over the process (a freed column's address is reused)
    std::mutex mu;              // guards lazy decode and the key statistics
};

struct DevTable {
    uint64_t id = 0;
    bool resident = false;
    uint64_t num_rows = 0;
    std::vector<std::unique_ptr<DevColumn>> cols;
    std::vector<std::vector<const void *>> host_pages;  // identity of each column
    std::shared_ptr<void> keepalive;                     // e.g. the mapping of a .tbl file
};

// Per-kernel hipEvent timing (enabled by cj_ctx_options::kernel_timing).
struct KernelAgg {
    uint64_t launches = 0;
    double ms = 0;
    uint64_t bytes = 0;
};
struct KernelTimer {
    struct Rec {
        std::string name;
        uint64_t bytes;
        hipEvent_t a, b;
    };
    bool on = false;
    hipStream_t s = nullptr;
    std::vector<Rec> recs;
    void begin(const char *name, uint64_t bytes);
    void end();
    void collect(std::map<std::string, KernelAgg> &acc);
};

// Phase clock of one execution (cj_timing's phases): a hipEvent on the
// execution stream at every phase change; the interval between two marks
// belongs to the phase that was current after the first.  Phases nest: an
// inner phase (a page decode inside a build) pauses the outer one, so no
// interval is counted twice.
struct PhaseClock {
    bool on = false;
    hipStream_t s = nullptr;
    class Context *ctx = nullptr;
    std::vector<std::pair<hipEvent_t, int>> marks;  // event, phase current after it (-1: none)
    std::vector<int> stack;
    void push(int phase);
    void pop();
    // adds each phase's milliseconds to ms[] (the stream must be idle)
    void collect(double *ms);
    ~PhaseClock();

  private:
    void mark();
};

struct ExecStream {
    hipStream_t s = nullptr;
    bool busy = false;
};

struct BatchTemplate;  // batch.cpp
struct BatchConfig {
    uint32_t load_factor;  // directory slots per build row (power-of-two rounded)
    uint32_t log_part;     // radix partitions of 2^log_part directory slots
    uint32_t chunk;        // probe tuples per join unit (the join tile)
    uint32_t join_queue;   // 1: join units from a device work queue, 0: strided by workgroup
    uint32_t grid;         // workgroups of every persistent launch; 0 = each kernel's resident grid
};

class Context {
  public:
    explicit Context(const cj_ctx_options &o);
    ~Context();

    int device = 0;
    cj_ctx_options opts{};
    DevicePool pool;
    PinnedPool pinned;
    std::string dev_name;
    int num_cus = 0;
    size_t max_lds = 65536;  // LDS bytes one workgroup may use

    // Fills dst with pages [p0, p0 + np) of column `col`; an empty function
    // copies them from the host page pointers of the descriptor.
    using PageFill = std::function<void(uint32_t col, uint32_t p0, uint32_t np, uint8_t *dst)>;
    uint64_t register_table(const cj_table_desc &t, const PageFill &fill = {}, std::shared_ptr<void> keepalive = {});
    void unregister_table(uint64_t id);

    void execute(const cj_plan_desc &plan, uint32_t flags, cj_result *out, cj_timing *timing);
    // Peak device bytes an execution of a plan took (its temporaries and
    // result pages, the blocks as the pool hands them out), recorded by plan
    // signature (structure and input pages: with resident tables, the same
    // plan); 0 if the plan has not run on this context.
    uint64_t learned_footprint(const cj_plan_desc &plan);
    std::mutex learned_mu;
    std::unordered_map<uint64_t, uint64_t> learned;
    // the same without allocating: temporaries in the GPU's fixed scratch,
    // result pages in caller buffers (cj_execute_noalloc)
    void execute_noalloc(const cj_plan_desc &plan, uint32_t flags, void *const *out_pages,
                         const uint32_t *out_cap_pages, uint32_t num_out, uint64_t *out_rows,
                         uint32_t *out_num_pages);
    // rows of one table passing a filter program (cj_filter)
    void filter(const cj_table_desc &t, const cj_filter_node *prog, uint32_t n, uint32_t flags, cj_result *out,
                cj_timing *timing);
    // many independent single-join plans in shared launches (batch.cpp)
    void execute_batch(const cj_plan_desc *plans, uint32_t n, uint32_t flags, cj_result *out, cj_batch_stats *stats);
    void release_batch_templates();
    std::map<uint64_t, BatchConfig> batch_tuning;                       // template class -> tuned config
    bool batch_pinned = false;                                          // cj_batch_config_set
    BatchConfig batch_pin{};
    std::map<uint64_t, std::shared_ptr<BatchTemplate>> batch_templates;  // workspaces + phase-1 graphs
    void release(cj_result *res);
    void to_host(cj_result *res);

    void kernel_stats(std::vector<cj_kernel_stat> &out);
    void kernel_stats_reset();

    // registry lookup: a host column matching a resident column
    DevColumn *find_resident(const cj_column_desc &c);

    hipStream_t acquire_stream();
    void release_stream(hipStream_t s);

    // device memory helpers used by the executor
    template <class T>
    T *dalloc(size_t n) {
        return static_cast<T *>(pool.alloc(n * sizeof(T) + 16));
    }
    void ensure_index(DevColumn &c, hipStream_t s);
    void ensure_decoded(DevColumn &c, hipStream_t s);
    void upload_column(DevColumn &c, const cj_column_desc &d, hipStream_t s,
                       const std::function<void(uint32_t p0, uint32_t np, uint8_t *dst)> &fill = {});
    void free_column(DevColumn &c);

    std::mutex mu;  // registry + stats
    std::map<uint64_t, std::unique_ptr<DevTable>> tables;
    std::map<std::pair<const void *, uint32_t>, std::pair<DevTable *, uint32_t>> by_first_page;
    uint64_t next_id = 1;
    std::map<std::string, KernelAgg> kstats;

    std::mutex smu;
    std::condition_variable scv;
    std::vector<ExecStream> streams;

    // timing events recycled across executions (phase clocks)
    hipEvent_t take_event();
    void give_event(hipEvent_t e);
    std::mutex emu;
    std::vector<hipEvent_t> event_pool;
};

// Result bookkeeping hidden behind cj_result::opaque.
struct ResultImpl {
    std::vector<cj_result_column> cols;
    std::vector<std::vector<void *>> host_page_ptrs;
    std::vector<uint8_t *> dev_pages;  // per column (pool memory)
    void *host_block = nullptr;        // pinned
};

}  // namespace cj
