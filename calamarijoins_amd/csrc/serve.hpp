// serve.hpp -- serving loop: plans submitted as they arrive are admitted by
// their estimated device footprint within a fixed budget and run
// concurrently, one worker per HIP stream, against the context's resident
// tables.  While one plan waits on a host read-back (the row count of a
// join), the kernels of the others keep the GPU busy.
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"

namespace cj {

// Owning copy of a cj_plan_desc (descriptor arrays only: the pages stay the
// caller's and must outlive the request).
struct PlanCopy {
    std::vector<cj_node_desc> nodes;
    std::vector<std::vector<cj_attr>> attrs;
    std::vector<cj_table_desc> inputs;
    std::vector<std::vector<cj_column_desc>> cols;
    std::vector<std::vector<std::vector<const void *>>> pages;
    cj_plan_desc desc{};
    explicit PlanCopy(const cj_plan_desc &p);
};

struct ServeRequest {
    uint64_t ticket = 0;
    std::unique_ptr<PlanCopy> plan;
    uint32_t flags = 0;
    uint64_t est_bytes = 0;
    uint32_t skipped = 0;  // times a later request was admitted first
    bool alone = false;    // ran out of device memory beside others: runs alone
    bool overlapped = false;  // another plan ran while this one did
    bool budget_waited = false;  // did not fit the budget at some admission
    std::chrono::steady_clock::time_point t_submit, t_start, t_done;
    bool done = false;
    int status = 0;
    std::string error;
    cj_result result{};
    cj_timing timing{};
};

// Device bytes a plan may hold at its peak: build structures, match and
// row-id lists, decoded columns and output pages, from the table sizes.
// A join's result is taken as the rows of its larger side, an upper bound
// for key / foreign-key joins (each row of the foreign-key side meets at
// most one key); a join whose result grows beyond that can run out of
// memory beside other plans, and is then run again alone.
// With a context, a column read from a resident table costs nothing when its
// decoded form is cached with the table or its INT32 pages are read in place,
// and a non-resident column adds its uploaded pages; and a plan that has run
// on the context is charged its recorded peak plus a quarter (and 8 MiB), if
// that is less.
uint64_t estimate_plan_bytes(const cj_plan_desc &p, Context *ctx = nullptr);

class Server {
  public:
    Server(Context &ctx, int workers, uint64_t admit_bytes);
    ~Server();
    uint64_t submit(const cj_plan_desc &plan, uint32_t flags);
    // CJ_OK with the result, CJ_ERR_BUSY on timeout (the request stays), or
    // the request's own error status.  A finished request is handed out once.
    int wait(uint64_t ticket, int64_t timeout_ms, cj_result *out, cj_timing *timing, std::string *err);
    void drain();

    cj_serve_stats stats() {
        std::lock_guard<std::mutex> g(mu_);
        return stats_;
    }

  private:
    void worker();
    ServeRequest *pick_locked();

    Context &ctx_;
    uint64_t budget_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<ServeRequest *> pending_;
    std::map<uint64_t, std::unique_ptr<ServeRequest>> all_;
    uint64_t admitted_ = 0;
    int running_ = 0;
    bool exclusive_ = false;  // a request that must run alone is running
    cj_serve_stats stats_{};
    std::vector<ServeRequest *> running_set_;
    bool stop_ = false;
    uint64_t next_ticket_ = 1;
    std::vector<std::thread> threads_;
};

}  // namespace cj
