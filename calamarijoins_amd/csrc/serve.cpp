// serve.cpp -- the serving loop (serve.hpp).
//
// The reference runs one plan at a time on a global thread pool
// (worker_pool.h:44-149, execute.cpp:311-325).  Here each worker thread owns
// one of the context's HIP streams and runs Context::execute; admission is
// by estimated device footprint so that concurrent plans never exceed the
// budget, taking requests in arrival order but letting a later request that
// fits pass one that does not (a request passed 8 times is admitted next).
#include "serve.hpp"

#include <algorithm>

namespace cj {

PlanCopy::PlanCopy(const cj_plan_desc &p) {
    nodes.assign(p.nodes, p.nodes + p.num_nodes);
    attrs.resize(p.num_nodes);
    for (uint32_t i = 0; i < p.num_nodes; ++i) {
        attrs[i].assign(p.nodes[i].output_attrs, p.nodes[i].output_attrs + p.nodes[i].num_output_attrs);
        nodes[i].output_attrs = attrs[i].data();
    }
    inputs.assign(p.inputs, p.inputs + p.num_inputs);
    cols.resize(p.num_inputs);
    pages.resize(p.num_inputs);
    for (uint32_t t = 0; t < p.num_inputs; ++t) {
        const cj_table_desc &td = p.inputs[t];
        cols[t].assign(td.columns, td.columns + td.num_columns);
        pages[t].resize(td.num_columns);
        for (uint32_t c = 0; c < td.num_columns; ++c) {
            pages[t][c].assign(td.columns[c].pages, td.columns[c].pages + td.columns[c].num_pages);
            cols[t][c].pages = pages[t][c].data();
        }
        inputs[t].columns = cols[t].data();
    }
    desc = p;
    desc.nodes = nodes.data();
    desc.inputs = inputs.data();
}

uint64_t estimate_plan_bytes(const cj_plan_desc &p, Context *ctx) {
    // rows a node produces: scans their table's rows, joins their probe side's
    std::vector<uint64_t> rows(p.num_nodes, 0);
    std::vector<uint32_t> scans(p.num_nodes, 0);  // base tables under a node: its row-id lists
    uint64_t bytes = 0;
    for (uint32_t i = 0; i < p.num_nodes; ++i) {  // children precede parents in well-formed plans
        const cj_node_desc &n = p.nodes[i];
        if (n.kind == CJ_NODE_SCAN) {
            if (n.base_table_id < p.num_inputs) {
                const cj_table_desc &t = p.inputs[n.base_table_id];
                rows[i] = t.num_rows;
                scans[i] = 1;
                for (uint32_t a = 0; a < n.num_output_attrs; ++a) {
                    const uint64_t ci = n.output_attrs[a].index;
                    DevColumn *dc = ctx && ci < t.num_columns ? ctx->find_resident(t.columns[ci]) : nullptr;
                    if (dc && (dc->dense_rows || ctx->opts.cache_decoded)) continue;  // in place, or cached once
                    // decoded column (dense values + validity), and its pages when uploaded
                    bytes += t.num_rows * (8ull + 1);
                    if (!dc && ci < t.num_columns) bytes += (uint64_t)t.columns[ci].num_pages * 8192;
                }
            }
            continue;
        }
        const uint64_t l = n.left < p.num_nodes ? rows[n.left] : 0;
        const uint64_t r = n.right < p.num_nodes ? rows[n.right] : 0;
        const uint64_t b = std::min(l, r), pr = std::max(l, r);
        rows[i] = pr;
        scans[i] = (n.left < p.num_nodes ? scans[n.left] : 0) + (n.right < p.num_nodes ? scans[n.right] : 0);
        // build structure (hash table or direct map), counts, match bitmap,
        // pair lists, one row-id list per base table that contributes columns
        const uint64_t lists = std::min<uint64_t>(scans[i], n.num_output_attrs);
        bytes += 40 * b + pr / 4 + 8 * pr + 4ull * pr * lists;
    }
    // result pages of the root
    if (p.root < p.num_nodes) bytes += rows[p.root] * 9ull * p.nodes[p.root].num_output_attrs;
    bytes += 64ull << 20;
    // a plan that has run on this context: its recorded peak, with a margin
    // (the bound above assumes every join's output as large as its larger
    // input; for key / foreign-key joins over filtered tables the real
    // footprint is tens to thousands of times smaller)
    if (ctx) {
        const uint64_t seen = ctx->learned_footprint(p);
        if (seen) bytes = std::min<uint64_t>(bytes, seen + seen / 4 + (8ull << 20));
    }
    return bytes;
}

Server::Server(Context &ctx, int workers, uint64_t admit_bytes) : ctx_(ctx), budget_(admit_bytes) {
    if (workers < 1) workers = 1;
    for (int i = 0; i < workers; ++i) threads_.emplace_back([this] { worker(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (auto &t : threads_) t.join();
    for (auto &kv : all_)
        if (kv.second->done && kv.second->status == CJ_OK) ctx_.release(&kv.second->result);
}

uint64_t Server::submit(const cj_plan_desc &plan, uint32_t flags) {
    auto r = std::make_unique<ServeRequest>();
    r->plan = std::make_unique<PlanCopy>(plan);
    r->flags = flags;
    r->est_bytes = estimate_plan_bytes(r->plan->desc, &ctx_);
    r->t_submit = std::chrono::steady_clock::now();
    uint64_t t;
    {
        std::lock_guard<std::mutex> g(mu_);
        if (stop_) throw BusyError("server is stopping");
        t = r->ticket = next_ticket_++;
        pending_.push_back(r.get());
        all_[t] = std::move(r);
    }
    cv_work_.notify_one();
    return t;
}

ServeRequest *Server::pick_locked() {
    if (exclusive_) return nullptr;  // a plan runs alone
    for (size_t i = 0; i < pending_.size(); ++i) {
        ServeRequest *r = pending_[i];
        if (r->alone) {  // waits at the head until nothing runs, and passes nothing
            if (running_ > 0) return nullptr;
            pending_.erase(pending_.begin() + (long)i);
            exclusive_ = true;
            return r;
        }
        const bool fits = admitted_ + r->est_bytes <= budget_ || running_ == 0;
        if (!fits && !r->budget_waited) {  // counted once per request
            r->budget_waited = true;
            ++stats_.budget_waits;
        }
        if (fits) {
            for (size_t k = 0; k < i; ++k) ++pending_[k]->skipped;
            pending_.erase(pending_.begin() + (long)i);
            return r;
        }
        if (r->skipped >= 8) break;  // it goes next: do not pass it again
    }
    return nullptr;
}

void Server::worker() {
    std::unique_lock<std::mutex> g(mu_);
    for (;;) {
        ServeRequest *r = nullptr;
        cv_work_.wait(g, [&] { return (stop_ && pending_.empty()) || (r = pick_locked()) != nullptr; });
        if (!r) return;
        admitted_ += r->est_bytes;
        ++running_;
        r->overlapped = !running_set_.empty();
        for (ServeRequest *o : running_set_) o->overlapped = true;
        running_set_.push_back(r);
        ++stats_.admitted;
        stats_.peak_admitted_bytes = std::max(stats_.peak_admitted_bytes, admitted_);
        stats_.peak_running = std::max(stats_.peak_running, (int32_t)running_);
        r->t_start = std::chrono::steady_clock::now();
        g.unlock();
        int status = CJ_OK;
        std::string err;
        bool retry = false;
        try {
            ctx_.execute(r->plan->desc, r->flags, &r->result, &r->timing);
        } catch (const InvalidError &e) {
            status = CJ_ERR_INVALID, err = e.what();
        } catch (const UnsupportedError &e) {
            status = CJ_ERR_UNSUPPORTED, err = e.what();
        } catch (const OomError &e) {
            status = CJ_ERR_OOM, err = e.what();
            retry = true;  // decided below: only if others ran beside it
        } catch (const DeviceError &e) {
            status = CJ_ERR_DEVICE, err = e.what();
        } catch (const std::bad_alloc &) {
            status = CJ_ERR_OOM, err = "host allocation failed";
        } catch (const std::exception &e) {
            status = CJ_ERR_INVALID, err = e.what();
        }
        g.lock();
        r->t_done = std::chrono::steady_clock::now();
        admitted_ -= r->est_bytes;
        --running_;
        running_set_.erase(std::find(running_set_.begin(), running_set_.end(), r));
        if (r->alone) exclusive_ = false;
        if (retry && !r->alone && r->overlapped) {  // out of memory beside others: once more, alone
            ++stats_.retried_alone;
            r->alone = true;
            pending_.push_front(r);
            cv_work_.notify_all();
            continue;
        }
        r->status = status;
        r->error = err;
        r->done = true;
        cv_done_.notify_all();
        cv_work_.notify_all();  // memory came back: another request may fit
    }
}

int Server::wait(uint64_t ticket, int64_t timeout_ms, cj_result *out, cj_timing *timing, std::string *err) {
    std::unique_lock<std::mutex> g(mu_);
    auto it = all_.find(ticket);
    if (it == all_.end()) throw InvalidError("unknown or already collected ticket");
    ServeRequest *r = it->second.get();
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0) cv_done_.wait(g, ready);
    else if (!cv_done_.wait_for(g, std::chrono::milliseconds(timeout_ms), ready)) return CJ_ERR_BUSY;
    const int status = r->status;
    if (status == CJ_OK) {
        *out = r->result;
        if (timing) {
            *timing = r->timing;
            timing->queue_ms = std::chrono::duration<double, std::milli>(r->t_start - r->t_submit).count();
            timing->service_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_start).count();
            timing->served_alone = r->alone ? 1 : 0;
        }
    } else if (err) {
        *err = r->error;
    }
    all_.erase(it);  // the caller owns the result now
    return status;
}

void Server::drain() {
    std::unique_lock<std::mutex> g(mu_);
    cv_done_.wait(g, [&] { return pending_.empty() && running_ == 0; });
}

}  // namespace cj
