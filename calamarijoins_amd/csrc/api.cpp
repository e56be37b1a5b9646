// api.cpp -- extern "C" entry points of libcalamarijoins_amd.so (declared in
// include/cj_api.h).  Exceptions never cross the boundary: they become a
// negative status plus a per-thread message for cj_last_error().
#include "engine.hpp"
#include "batch.hpp"
#include "serve.hpp"
#include "tbl.hpp"

#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>

namespace {
thread_local std::string g_err;

template <class F>
int guard(F &&f) {
    try {
        f();
        return CJ_OK;
    } catch (const cj::UnsupportedError &e) {
        g_err = e.what();
        return CJ_ERR_UNSUPPORTED;
    } catch (const cj::InvalidError &e) {
        g_err = e.what();
        return CJ_ERR_INVALID;
    } catch (const cj::OomError &e) {
        g_err = e.what();
        return CJ_ERR_OOM;
    } catch (const cj::BusyError &e) {
        g_err = e.what();
        return CJ_ERR_BUSY;
    } catch (const std::bad_alloc &e) {
        g_err = "host allocation failed";
        return CJ_ERR_OOM;
    } catch (const cj::DeviceError &e) {
        g_err = e.what();
        return CJ_ERR_DEVICE;
    } catch (const std::exception &e) {
        g_err = e.what();
        return CJ_ERR_INVALID;
    } catch (...) {
        g_err = "unknown error";
        return CJ_ERR_INVALID;
    }
}

cj::Context *C(cj_ctx *c) { return reinterpret_cast<cj::Context *>(c); }

// serving loops by context
std::mutex g_srv_mu;
std::map<cj::Context *, std::shared_ptr<cj::Server>> g_servers;

std::shared_ptr<cj::Server> server_of(cj_ctx *ctx) {
    std::lock_guard<std::mutex> g(g_srv_mu);
    auto it = g_servers.find(C(ctx));
    if (it == g_servers.end()) throw cj::InvalidError("serving loop not started on this context");
    return it->second;
}

void stop_server(cj::Context *c) {
    std::shared_ptr<cj::Server> s;
    {
        std::lock_guard<std::mutex> g(g_srv_mu);
        auto it = g_servers.find(c);
        if (it == g_servers.end()) return;
        s = it->second;
        g_servers.erase(it);
    }
    s->drain();
}
}  // namespace

void cj::set_last_error(const std::string &msg) { g_err = msg; }

struct cj_tbl {
    std::shared_ptr<cj::TblFile> f;
};

extern "C" {

const char *cj_last_error(void) { return g_err.c_str(); }

int cj_ctx_create(const cj_ctx_options *opts, cj_ctx **out) {
    if (!out) {
        g_err = "out is null";
        return CJ_ERR_INVALID;
    }
    *out = nullptr;
    return guard([&] {
        cj_ctx_options o{};
        if (opts) o = *opts;
        *out = reinterpret_cast<cj_ctx *>(new cj::Context(o));
    });
}

int cj_ctx_destroy(cj_ctx *ctx) {
    return guard([&] {
        stop_server(C(ctx));
        delete C(ctx);
    });
}

int cj_ctx_device_info(cj_ctx *ctx, char *name, size_t name_cap, int32_t *num_cus) {
    return guard([&] {
        if (!ctx) throw cj::InvalidError("null context");
        if (name && name_cap) std::snprintf(name, name_cap, "%s", C(ctx)->dev_name.c_str());
        if (num_cus) *num_cus = C(ctx)->num_cus;
    });
}

int cj_ctx_memory(cj_ctx *ctx, int32_t reset_peak, uint64_t *reserved, uint64_t *in_use, uint64_t *peak_in_use,
                  uint64_t *budget) {
    return guard([&] {
        if (!ctx) throw cj::InvalidError("null context");
        cj::DevicePool &p = C(ctx)->pool;
        if (reserved) *reserved = p.reserved();
        if (in_use) *in_use = p.in_use();
        if (peak_in_use) *peak_in_use = p.peak_in_use();
        if (budget) *budget = p.budget();
        if (reset_peak) p.reset_peak();
    });
}

int cj_table_register(cj_ctx *ctx, const cj_table_desc *table, uint64_t *out_id) {
    return guard([&] {
        if (!ctx || !table || !out_id) throw cj::InvalidError("null argument");
        CJ_HIP(hipSetDevice(C(ctx)->device));
        *out_id = C(ctx)->register_table(*table);
    });
}

int cj_table_unregister(cj_ctx *ctx, uint64_t id) {
    return guard([&] {
        if (!ctx) throw cj::InvalidError("null context");
        CJ_HIP(hipSetDevice(C(ctx)->device));
        C(ctx)->unregister_table(id);
    });
}

int cj_execute(cj_ctx *ctx, const cj_plan_desc *plan, uint32_t flags, cj_result *out, cj_timing *timing) {
    return guard([&] {
        if (!ctx || !plan || !out) throw cj::InvalidError("null argument");
        CJ_HIP(hipSetDevice(C(ctx)->device));
        C(ctx)->execute(*plan, flags, out, timing);
    });
}

int cj_execute_noalloc(cj_ctx *ctx, const cj_plan_desc *plan, uint32_t flags, void *const *out_pages,
                       const uint32_t *out_cap_pages, uint32_t num_out, uint64_t *out_rows,
                       uint32_t *out_num_pages) {
    return guard([&] {
        if (!ctx || !plan || !out_rows || (num_out && (!out_pages || !out_cap_pages || !out_num_pages)))
            throw cj::InvalidError("null argument");
        CJ_HIP(hipSetDevice(C(ctx)->device));
        C(ctx)->execute_noalloc(*plan, flags, out_pages, out_cap_pages, num_out, out_rows, out_num_pages);
    });
}

int cj_result_release(cj_ctx *ctx, cj_result *res) {
    return guard([&] {
        if (!ctx) throw cj::InvalidError("null context");
        C(ctx)->release(res);
    });
}

int cj_result_to_host(cj_ctx *ctx, cj_result *res) {
    return guard([&] {
        if (!ctx) throw cj::InvalidError("null context");
        CJ_HIP(hipSetDevice(C(ctx)->device));
        C(ctx)->to_host(res);
    });
}

int cj_kernel_stats(cj_ctx *ctx, cj_kernel_stat *out, uint32_t cap, uint32_t *count) {
    return guard([&] {
        if (!ctx || !count) throw cj::InvalidError("null argument");
        std::vector<cj_kernel_stat> v;
        C(ctx)->kernel_stats(v);
        *count = (uint32_t)v.size();
        for (uint32_t i = 0; i < cap && i < v.size(); ++i) out[i] = v[i];
    });
}

int cj_kernel_stats_reset(cj_ctx *ctx) {
    return guard([&] {
        if (!ctx) throw cj::InvalidError("null context");
        C(ctx)->kernel_stats_reset();
    });
}

int cj_serve_start(cj_ctx *ctx, const cj_serve_options *opts) {
    return guard([&] {
        if (!ctx || !opts) throw cj::InvalidError("null argument");
        cj::Context *c = C(ctx);
        const int workers = std::max(1, std::min<int>(opts->workers, (int)c->streams.size()));
        std::lock_guard<std::mutex> g(g_srv_mu);
        if (g_servers.count(c)) throw cj::InvalidError("serving loop already started");
        g_servers[c] = std::make_shared<cj::Server>(*c, workers, opts->admit_bytes);
    });
}

int cj_submit(cj_ctx *ctx, const cj_plan_desc *plan, uint32_t flags, uint64_t *ticket) {
    return guard([&] {
        if (!ctx || !plan || !ticket) throw cj::InvalidError("null argument");
        *ticket = server_of(ctx)->submit(*plan, flags);
    });
}

int cj_wait(cj_ctx *ctx, uint64_t ticket, int64_t timeout_ms, cj_result *out, cj_timing *timing) {
    if (!ctx || !out) {
        g_err = "null argument";
        return CJ_ERR_INVALID;
    }
    int st = CJ_OK;
    const int rc = guard([&] {
        std::memset(out, 0, sizeof *out);
        std::string err;
        st = server_of(ctx)->wait(ticket, timeout_ms, out, timing, &err);
        if (st != CJ_OK && st != CJ_ERR_BUSY) g_err = err;
        if (st == CJ_ERR_BUSY) g_err = "timed out";
    });
    return rc != CJ_OK ? rc : st;
}

int cj_serve_stats_get(cj_ctx *ctx, cj_serve_stats *out) {
    return guard([&] {
        if (!ctx || !out) throw cj::InvalidError("null argument");
        *out = server_of(ctx)->stats();
    });
}

int cj_serve_stop(cj_ctx *ctx) {
    return guard([&] {
        if (!ctx) throw cj::InvalidError("null context");
        stop_server(C(ctx));
    });
}

int cj_batch_config_set(cj_ctx *ctx, const cj_batch_config *cfg) {
    return guard([&] {
        if (!ctx) throw cj::InvalidError("null context");
        cj::Context *c = C(ctx);
        std::lock_guard<std::mutex> g(c->mu);
        if (!cfg) {
            c->batch_pinned = false;
            return;
        }
        if ((cfg->load_factor != 2 && cfg->load_factor != 4) || cfg->log_part < 8 ||
            cfg->log_part > cj::kBatchMaxLogPart || cfg->chunk == 0 || cfg->chunk > cj::kBatchMaxChunk ||
            cfg->join_queue > 1 || cfg->grid > 65536)
            throw cj::InvalidError("batch configuration out of range");
        c->batch_pin = cj::BatchConfig{cfg->load_factor, cfg->log_part, cfg->chunk, cfg->join_queue, cfg->grid};
        c->batch_pinned = true;
    });
}

int cj_execute_batch(cj_ctx *ctx, const cj_plan_desc *plans, uint32_t n, uint32_t flags, cj_result *out,
                     cj_batch_stats *stats) {
    return guard([&] {
        if (!ctx || (n && (!plans || !out))) throw cj::InvalidError("null argument");
        CJ_HIP(hipSetDevice(C(ctx)->device));
        C(ctx)->execute_batch(plans, n, flags, out, stats);
    });
}

int cj_filter(cj_ctx *ctx, const cj_table_desc *table, const cj_filter_node *program, uint32_t num_nodes,
              uint32_t flags, cj_result *out, cj_timing *timing) {
    return guard([&] {
        if (!ctx || !table || !program || !out) throw cj::InvalidError("null argument");
        CJ_HIP(hipSetDevice(C(ctx)->device));
        C(ctx)->filter(*table, program, num_nodes, flags, out, timing);
    });
}

int cj_tbl_write(const char *path, const cj_table_desc *table) {
    return guard([&] {
        if (!path || !table) throw cj::InvalidError("null argument");
        cj::tbl_write(path, *table);
    });
}

int cj_tbl_open(const char *path, cj_tbl **out, cj_table_desc *desc) {
    if (out) *out = nullptr;
    return guard([&] {
        if (!path || !out || !desc) throw cj::InvalidError("null argument");
        auto h = std::make_unique<cj_tbl>();
        h->f = cj::TblFile::open(path, /*populate*/ true);
        *desc = h->f->desc();
        *out = h.release();
    });
}

int cj_tbl_load(cj_ctx *ctx, const char *path, cj_tbl **out, cj_table_desc *desc, uint64_t *table_id) {
    if (out) *out = nullptr;
    return guard([&] {
        if (!ctx || !path || !out || !desc || !table_id) throw cj::InvalidError("null argument");
        cj::Context *c = C(ctx);
        CJ_HIP(hipSetDevice(c->device));
        auto h = std::make_unique<cj_tbl>();
        // the mapping gives the host page addresses plans use; the pages
        // themselves go from the file to HBM through pinned buffers
        h->f = cj::TblFile::open(path, /*populate*/ false);
        const int threads = c->opts.host_threads > 0
                                ? c->opts.host_threads
                                : (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
        std::shared_ptr<cj::TblFile> f = h->f;
        *table_id = c->register_table(
            f->desc(),
            [f, threads](uint32_t col, uint32_t p0, uint32_t np, uint8_t *dst) {
                cj::tbl_read_pages(*f, col, p0, np, dst, threads);
            },
            f);
        *desc = f->desc();
        *out = h.release();
    });
}

int cj_tbl_close(cj_tbl *t) {
    return guard([&] { delete t; });
}

int cj_tbl_cache_name(const char *table, const char *filter_text, char *out, size_t cap) {
    return guard([&] {
        if (!table || !out || !cap) throw cj::InvalidError("null argument");
        const std::string n = cj::tbl_cache_name(table, filter_text);
        if (n.size() + 1 > cap) throw cj::InvalidError("name buffer too small");
        std::memcpy(out, n.c_str(), n.size() + 1);
    });
}

int cj_plan_footprint(cj_ctx *ctx, const cj_plan_desc *plan, uint64_t *bytes) {
    return guard([&] {
        if (!plan || !bytes) throw cj::InvalidError("null argument");
        *bytes = cj::estimate_plan_bytes(*plan, ctx ? C(ctx) : nullptr);
    });
}

}  // extern "C"
