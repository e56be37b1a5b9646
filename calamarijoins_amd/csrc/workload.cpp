// workload.cpp -- synthetic tables (benchmark and tests) and the host page
// writer; see include/cj_workload.h.
#include "../../include/cj_workload.h"
#include "../../include/cj_api.h"
#include "cj_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr uint64_t kPhi = 0x9E3779B97F4A7C15ull;
constexpr size_t kPageSz = CJ_PAGE_SIZE;
constexpr uint64_t kKeepStream = 0x6b656570ull;  // "keep"

inline uint64_t mix64(uint64_t z) {
    z ^= z >> 30;
    z *= 0xbf58476d1ce4e5b9ull;
    z ^= z >> 27;
    z *= 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
inline uint64_t u_at(uint64_t seed, uint64_t stream, uint64_t row) {
    return mix64((seed ^ (stream * kPhi)) + (row + 1) * kPhi);
}
inline double unit(uint64_t u) { return (double)(u >> 11) * (1.0 / 9007199254740992.0); }

template <class F>
void parallel_for(uint64_t n, int threads, F &&f) {
    if (threads <= 1 || n < 65536) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    uint64_t per = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        uint64_t a = t * per, b = std::min(n, a + per);
        if (a < b) th.emplace_back([&f, a, b] { f(a, b); });
    }
    for (auto &x : th) x.join();
}

// Seeded Fisher-Yates permutation of [0, n).
std::vector<int32_t> make_perm(uint64_t n, uint64_t seed, uint64_t stream) {
    std::vector<int32_t> p(n);
    for (uint64_t i = 0; i < n; ++i) p[i] = (int32_t)i;
    for (uint64_t i = n; i > 1; --i) {
        uint64_t j = u_at(seed, stream, i - 1) % i;
        std::swap(p[i - 1], p[j]);
    }
    return p;
}

// Cumulative Zipf(s) weights over ranks 1..n (unnormalised).
std::vector<double> make_zipf_cdf(uint64_t n, double s) {
    std::vector<double> c(n);
    double acc = 0;
    for (uint64_t k = 0; k < n; ++k) {
        acc += std::pow((double)(k + 1), -s);
        c[k] = acc;
    }
    return c;
}

// ---- page writer: the canonical capacity rules -------------------------
struct FixedWriter {
    uint8_t *out;
    uint64_t max_pages, npages = 0;
    size_t w, begin;
    uint32_t rows = 0, vals = 0;
    uint8_t stage[kPageSz];
    uint8_t bitmap[kPageSz];
    bool overflow = false;
    FixedWriter(uint8_t *o, uint64_t mp, size_t width) : out(o), max_pages(mp), w(width), begin(width == 4 ? 4 : 8) {
        std::memset(bitmap, 0, sizeof bitmap);
    }
    void flush() {
        if (npages >= max_pages) {
            overflow = true;
            return;
        }
        uint8_t *p = out + npages * kPageSz;
        std::memset(p, 0, kPageSz);
        uint16_t r = (uint16_t)rows, v = (uint16_t)vals;
        std::memcpy(p, &r, 2);
        std::memcpy(p + 2, &v, 2);
        std::memcpy(p + begin, stage, (size_t)vals * w);
        size_t nb = (rows + 7) / 8;
        std::memcpy(p + kPageSz - nb, bitmap, nb);
        std::memset(bitmap, 0, nb);
        ++npages;
        rows = vals = 0;
    }
    void add(const uint8_t *v, bool ok) {
        if (ok) {
            if (begin + (vals + 1) * w + rows / 8 + 1 > kPageSz) flush();
            std::memcpy(stage + (size_t)vals * w, v, w);
            bitmap[rows / 8] |= (uint8_t)(1u << (rows % 8));
            ++vals;
        } else {
            if (begin + vals * w + rows / 8 + 1 > kPageSz) flush();
        }
        ++rows;
    }
    void finish() {
        if (rows) flush();
    }
};

struct VarcharWriter {
    uint8_t *out;
    uint64_t max_pages, npages = 0;
    uint32_t rows = 0, nchars = 0;
    std::vector<uint16_t> offs;
    uint8_t chars[kPageSz];
    uint8_t bitmap[kPageSz];
    bool overflow = false;
    VarcharWriter(uint8_t *o, uint64_t mp) : out(o), max_pages(mp) { std::memset(bitmap, 0, sizeof bitmap); }
    uint8_t scratch[kPageSz];  // the page of a counting pass (out == nullptr)
    uint8_t *new_page() {
        if (!out) {  // counting pass: pages are counted, not kept
            ++npages;
            return scratch;
        }
        if (npages >= max_pages) {
            overflow = true;
            return nullptr;
        }
        uint8_t *p = out + npages * kPageSz;
        std::memset(p, 0, kPageSz);
        ++npages;
        return p;
    }
    void flush() {
        uint8_t *p = new_page();
        if (p) {
            uint16_t r = (uint16_t)rows, v = (uint16_t)offs.size();
            std::memcpy(p, &r, 2);
            std::memcpy(p + 2, &v, 2);
            std::memcpy(p + 4, offs.data(), offs.size() * 2);
            std::memcpy(p + 4 + offs.size() * 2, chars, nchars);
            size_t nb = (rows + 7) / 8;
            std::memcpy(p + kPageSz - nb, bitmap, nb);
        }
        std::memset(bitmap, 0, (rows + 7) / 8 + 1);
        offs.clear();
        rows = nchars = 0;
    }
    void add_long(const char *s, size_t len) {
        bool first = true;
        for (size_t off = 0; off < len && !overflow;) {
            uint8_t *p = new_page();
            if (!p) return;
            uint16_t tag = first ? 0xffff : 0xfffe;
            uint16_t n = (uint16_t)std::min(len - off, kPageSz - 4);
            std::memcpy(p, &tag, 2);
            std::memcpy(p + 2, &n, 2);
            std::memcpy(p + 4, s + off, n);
            off += n;
            first = false;
        }
    }
    void add(const char *s, size_t len, bool ok) {
        if (!ok) {
            if (4 + offs.size() * 2 + nchars + rows / 8 + 1 > kPageSz) flush();
            ++rows;
            return;
        }
        if (len > kPageSz - 7) {
            if (rows) flush();
            add_long(s, len);
            return;
        }
        if (4 + (offs.size() + 1) * 2 + nchars + len + rows / 8 + 1 > kPageSz) flush();
        std::memcpy(chars + nchars, s, len);
        nchars += (uint32_t)len;
        offs.push_back((uint16_t)nchars);
        bitmap[rows / 8] |= (uint8_t)(1u << (rows % 8));
        ++rows;
    }
    void finish() {
        if (rows) flush();
    }
};

// Pages the VARCHAR writer needs for these rows (a counting pass, so the
// page block is allocated exactly rather than for one page per row).
uint64_t varchar_page_count(uint64_t n, const uint64_t *offs, const char *bytes, const uint8_t *valid) {
    auto wr = std::make_unique<VarcharWriter>(nullptr, 0);
    for (uint64_t i = 0; i < n; ++i) wr->add(bytes + offs[i], offs[i + 1] - offs[i], valid[i] != 0);
    wr->finish();
    return wr->npages;
}

uint8_t *alloc_pages(uint64_t n) {
    void *p = std::aligned_alloc(4096, std::max<uint64_t>(n, 1) * kPageSz);
    if (!p) throw std::bad_alloc();
    return static_cast<uint8_t *>(p);
}

struct GenOwned {
    std::vector<int32_t> types;
    std::vector<uint32_t> num_pages;
    std::vector<uint8_t *> pages;
    ~GenOwned() {
        for (auto *p : pages) std::free(p);
    }
};



}  // namespace

extern "C" {

int cj_pages_encode_fixed(int32_t type, uint64_t n_rows, const void *values, const uint8_t *valid, void *out,
                          uint64_t max_pages, uint64_t *num_pages) {
    if (type != CJ_INT32 && type != CJ_INT64 && type != CJ_FP64) return CJ_ERR_INVALID;
    size_t w = type == CJ_INT32 ? 4 : 8;
    auto wr = std::make_unique<FixedWriter>(static_cast<uint8_t *>(out), max_pages, w);
    const uint8_t *v = static_cast<const uint8_t *>(values);
    for (uint64_t i = 0; i < n_rows && !wr->overflow; ++i) wr->add(v + i * w, valid ? valid[i] != 0 : true);
    wr->finish();
    if (wr->overflow) return CJ_ERR_INVALID;
    *num_pages = wr->npages;
    return CJ_OK;
}

int cj_pages_encode_varchar(uint64_t n_rows, const uint64_t *offsets, const char *bytes, const uint8_t *valid,
                            void *out, uint64_t max_pages, uint64_t *num_pages) {
    auto wr = std::make_unique<VarcharWriter>(static_cast<uint8_t *>(out), max_pages);
    for (uint64_t i = 0; i < n_rows && !wr->overflow; ++i)
        wr->add(bytes + offsets[i], offsets[i + 1] - offsets[i], valid ? valid[i] != 0 : true);
    wr->finish();
    if (wr->overflow) return CJ_ERR_INVALID;
    *num_pages = wr->npages;
    return CJ_OK;
}

int cj_gen_table_create(uint64_t candidate_rows, double keep_frac, uint64_t seed, uint32_t num_columns,
                        const cj_colgen *cols, int32_t threads, cj_gen_table *out) {
    try {
        std::memset(out, 0, sizeof *out);
        if (threads <= 0) threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        // surviving candidate rows
        std::vector<uint64_t> kept;
        if (keep_frac >= 1.0) {
            kept.resize(candidate_rows);
            for (uint64_t i = 0; i < candidate_rows; ++i) kept[i] = i;
        } else {
            for (uint64_t i = 0; i < candidate_rows; ++i)
                if (unit(u_at(seed, kKeepStream, i)) < keep_frac) kept.push_back(i);
        }
        const uint64_t n = kept.size();
        auto own = std::make_unique<GenOwned>();
        for (uint32_t c = 0; c < num_columns; ++c) {
            const cj_colgen &g = cols[c];
            const uint64_t sv = g.stream * 4, sn = g.stream * 4 + 1, sa = g.stream * 4 + 2;
            std::vector<uint8_t> valid(n);
            parallel_for(n, threads, [&](uint64_t a, uint64_t b) {
                for (uint64_t r = a; r < b; ++r)
                    valid[r] = !(g.null_frac > 0 && unit(u_at(seed, sn, kept[r])) < g.null_frac);
            });
            uint8_t *pages = nullptr;
            uint64_t np = 0;
            if (g.kind == CJ_GEN_VARCHAR_VOCAB) {
                // words of the vocabulary, each ended by a NUL byte
                std::vector<std::pair<uint64_t, uint32_t>> words;
                if (!g.vocab || !g.vocab_bytes) return CJ_ERR_INVALID;
                for (uint64_t a = 0, b = 0; b < g.vocab_bytes; ++b)
                    if (g.vocab[b] == '\0') {
                        words.push_back({a, (uint32_t)(b - a)});
                        a = b + 1;
                    }
                if (words.empty()) return CJ_ERR_INVALID;
                std::vector<uint32_t> pick(n);
                std::vector<uint64_t> offs(n + 1, 0);
                parallel_for(n, threads, [&](uint64_t a, uint64_t b) {
                    for (uint64_t r = a; r < b; ++r) pick[r] = (uint32_t)(u_at(seed, sv, kept[r]) % words.size());
                });
                for (uint64_t r = 0; r < n; ++r) offs[r + 1] = offs[r] + (valid[r] ? words[pick[r]].second : 0);
                std::vector<char> bytes(offs[n] + 8);
                parallel_for(n, threads, [&](uint64_t a, uint64_t b) {
                    for (uint64_t r = a; r < b; ++r)
                        if (valid[r]) std::memcpy(bytes.data() + offs[r], g.vocab + words[pick[r]].first, words[pick[r]].second);
                });
                uint64_t maxp = varchar_page_count(n, offs.data(), bytes.data(), valid.data());
                pages = alloc_pages(maxp);
                int rc = cj_pages_encode_varchar(n, offs.data(), bytes.data(), valid.data(), pages, maxp, &np);
                if (rc) { std::free(pages); throw std::runtime_error("varchar encode overflow"); }
            } else if (g.kind == CJ_GEN_VARCHAR) {
                std::vector<uint64_t> offs(n + 1, 0);
                std::vector<uint32_t> lens(n);
                parallel_for(n, threads, [&](uint64_t a, uint64_t b) {
                    for (uint64_t r = a; r < b; ++r) {
                        if (!valid[r]) { lens[r] = 0; continue; }
                        uint64_t u = u_at(seed, sv, kept[r]);
                        uint64_t u2 = u_at(seed, sa, kept[r]);
                        if (g.long_frac > 0 && unit(u2) < g.long_frac)
                            lens[r] = (uint32_t)(g.long_lo + (int64_t)(mix64(u2) % (uint64_t)(g.long_hi - g.long_lo + 1)));
                        else
                            lens[r] = (uint32_t)(g.lo + (int64_t)(u % (uint64_t)(g.hi - g.lo + 1)));
                    }
                });
                for (uint64_t r = 0; r < n; ++r) offs[r + 1] = offs[r] + lens[r];
                std::vector<char> bytes(offs[n] + 8);
                parallel_for(n, threads, [&](uint64_t a, uint64_t b) {
                    for (uint64_t r = a; r < b; ++r) {
                        uint64_t h = u_at(seed, sv ^ 0x5a5a5a5aull, kept[r]);
                        for (uint64_t i = 0; i < lens[r]; ++i) {
                            if ((i & 7) == 0) h = mix64(h + kPhi);
                            bytes[offs[r] + i] = (char)(32 + ((h >> (8 * (i & 7))) & 0xff) % 95);
                        }
                    }
                });
                uint64_t maxp = varchar_page_count(n, offs.data(), bytes.data(), valid.data());
                pages = alloc_pages(maxp);
                int rc = cj_pages_encode_varchar(n, offs.data(), bytes.data(), valid.data(), pages, maxp, &np);
                if (rc) { std::free(pages); throw std::runtime_error("varchar encode overflow"); }
            } else {
                size_t w = g.type == CJ_INT32 ? 4 : 8;
                std::vector<uint8_t> vals(n * w + 8);
                std::vector<int32_t> perm;
                std::vector<double> cdf;
                if (g.kind == CJ_GEN_PERM || g.kind == CJ_GEN_ZIPF_FK || g.kind == CJ_GEN_UNIFORM_FK)
                    perm = make_perm(g.domain, seed, g.perm_stream);
                if (g.kind == CJ_GEN_ZIPF_FK) cdf = make_zipf_cdf(g.domain, g.zipf_s);
                parallel_for(n, threads, [&](uint64_t a, uint64_t b) {
                    for (uint64_t r = a; r < b; ++r) {
                        uint64_t row = kept[r];
                        uint64_t u = u_at(seed, sv, row);
                        uint8_t *dst = vals.data() + r * w;
                        switch (g.kind) {
                        case CJ_GEN_PERM: {
                            int32_t v = perm[row % g.domain];
                            std::memcpy(dst, &v, 4);
                            break;
                        }
                        case CJ_GEN_UNIFORM: {
                            int32_t v = (int32_t)(g.lo + (int64_t)(u % (uint64_t)(g.hi - g.lo)));
                            std::memcpy(dst, &v, 4);
                            break;
                        }
                        case CJ_GEN_UNIFORM_FK: {
                            int32_t v = perm[u % g.domain];
                            std::memcpy(dst, &v, 4);
                            break;
                        }
                        case CJ_GEN_ZIPF_FK: {
                            uint64_t ua = u_at(seed, sa, row);
                            int32_t v;
                            if (g.miss_frac > 0 && unit(ua) < g.miss_frac) {
                                v = (int32_t)(g.miss_lo + (int64_t)(mix64(ua) % (uint64_t)(g.miss_hi - g.miss_lo)));
                            } else {
                                double t = unit(u) * cdf.back();
                                uint64_t rank = (uint64_t)(std::lower_bound(cdf.begin(), cdf.end(), t) - cdf.begin());
                                if (rank >= g.domain) rank = g.domain - 1;
                                v = perm[rank];
                            }
                            std::memcpy(dst, &v, 4);
                            break;
                        }
                        case CJ_GEN_I32_FULL: {
                            uint32_t v = (uint32_t)(u >> 32);
                            std::memcpy(dst, &v, 4);
                            break;
                        }
                        case CJ_GEN_I64_FULL:
                            std::memcpy(dst, &u, 8);
                            break;
                        case CJ_GEN_F64_SCALED: {
                            double d = unit(u) * g.scale;
                            std::memcpy(dst, &d, 8);
                            break;
                        }
                        default:
                            break;
                        }
                    }
                });
                uint64_t per = w == 4 ? 1984 : 1007;
                uint64_t maxp = n / per + n / 60000 + 4;
                pages = alloc_pages(maxp);
                int rc = cj_pages_encode_fixed(g.type, n, vals.data(), valid.data(), pages, maxp, &np);
                if (rc) { std::free(pages); throw std::runtime_error("fixed encode overflow"); }
            }
            own->types.push_back(g.type);
            own->num_pages.push_back((uint32_t)np);
            own->pages.push_back(pages);
        }
        out->num_rows = n;
        out->num_columns = num_columns;
        out->types = own->types.data();
        out->num_pages = own->num_pages.data();
        out->pages = own->pages.data();
        out->opaque = own.release();
        return CJ_OK;
    } catch (const std::exception &e) {
        cj::set_last_error(e.what());
        return CJ_ERR_INVALID;
    }
}

void cj_gen_table_free(cj_gen_table *t) {
    if (!t || !t->opaque) return;
    delete static_cast<GenOwned *>(t->opaque);
    std::memset(t, 0, sizeof *t);
}

}  // extern "C"
