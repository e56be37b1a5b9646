/*
 * cj_oracle.c -- CPU oracle for the join pipeline.  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
 * never by the product library.
 *
 * A plain-C restatement of the reference's execute() path (all citations are
 * to files of the reference repository):
 *   - page codec .............. src/build_table.cpp:520-924
 *   - random page reads ....... include/data_access/columnar_reader.h:120-474
 *   - side choice / remap ..... include/join_execution/join_setup.h:108-173
 *   - hash table .............. include/join_execution/hashtable.h:63-505
 *   - Bloom tags .............. include/join_execution/bloom_tags.h:35-263
 *   - probe ................... include/join_execution/hash_join.h:87-244
 *   - nested loop ............. include/join_execution/nested_loop.h:50-244
 *   - materialize ............. include/materialization/materialize.h:85-321
 *   - intermediates ........... include/materialization/construct_intermediate.h:121-220
 *   - recursion ............... src/execute.cpp:84-300
 *
 * Parity pinning of each part is documented in the test files that exercise
 * it (tests/test_oracle_*.py): golden vectors from the reference's unit
 * tests, the CRC-32C check value, the closed form of the Bloom tag table,
 * brute-force joins on small inputs and codec round trips.
 *
 * Deliberate extension beyond the reference: INT64 and FP64 output columns
 * (undefined behaviour in the reference, which routes them through the
 * VARCHAR builder) are copied bit-for-bit with their NULL bitmap.
 */
#define _GNU_SOURCE
#include "cj_oracle.h"

#include <pthread.h>
#include <stdarg.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define PAGE ((size_t)CJ_PAGE_SIZE)
#define I32_NULL ((int32_t)0x80000000)
#define LONG_SLOT 0x1FFF

static __thread char g_err[512];
const char *cjo_last_error(void) { return g_err; }
static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

static inline uint16_t rd16(const uint8_t *p) { uint16_t v; memcpy(&v, p, 2); return v; }
static inline void wr16(uint8_t *p, uint16_t v) { memcpy(p, &v, 2); }

/* ======================================================================
 * Hashing: CRC-32C (Castagnoli, reflected polynomial 0x82F63B78) without
 * pre/post inversion, which is what the SSE4.2 crc32 instruction used by
 * the reference computes (hashtable.h:137-154); multiplicative mix with
 * ((0x8648DBDB << 32) + 1) (hashtable.h:146-155).
 * ====================================================================== */
static uint32_t crc_tab[256];
static uint16_t bloom_tab[2048];
static pthread_once_t tab_once = PTHREAD_ONCE_INIT;

static void init_tables(void) {
    for (uint32_t i = 0; i < 256; ++i) {
        uint32_t c = i;
        for (int k = 0; k < 8; ++k) c = (c & 1) ? (c >> 1) ^ 0x82F63B78u : c >> 1;
        crc_tab[i] = c;
    }
    /* BLOOM_TAGS closed form: the 2048 smallest 16-bit integers with
     * exactly five bits set, ascending (bloom_tags.h:35-263). */
    uint32_t n = 0;
    for (uint32_t v = 0; v < 65536 && n < 2048; ++v)
        if (__builtin_popcount(v) == 5) bloom_tab[n++] = (uint16_t)v;
}

uint32_t cjo_crc32c_bytes(uint32_t crc, const uint8_t *p, size_t n) {
    pthread_once(&tab_once, init_tables);
    for (size_t i = 0; i < n; ++i) crc = crc_tab[(crc ^ p[i]) & 0xFF] ^ (crc >> 8);
    return crc;
}
uint32_t cjo_crc32c_u32(uint32_t crc, uint32_t v) {
    uint8_t b[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
    return cjo_crc32c_bytes(crc, b, 4);
}
uint16_t cjo_bloom_tag(uint32_t idx) {
    pthread_once(&tab_once, init_tables);
    return bloom_tab[idx & 0x7FF];
}

#if defined(__x86_64__)
__attribute__((target("sse4.2"))) static uint32_t crc_hw(uint32_t v) {
    return __builtin_ia32_crc32si(0, v);
}
static int have_sse42(void) { return __builtin_cpu_supports("sse4.2"); }
#else
static uint32_t crc_hw(uint32_t v) { return cjo_crc32c_u32(0, v); }
static int have_sse42(void) { return 0; }
#endif
static int g_use_hw = -1;
static inline uint64_t hash_key(int32_t key) {
    uint32_t c = g_use_hw ? crc_hw((uint32_t)key) : cjo_crc32c_u32(0, (uint32_t)key);
    return (uint64_t)c * ((0x8648DBDBull << 32) + 1);
}
static inline uint16_t bloom_of(uint64_t h) { return bloom_tab[(h >> 32) & 0x7FF]; }

/* ======================================================================
 * Canonical page codec (Table::to_columnar, build_table.cpp:679-924).
 * ====================================================================== */
typedef struct {
    uint8_t *out;
    size_t max_pages, npages;
    uint8_t page[PAGE]; /* value/offset area being filled          */
    uint8_t bitmap[PAGE];
    uint8_t chars[PAGE];
    uint32_t num_rows, nvals, nchars;
    int overflow;
} enc_t;

static uint8_t *enc_new_page(enc_t *e) {
    if (e->npages >= e->max_pages) { e->overflow = 1; return NULL; }
    uint8_t *p = e->out + e->npages * PAGE;
    memset(p, 0, PAGE);
    e->npages++;
    return p;
}

static void enc_save_fixed(enc_t *e, size_t width) {
    uint8_t *p = enc_new_page(e);
    if (p) {
        size_t bm = (e->num_rows + 7) / 8;
        wr16(p, (uint16_t)e->num_rows);
        wr16(p + 2, (uint16_t)e->nvals);
        size_t begin = width == 4 ? 4 : 8;
        memcpy(p + begin, e->page, (size_t)e->nvals * width);
        memcpy(p + PAGE - bm, e->bitmap, bm);
    }
    e->num_rows = e->nvals = 0;
    memset(e->bitmap, 0, sizeof e->bitmap);
}

size_t cjo_encode_fixed(int32_t type, uint64_t n_rows, const void *values,
                        const uint8_t *valid, uint8_t *out, size_t max_pages) {
    if (type != CJ_INT32 && type != CJ_INT64 && type != CJ_FP64) return (size_t)-1;
    size_t w = type == CJ_INT32 ? 4 : 8, begin = type == CJ_INT32 ? 4 : 8;
    enc_t *e = calloc(1, sizeof *e);
    e->out = out; e->max_pages = max_pages;
    const uint8_t *src = values;
    for (uint64_t i = 0; i < n_rows && !e->overflow; ++i) {
        if (valid[i]) {
            if (begin + (e->nvals + 1) * w + (e->num_rows / 8 + 1) > PAGE) enc_save_fixed(e, w);
            e->bitmap[e->num_rows / 8] |= (uint8_t)(1u << (e->num_rows % 8));
            memcpy(e->page + (size_t)e->nvals * w, src + i * w, w);
            e->nvals++;
            e->num_rows++;
        } else {
            if (begin + e->nvals * w + (e->num_rows / 8 + 1) > PAGE) enc_save_fixed(e, w);
            e->num_rows++;
        }
    }
    if (e->num_rows != 0 && !e->overflow) enc_save_fixed(e, w);
    size_t r = e->overflow ? (size_t)-1 : e->npages;
    free(e);
    return r;
}

static void enc_save_varchar(enc_t *e) {
    uint8_t *p = enc_new_page(e);
    if (p) {
        size_t bm = (e->num_rows + 7) / 8;
        wr16(p, (uint16_t)e->num_rows);
        wr16(p + 2, (uint16_t)e->nvals);
        memcpy(p + 4, e->page, (size_t)e->nvals * 2);
        memcpy(p + 4 + (size_t)e->nvals * 2, e->chars, e->nchars);
        memcpy(p + PAGE - bm, e->bitmap, bm);
    }
    e->num_rows = e->nvals = e->nchars = 0;
    memset(e->bitmap, 0, sizeof e->bitmap);
}

static void enc_long_string(enc_t *e, const char *s, size_t len) {
    size_t off = 0;
    int first = 1;
    while (off < len && !e->overflow) {
        uint8_t *p = enc_new_page(e);
        if (!p) return;
        wr16(p, first ? 0xffff : 0xfffe);
        first = 0;
        size_t n = len - off < PAGE - 4 ? len - off : PAGE - 4;
        wr16(p + 2, (uint16_t)n);
        memcpy(p + 4, s + off, n);
        off += n;
    }
}

size_t cjo_encode_varchar(uint64_t n_rows, const uint64_t *offsets,
                          const char *bytes, const uint8_t *valid, uint8_t *out,
                          size_t max_pages) {
    enc_t *e = calloc(1, sizeof *e);
    e->out = out; e->max_pages = max_pages;
    for (uint64_t i = 0; i < n_rows && !e->overflow; ++i) {
        if (valid[i]) {
            size_t len = offsets[i + 1] - offsets[i];
            const char *s = bytes + offsets[i];
            if (len > PAGE - 7) {
                if (e->num_rows > 0) enc_save_varchar(e);
                enc_long_string(e, s, len);
            } else {
                if (4 + (e->nvals + 1) * 2 + (e->nchars + len) + (e->num_rows / 8 + 1) > PAGE)
                    enc_save_varchar(e);
                e->bitmap[e->num_rows / 8] |= (uint8_t)(1u << (e->num_rows % 8));
                memcpy(e->chars + e->nchars, s, len);
                e->nchars += (uint32_t)len;
                wr16(e->page + (size_t)e->nvals * 2, (uint16_t)e->nchars);
                e->nvals++;
                e->num_rows++;
            }
        } else {
            if (4 + e->nvals * 2 + e->nchars + (e->num_rows / 8 + 1) > PAGE) enc_save_varchar(e);
            e->num_rows++;
        }
    }
    if (e->num_rows != 0 && !e->overflow) enc_save_varchar(e);
    size_t r = e->overflow ? (size_t)-1 : e->npages;
    free(e);
    return r;
}

/* ======================================================================
 * Decoders (Table::from_columnar, build_table.cpp:520-659).
 * ====================================================================== */
int64_t cjo_decode_fixed(int32_t type, const void *const *pages, uint32_t num_pages,
                         void *values, uint8_t *valid) {
    size_t w = type == CJ_INT32 ? 4 : 8, begin = type == CJ_INT32 ? 4 : 8;
    if (type == CJ_VARCHAR) return -1;
    int64_t row = 0;
    for (uint32_t pg = 0; pg < num_pages; ++pg) {
        const uint8_t *p = pages[pg];
        uint32_t n = rd16(p);
        const uint8_t *bm = p + PAGE - (n + 7) / 8;
        uint32_t vi = 0;
        for (uint32_t i = 0; i < n; ++i, ++row) {
            int ok = (bm[i / 8] >> (i % 8)) & 1;
            if (valid) valid[row] = (uint8_t)ok;
            if (ok) {
                if (values) memcpy((uint8_t *)values + row * w, p + begin + (size_t)vi * w, w);
                vi++;
            } else if (values) {
                memset((uint8_t *)values + row * w, 0, w);
            }
        }
    }
    return row;
}

int64_t cjo_decode_varchar(const void *const *pages, uint32_t num_pages,
                           uint64_t *offsets, char *bytes, uint64_t bytes_cap,
                           uint8_t *valid, uint64_t *total_bytes) {
    int64_t row = 0;
    uint64_t nb = 0;
    if (offsets) offsets[0] = 0;
    for (uint32_t pg = 0; pg < num_pages; ++pg) {
        const uint8_t *p = pages[pg];
        uint32_t n = rd16(p);
        if (n == 0xffff || n == 0xfffe) {
            uint32_t nc = rd16(p + 2);
            if (n == 0xfffe && row == 0) return -1; /* must follow a string */
            if (n == 0xffff) {
                if (valid) valid[row] = 1;
                row++;
            }
            if (bytes && nb + nc <= bytes_cap) memcpy(bytes + nb, p + 4, nc);
            nb += nc;
            if (offsets) offsets[row] = nb;
            continue;
        }
        uint32_t nv = rd16(p + 2);
        const uint8_t *offs = p + 4;
        const uint8_t *chars = p + 4 + (size_t)nv * 2;
        const uint8_t *bm = p + PAGE - (n + 7) / 8;
        uint32_t vi = 0, start = 0;
        for (uint32_t i = 0; i < n; ++i) {
            int ok = (bm[i / 8] >> (i % 8)) & 1;
            if (valid) valid[row] = (uint8_t)ok;
            if (ok) {
                uint32_t end = rd16(offs + 2 * vi);
                vi++;
                if (end < start) return -1;
                uint32_t len = end - start;
                if (bytes && nb + len <= bytes_cap) memcpy(bytes + nb, chars + start, len);
                nb += len;
                start = end;
            }
            row++;
            if (offsets) offsets[row] = nb;
        }
    }
    if (total_bytes) *total_bytes = nb;
    return row;
}

/* ---- order-free row hashing -------------------------------------------- */
static inline uint64_t mix64(uint64_t z) {
    z ^= z >> 30; z *= 0xbf58476d1ce4e5b9ull;
    z ^= z >> 27; z *= 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static uint64_t hash_bytes(const uint8_t *p, size_t n) {
    uint64_t h = 0x9e3779b97f4a7c15ull ^ n;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) { uint64_t v; memcpy(&v, p + i, 8); h = mix64(h ^ v) + i; }
    uint64_t t = 0;
    memcpy(&t, p + i, n - i);
    return mix64(h ^ t ^ ((uint64_t)(n - i) << 56));
}

int cjo_row_hashes(const cj_table_desc *t, uint64_t *out) {
    uint64_t n = t->num_rows;
    for (uint64_t r = 0; r < n; ++r) out[r] = 0x51ed270b27a5c3d1ull;
    for (uint32_t c = 0; c < t->num_columns; ++c) {
        const cj_column_desc *col = &t->columns[c];
        uint8_t *valid = malloc(n ? n : 1);
        uint64_t salt = mix64(0x1234 + c * 7919 + (uint64_t)col->type);
        if (col->type == CJ_VARCHAR) {
            uint64_t total = 0;
            if (cjo_decode_varchar(col->pages, col->num_pages, NULL, NULL, 0, NULL, &total) != (int64_t)n) {
                free(valid);
                return fail(CJ_ERR_INVALID, "column %u: row count mismatch", c);
            }
            uint64_t *offs = malloc((n + 1) * 8);
            char *bytes = malloc(total ? total : 1);
            cjo_decode_varchar(col->pages, col->num_pages, offs, bytes, total, valid, NULL);
            for (uint64_t r = 0; r < n; ++r) {
                uint64_t h = valid[r] ? hash_bytes((uint8_t *)bytes + offs[r], offs[r + 1] - offs[r]) : 0xdeadull;
                out[r] = mix64(out[r] * 0x100000001b3ull ^ (h + salt));
            }
            free(offs); free(bytes);
        } else {
            size_t w = col->type == CJ_INT32 ? 4 : 8;
            uint8_t *vals = malloc(n * w + 8);
            if (cjo_decode_fixed(col->type, col->pages, col->num_pages, vals, valid) != (int64_t)n) {
                free(vals); free(valid);
                return fail(CJ_ERR_INVALID, "column %u: row count mismatch", c);
            }
            for (uint64_t r = 0; r < n; ++r) {
                uint64_t v = 0;
                memcpy(&v, vals + r * w, w);
                uint64_t h = valid[r] ? mix64(v + 0x77) : 0xdeadull;
                out[r] = mix64(out[r] * 0x100000001b3ull ^ (h + salt));
            }
            free(vals);
        }
        free(valid);
    }
    return 0;
}

/* ======================================================================
 * Thread fan-out (stands in for WorkerThreadPool::execute,
 * platform/worker_pool.h:129-142).
 * ====================================================================== */
typedef void (*task_fn)(void *arg, int tid, int nt);
typedef struct { task_fn fn; void *arg; int tid, nt; } task_t;
static void *task_tramp(void *p) { task_t *t = p; t->fn(t->arg, t->tid, t->nt); return NULL; }
static void run_parallel(int nt, task_fn fn, void *arg) {
    if (nt <= 1) { fn(arg, 0, 1); return; }
    pthread_t th[256];
    task_t tk[256];
    if (nt > 256) nt = 256;
    for (int i = 0; i < nt; ++i) {
        tk[i] = (task_t){fn, arg, i, nt};
        if (i) pthread_create(&th[i], NULL, task_tramp, &tk[i]);
    }
    fn(arg, 0, nt);
    for (int i = 1; i < nt; ++i) pthread_join(th[i], NULL);
}

/* growable u32 vector */
typedef struct { uint32_t *d; size_t n, cap; } vec32;
static inline void v32_push(vec32 *v, uint32_t x) {
    if (v->n == v->cap) { v->cap = v->cap ? v->cap * 2 : 1024; v->d = realloc(v->d, v->cap * 4); }
    v->d[v->n++] = x;
}

/* ======================================================================
 * Page index + random reads (ColumnarReader, columnar_reader.h:120-474).
 * ====================================================================== */
typedef struct {
    const cj_column_desc *col;
    uint32_t *cum;        /* cumulative rows after each page (0xffff=1, 0xfffe=0) */
    uint32_t **prefix;    /* per sparse page: popcount before each 64-row word   */
    int dense_i32;        /* all pages dense with equal rows (but the last)      */
    uint32_t rows_per_page;
} pidx_t;

static void pidx_build(pidx_t *ix, const cj_column_desc *col) {
    ix->col = col;
    ix->cum = malloc(((size_t)col->num_pages + 1) * 4);
    ix->prefix = calloc(col->num_pages + 1, sizeof(uint32_t *));
    uint32_t total = 0, first = 0;
    int cand = col->type == CJ_INT32;
    for (uint32_t pg = 0; pg < col->num_pages; ++pg) {
        const uint8_t *p = col->pages[pg];
        uint32_t n = rd16(p), nv = rd16(p + 2);
        if (n == 0xfffe) cand = 0;
        else if (n == 0xffff) { cand = 0; total += 1; }
        else total += n;
        ix->cum[pg] = total;
        if (n != 0xfffe && n != 0xffff && n != nv) {
            cand = 0;
            size_t bm_size = (n + 7) / 8, words = (n + 63) / 64;
            const uint8_t *bm = p + PAGE - bm_size;
            uint32_t *ps = malloc(words * 4), sum = 0;
            for (size_t w = 0; w < words; ++w) {
                ps[w] = sum;
                uint64_t word = 0;
                size_t off = w * 8, rem = bm_size > off ? bm_size - off : 0;
                memcpy(&word, bm + off, rem < 8 ? rem : 8);
                sum += (uint32_t)__builtin_popcountll(word);
            }
            ix->prefix[pg] = ps;
        }
        if (cand) {
            if (pg == 0) first = n;
            else if (pg + 1 < col->num_pages && n != first) cand = 0;
        }
    }
    ix->dense_i32 = cand && col->num_pages > 0 && first > 0;
    ix->rows_per_page = first;
}
static void pidx_free(pidx_t *ix) {
    if (!ix->cum) return;
    for (uint32_t pg = 0; pg < ix->col->num_pages; ++pg) free(ix->prefix[pg]);
    free(ix->prefix); free(ix->cum);
    ix->cum = NULL;
}

typedef struct { uint32_t start, end, page; int valid; } cursor_t;

static inline uint32_t find_page(const pidx_t *ix, uint32_t rid, cursor_t *cur) {
    if (cur->valid && rid >= cur->start && rid < cur->end) return cur->page;
    uint32_t lo = 0, hi = ix->col->num_pages;
    if (cur->valid && rid == cur->end && cur->page + 1 < hi) lo = cur->page + 1;
    while (lo < hi) { /* upper_bound(cum, rid) */
        uint32_t mid = (lo + hi) / 2;
        if (ix->cum[mid] <= rid) lo = mid + 1; else hi = mid;
    }
    cur->page = lo;
    cur->start = lo ? ix->cum[lo - 1] : 0;
    cur->end = ix->cum[lo];
    cur->valid = 1;
    return lo;
}

/* Value index of local row r in page pg, or -1 if NULL (read_sparse, 452-474). */
static inline int32_t value_index(const pidx_t *ix, uint32_t pg, uint32_t r) {
    const uint8_t *p = ix->col->pages[pg];
    uint32_t n = rd16(p), nv = rd16(p + 2);
    if (n == nv) return (int32_t)r;
    const uint8_t *bm = p + PAGE - (n + 7) / 8;
    if (!((bm[r >> 3] >> (r & 7)) & 1)) return -1;
    uint64_t word = 0;
    size_t bm_size = (n + 7) / 8, off = (r >> 6) * 8, rem = bm_size - off;
    memcpy(&word, bm + off, rem < 8 ? rem : 8);
    uint32_t bit = r & 63;
    uint64_t mask = bit ? (word & ((1ull << bit) - 1)) : 0;
    return (int32_t)(ix->prefix[pg][r >> 6] + (uint32_t)__builtin_popcountll(mask));
}

/* value_t read (INT32 value or VARCHAR ref; NULL = INT32_MIN). */
static inline int32_t read_v32(const pidx_t *ix, uint32_t rid, cursor_t *cur) {
    if (ix->dense_i32) {
        uint32_t pg = rid / ix->rows_per_page, r = rid - pg * ix->rows_per_page;
        int32_t v;
        memcpy(&v, (const uint8_t *)ix->col->pages[pg] + 4 + 4 * (size_t)r, 4);
        return v;
    }
    uint32_t pg = find_page(ix, rid, cur);
    const uint8_t *p = ix->col->pages[pg];
    if (rd16(p) == 0xffff) return (int32_t)((LONG_SLOT << 19) | (pg & 0x7FFFF));
    int32_t vi = value_index(ix, pg, rid - cur->start);
    if (vi < 0) return I32_NULL;
    if (ix->col->type == CJ_INT32) {
        int32_t v;
        memcpy(&v, p + 4 + 4 * (size_t)vi, 4);
        return v;
    }
    return (int32_t)(((uint32_t)vi << 19) | (pg & 0x7FFFF));
}
/* 8-byte read (extension for INT64/FP64). */
static inline int read_v64(const pidx_t *ix, uint32_t rid, cursor_t *cur, int64_t *out) {
    uint32_t pg = find_page(ix, rid, cur);
    int32_t vi = value_index(ix, pg, rid - cur->start);
    if (vi < 0) return 0;
    memcpy(out, (const uint8_t *)ix->col->pages[pg] + 8 + 8 * (size_t)vi, 8);
    return 1;
}

/* ======================================================================
 * Intermediates (mema::column_t / ExecuteResult, intermediate.h:40-172).
 * INT32 and VARCHAR cells are value_t; INT64/FP64 use 8-byte cells.
 * ====================================================================== */
typedef struct {
    int32_t type;
    int32_t *v;       /* value_t cells                       */
    int64_t *w;       /* INT64/FP64 cells                    */
    uint8_t *wv;      /* INT64/FP64 validity                 */
    uint8_t src_table, src_col;
} icol_t;
typedef struct { uint64_t n; uint32_t ncols; icol_t *cols; } inter_t;

static void inter_free(inter_t *r) {
    for (uint32_t c = 0; c < r->ncols; ++c) { free(r->cols[c].v); free(r->cols[c].w); free(r->cols[c].wv); }
    free(r->cols);
    memset(r, 0, sizeof *r);
}

typedef struct {
    int columnar;
    const cj_table_desc *table;
    uint8_t table_id;
    inter_t inter;
    const cj_node_desc *node;
} input_t;

static uint64_t in_rows(const input_t *in) { return in->columnar ? in->table->num_rows : in->inter.n; }
static uint32_t in_width(const input_t *in) { return in->node->num_output_attrs; }

/* ======================================================================
 * Unchained hash table (hashtable.h:63-505).
 * ====================================================================== */
#define TUPLES_PER_PARTITION 262144u /* (128 MiB LLC / 32 threads / 2) / 8 B */

typedef struct { int32_t key; uint32_t rid; } tuple_t;
typedef struct { tuple_t *d; size_t n, cap; } tvec;
static inline void tv_push(tvec *v, tuple_t t) {
    if (v->n == v->cap) { v->cap = v->cap ? v->cap * 2 : 256; v->d = realloc(v->d, v->cap * sizeof(tuple_t)); }
    v->d[v->n++] = t;
}

typedef struct {
    uint64_t *dir;
    size_t nslots;
    int shift;
    int32_t *keys;
    uint32_t *rids;
    size_t size;
} htab_t;

static void ht_init(htab_t *ht, size_t build_rows) {
    size_t pow2 = 2048;
    while (pow2 < build_rows) pow2 <<= 1;
    ht->nslots = pow2;
    ht->dir = calloc(pow2, 8);
    ht->shift = __builtin_ctzll(pow2);
    ht->keys = NULL; ht->rids = NULL; ht->size = 0;
}
static void ht_free(htab_t *ht) { free(ht->dir); free(ht->keys); free(ht->rids); }
static inline size_t slot_for(const htab_t *ht, uint64_t h) { return (size_t)(h >> (64 - ht->shift)); }

static size_t num_partitions(const htab_t *ht, size_t n, int nt) {
    size_t minp = (n + TUPLES_PER_PARTITION - 1) / TUPLES_PER_PARTITION;
    if (minp < (size_t)nt) minp = (size_t)nt;
    size_t p = 1;
    while (p < minp) p <<= 1;
    return p < ht->nslots ? p : ht->nslots;
}

typedef struct {
    htab_t *ht;
    tvec *parts; /* [nt][P] */
    int nt;
    size_t P;
    int pbits;
    size_t *goff;
    /* partition-pass sources */
    const cj_column_desc *col;
    const uint32_t *page_off;
    const inter_t *inter;
    uint32_t key_col;
} build_job;

/* build_partition (hashtable.h:197-246) */
static void build_partition(build_job *j, size_t p) {
    htab_t *ht = j->ht;
    size_t spp = ht->nslots / j->P, s0 = p * spp, base = j->goff[p];
    if (j->goff[p + 1] == base) {
        for (size_t s = 0; s < spp; ++s) ht->dir[s0 + s] = (uint64_t)base << 16;
        return;
    }
    uint32_t *cnt = calloc(spp, 4), *off = malloc(spp * 4);
    for (int t = 0; t < j->nt; ++t) {
        tvec *v = &j->parts[(size_t)t * j->P + p];
        for (size_t i = 0; i < v->n; ++i) cnt[slot_for(ht, hash_key(v->d[i].key)) - s0]++;
    }
    uint32_t run = (uint32_t)base;
    for (size_t s = 0; s < spp; ++s) { off[s] = run; run += cnt[s]; cnt[s] = 0; }
    for (int t = 0; t < j->nt; ++t) {
        tvec *v = &j->parts[(size_t)t * j->P + p];
        for (size_t i = 0; i < v->n; ++i) {
            uint64_t h = hash_key(v->d[i].key);
            size_t ls = slot_for(ht, h) - s0;
            uint32_t idx = off[ls] + cnt[ls]++;
            ht->keys[idx] = v->d[i].key;
            ht->rids[idx] = v->d[i].rid;
            ht->dir[s0 + ls] |= bloom_of(h);
        }
    }
    for (size_t s = 0; s < spp; ++s) {
        uint64_t end = (uint64_t)off[s] + cnt[s];
        ht->dir[s0 + s] = (end << 16) | (ht->dir[s0 + s] & 0xFFFF);
    }
    free(cnt); free(off);
}

static inline void part_append(build_job *j, int t, int32_t key, uint32_t rid) {
    uint64_t h = hash_key(key);
    size_t p = j->pbits == 0 ? 0 : (size_t)(h >> (64 - j->pbits));
    tv_push(&j->parts[(size_t)t * j->P + p], (tuple_t){key, rid});
}

static void partition_columnar_task(void *arg, int t, int nt) {
    build_job *j = arg;
    size_t np = j->col->num_pages, batch = (np + nt - 1) / nt;
    size_t a = (size_t)t * batch, b = a + batch < np ? a + batch : np;
    for (size_t pg = a; pg < b; ++pg) {
        const uint8_t *p = j->col->pages[pg];
        uint32_t n = rd16(p), nv = rd16(p + 2), base = j->page_off[pg];
        const uint8_t *vals = p + 4;
        if (n == nv) {
            for (uint32_t r = 0; r < n; ++r) {
                int32_t k; memcpy(&k, vals + 4 * r, 4);
                part_append(j, t, k, base + r);
            }
        } else {
            const uint8_t *bm = p + PAGE - (n + 7) / 8;
            uint32_t vi = 0;
            for (uint32_t r = 0; r < n; ++r)
                if ((bm[r / 8] >> (r % 8)) & 1) {
                    int32_t k; memcpy(&k, vals + 4 * vi++, 4);
                    part_append(j, t, k, base + r);
                }
        }
    }
}
static void partition_inter_task(void *arg, int t, int nt) {
    build_job *j = arg;
    size_t n = j->inter->n, batch = (n + nt - 1) / nt;
    size_t a = (size_t)t * batch, b = a + batch < n ? a + batch : n;
    const int32_t *v = j->inter->cols[j->key_col].v;
    for (size_t i = a; i < b; ++i) part_append(j, t, v[i], (uint32_t)i); /* no NULL filter (hashtable.h:357) */
}
static void build_partitions_task(void *arg, int t, int nt) {
    build_job *j = arg;
    for (size_t p = (size_t)t; p < j->P; p += (size_t)nt) build_partition(j, p);
}

static void ht_finish_build(build_job *j, int nt_pool) {
    j->goff = calloc(j->P + 1, sizeof(size_t));
    for (size_t p = 0; p < j->P; ++p) {
        size_t s = 0;
        for (int t = 0; t < j->nt; ++t) s += j->parts[(size_t)t * j->P + p].n;
        j->goff[p + 1] = j->goff[p] + s;
    }
    size_t total = j->goff[j->P];
    if (total) {
        j->ht->keys = malloc(total * 4);
        j->ht->rids = malloc(total * 4);
        j->ht->size = total;
        int nt = j->nt < nt_pool ? j->nt : nt_pool;
        run_parallel(nt, build_partitions_task, j);
    }
    for (size_t i = 0; i < (size_t)j->nt * j->P; ++i) free(j->parts[i].d);
    free(j->parts); free(j->goff);
}

/* build_columnar (hashtable.h:405-505) with the thread hint of 8 (hash_join.h:55). */
static void ht_build_columnar(htab_t *ht, const cj_column_desc *col, int pool) {
    if (col->num_pages == 0) return;
    uint32_t *page_off = malloc(col->num_pages * 4u);
    size_t total = 0;
    for (uint32_t pg = 0; pg < col->num_pages; ++pg) { page_off[pg] = (uint32_t)total; total += rd16(col->pages[pg]); }
    if (total == 0) { free(page_off); return; }
    int nt = 8 < pool ? 8 : pool;
    if (nt < 1) nt = 1;
    if (col->num_pages < 16) nt = 1;
    build_job j = {0};
    j.ht = ht; j.nt = nt; j.col = col; j.page_off = page_off;
    j.P = num_partitions(ht, total, nt);
    j.pbits = __builtin_ctzll(j.P);
    j.parts = calloc((size_t)nt * j.P, sizeof(tvec));
    run_parallel(nt, partition_columnar_task, &j);
    ht_finish_build(&j, nt);
    free(page_off);
}

/* build_intermediate (hashtable.h:326-390). */
static void ht_build_inter(htab_t *ht, const inter_t *in, uint32_t key_col, int pool) {
    if (in->n == 0) return;
    int nt = in->n < 1000 ? 1 : pool;
    build_job j = {0};
    j.ht = ht; j.nt = nt; j.inter = in; j.key_col = key_col;
    j.P = num_partitions(ht, in->n, nt);
    j.pbits = __builtin_ctzll(j.P);
    j.parts = calloc((size_t)nt * j.P, sizeof(tvec));
    run_parallel(nt, partition_inter_task, &j);
    ht_finish_build(&j, nt);
}

/* find_indices (hashtable.h:298-313) */
static inline void ht_find(const htab_t *ht, int32_t key, uint64_t *s, uint64_t *e) {
    *s = *e = 0;
    if (ht->size == 0) return;
    uint64_t h = hash_key(key);
    size_t slot = slot_for(ht, h);
    uint64_t entry = ht->dir[slot];
    uint16_t tag = bloom_of(h);
    if ((entry & tag) != tag) return;
    *e = entry >> 16;
    *s = slot == 0 ? 0 : ht->dir[slot - 1] >> 16;
}

/* ======================================================================
 * Probe (hash_join.h:87-244) and nested loop (nested_loop.h:107-244).
 * Matches go to per-thread (build_rid, probe_rid) buffers
 * (ThreadLocalMatchBuffer, match_collector.h:66-221).
 * ====================================================================== */
typedef struct { vec32 l, r; } mbuf_t;

typedef struct {
    const htab_t *ht;
    const cj_column_desc *pcol;   /* columnar probe */
    const uint32_t *page_off;
    const int32_t *pvals;         /* intermediate probe */
    uint64_t pn;
    mbuf_t *bufs;
    atomic_size_t next;
    /* nested loop */
    int nl;
    int32_t bv[8];
    uint32_t bid[8];
    int bcount;
} probe_job;

static inline void emit(probe_job *j, mbuf_t *b, int32_t key, uint32_t prid) {
    if (j->nl) {
        for (int k = 0; k < j->bcount; ++k)
            if (j->bv[k] == key) { v32_push(&b->l, j->bid[k]); v32_push(&b->r, prid); }
        return;
    }
    uint64_t s, e;
    ht_find(j->ht, key, &s, &e);
    for (uint64_t i = s; i < e; ++i)
        if (j->ht->keys[i] == key) { v32_push(&b->l, j->ht->rids[i]); v32_push(&b->r, prid); }
}

static void probe_columnar_task(void *arg, int t, int nt) {
    (void)nt;
    probe_job *j = arg;
    mbuf_t *b = &j->bufs[t];
    for (;;) {
        size_t pg = atomic_fetch_add(&j->next, 1);
        if (pg >= j->pcol->num_pages) break;
        const uint8_t *p = j->pcol->pages[pg];
        uint32_t n = rd16(p), nv = rd16(p + 2), prid = j->page_off[pg];
        const uint8_t *vals = p + 4;
        if (n == nv) {
            for (uint32_t r = 0; r < n; ++r, ++prid) {
                int32_t k; memcpy(&k, vals + 4 * r, 4);
                emit(j, b, k, prid);
            }
        } else {
            const uint8_t *bm = p + PAGE - (n + 7) / 8;
            uint32_t vi = 0;
            for (uint32_t r = 0; r < n; ++r, ++prid)
                if ((bm[r / 8] >> (r % 8)) & 1) {
                    int32_t k; memcpy(&k, vals + 4 * vi++, 4);
                    emit(j, b, k, prid);
                }
        }
    }
}
static void probe_inter_task(void *arg, int t, int nt) {
    probe_job *j = arg;
    mbuf_t *b = &j->bufs[t];
    /* 16 KiB IR pages of 4096 cells are the unit of work (hash_join.h:104-109) */
    size_t npages = (j->pn + 4095) / 4096;
    (void)nt;
    for (;;) {
        size_t pg = atomic_fetch_add(&j->next, 1);
        if (pg >= npages) break;
        uint64_t a = pg * 4096, e = a + 4096 < j->pn ? a + 4096 : j->pn;
        for (uint64_t i = a; i < e; ++i)
            if (j->pvals[i] != I32_NULL) emit(j, b, j->pvals[i], (uint32_t)i);
    }
}

/* ======================================================================
 * Output pages (page builders, page_builders.h:92-391): each thread packs
 * its own rows with the canonical capacity rules; page lists are
 * concatenated in thread order (materialize.h:173-185).
 * ====================================================================== */
typedef struct { uint8_t *pages; size_t npages, cap; } plist_t;

static void pl_reserve(plist_t *pl, size_t extra) {
    if (pl->npages + extra > pl->cap) {
        size_t nc = pl->cap ? pl->cap * 2 : 16;
        while (nc < pl->npages + extra) nc *= 2;
        pl->pages = realloc(pl->pages, nc * PAGE);
        pl->cap = nc;
    }
}

typedef struct {
    const cj_plan_desc *plan;
    const input_t *bin, *pin;
    const mbuf_t *bufs;
    int nbufs;
    /* column source */
    int from_build;
    int columnar;
    const pidx_t *ix;
    const icol_t *ic;
    const cj_column_desc *str_src; /* VARCHAR bytes source column */
    pidx_t *str_ix;
    int32_t type;
    plist_t *out;   /* per thread */
    int err;
} mat_job;

static void mat_task(void *arg, int t, int nt) {
    (void)nt;
    mat_job *j = arg;
    if (t >= j->nbufs) return;
    const mbuf_t *b = &j->bufs[t];
    const vec32 *ids = j->from_build ? &b->l : &b->r;
    size_t n = ids->n;
    if (!n) return;
    cursor_t cur = {0}, scur = {0};
    uint8_t *valid = malloc(n);
    plist_t *pl = &j->out[t];
    if (j->type == CJ_VARCHAR) {
        uint64_t *offs = malloc((n + 1) * 8);
        size_t cap = 1 << 16, nb = 0;
        char *bytes = malloc(cap);
        offs[0] = 0;
        for (size_t i = 0; i < n; ++i) {
            uint32_t rid = ids->d[i];
            int32_t v = j->columnar ? read_v32(j->ix, rid, &cur) : j->ic->v[rid];
            valid[i] = v != I32_NULL;
            if (valid[i]) {
                uint32_t pg = (uint32_t)v & 0x7FFFF, slot = ((uint32_t)v >> 19) & 0x1FFF;
                const cj_column_desc *sc = j->str_src;
                const uint8_t *p = sc->pages[pg];
                if (slot == LONG_SLOT) {
                    /* copy_long_string_pages (page_builders.h:344-360) */
                    for (uint32_t q = pg; q < sc->num_pages; ++q) {
                        const uint8_t *pp = sc->pages[q];
                        if (q != pg && rd16(pp) != 0xfffe) break;
                        size_t len = rd16(pp + 2);
                        while (nb + len > cap) { cap *= 2; bytes = realloc(bytes, cap); }
                        memcpy(bytes + nb, pp + 4, len);
                        nb += len;
                    }
                } else {
                    /* get_string_view (page_builders.h:26-37) */
                    uint32_t nv = rd16(p + 2);
                    uint32_t end = rd16(p + 4 + 2 * slot), start = slot ? rd16(p + 4 + 2 * (slot - 1)) : 0;
                    size_t len = end - start;
                    while (nb + len > cap) { cap *= 2; bytes = realloc(bytes, cap); }
                    memcpy(bytes + nb, p + 4 + 2 * (size_t)nv + start, len);
                    nb += len;
                }
            }
            offs[i + 1] = nb;
        }
        size_t maxp = n + nb / (PAGE - 8) + 2;
        pl_reserve(pl, maxp);
        size_t got = cjo_encode_varchar(n, offs, bytes, valid, pl->pages + pl->npages * PAGE, maxp);
        if (got == (size_t)-1) j->err = 1; else pl->npages += got;
        free(offs); free(bytes);
    } else {
        size_t w = j->type == CJ_INT32 ? 4 : 8;
        uint8_t *vals = malloc(n * w);
        for (size_t i = 0; i < n; ++i) {
            uint32_t rid = ids->d[i];
            if (j->type == CJ_INT32) {
                int32_t v = j->columnar ? read_v32(j->ix, rid, &cur) : j->ic->v[rid];
                /* value_t NULL sentinel: INT32_MIN reads as NULL (page_builders.h:123) */
                valid[i] = v != I32_NULL;
                memcpy(vals + 4 * i, &v, 4);
            } else {
                int64_t v = 0;
                if (j->columnar) valid[i] = (uint8_t)read_v64(j->ix, rid, &cur, &v);
                else { valid[i] = j->ic->wv[rid]; v = j->ic->w[rid]; }
                memcpy(vals + 8 * i, &v, 8);
            }
        }
        size_t per = w == 4 ? 1984 : 1007;
        size_t maxp = n / per + n / 60000 + 4;
        pl_reserve(pl, maxp);
        size_t got = cjo_encode_fixed(j->type, n, vals, valid, pl->pages + pl->npages * PAGE, maxp);
        if (got == (size_t)-1) j->err = 1; else pl->npages += got;
        free(vals);
    }
    (void)scur;
    free(valid);
}

/* intermediate gather (construct_intermediate.h:167-218) */
typedef struct {
    const mbuf_t *bufs;
    const size_t *starts;
    int nbufs;
    int from_build, columnar;
    const pidx_t *ix;
    const icol_t *ic;
    icol_t *dst;
} gather_job;

static void gather_task(void *arg, int t, int nt) {
    (void)nt;
    gather_job *j = arg;
    if (t >= j->nbufs) return;
    const mbuf_t *b = &j->bufs[t];
    const vec32 *ids = j->from_build ? &b->l : &b->r;
    size_t k = j->starts[t];
    cursor_t cur = {0};
    for (size_t i = 0; i < ids->n; ++i, ++k) {
        uint32_t rid = ids->d[i];
        if (j->dst->type == CJ_INT64 || j->dst->type == CJ_FP64) {
            int64_t v = 0;
            int ok;
            if (j->columnar) ok = read_v64(j->ix, rid, &cur, &v);
            else { ok = j->ic->wv[rid]; v = j->ic->w[rid]; }
            j->dst->w[k] = v;
            j->dst->wv[k] = (uint8_t)ok;
        } else {
            j->dst->v[k] = j->columnar ? read_v32(j->ix, rid, &cur) : j->ic->v[rid];
        }
    }
}

/* ======================================================================
 * execute_impl (execute.cpp:207-300)
 * ====================================================================== */
typedef struct {
    const cj_plan_desc *plan;
    int threads;
    cjo_result *root_out;
} exec_t;

static int exec_node(exec_t *x, uint64_t node_idx, int is_root, inter_t *out);

/* resolve_join_input (execute.cpp:84-99) */
static int resolve(exec_t *x, uint64_t idx, input_t *in) {
    const cj_plan_desc *plan = x->plan;
    if (idx >= plan->num_nodes) return fail(CJ_ERR_INVALID, "node index %llu out of range", (unsigned long long)idx);
    memset(in, 0, sizeof *in);
    in->node = &plan->nodes[idx];
    if (in->node->kind == CJ_NODE_SCAN) {
        if (in->node->base_table_id >= plan->num_inputs) return fail(CJ_ERR_INVALID, "bad base_table_id");
        in->columnar = 1;
        in->table = &plan->inputs[in->node->base_table_id];
        in->table_id = (uint8_t)in->node->base_table_id;
        for (uint32_t a = 0; a < in->node->num_output_attrs; ++a) {
            const cj_attr *at = &in->node->output_attrs[a];
            if (at->index >= in->table->num_columns) return fail(CJ_ERR_INVALID, "scan attr out of range");
            if (in->table->columns[at->index].type != at->type) return fail(CJ_ERR_UNSUPPORTED, "scan attr type mismatch");
        }
        return 0;
    }
    return exec_node(x, idx, 0, &in->inter);
}

static int32_t src_type(const input_t *in, uint32_t local) {
    if (in->columnar) return in->table->columns[in->node->output_attrs[local].index].type;
    return in->inter.cols[local].type;
}

static int exec_node(exec_t *x, uint64_t node_idx, int is_root, inter_t *out) {
    const cj_plan_desc *plan = x->plan;
    const cj_node_desc *node = &plan->nodes[node_idx];
    if (node->kind != CJ_NODE_JOIN) return fail(CJ_ERR_UNSUPPORTED, "root must be a join node");
    int rc;
    input_t L, R;
    if ((rc = resolve(x, node->left, &L))) return rc;
    if ((rc = resolve(x, node->right, &R))) { if (!L.columnar) inter_free(&L.inter); return rc; }
    if (node->left_attr >= in_width(&L) || node->right_attr >= in_width(&R)) {
        rc = fail(CJ_ERR_INVALID, "join attr out of range");
        goto done_inputs;
    }
    if (src_type(&L, (uint32_t)node->left_attr) != CJ_INT32 || src_type(&R, (uint32_t)node->right_attr) != CJ_INT32) {
        rc = fail(CJ_ERR_UNSUPPORTED, "join keys must be INT32");
        goto done_inputs;
    }
    /* select_build_probe_side (join_setup.h:108-136): smaller side builds,
     * ties go left; JoinNode::build_left is not consulted. */
    int build_left = in_rows(&L) <= in_rows(&R);
    input_t *B = build_left ? &L : &R, *P = build_left ? &R : &L;
    uint32_t battr = (uint32_t)(build_left ? node->left_attr : node->right_attr);
    uint32_t pattr = (uint32_t)(build_left ? node->right_attr : node->left_attr);
    uint32_t lw = in_width(&L), bw = in_width(B);
    uint32_t nout = node->num_output_attrs;
    uint32_t *remap = malloc((nout ? nout : 1) * 4);
    for (uint32_t a = 0; a < nout; ++a) {
        uint64_t ci = node->output_attrs[a].index;
        if (ci >= (uint64_t)lw + in_width(&R)) { free(remap); rc = fail(CJ_ERR_INVALID, "output attr out of range"); goto done_inputs; }
        remap[a] = build_left ? (uint32_t)ci : (ci < lw ? bw + (uint32_t)ci : (uint32_t)(ci - lw));
    }
    int pool = x->threads;
    uint64_t brows = in_rows(B);
    int use_nl = brows < 8; /* execute.cpp:236-238 */

    probe_job pj = {0};
    pj.bufs = calloc(pool, sizeof(mbuf_t));
    htab_t ht = {0};
    if (!use_nl) {
        ht_init(&ht, brows);
        if (B->columnar) ht_build_columnar(&ht, &B->table->columns[B->node->output_attrs[battr].index], pool);
        else ht_build_inter(&ht, &B->inter, battr, pool);
        pj.ht = &ht;
    } else {
        /* visit_rows (nested_loop.h:50-95): up to 8 non-NULL build keys */
        pj.nl = 1;
        if (B->columnar) {
            const cj_column_desc *c = &B->table->columns[B->node->output_attrs[battr].index];
            uint32_t rid = 0;
            for (uint32_t pg = 0; pg < c->num_pages; ++pg) {
                const uint8_t *p = c->pages[pg];
                uint32_t n = rd16(p), nv = rd16(p + 2), vi = 0;
                const uint8_t *bm = p + PAGE - (n + 7) / 8;
                for (uint32_t r = 0; r < n; ++r, ++rid) {
                    int ok = n == nv || ((bm[r / 8] >> (r % 8)) & 1);
                    if (!ok) continue;
                    int32_t k; memcpy(&k, p + 4 + 4 * (size_t)(n == nv ? r : vi), 4);
                    vi++;
                    if (pj.bcount < 8) { pj.bv[pj.bcount] = k; pj.bid[pj.bcount] = rid; pj.bcount++; }
                }
            }
        } else {
            for (uint64_t i = 0; i < B->inter.n; ++i) {
                int32_t k = B->inter.cols[battr].v[i];
                if (k != I32_NULL && pj.bcount < 8) { pj.bv[pj.bcount] = k; pj.bid[pj.bcount] = (uint32_t)i; pj.bcount++; }
            }
        }
    }
    uint32_t *poff = NULL;
    if (!(use_nl && (brows == 0 || in_rows(P) == 0))) {
        if (P->columnar) {
            const cj_column_desc *c = &P->table->columns[P->node->output_attrs[pattr].index];
            poff = malloc(((size_t)c->num_pages + 1) * 4);
            uint32_t run = 0;
            for (uint32_t pg = 0; pg < c->num_pages; ++pg) { poff[pg] = run; run += rd16(c->pages[pg]); }
            pj.pcol = c; pj.page_off = poff;
            run_parallel(pool, probe_columnar_task, &pj);
        } else {
            pj.pvals = P->inter.cols[pattr].v;
            pj.pn = P->inter.n;
            run_parallel(pool, probe_inter_task, &pj);
        }
    }
    free(poff);
    if (!use_nl) ht_free(&ht);

    size_t total = 0;
    size_t *starts = malloc((size_t)pool * sizeof(size_t));
    for (int t = 0; t < pool; ++t) { starts[t] = total; total += pj.bufs[t].l.n; }

    rc = 0;
    pidx_t *bix = calloc(in_width(B) + 1, sizeof(pidx_t)), *pix = calloc(in_width(P) + 1, sizeof(pidx_t));
    if (is_root) {
        cjo_result *res = x->root_out;
        res->num_rows = total;
        res->num_columns = nout;
        res->types = calloc(nout + 1, 4);
        res->num_pages = calloc(nout + 1, 4);
        res->pages = calloc(nout + 1, sizeof(void **));
        for (uint32_t a = 0; a < nout && !rc; ++a) {
            res->types[a] = node->output_attrs[a].type;
            if (total == 0) continue; /* create_empty_result (materialize.h:61-69) */
            uint32_t ci = remap[a];
            int fb = ci < bw;
            input_t *S = fb ? B : P;
            uint32_t local = fb ? ci : ci - bw;
            pidx_t *ixs = fb ? bix : pix;
            if (src_type(S, local) != node->output_attrs[a].type) { rc = fail(CJ_ERR_UNSUPPORTED, "output attr type mismatch"); break; }
            mat_job mj = {0};
            mj.plan = plan; mj.bufs = pj.bufs; mj.nbufs = pool; mj.from_build = fb; mj.columnar = S->columnar;
            mj.type = node->output_attrs[a].type;
            if (S->columnar) {
                const cj_column_desc *c = &S->table->columns[S->node->output_attrs[local].index];
                if (!ixs[local].cum) pidx_build(&ixs[local], c);
                mj.ix = &ixs[local];
                mj.str_src = c;
            } else {
                mj.ic = &S->inter.cols[local];
                if (mj.type == CJ_VARCHAR) mj.str_src = &plan->inputs[mj.ic->src_table].columns[mj.ic->src_col];
            }
            mj.out = calloc(pool, sizeof(plist_t));
            run_parallel(pool, mat_task, &mj);
            size_t np = 0;
            for (int t = 0; t < pool; ++t) np += mj.out[t].npages;
            uint8_t *blk = malloc(np * PAGE + 1);
            res->pages[a] = malloc((np + 1) * sizeof(void *));
            size_t k = 0;
            for (int t = 0; t < pool; ++t) {
                memcpy(blk + k * PAGE, mj.out[t].pages, mj.out[t].npages * PAGE);
                k += mj.out[t].npages;
                free(mj.out[t].pages);
            }
            for (size_t q = 0; q < np; ++q) res->pages[a][q] = blk + q * PAGE;
            if (np == 0) free(blk);
            res->num_pages[a] = (uint32_t)np;
            free(mj.out);
            if (mj.err) rc = fail(CJ_ERR_INVALID, "page overflow");
        }
    } else {
        out->n = total;
        out->ncols = nout;
        out->cols = calloc(nout + 1, sizeof(icol_t));
        for (uint32_t a = 0; a < nout && !rc; ++a) {
            uint32_t ci = remap[a];
            int fb = ci < bw;
            input_t *S = fb ? B : P;
            uint32_t local = fb ? ci : ci - bw;
            icol_t *d = &out->cols[a];
            d->type = src_type(S, local);
            /* initialize_output_columns (join_setup.h:180-215): provenance */
            if (S->columnar) { d->src_table = S->table_id; d->src_col = (uint8_t)S->node->output_attrs[local].index; }
            else { d->src_table = S->inter.cols[local].src_table; d->src_col = S->inter.cols[local].src_col; }
            if (d->type == CJ_INT64 || d->type == CJ_FP64) { d->w = malloc(total * 8 + 8); d->wv = malloc(total + 1); }
            else d->v = malloc(total * 4 + 4);
            if (!total) continue;
            gather_job gj = {0};
            gj.bufs = pj.bufs; gj.starts = starts; gj.nbufs = pool; gj.from_build = fb; gj.columnar = S->columnar; gj.dst = d;
            if (S->columnar) {
                const cj_column_desc *c = &S->table->columns[S->node->output_attrs[local].index];
                pidx_t *ixs = fb ? bix : pix;
                if (!ixs[local].cum) pidx_build(&ixs[local], c);
                gj.ix = &ixs[local];
            } else gj.ic = &S->inter.cols[local];
            run_parallel(pool, gather_task, &gj);
        }
        if (rc) inter_free(out);
    }
    for (uint32_t a = 0; a < in_width(B); ++a) pidx_free(&bix[a]);
    for (uint32_t a = 0; a < in_width(P); ++a) pidx_free(&pix[a]);
    free(bix); free(pix); free(starts);
    for (int t = 0; t < pool; ++t) { free(pj.bufs[t].l.d); free(pj.bufs[t].r.d); }
    free(pj.bufs);
    free(remap);
done_inputs:
    if (!L.columnar) inter_free(&L.inter);
    if (!R.columnar) inter_free(&R.inter);
    return rc;
}

void cjo_result_free(cjo_result *res) {
    if (!res) return;
    for (uint32_t a = 0; a < res->num_columns; ++a) {
        if (res->pages && res->pages[a]) {
            if (res->num_pages[a]) free(res->pages[a][0]);
            free(res->pages[a]);
        }
    }
    free(res->types); free(res->num_pages); free(res->pages);
    memset(res, 0, sizeof *res);
}

int cjo_execute(const cj_plan_desc *plan, int threads, cjo_result *out) {
    pthread_once(&tab_once, init_tables);
    if (g_use_hw < 0) g_use_hw = have_sse42();
    memset(out, 0, sizeof *out);
    if (!plan || plan->root >= plan->num_nodes) return fail(CJ_ERR_INVALID, "bad plan root");
    if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (threads > 256) threads = 256;
    exec_t x = {plan, threads, out};
    int rc = exec_node(&x, plan->root, 1, NULL);
    if (rc) cjo_result_free(out);
    return rc;
}

/* Self-check hook for tests: count mismatches between the table-driven
 * CRC-32C and the SSE4.2 instruction over n pseudo-random words
 * (-1 if the instruction is unavailable). */
int64_t cjo_crc32c_hw_mismatches(uint64_t n) {
    pthread_once(&tab_once, init_tables);
    if (!have_sse42()) return -1;
    int64_t bad = 0;
    uint64_t s = 12345;
    for (uint64_t i = 0; i < n; ++i) {
        s = mix64(s + 0x9e3779b97f4a7c15ull);
        uint32_t v = (uint32_t)s;
        if (crc_hw(v) != cjo_crc32c_u32(0, v)) bad++;
    }
    return bad;
}
