/*
 * cj_workload.h -- synthetic tables for the benchmark and tests, and a
 * host-side page writer (libcalamarijoins_amd.so).
 *
 * Not part of the drop-in boundary (no reference counterpart beyond the
 * page writer ColumnInserter, reference include/data_model/plan.h:187-366).
 * Every value is a pure function of (seed, column stream, row), so CPU and
 * GPU runs regenerate identical tables:
 *     u(row) = mix64(seed ^ stream * 0x9E3779B97F4A7C15 + (row + 1) * 0x9E3779B97F4A7C15)
 * with mix64 the SplitMix64 finaliser.
 */
#ifndef CJ_WORKLOAD_H
#define CJ_WORKLOAD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Column generators. */
enum {
    CJ_GEN_PERM = 0,      /* INT32 key = perm_stream(row_id) over [0, domain)     */
    CJ_GEN_UNIFORM = 1,   /* INT32 uniform in [lo, hi)                            */
    CJ_GEN_ZIPF_FK = 2,   /* INT32 = perm(rank), rank ~ Zipf(s) over [0, domain);
                             w.p. miss_frac a miss uniform in [miss_lo, miss_hi)  */
    CJ_GEN_I32_FULL = 3,  /* INT32 full range                                     */
    CJ_GEN_I64_FULL = 4,  /* INT64 full range                                     */
    CJ_GEN_F64_SCALED = 5,/* FP64 = (u53 / 2^53) * scale                          */
    CJ_GEN_VARCHAR = 6,   /* printable ASCII, length U[lo, hi]; w.p. long_frac a
                             long string of length U[long_lo, long_hi]            */
    CJ_GEN_UNIFORM_FK = 7,/* INT32 = perm(uniform rank in [0, domain))            */
    CJ_GEN_VARCHAR_VOCAB = 8 /* VARCHAR word of `vocab`, chosen uniformly          */
};

typedef struct cj_colgen {
    int32_t kind;
    int32_t type;          /* output DataType (CJ_INT32 ...)                 */
    uint64_t stream;       /* column stream id                               */
    double null_frac;      /* NULL probability                               */
    uint64_t domain;       /* PERM / ZIPF_FK / UNIFORM_FK domain size        */
    uint64_t perm_stream;  /* stream id of the permutation (PERM / FK kinds) */
    double zipf_s;
    double miss_frac;
    int64_t lo, hi;        /* UNIFORM range; VARCHAR length range            */
    int64_t miss_lo, miss_hi;
    double scale;
    double long_frac;
    int64_t long_lo, long_hi;
    const char *vocab;     /* VARCHAR_VOCAB: words, each ended by a NUL byte  */
    uint64_t vocab_bytes;
} cj_colgen;

/* A generated table: page blocks owned by the library. */
typedef struct cj_gen_table {
    uint64_t num_rows;
    uint32_t num_columns;
    int32_t *types;
    uint32_t *num_pages;
    uint8_t **pages;       /* per column: num_pages * 8192 contiguous bytes  */
    void *opaque;
} cj_gen_table;

/* Rows i in [0, candidate_rows) are kept w.p. keep_frac (seeded); the
 * generators see the candidate row index, so a filtered dimension keeps
 * the keys perm(i) of its surviving rows. */
int cj_gen_table_create(uint64_t candidate_rows, double keep_frac, uint64_t seed, uint32_t num_columns,
                        const cj_colgen *cols, int32_t threads, cj_gen_table *out);
void cj_gen_table_free(cj_gen_table *t);

/* Host page writer.  Same capacity rules as the API's ColumnInserter;
 * writes pages back to back into out (room for max_pages pages). */
int cj_pages_encode_fixed(int32_t type, uint64_t n_rows, const void *values, const uint8_t *valid, void *out,
                          uint64_t max_pages, uint64_t *num_pages);
int cj_pages_encode_varchar(uint64_t n_rows, const uint64_t *offsets, const char *bytes, const uint8_t *valid,
                            void *out, uint64_t max_pages, uint64_t *num_pages);

#ifdef __cplusplus
}
#endif
#endif
