/*
 * cj_api.h -- C ABI of the MI355X join engine (libcalamarijoins_amd.so).
 *
 * This is the drop-in boundary for the join pipeline behind the contest API
 *
 *     void*         Contest::build_context();
 *     void          Contest::destroy_context(void*);
 *     ColumnarTable Contest::execute(const Plan&, void*, TimingStats*);
 *
 * (reference: include/data_model/plan.h:380-385, src/execute.cpp:311-329).
 * The C++ header include/cj/plan.h re-declares that API on top of these
 * functions; any other language binds the functions below directly.
 *
 * All signatures use plain pointers and sizes.  Pages are the contest's
 * 8192-byte columnar pages (reference README.md:65-104, plan.h:78-81):
 *   u16 n_rows | u16 n_vals | values (INT32 at +4, INT64/FP64 at +8) or
 *   u16 VARCHAR end offsets at +4 with chars at +4+2*n_vals | NULL bitmap
 *   in the last ceil(n_rows/8) bytes (LSB first).  n_rows 0xffff / 0xfffe
 *   mark the first / continuation page of a long string (u16 char count at
 *   +2, chars at +4).
 *
 * Error convention: every function returns CJ_OK (0) or a negative status;
 * cj_last_error() gives the message of the calling thread's last failure.
 */
#ifndef CJ_API_H
#define CJ_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CJ_PAGE_SIZE 8192

/* Status codes. */
enum {
    CJ_OK = 0,
    CJ_ERR_INVALID = -1,     /* malformed plan / table / argument          */
    CJ_ERR_UNSUPPORTED = -2, /* shape the engine rejects (e.g. non-INT32 key) */
    CJ_ERR_DEVICE = -3,      /* HIP runtime failure                         */
    CJ_ERR_OOM = -4,         /* device memory budget exhausted              */
    CJ_ERR_BUSY = -5         /* serving queue full / not admitted           */
};

/* DataType, same numbering as the reference enum (foundation/attribute.h:7-12). */
enum { CJ_INT32 = 0, CJ_INT64 = 1, CJ_FP64 = 2, CJ_VARCHAR = 3 };

/* Plan node kinds (reference NodeType, plan.h:46-49: HashJoin, Scan). */
enum { CJ_NODE_JOIN = 0, CJ_NODE_SCAN = 1 };

typedef struct cj_ctx cj_ctx;

/* One column of a ColumnarTable: reference Column{type, pages} (plan.h:84-136). */
typedef struct cj_column_desc {
    int32_t type;             /* CJ_INT32 .. CJ_VARCHAR                 */
    uint32_t num_pages;
    const void *const *pages; /* host pointers to 8 KiB pages, 8-aligned */
} cj_column_desc;

/* Reference ColumnarTable{num_rows, columns} (plan.h:139-142). */
typedef struct cj_table_desc {
    uint64_t num_rows;
    uint32_t num_columns;
    const cj_column_desc *columns;
} cj_table_desc;

/* One entry of PlanNode::output_attrs: tuple<size_t, DataType>. */
typedef struct cj_attr {
    uint64_t index;
    int32_t type;
} cj_attr;

/* Reference PlanNode{variant<ScanNode, JoinNode>, output_attrs} (plan.h:51-70). */
typedef struct cj_node_desc {
    int32_t kind;           /* CJ_NODE_SCAN or CJ_NODE_JOIN           */
    int32_t build_left;     /* JoinNode::build_left (a hint; see below) */
    uint64_t base_table_id; /* ScanNode::base_table_id                */
    uint64_t left, right;   /* JoinNode children (node indices)       */
    uint64_t left_attr, right_attr;
    uint32_t num_output_attrs;
    const cj_attr *output_attrs;
} cj_node_desc;

/* Reference Plan{nodes, inputs, root} (plan.h:149-185). */
typedef struct cj_plan_desc {
    uint32_t num_nodes;
    const cj_node_desc *nodes;
    uint32_t num_inputs;
    const cj_table_desc *inputs;
    uint64_t root;
} cj_plan_desc;

/* Reference Contest::TimingStats (plan.h:370-378), plus device detail.
 * The int64 phase fields are whole milliseconds, truncated as the reference
 * truncates them (execute.cpp:123-190, 224-266); phase_ms holds the same
 * phases untruncated.  Phases are measured with hipEvents on the execution
 * stream (upload of non-resident pages with the host clock) and do not
 * overlap, so their sum never exceeds total_execution_ms:
 *   setup            upload of non-resident input pages, page decode, page indexes
 *   hashtable_build  direct-map / hash-table build of a join's build side
 *   hash_join_probe  counting pass, offsets, (build, probe) row-id pairs
 *   nested_loop_join the same for build sides of fewer than 8 rows
 *   materialize      root result pages, and their copy to host memory
 *   intermediate     row-id lists of non-root join results */
enum {
    CJ_PHASE_HASHTABLE_BUILD = 0,
    CJ_PHASE_HASH_JOIN_PROBE = 1,
    CJ_PHASE_NESTED_LOOP_JOIN = 2,
    CJ_PHASE_MATERIALIZE = 3,
    CJ_PHASE_SETUP = 4,
    CJ_PHASE_INTERMEDIATE = 5,
    CJ_NUM_PHASES = 6
};
typedef struct cj_timing {
    int64_t hashtable_build_ms;
    int64_t hash_join_probe_ms;
    int64_t nested_loop_join_ms;
    int64_t materialize_ms;
    int64_t setup_ms;
    int64_t total_execution_ms;
    int64_t intermediate_ms;
    double device_ms;   /* first kernel .. last output page in HBM (hipEvents) */
    double h2d_ms;      /* page upload of non-resident inputs              */
    double d2h_ms;      /* result page download (host output only)         */
    uint64_t bytes_alg; /* algorithmic bytes: input pages read + output written */
    double queue_ms;    /* served plans: submit .. admitted (0 for cj_execute) */
    double service_ms;  /* served plans: admitted .. result ready              */
    double phase_ms[CJ_NUM_PHASES]; /* CJ_PHASE_* order, untruncated             */
    int32_t served_alone; /* served plans: 1 when the plan ran out of device
                           * memory beside others and was run again alone */
} cj_timing;

/* A result column. Host output: pages[i] are host pointers valid until
 * cj_result_release.  Device output: device_pages is the HBM address of
 * num_pages contiguous 8 KiB pages (pages == NULL). */
typedef struct cj_result_column {
    int32_t type;
    uint32_t num_pages;
    void **pages;
    void *device_pages;
} cj_result_column;

typedef struct cj_result {
    uint64_t num_rows;
    uint32_t num_columns;
    cj_result_column *columns;
    void *opaque; /* owned by the library */
} cj_result;

/* Context options (zero-initialise, then set what you need). */
typedef struct cj_ctx_options {
    int32_t device;             /* HIP device ordinal                       */
    uint64_t device_budget;     /* bytes of HBM the context may use; 0=auto */
    int32_t num_streams;        /* concurrent execution slots; 0 = 4        */
    int32_t cache_decoded;      /* keep decoded columns of resident tables  */
    int32_t kernel_timing;      /* record per-kernel hipEvents (profiling)  */
    int32_t host_threads;       /* threads for host page staging; 0 = auto  */
} cj_ctx_options;

/* Execution flags. */
enum {
    CJ_EXEC_OUTPUT_DEVICE = 1, /* keep result pages in HBM (no D2H)          */
    CJ_EXEC_NO_FAST_PATH = 2,  /* force the general (row-id list) path        */
    CJ_EXEC_NO_DIRECT = 4,     /* never use the direct-mapped build side      */
    CJ_EXEC_NO_ALIGNED = 8     /* never write page-aligned root output        */
};

/* ---- context: replaces Contest::build_context / destroy_context
 *      (reference src/execute.cpp:327-329). ------------------------------- */
int cj_ctx_create(const cj_ctx_options *opts, cj_ctx **out);
int cj_ctx_destroy(cj_ctx *ctx);
const char *cj_last_error(void);
/* Device name and CU count of the context's GPU (diagnostics). */
int cj_ctx_device_info(cj_ctx *ctx, char *name, size_t name_cap, int32_t *num_cus);
/* Device memory of the context's pool (diagnostics): bytes obtained from the
 * driver, bytes in use now, the most in use at once since the last reset
 * (reset_peak != 0 resets it to the current use), and the budget. */
int cj_ctx_memory(cj_ctx *ctx, int32_t reset_peak, uint64_t *reserved, uint64_t *in_use, uint64_t *peak_in_use,
                  uint64_t *budget);

/* ---- HBM-resident base tables.  A registered table's pages are uploaded
 *      once; later plans whose inputs use the same page pointers read the
 *      HBM copy.  The caller must not modify or free registered pages until
 *      cj_table_unregister.  (No reference counterpart: the reference reads
 *      host pages in place, plan.h:84-142.) ------------------------------- */
int cj_table_register(cj_ctx *ctx, const cj_table_desc *table, uint64_t *out_id);
int cj_table_unregister(cj_ctx *ctx, uint64_t id);

/* ---- execute: replaces Contest::execute (reference src/execute.cpp:311-325).
 *      Synchronous.  The result is owned by the caller until
 *      cj_result_release.  timing may be NULL. ----------------------------- */
int cj_execute(cj_ctx *ctx, const cj_plan_desc *plan, uint32_t flags,
               cj_result *out, cj_timing *timing);
int cj_result_release(cj_ctx *ctx, cj_result *res);

/* ---- execute without allocating.  No counterpart in the reference: for
 *      callers that cannot allocate device memory (code under HIP graph
 *      capture, tight serving loops, harnesses that forbid hipMalloc).  The
 *      same plans and results as cj_execute with CJ_EXEC_OUTPUT_DEVICE, but
 *      the call allocates nothing: every temporary lives in a fixed device
 *      array of CJ_NOALLOC_SCRATCH_BYTES that the library holds on each GPU
 *      it runs on until the process ends, and result column i is written
 *      to the caller's device buffer out_pages[i] (room for
 *      out_cap_pages[i] pages of 8 KiB).  Limits: every input must be a
 *      registered (resident) table; a plan whose temporaries exceed the
 *      scratch, or a column whose pages exceed its buffer, fails with
 *      CJ_ERR_OOM and a message naming the size it needed (a column of
 *      8-byte values may need one page of room beyond the pages it fills;
 *      the root uses the block or the general writers, whose page counts
 *      follow from the result size).  The
 *      call is synchronous.  All calls on one GPU share its scratch, which
 *      keeps the last call's contents (each call clears what it needs):
 *      the library runs them one at a time, so calls that would overlap on
 *      one GPU wait for each other.  On success *out_rows and
 *      out_num_pages[0 .. num_out) give the result. */
#define CJ_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int cj_execute_noalloc(cj_ctx *ctx, const cj_plan_desc *plan, uint32_t flags, void *const *out_pages,
                       const uint32_t *out_cap_pages, uint32_t num_out, uint64_t *out_rows,
                       uint32_t *out_num_pages);
/* Copy a device-output result into host pages (fills columns[i].pages). */
int cj_result_to_host(cj_ctx *ctx, cj_result *res);

/* Per-kernel device time accumulated while kernel_timing is on. */
typedef struct cj_kernel_stat {
    char name[48];
    uint64_t launches;
    double total_ms;
    uint64_t bytes; /* algorithmic bytes moved by those launches */
} cj_kernel_stat;
int cj_kernel_stats(cj_ctx *ctx, cj_kernel_stat *out, uint32_t cap, uint32_t *count);
int cj_kernel_stats_reset(cj_ctx *ctx);

/* ---- serving loop (no counterpart in the reference, which runs one plan
 *      at a time: execute.cpp:311-325).  Plans submitted as they arrive are
 *      admitted by estimated device footprint within admit_bytes and run
 *      concurrently, one worker per stream (at most num_streams of the
 *      context).  A plan that runs out of device memory while others run is
 *      queued again at the head and run alone (nothing else is admitted
 *      until it finishes; cj_timing.served_alone says so).  The plan's
 *      descriptor arrays are copied at submit; its pages must stay valid
 *      until cj_wait returns. --------------------------------------------- */
typedef struct cj_serve_options {
    int32_t workers;      /* concurrent plans (<= the context's num_streams)  */
    uint64_t admit_bytes; /* footprint budget of the admitted plans           */
} cj_serve_options;
int cj_serve_start(cj_ctx *ctx, const cj_serve_options *opts);
int cj_submit(cj_ctx *ctx, const cj_plan_desc *plan, uint32_t flags, uint64_t *ticket);
/* CJ_OK and the result (owned by the caller), CJ_ERR_BUSY when timeout_ms
 * (>= 0) passes first, or the plan's own error.  timeout_ms < 0 waits. */
int cj_wait(cj_ctx *ctx, uint64_t ticket, int64_t timeout_ms, cj_result *out, cj_timing *timing);
/* Finishes every submitted plan, then stops the workers.  Results not yet
 * collected are released. */
int cj_serve_stop(cj_ctx *ctx);
/* Counters of the running serving loop since cj_serve_start. */
typedef struct cj_serve_stats {
    uint64_t admitted;             /* plans started (a plan run again alone counts twice) */
    uint64_t budget_waits;         /* plans that waited at least once because the budget was full */
    uint64_t retried_alone;        /* plans run again alone after running out of memory */
    uint64_t peak_admitted_bytes;  /* largest sum of admitted footprints */
    int32_t peak_running;          /* most plans running at once */
} cj_serve_stats;
int cj_serve_stats_get(cj_ctx *ctx, cj_serve_stats *out);
/* Estimated device footprint the server of ctx charges for a plan; ctx may
 * be NULL (then as if no input were resident).  An upper bound from the plan's
 * shape (every join's output as large as its larger input), or, once the plan
 * has run on ctx, its recorded peak plus a quarter when that is less: the
 * peak device bytes (temporaries and result pages) of its executions. */
int cj_plan_footprint(cj_ctx *ctx, const cj_plan_desc *plan, uint64_t *bytes);

/* ---- batched execution of many independent plans (no counterpart in the
 *      reference).  Plans that are one join of two resident base-table scans
 *      with INT32 keys and fixed-width outputs run together: every join phase
 *      is one persistent-kernel launch over a device work list covering all
 *      of them, the first phase sequence is a hipGraph replayed while the
 *      batch shape repeats, and the result sizes are read back once.  Other
 *      plans run one at a time.  out[i] is plan i's result (each released
 *      with cj_result_release); results equal cj_execute's. */
enum {
    CJ_BATCH_AUTOTUNE = 16 /* time the candidate configurations on the first
                            * batch of a template class and keep the best */
};
typedef struct cj_batch_stats {
    uint32_t plans_batched;  /* plans run in the shared launches           */
    uint32_t plans_single;   /* plans run one at a time                     */
    double device_ms;        /* shared launches: phase-1 graph + output phase */
    double total_ms;         /* the whole call                              */
    uint64_t output_rows;    /* rows of the batched plans                   */
    uint32_t load_factor;    /* directory slots per build row (chosen)      */
    uint32_t grid;           /* workgroups of the join write launch (chosen) */
    int32_t tuned;           /* this call ran the autotuner                 */
    int32_t graph_replayed;  /* the phase-1 graph of an earlier call ran    */
    uint32_t log_part;       /* slots per radix partition = 2^log_part (chosen) */
    uint32_t chunk;          /* probe tuples per join unit (chosen)         */
    uint32_t join_queue;     /* join units from a device work queue (1) or strided (0), chosen */
} cj_batch_stats;
int cj_execute_batch(cj_ctx *ctx, const cj_plan_desc *plans, uint32_t n, uint32_t flags, cj_result *out,
                     cj_batch_stats *stats);
/* Pin the configuration of later cj_execute_batch calls on ctx (the
 * autotuner and the default step aside), or unpin it with cfg == NULL.
 * load_factor 2 or 4; log_part 8..12; chunk 1..16384; join_queue 0 or 1;
 * grid 0 = each kernel's resident grid, else the workgroups of every
 * persistent launch (1..65536).  Results do not depend on it. */
typedef struct cj_batch_config {
    uint32_t load_factor;  /* directory slots per build row                */
    uint32_t log_part;     /* slots per radix partition = 2^log_part       */
    uint32_t chunk;        /* probe tuples per join unit                   */
    uint32_t join_queue;   /* join units from a device work queue (1)      */
    uint32_t grid;         /* workgroups per persistent launch, 0 = resident */
} cj_batch_config;
int cj_batch_config_set(cj_ctx *ctx, const cj_batch_config *cfg);

/* ---- filters on the GPU: replace the filter step of the harness's scans,
 *      Table::from_csv(attributes, path, filter) (reference
 *      src/build_table.cpp:201-360) with its Statement evaluation
 *      (include/data_model/statement.h, src/statement.cpp:47-249,
 *      include/data_model/inner_column.h).  The filter is a postfix program
 *      of comparison and logical nodes over the table's columns; the result
 *      holds the rows that pass, in their original order, with every column
 *      of the table.  Column-wise semantics of the reference: a comparison is
 *      false on NULL; IS NULL is its own test; NOT inverts (NOT over a NULL
 *      comparison is true); INT32 columns compare with the integer literal
 *      cast to int32; INT64 with the integer literal; FP64 with the float
 *      literal; VARCHAR byte-wise with the string literal, LIKE / NOT LIKE
 *      with % and _ as the reference's regular expression reads them (_ is
 *      one UTF-8 character, neither matches a newline). ------------------- */
enum { CJ_F_CMP = 0, CJ_F_AND = 1, CJ_F_OR = 2, CJ_F_NOT = 3 };
enum {
    CJ_OP_EQ = 0, CJ_OP_NEQ = 1, CJ_OP_LT = 2, CJ_OP_GT = 3, CJ_OP_LEQ = 4, CJ_OP_GEQ = 5,
    CJ_OP_LIKE = 6, CJ_OP_NOT_LIKE = 7, CJ_OP_IS_NULL = 8, CJ_OP_IS_NOT_NULL = 9
};  /* Comparison::Op numbering (statement.h) */
enum { CJ_LIT_NONE = 0, CJ_LIT_INT = 1, CJ_LIT_FLOAT = 2, CJ_LIT_STR = 3 };
typedef struct cj_filter_node {
    int32_t kind;      /* CJ_F_*; AND / OR pop two results, NOT one       */
    int32_t op;        /* CJ_OP_* (CJ_F_CMP)                               */
    uint32_t column;   /* column of the table (CJ_F_CMP)                   */
    int32_t lit_kind;  /* CJ_LIT_*                                         */
    int64_t ival;
    double fval;
    const char *sval;  /* string literal bytes (need not end in NUL)       */
    uint32_t slen;
} cj_filter_node;
int cj_filter(cj_ctx *ctx, const cj_table_desc *table, const cj_filter_node *program, uint32_t num_nodes,
              uint32_t flags, cj_result *out, cj_timing *timing);

/* ---- .tbl table cache files: replace the harness's DumpTable::dump and
 *      Table::from_cache (reference include/data_access/table.h:30-37,
 *      208-221; src/build_table.cpp:140-197).  Layout: one 8192-byte header
 *      page {u64 num_rows; u64 num_cols; i32 types[16]; u64 num_pages[16];
 *      zero filler}, then the 8 KiB pages of each column, column after column.
 *      At most 16 columns. ------------------------------------------------- */
typedef struct cj_tbl cj_tbl;
/* Write `table` to `path` (through a temporary file renamed over it). */
int cj_tbl_write(const char *path, const cj_table_desc *table);
/* Map `path` read-only: desc's page pointers point into the mapping (valid
 * until cj_tbl_close), as Table::from_cache does.  Size, column count and
 * types are checked against the header. */
int cj_tbl_open(const char *path, cj_tbl **out, cj_table_desc *desc);
/* Open `path` and stream its pages from the file into HBM as a resident
 * table (cj_table_register without a host copy).  Plans use desc (pointers
 * into the mapping); the mapping lives until both cj_tbl_close and
 * cj_table_unregister(table_id) happened. */
int cj_tbl_load(cj_ctx *ctx, const char *path, cj_tbl **out, cj_table_desc *desc, uint64_t *table_id);
int cj_tbl_close(cj_tbl *t);
/* The harness's cache file name of a scan: "<table>_<FNV-1a 64 of the
 * filter text, decimal>.tbl", filter_text NULL = "No filter"
 * (reference tests/read_sql.cpp:1154-1158, include/data_access/table.h:15-27). */
int cj_tbl_cache_name(const char *table, const char *filter_text, char *out, size_t cap);

#ifdef __cplusplus
}
#endif

#endif /* CJ_API_H */
