/* C ABI of the MI355X join-pipeline library (libgpujoin.so).
 *
 * Everything crosses this boundary as plain C: opaque handles, POD
 * descriptors, host or device pointers and sizes, and an int status
 * (0 = success; otherwise cgj_last_error() names the failure). No C++ or
 * framework types appear here. The C++ contest surface
 * (Contest::build_context / destroy_context / execute, include/contest/plan.h)
 * is a thin adapter over these calls.
 *
 * Type codes follow the contest's DataType order (reference
 * include/attribute.h:8-13).
 */
#ifndef CGJ_GPUJOIN_H
#define CGJ_GPUJOIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CGJ_INT32   = 0,
    CGJ_INT64   = 1,
    CGJ_FP64    = 2,
    CGJ_VARCHAR = 3,
};

enum {
    CGJ_NODE_JOIN = 0,
    CGJ_NODE_SCAN = 1,
};

/* Bytes per contest page (reference include/plan.h:54). */
#define CGJ_PAGE_SIZE 8192

typedef struct cgj_context cgj_context;
typedef struct cgj_dtable  cgj_dtable;
typedef struct cgj_dresult cgj_dresult;

/* ---- descriptors: POD mirror of the contest Plan (reference include/plan.h:32-149) ---- */

/* One column of a host-resident table: `num_pages` pointers to 8 KiB pages. */
typedef struct {
    const void* const* pages;
    uint64_t           num_pages;
    int32_t            type;
} cgj_host_column;

typedef struct {
    uint64_t               num_rows;
    const cgj_host_column* columns;
    uint32_t               num_columns;
} cgj_host_table;

typedef struct {
    uint64_t index; /* column of the child (scan: base table column) */
    int32_t  type;
} cgj_attr;

/* A scan node (kind = CGJ_NODE_SCAN, base_table_id) or a join node
 * (kind = CGJ_NODE_JOIN, left/right node ids, left_attr/right_attr, build_left).
 * Mirrors ScanNode / JoinNode / PlanNode (reference include/plan.h:32-52). */
typedef struct {
    int32_t         kind;
    int32_t         build_left;
    uint64_t        base_table_id;
    uint64_t        left;
    uint64_t        right;
    uint64_t        left_attr;
    uint64_t        right_attr;
    const cgj_attr* output_attrs;
    uint32_t        num_output_attrs;
} cgj_node;

typedef struct {
    const cgj_node* nodes;
    uint32_t        num_nodes;
    uint64_t        root;
} cgj_plan;

/* Result columns own their pages. Each page is a separate `new Page`-compatible
 * allocation (8192 bytes, 8-byte aligned, from C++ operator new) so a contest
 * ColumnarTable can adopt the pointers and release them with `delete`. */
typedef struct {
    int32_t  type;
    uint64_t num_pages;
    void**   pages;
} cgj_result_column;

typedef struct {
    uint64_t           num_rows;
    uint32_t           num_columns;
    cgj_result_column* columns;
} cgj_result;

/* Per-call statistics of the last execute on a context (device-side timing). */
typedef struct {
    double   ms_total;       /* device time of all kernels of the query        */
    double   ms_build;       /* hash-table build pipelines                     */
    double   ms_probe;       /* output pipeline                                */
    uint64_t input_pages;    /* 8 KiB input pages referenced by the plan       */
    uint64_t output_pages;   /* 8 KiB output pages written                     */
    uint64_t build_tuples;   /* tuples inserted into hash tables               */
    uint64_t probe_rows;     /* rows streamed by the output pipeline's scan    */
    uint64_t output_rows;
    uint32_t pipelines;      /* number of pipeline launches                    */
    uint32_t retries;        /* re-runs: hash-table builds that changed layout, staging or
                                page-pool overflows                             */
    uint64_t probe_input_pages; /* input pages scanned by the output pipeline  */
    uint64_t output_ref_pages;  /* output pages at the reference's packing:
                                   ceil((rows + width*nonnull)/bit budget) per
                                   fixed-width column, pages written otherwise */
    uint64_t build_bytes;       /* compact hash-table tuples: per tuple 4 bytes per INT32 word, 8 per other word */
    double   ms_probe_kernel;   /* HIP-event time of the output pipeline kernel */
    double   ms_probe_partition;/* HIP-event time of the probe side's radix partitioning */
    uint32_t probe_kernel;      /* output pipeline ran on: 0 the general kernel, 1 the lean one-probe kernel */
    uint32_t probe_waves;       /* waves requested for the output pipeline */
    double   ms_plan;           /* host time of cardinality estimation + join ordering,
                                   statistics counts excluded */
    double   ms_stats;          /* time of the 2-way counts run for the statistics catalog */
    uint32_t est_sets;          /* connected relation sets estimated by the network */
    uint32_t est_fallbacks;     /* ... of them, by the sampling estimate (low confidence) */
    uint32_t est_stat_joins;    /* 2-way counts run (pairs not yet in the catalog) */
    uint32_t retries_build;     /* ... of them, builds that changed layout (walk cap, region overflow) */
    double   ms_host_upload;    /* cgj_execute only: wall time of copying the scanned host pages in */
    double   ms_host_download;  /* cgj_execute only: wall time of copying the result pages out into
                                   one operator-new block per page */
    uint32_t replans;           /* join orderings computed (one per hash-table build of the
                                   incremental loop, plus the output pipeline's) */
    uint32_t singletons;        /* one-row relations removed before planning (their values
                                   restrict the other members of their classes) */
    uint32_t restrictions;      /* scan restrictions applied (singleton constants; keys of
                                   built tables of at most 32 tuples; semi-join tables) */
    uint32_t semijoins;         /* built tables used only as an exact restriction of their
                                   key's producers (duplicate-free, key only) */
    uint32_t probe_packed;      /* the output pipeline's lean kernel read its hash table packed
                                   (INT32 key | INT32 attribute << 32, 8 bytes per slot) */
    uint32_t pools_compacted;   /* result page pools moved into a pool of exactly their pages
                                   (the pool, sized for an upper bound, was mostly unused) */
    uint32_t blooms_skipped;    /* probes that left out their table's bloom filter: at least
                                   half of the driving relation's sampled keys matched */
} cgj_stats;

/* Execution options of a context (0 = success, -1 = unknown name):
 *   "radix_min_table_bytes" / "radix_min_rows": the probe side is radix-partitioned
 *       when the first probed hash table and the driving table exceed both (default: never);
 *   "max_load": hash-table load factor (0 = automatic);
 *   "bloom_bits": bloom bits per key (-1 = automatic, 0 = none);
 *   "partitioned_build": radix-partitioned LDS build (-1 automatic, 0 off, 1 on);
 *   "lean_kernel": 1 (default) runs one-probe pipelines on the lean kernel, 0 on the general one;
 *   "lean_waves_per_cu": grid of the lean kernel in waves per compute unit (4..64, default 24);
 *   "aligned_pairs": the lean kernel's probe steps read 2-slot groups aligned to two
 *       slots (1, default) or the next two slots (0);
 *   "build_threads": threads per workgroup of the partitioned build kernels (256, 512,
 *       1024 = default);
 *   "group_walk": a hash-table insert that passes this many slots holding its own key
 *       (a heavily repeated key) rebuilds the table grouped by key (default 32; 0 = never);
 *   "force_grouped": build every hash table grouped by key (1) or only when needed (0);
 *   "scan_build": hash tables on a plain scan are partitioned straight from its pages (1,
 *       default) or from staged tuples (0);
 *   "scan_workgroups": workgroups of that partition (16..4096, default 512);
 *   "block_prefilter": a pipeline over pages tests its early bloom filters on eight 64-row
 *       blocks at a time before the full per-block work (1, default) or not (0);
 *   "scan_insert_max_bytes": a hash table on a plain scan of at most this many bytes is
 *       inserted straight from the pages by one kernel (default 8 MiB, 0 = never);
 *   "incremental": re-plan after every hash-table build, with the tables built so far as
 *       fixed build sides and the samples re-filtered through them (1, default), or plan
 *       once (0);
 *   "restrictions": in the incremental loop, a built table of at most 32 tuples restricts
 *       the unjoined producers of its key by its keys (Eq, Eq2, range or tiny table), and a
 *       duplicate-free table carrying only its key becomes an exact restriction of them
 *       instead of a join input (reference engine/query/QueryPlan.cpp:602-625,
 *       engine/storage/RestrictionLogic.cpp:397-447) (1, default) or neither (0);
 *   "singletons": one-row relations are removed before planning and their values
 *       restrict the other members of their classes (QueryPlan.cpp:680-717) (1, default)
 *       or kept as relations (0);
 *   "packed_tables": the one-probe kernel reads a hash table of an INT32 key and one INT32
 *       attribute as one 8-byte word per slot (1, default) or as two words (0). */
int cgj_set_option(cgj_context* ctx, const char* name, double value);

/* sizeof(cgj_stats) as compiled into the library (binding sanity check). */
size_t cgj_stats_size(void);

/* ---- context (reference engine/Execute.cpp:42-48) ---- */

/* Replaces Contest::build_context(): binds `device`, creates the HIP stream,
 * the page pool and the tuning state. NULL on failure. */
cgj_context* cgj_context_create(int device);
/* Replaces Contest::destroy_context(void*). */
void         cgj_context_destroy(cgj_context* ctx);
/* Message of the last failed call on `ctx` (empty string if none). */
const char*  cgj_last_error(const cgj_context* ctx);

/* ---- whole-query execution (reference engine/Execute.cpp:28-40) ---- */

/* Replaces Contest::execute(const Plan&, void*): uploads the referenced input
 * columns, runs the join pipelines on the GPU and returns the result pages in
 * host memory. Row order is unspecified (as in the reference). */
int  cgj_execute(cgj_context* ctx, const cgj_plan* plan, const cgj_host_table* inputs,
                 uint32_t num_inputs, cgj_result* out);
/* Releases pages and arrays of a result produced by this library. */
void cgj_result_release(cgj_result* res);

/* Device-resident tables: pages live in HBM (reference engine/query/PlanImport.cpp:91-101
 * imports host pages zero-copy; here the import is one upload per column). */
cgj_dtable* cgj_dtable_upload(cgj_context* ctx, const cgj_host_table* table);
/* Tables and device results keep their memory pool alive: they may be freed
 * before or after their context is destroyed (`ctx` may then be NULL). */
void        cgj_dtable_free(cgj_context* ctx, cgj_dtable* t);
uint64_t    cgj_dtable_num_rows(const cgj_dtable* t);
/* Rebuilds every column's rows-per-page prefix and 64-row block index, the
 * per-query work of TableScan::prepareColumn / makeSearchVector (reference
 * engine/op/TableScan.cpp:546-582, run from QueryPlan.cpp:195-198 inside every
 * execute()). Tables are indexed when uploaded or generated; this entry point
 * repeats that work so that a per-query measurement can include it. */
int         cgj_dtable_prepare(cgj_context* ctx, cgj_dtable* t);
/* Copies one column's pages of a device table to host memory (the inverse of
 * cgj_dtable_upload), one operator-new block per page; release with
 * cgj_column_release. Lets a checker read exactly the pages a query scanned
 * (for example tables generated in HBM). */
int         cgj_dtable_download(cgj_context* ctx, const cgj_dtable* t, uint32_t column, cgj_result_column* out);
uint64_t    cgj_dtable_num_pages(const cgj_dtable* t, uint32_t column);

/* Runs `num_plans` independent plans over device-resident inputs concurrently:
 * plan i reads the next num_inputs[i] handles of `inputs` (concatenated), and
 * out[i] receives its result (free each with cgj_dresult_free). `workers`
 * host threads (1..16), each with its own HIP stream, take plans in turn, so
 * one plan's host round trips overlap another's kernels; each worker plans
 * its queries with the context's estimator (cgj_set_estimator). Released
 * device blocks of 4 MiB and more are shared by the workers (reuse on another
 * stream waits for the releasing stream). HIP maps streams onto
 * GPU_MAX_HW_QUEUES hardware queues per process (default 4): with fewer
 * queues than workers + 1, workers sharing a queue run their kernels one
 * after another, so set it before the process first uses HIP. The contest
 * harness calls Contest::execute once per query (reference
 * tests/read_sql.cpp:1254-1256); this is that loop, batched. Results are
 * those of cgj_execute_device. */
int      cgj_execute_many(cgj_context* ctx, const cgj_plan* plans, uint32_t num_plans, cgj_dtable* const* inputs,
                          const uint32_t* num_inputs, uint32_t workers, cgj_dresult** out);

/* Runs the plan over device-resident inputs; the result stays in HBM. This is
 * the device-resident form of Contest::execute used for roofline timing. */
int      cgj_execute_device(cgj_context* ctx, const cgj_plan* plan, cgj_dtable* const* inputs,
                            uint32_t num_inputs, cgj_dresult** out);
uint64_t cgj_dresult_num_rows(const cgj_dresult* r);
uint64_t cgj_dresult_num_pages(const cgj_dresult* r);
/* Copies a device result's pages to host memory in contest layout. */
int      cgj_dresult_download(cgj_context* ctx, const cgj_dresult* r, cgj_result* out);
void     cgj_dresult_free(cgj_context* ctx, cgj_dresult* r);

int      cgj_last_stats(const cgj_context* ctx, cgj_stats* out);

/* Counts the result rows of the plan without materialising pages: the same
 * builds and probe chain, with a counting target. Used for exact sub-plan
 * cardinalities (the estimator's training labels and its q-error reference). */
int      cgj_count_device(cgj_context* ctx, const cgj_plan* plan, cgj_dtable* const* inputs, uint32_t num_inputs,
                          uint64_t* rows);

/* No counterpart in the reference: cgj_count_device for a plan that joins two
 * scans on one key, for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops). It calls no hipMalloc/hipFree: the
 * hash table (key-only slots at load <= 0.5) and its counters live in a fixed
 * device array of CGJ_NOALLOC_TABLE_BYTES (plus 4 KiB) that each GPU the
 * library runs on holds until the process ends. A build side (the smaller
 * input) whose table would not fit is rejected (-1, cgj_last_error names the
 * limit), as is any other plan shape. Calls on one GPU share that array and
 * each clears it first: calls that overlap on one GPU race. */
#define CGJ_NOALLOC_TABLE_BYTES (64u << 20)
int      cgj_count_device_noalloc(cgj_context* ctx, const cgj_plan* plan, cgj_dtable* const* inputs,
                                  uint32_t num_inputs, uint64_t* rows);

/* Bandwidth probe: copies `bytes` bytes between two HBM buffers with a plain
 * streaming kernel, `reps` times; returns the mean kernel time in ms (or a
 * negative value on error). Calibrates HBM byte counters and reports the
 * achievable copy bandwidth next to the roofline peak. */
double   cgj_copy_probe(cgj_context* ctx, uint64_t bytes, uint32_t reps);

/* ---- batched joins ----
 * Many independent plans, each one equi-join of two scans with up to four
 * fixed-width output columns (the reference runs such queries one
 * Contest::execute at a time, engine/Execute.cpp:28-40 → QueryPlan.cpp:414-678
 * per join). A batch runs them all in one hipGraph launch: table fills plus one
 * persistent kernel per output width, fed by a device queue of build and
 * probe tasks. No reference interface is replaced; this is the batched form of
 * cgj_execute_device over the same plans, with the same results per plan. */
typedef struct cgj_batch cgj_batch;

typedef struct {
    double   ms_last;  /* device time of the last cgj_batch_run (graph launch) */
    uint64_t tasks;    /* queue entries (build + probe tasks)                  */
    uint32_t waves;    /* persistent grid (resident waves)                     */
    uint32_t launches; /* kernel nodes in the graph                            */
    uint32_t retries;  /* re-runs after a job's page pool overflowed           */
    uint32_t jobs;
} cgj_batch_stats;

/* `inputs` holds num_inputs[0] tables of plan 0, then those of plan 1, ...
 * Plans of any other shape are rejected (NULL; cgj_last_error says why). */
cgj_batch* cgj_batch_create(cgj_context* ctx, const cgj_plan* plans, uint32_t num_plans, cgj_dtable* const* inputs,
                            const uint32_t* num_inputs);
int        cgj_batch_run(cgj_context* ctx, cgj_batch* b);
/* Tunes task sizes, load factor and bloom bits per (build rows, probe rows,
 * key width) class by timing candidate settings on this batch (several runs);
 * the winners stay in the context for later batches. */
int        cgj_batch_autotune(cgj_context* ctx, cgj_batch* b);
int        cgj_batch_info(const cgj_batch* b, cgj_batch_stats* out);
uint64_t   cgj_batch_num_rows(const cgj_batch* b, uint32_t plan);
int        cgj_batch_download(cgj_context* ctx, cgj_batch* b, uint32_t plan, cgj_result* out);
void       cgj_batch_free(cgj_context* ctx, cgj_batch* b);

/* ---- learned cardinality estimator ----
 * A 3-layer MLP (in features -> 256 -> 256 -> 2, ReLU) that predicts the
 * log10 cardinality of a connected set of relations and the log of its
 * standard deviation. It stands in for the reference's sampling estimate
 * (engine/query/QueryPlan.cpp:66-159, 719-746; QueryGraph.cpp:53-68) in join
 * ordering. Training runs on the GPU (FP32 forward, backward and Adam);
 * inference takes BF16 inputs with FP32 accumulation on the matrix cores.
 * Parameters are one flat FP32 array: W1 (in x 256), b1, W2 (256 x 256), b2,
 * W3 (256 x 2), b3; row-major, input index first. Host arrays throughout. */
typedef struct cgj_mlp cgj_mlp;

cgj_mlp* cgj_mlp_create(cgj_context* ctx, uint32_t in_features, uint64_t seed);
uint64_t cgj_mlp_num_params(const cgj_mlp* m);
int      cgj_mlp_set_params(cgj_context* ctx, cgj_mlp* m, const float* params);
int      cgj_mlp_get_params(cgj_context* ctx, const cgj_mlp* m, float* params);
/* Mean heteroscedastic Gaussian loss, 0.5 (y - mu)^2 exp(-2 s) + s, over the n
 * rows of x (n x in) against y, and its gradient (num_params floats). */
int      cgj_mlp_loss_grad(cgj_context* ctx, cgj_mlp* m, const float* x, const float* y, uint32_t n, float* loss,
                           float* grads);
int      cgj_mlp_train(cgj_context* ctx, cgj_mlp* m, const float* x, const float* y, uint32_t n, uint32_t epochs,
                       uint32_t batch, float lr, float weight_decay,
                       uint64_t seed, float* final_loss);
/* cgj_mlp_train: Adam (beta 0.9/0.999, eps 1e-8) with `weight_decay` x params
 * added to each gradient, over shuffled batches; *final_loss = the last
 * epoch's mean loss. */
/* out: n x 2 (mu, log sigma). _predict is the BF16 MFMA path, _predict_f32 the FP32 forward. */
int      cgj_mlp_predict(cgj_context* ctx, cgj_mlp* m, const float* x, uint32_t n, float* out);
int      cgj_mlp_predict_f32(cgj_context* ctx, cgj_mlp* m, const float* x, uint32_t n, float* out);
void     cgj_mlp_free(cgj_context* ctx, cgj_mlp* m);

/* ---- cardinality estimation for join ordering ----
 * The reference plans with DPccp under the rule "a join is as large as its
 * largest input" (engine/query/QueryGraph.cpp:53-68) over sampled relation
 * cardinalities (engine/query/QueryPlan.cpp:66-72, 719-746): up to 64 sample
 * rows per relation, cardinality = rows x passing sample rows / sample rows
 * (rows / (2 x sample rows) when none passes), re-filtered through every hash
 * table it builds, with a re-plan after each build (QueryPlan.cpp:670-676,
 * 748-782). That is CGJ_EST_SAMPLE, the default; samples are gathered on the
 * device and re-filtered by probing each new hash table. CGJ_EST_MAX applies
 * the max rule to exact relation sizes. CGJ_EST_INDEPENDENCE multiplies
 * relation sizes by exact pairwise join selectivities; CGJ_EST_LEARNED adds
 * the network's output mu (a log10 correction; the network must have
 * cgj_num_features() inputs) to the independence estimate (log10, at least 0)
 * of every connected relation set, all sets in one batched inference, and uses
 * the sampling estimate where the predicted spread sigma exceeds `max_sigma`
 * (log10 units). Pairwise selectivities come from exact 2-way counts on the
 * device, cached per (table, column) pair for the context's lifetime. Every
 * mode re-plans after each hash-table build (option "incremental"), pricing a
 * built table at its exact size. The estimator only changes the join order:
 * results are identical in every mode. The context keeps a reference to the
 * network, so its handle may be freed. */
enum { CGJ_EST_MAX = 0, CGJ_EST_INDEPENDENCE = 1, CGJ_EST_LEARNED = 2, CGJ_EST_SAMPLE = 3 };
uint32_t cgj_num_features(void);
int      cgj_set_estimator(cgj_context* ctx, uint32_t mode, cgj_mlp* mlp, float max_sigma);
/* Empties the context's catalog of pairwise join counts (the next plans under
 * CGJ_EST_INDEPENDENCE or CGJ_EST_LEARNED count their pairs again). No
 * counterpart in the reference, which keeps no catalog: a caller that times
 * planning from cold clears it first. */
int      cgj_clear_statistics(cgj_context* ctx);
/* Relations of a plan are numbered in its depth-first scan order (left
 * input first). Writes, for up to `cap` connected sets of two or more
 * relations, the set (bit r = relation r) and its cgj_num_features() features;
 * returns the number of such sets, or -1 on error. */
int64_t  cgj_plan_features(cgj_context* ctx, const cgj_plan* plan, cgj_dtable* const* inputs, uint32_t num_inputs,
                           uint64_t* sets, float* features, uint64_t cap);
/* The sampling estimate of each relation of the plan (numbered as for
 * cgj_plan_features): rows x the fraction of its (up to 64) sample rows whose
 * join columns are all non-NULL, rows / (2 x sample rows) when none is
 * (reference engine/query/QueryPlan.cpp:66-72, 719-746). The estimate of a
 * set is the largest of its relations' (QueryGraph.cpp:53-68). `cards` holds
 * one double per relation; *num_rels receives their number. */
int      cgj_sample_cardinalities(cgj_context* ctx, const cgj_plan* plan, cgj_dtable* const* inputs,
                                  uint32_t num_inputs, double* cards, uint32_t cap, uint32_t* num_rels);
/* Exact result rows of the join of a connected relation set of the plan. */
int      cgj_count_subset(cgj_context* ctx, const cgj_plan* plan, cgj_dtable* const* inputs, uint32_t num_inputs,
                          uint64_t relset, uint64_t* rows);
/* The join order the planner picks under the current estimator: per join
 * (children first), the relation sets of the build and probe sides and the
 * estimated result cardinality. *num_joins = number of joins (at most `cap`
 * are written). */
int      cgj_plan_joins(cgj_context* ctx, const cgj_plan* plan, cgj_dtable* const* inputs, uint32_t num_inputs,
                        uint64_t* build_sets, uint64_t* probe_sets, double* cards, uint32_t cap,
                        uint32_t* num_joins);

/* ---- operator-level entry points ---- */

/* Page decode of one device column into 64-bit slots, NULL = all ones, INT32
 * zero-extended, FP64 raw bits, VARCHAR as a string handle. Replaces
 * TableScan::Reader::skipTo/step64 (reference engine/op/TableScan.cpp:82-324).
 * The kernel writes the slots to HBM; they are then copied to `slots`, a host
 * array of at least cgj_dtable_num_rows() entries. */
int cgj_scan_decode(cgj_context* ctx, const cgj_dtable* t, uint32_t column, uint64_t* slots);

/* Encode host values into contest pages with the contest page writer
 * (reference include/plan.h:151-335). values: int32/int64/double array of
 * `n` entries (or, for VARCHAR, `n`+1 u64 byte offsets into `chars`);
 * `valid` = NULL for no NULLs, else one byte per row (0 = NULL).
 * The pages are returned in `out` (type, num_pages, pages) and released with
 * cgj_result_release semantics via cgj_column_release. */
int  cgj_encode_column(int32_t type, const void* values, const char* chars, const uint8_t* valid,
                       uint64_t n, cgj_result_column* out);
void cgj_column_release(cgj_result_column* col);

/* ---- synthetic generators (device-resident inputs for benchmarks) ---- */

/* Column recipe for cgj_dtable_generate. Row r of the column draws its value
 * and its NULL flag from SplitMix64 streams keyed by (seed, stream, r). */
typedef struct {
    int32_t  type;        /* CGJ_INT32 / CGJ_INT64 / CGJ_FP64                           */
    int32_t  kind;        /* CGJ_GEN_* below                                             */
    uint64_t stream;      /* stream id for values                                        */
    uint64_t null_stream; /* stream id for NULL draws                                    */
    double   null_frac;   /* Bernoulli NULL probability                                  */
    uint64_t domain;      /* key domain size (unique / fk kinds)                         */
    double   zipf_s;      /* Zipf exponent for CGJ_GEN_FK_ZIPF                           */
    double   miss_frac;   /* fraction of non-matching keys for FK kinds                  */
    uint64_t lo, hi;      /* range for CGJ_GEN_UNIFORM (inclusive lo, exclusive hi)      */
} cgj_gen_column;

enum {
    CGJ_GEN_UNIQUE_KEY = 0, /* key = bijective mix of r (unique, shuffled)               */
    CGJ_GEN_FK_ZIPF    = 1, /* key of a Zipf-drawn rank in [0, domain), or a miss key    */
    CGJ_GEN_UNIFORM    = 2, /* integer uniform in [lo, hi)                               */
    CGJ_GEN_UNIT_F64   = 3, /* double uniform in [0, 1)                                  */
    CGJ_GEN_PERM_KEY   = 4, /* INT32 key: permutation of 0..domain-1 (r < domain)        */
    CGJ_GEN_FK_UNIFORM = 5, /* key = permutation value of a uniform rank, or a miss key  */
    CGJ_GEN_ASCII      = 6, /* VARCHAR, host only: length uniform in [lo, hi], a fraction
                               miss_frac of rows gets a long string of 8186..16000 bytes */
};

/* Generates a table of `num_rows` rows directly in HBM, paged exactly as the
 * contest page writer would page it. */
cgj_dtable* cgj_dtable_generate(cgj_context* ctx, uint64_t seed, uint64_t num_rows,
                                const cgj_gen_column* columns, uint32_t num_columns);
/* Same recipe evaluated on the host (for CPU baselines and tests): writes
 * `num_rows` rows starting at `first_row` as contest pages. */
int cgj_host_generate(uint64_t seed, uint64_t first_row, uint64_t num_rows,
                      const cgj_gen_column* column, cgj_result_column* out);

#ifdef __cplusplus
}
#endif

#endif /* CGJ_GPUJOIN_H */
