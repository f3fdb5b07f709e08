// Host-visible launch interface of the HIP kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.hpp"

namespace cgj {

constexpr int kMaxScan   = 16; // scanned columns per pipeline
constexpr int kMaxProbes = 4;  // hash-table probes fused into one pipeline
constexpr int kMaxChecks = 4;  // residual equalities per probe
constexpr int kMaxOut    = 16; // output columns
constexpr int kMaxWords  = 16; // words per hash-table entry / staged tuple

struct ProbeSpec {
    const uint64_t* ht;        // cap x words entries; word 0 is the key
    uint32_t        shift;     // slot = hash_key(key) >> shift
    uint32_t        words;     // words per entry
    uint64_t        mask;      // region slots - 1: linear probing wraps inside a region
    uint32_t        key_slot;  // tuple slot holding the probe key
    uint32_t        unique;    // keys are unique: stop at the first match
    uint32_t        nattrs;    // entry words copied into tuple slots
    uint32_t        attr_word[kMaxWords];
    uint32_t        attr_slot[kMaxWords];
    uint32_t        nchecks;   // residual equalities: entry word == tuple slot
    uint32_t        check_word[kMaxChecks];
    uint32_t        check_slot[kMaxChecks];
    const uint64_t* bloom;     // blocked bloom filter (nullptr = none)
    uint32_t        bloom_log; // log2(words)
    uint32_t        early;     // key is a scanned column: filter before probing
    // Packed copy of a two-word table (the one-probe kernel only; nullptr =
    // none): one word per slot, key | attr << 32, same slots as `ht`; used
    // unless *pack_flag is nonzero (an entry did not fit 32 + 32 bits).
    const uint64_t* ht_packed;
    const uint32_t* pack_flag;
    // Grouped table: ht holds 2-word slots [key, start << 24 | count] and the
    // matching tuples' words 1.. (attr_word / check_word - 1) are payload rows
    // start .. start + count - 1 of `pwords` words.
    uint32_t        grouped;
    uint32_t        pwords;
    const uint64_t* payload;
};

struct OutSpec {
    uint32_t slot;
    int32_t  type;
    uint32_t vc;   // index among VARCHAR output columns (staging area)
    uint32_t pad;
};

enum : uint32_t { TARGET_STAGE = 0, TARGET_PAGES = 1, TARGET_COUNT = 2 };
enum : uint32_t { DRIVER_PAGES = 0, DRIVER_TUPLES = 1 };

// Counters written by a pipeline launch (device memory).
// The counters that waves update while a kernel runs (reservations, the
// morsel queue, the closing sums) each sit on their own 256-byte line, so
// their atomics do not queue behind each other's.
struct PipeCounters {
    unsigned long long tuples;    // staged tuples / output rows / counted tuples
    unsigned long long pad0[31];
    unsigned long long pages;     // pool pages (or staging slots) reserved by atomics
    unsigned long long pad1[31];
    unsigned long long morsel;    // morsels handed out beyond each wave's first (dynamic schedule)
    unsigned long long pad2[31];
    unsigned long long overflow;  // nonzero: a buffer was too small, re-run
    unsigned long long fault;     // nonzero: a hash-table walk met no empty slot within its bound
    unsigned long long packed;    // one-probe kernel: nonzero when it read the packed table
    // one-probe kernel: pool pages 0 .. static_pages - 1 were given to the
    // waves before any atomic (`page_chunk` each); atomics reserve after them
    unsigned long long static_pages;
    unsigned long long waves;     // one-probe kernel with per-wave sums: waves of the grid
    unsigned long long pad3[27];
    unsigned long long nnz[kMaxOut];       // non-NULL values written per output column
    unsigned long long col_pages[kMaxOut]; // pages written per output column

    unsigned long long reserved() const { return static_pages + pages; }
};

struct PipeArgs {
    // driver
    uint32_t        driver;
    uint32_t        nscan;
    uint64_t        nrows;
    uint32_t        morsel_rows; // multiple of 64
    uint32_t        not_null;    // bitmask of tuple slots that must be non-NULL
    uint64_t        nmorsels;
    uint32_t        xcd_order;   // morsel range split across the 8 XCDs (grid multiple of 8 blocks)
    ScanCol         scan[kMaxScan];
    const uint64_t* tuples;      // DRIVER_TUPLES: rows x tuple_words
    uint32_t        tuple_words;
    uint32_t        tuple_word[kMaxScan];
    // probes
    uint32_t        nprobes;
    ProbeSpec       probes[kMaxProbes];
    // target
    uint32_t        target;
    uint32_t        stage_words;
    uint32_t        stage_slot[kMaxWords];
    uint64_t*       stage;
    uint64_t        stage_cap;   // tuple slots; unused slots carry key kEmpty
    uint64_t        stage_chunk; // slots reserved per atomic (>= 64)
    uint32_t        nout;
    uint32_t        nvarchar;
    OutSpec         out[kMaxOut];
    uint8_t*        pool;        // output page pool
    uint64_t        pool_pages;
    uint32_t*       dir;         // per pool page: column, wave stream, sequence, valid
    uint64_t        dir_cap;     // unused (the directory has pool_pages entries)
    uint8_t*        str_stage;   // per wave x VARCHAR column: 8 KiB
    const uint8_t* const* col_base;  // pages of column id (string handles)
    const uint32_t* col_npages;
    PipeCounters*   counters;
    uint32_t        page_chunk;  // pool pages reserved per atomic (1..kPageChunk)
    uint32_t        grid_per_cu; // host only: waves per CU of the launch (0 = default)
    uint32_t        prefilter;   // pages driver: test early bloom filters four blocks at a time first
    const uint64_t* sel;         // pages driver: per 64-row block, the rows that pass the scan's
                                 // restrictions (launch_filter_select), nullptr = all
    uint64_t        abort_pages; // a wave whose reservation starts at or past this many pool pages (or
                                 // staging slots) stops: the result cannot fit device memory
    // Per-wave closing sums instead of atomics (as LeanArgs::wave_sums).
    unsigned long long* wave_sums;
    uint32_t            wave_sums_stride;
    // Nonzero: each wave owns pool pages (or staging slots) [wave * chunk,
    // (wave + 1) * chunk) from the start, and atomics on counters->pages
    // reserve after the grid's static part (as LeanArgs::static_chunk).
    uint32_t            static_chunk;
};

// ---- lean one-probe pipeline ----
constexpr int      kLeanOut   = 4;  // fixed-width output columns
constexpr uint32_t kPageChunk = 32; // pool pages reserved per atomic (upper bound)
// Fields of a wave's closing sums: rows, non-NULL values of output column c
// (field 1 + c), pages of output column c (field 1 + kMaxOut + c).
constexpr uint32_t kSumFields = 1 + 2 * kMaxOut;

// Where an output column's value comes from.
enum : uint32_t { SRC_KEY = 0, SRC_SCAN = 1, SRC_HT = 2 };

struct LeanOutCol {
    ScanCol  col;   // SRC_SCAN over pages
    uint32_t src;   // SRC_*
    uint32_t word;  // SRC_SCAN over tuples: tuple word; SRC_HT: entry word
    int32_t  type;  // output type (fixed width)
    uint32_t pad;
};

struct LeanArgs {
    ScanCol         key;        // probe key column (DRIVER_PAGES)
    const uint64_t* tuples;     // DRIVER_TUPLES: rows x tuple_words, word 0 never NULL but in holes
    uint32_t        tuple_words;
    uint32_t        key_word;   // DRIVER_TUPLES: tuple word of the probe key
    uint32_t        driver;
    uint32_t        morsel_rows;
    uint32_t        xcd_order;
    uint32_t        unique;
    uint64_t        nrows;
    uint64_t        nmorsels;
    const uint64_t* ht;
    uint64_t        mask;
    uint32_t        shift;
    uint32_t        words;
    const uint64_t* bloom;
    uint32_t        bloom_log;
    uint32_t        nout;
    LeanOutCol      out[kLeanOut];
    uint8_t*        pool;
    uint64_t        pool_pages;
    uint32_t*       dir;
    PipeCounters*   counters;
    uint32_t        page_chunk; // pool pages reserved per atomic (1..kPageChunk)
    // Probe steps read the 2-slot group {q & ~1, q | 1} holding the next slot q
    // (slot q & ~1 ignored when q is odd) instead of {q, q + 1}: with 16-byte
    // entries a step then never spans two 64-byte lines. Entries of <= 2 words.
    uint32_t        aligned_pairs;
    uint64_t        abort_pages; // as PipeArgs::abort_pages
    const uint64_t* sel;         // as PipeArgs::sel
    const uint64_t* ht_packed;   // as ProbeSpec::ht_packed (nullptr = none)
    const uint32_t* pack_flag;
    // Nonzero: each wave of the launch owns pool pages [wave * page_chunk,
    // (wave + 1) * page_chunk) from the start, and atomics on counters->pages
    // reserve after the grid's static pages (k_lean; the batch kernel leaves it 0).
    uint32_t        static_chunk;
    // Nonzero: instead of atomics on the counters, each wave stores its closing
    // sums as wave_sums[field * wave_sums_stride + wave] (fields as kSumFields)
    // and launch_fold_sums adds them up after the kernel; wave 0 stores the
    // grid's wave count in counters->waves.
    unsigned long long* wave_sums;
    uint32_t            wave_sums_stride;
};

// ---- batched joins: one persistent launch over a device work queue ----
constexpr int kBatchAttrs = 3; // build-side words besides the key

// Build side of one batched join: the build table's key and attribute columns
// are inserted into `ht` (global compare-and-swap, linear probing over the
// whole table) in tasks of `morsel_rows` rows.
struct BatchBuild {
    ScanCol   key;
    ScanCol   attr[kBatchAttrs];
    uint32_t  nattr;
    uint32_t  morsel_rows;
    uint64_t  rows;
    uint32_t  morsels;
    uint32_t  bloom_log;
    uint64_t* ht;
    uint64_t  mask;  // capacity - 1
    uint32_t  shift; // slot = hash >> shift
    uint32_t  words;
    uint64_t* bloom;
};

// Per-job flags and timing, in device memory (zeroed before every run).
struct BatchSync {
    unsigned int       dup;          // bit 0: a duplicate build key was seen; bit 1: an insert found no
                                     // empty slot (the table is incomplete)
    unsigned int       pad[3];
    unsigned long long build_clocks; // wall-clock ticks spent in build tasks
    unsigned long long probe_clocks; // ... in probe tasks
};

struct BatchJob {
    LeanArgs   probe; // nmorsels = probe tasks of morsel_rows rows each
    BatchBuild build;
    BatchSync* sync;
    uint64_t   pad;
};

// A queue entry: build task `index` or probe task `index` of job `job`.
struct BatchTask {
    uint32_t job;
    uint32_t kind; // 0 build, 1 probe
    uint64_t index;
};

// `nout` value that selects the build kernel in launch_batch / batch_resident_waves.
constexpr uint32_t kBatchBuildKernel = 0xffffffffu;
// Runs `ntasks` tasks of one kind with a persistent grid of `max_waves` waves
// (0: as many as fit the GPU); `queue` is the zeroed task counter. With
// nout = kBatchBuildKernel the tasks are build tasks; otherwise they are probe
// tasks of jobs with `nout` output columns, launched after the build launch.
hipError_t launch_batch(const BatchJob* jobs, const BatchTask* tasks, uint64_t ntasks, unsigned long long* queue,
                        uint32_t nout, uint32_t max_waves, hipStream_t s);
// Waves of the build kernel / the probe kernel for `nout` outputs that fit the GPU at once.
uint32_t batch_resident_waves(uint32_t nout);
// Ticks per second of the clock behind BatchSync's *_clocks.
double batch_clock_hz();

// Radix partitioning of a driving table by the hash of its probe key.
struct RadixArgs {
    ScanCol   scan[kMaxScan]; // column 0 is the probe key
    uint32_t  nscan;
    uint32_t  not_null;       // columns that must be non-NULL (join columns)
    uint64_t  nrows;
    uint32_t  rshift;         // partition = hash_key(key) >> rshift
    uint32_t  nparts;
    uint64_t  blocks_per_wg;  // 64-row blocks per workgroup
    uint32_t* bh;             // nparts x workgroups counters, scanned into offsets
    uint64_t* out;            // tuples of nscan words, in partition order
};
// Histogram, scan and scatter; `total` receives the number of tuples written.
hipError_t launch_radix_partition(const RadixArgs& a, uint32_t wgs, uint32_t* scan_scratch, uint64_t* total,
                                  hipStream_t s);

// Column preparation: rows per page prefix and the block index. `dense`
// (two words, zeroed by the caller) receives 1 in word 0 unless every page is
// NULL-free with as many rows as page 0 (the last at most as many), and the
// rows of page 0 in word 1.
hipError_t launch_prepare_column(const uint8_t* pages, uint64_t npages, int32_t type, uint64_t nrows,
                                 uint32_t* row_start, BlockIdx* blocks, uint32_t* scratch, uint64_t* total_rows_dev,
                                 uint32_t* dense, hipStream_t s);
size_t     prepare_scratch_words(uint64_t npages);

hipError_t launch_decode(const ScanCol& c, uint64_t nrows, uint64_t* out, hipStream_t s);
hipError_t launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s);
hipError_t launch_copy(void* dst, const void* src, uint64_t n16, hipStream_t s);
// Valid pages of a result pool of `src_pages` slots (directory after the
// slots, `reserved` entries to scan) moved into a pool of `dst_pages` slots;
// *counter must start at 0, *fault is set when more pages are valid than
// `dst_pages`.
hipError_t launch_compact_pages(const uint8_t* src, uint64_t src_pages, uint64_t reserved, uint8_t* dst,
                                uint64_t dst_pages, unsigned long long* counter, unsigned long long* fault,
                                hipStream_t s);
// flags[0] = duplicate keys, flags[2] = an insert passed walk_cap slots
// holding its own key (0 = no cap) and stopped (the table is incomplete).
hipError_t launch_ht_insert(const uint64_t* stage, uint64_t n, uint32_t words, uint64_t* ht, uint32_t shift,
                            uint64_t mask, uint64_t* bloom, uint32_t bloom_log, uint32_t* flags, uint32_t walk_cap,
                            hipStream_t s);
size_t     scan_scratch_words(uint64_t n);
hipError_t launch_exclusive_scan(uint32_t* data, uint64_t n, uint32_t* scratch, uint64_t* total, hipStream_t s);
// Radix-partitioned build of a table of 2^log_cap slots split into
// 2^log_regions regions; `bh` holds blocks << log_regions counters, `parts`
// n x words tuples. flags[0] = duplicate keys, flags[1] = a region overflowed.
// `n` = staging slots (slots with key kEmpty are skipped), `n_real` = tuples.
hipError_t launch_part_build(const uint64_t* stage, uint64_t n, uint64_t n_real, uint32_t words, uint64_t* parts,
                             uint32_t* bh,
                             uint32_t blocks, uint64_t chunk, uint32_t* scan_scratch, uint64_t* total,
                             uint64_t* ht, uint32_t log_cap, uint32_t log_regions, uint64_t* bloom,
                             uint32_t bloom_log, uint32_t* flags, uint32_t threads, uint32_t walk_cap,
                             hipStream_t s);
// Grouped table of 2^log_cap two-word slots [key, start << 24 | count] in
// 2^log_regions regions, with the key's tuples (staged words 1..) in `payload`
// grouped by key; same partition buffers as launch_part_build. flags[1] = a
// region's distinct keys overflowed it, flags[3] = a group exceeded 2^24 - 1.
hipError_t launch_group_build(const uint64_t* stage, uint64_t n, uint64_t n_real, uint32_t words, uint64_t* parts,
                              uint32_t* bh, uint32_t blocks, uint64_t chunk, uint32_t* scan_scratch,
                              uint64_t* total, uint64_t* ht, uint64_t* payload, uint32_t log_cap,
                              uint32_t log_regions, uint64_t* bloom, uint32_t bloom_log, uint32_t* flags,
                              uint32_t threads, hipStream_t s);
size_t     group_build_lds(uint32_t rslots, uint32_t threads);
// Partitioning straight from a base table's pages (a hash table built on a
// plain scan): tuple word w is column col[w]; rows with NULL in a `notnull`
// word are dropped. Same histogram layout (bh[region * blocks + block]) and
// region-ordered `parts` as the staged path; workgroup b takes 64-row blocks
// [b * blocks_per_wg, ...).
struct ScanBuildArgs {
    ScanCol  col[kMaxWords];
    uint32_t words;
    uint32_t notnull;
    uint64_t nrows;
    uint64_t blocks_per_wg;
    uint32_t rshift;   // region = hash >> rshift (64: one region)
    uint32_t nregions;
    const uint64_t* sel; // as PipeArgs::sel
};
hipError_t launch_scan_partition(const ScanBuildArgs& a, uint32_t blocks, uint32_t* bh, uint32_t* scan_scratch,
                                 uint64_t* total, uint64_t* parts, uint32_t threads, hipStream_t s);
// Small hash table (2^log_cap slots of a.words words, filled with kEmpty by
// the caller) inserted straight from the scan's pages: same layout and flags
// as launch_ht_insert; *count += rows inserted (zeroed by the caller).
hipError_t launch_scan_insert(const ScanBuildArgs& a, uint64_t* ht, uint32_t log_cap, uint64_t* bloom,
                              uint32_t bloom_log, uint32_t* flags, uint32_t walk_cap, unsigned long long* count,
                              hipStream_t s);
// The LDS region build alone, over an already partitioned `parts` holding
// *n_dev tuples (read on the device: no host round trip in between).
hipError_t launch_region_build(const uint64_t* parts, const uint32_t* bh, uint32_t blocks, const uint64_t* n_dev,
                               uint32_t words, uint64_t* ht, uint32_t log_cap, uint32_t log_regions,
                               uint64_t* bloom, uint32_t bloom_log, uint32_t* flags, uint32_t threads,
                               uint32_t walk_cap, hipStream_t s);
// Launches the fused pipeline. `slots` = tuple width (selects the kernel variant).
hipError_t launch_pipeline(const PipeArgs& a, uint32_t slots, uint32_t grid_waves, hipStream_t s);
// Lean variant for one probe without residual predicates and a count or
// fixed-width page target (<= 4 scanned, <= 4 output columns, <= 8 slots).
hipError_t launch_probe1(const PipeArgs& a, uint32_t slots, uint32_t grid_waves, bool aligned_pairs, hipStream_t s);
bool       probe1_applicable(const PipeArgs& a, uint32_t slots);
uint32_t   pipeline_waves_per_block();
// Packs a two-word table of `cap` slots into one word per slot,
// key | attr << 32 (empty slots stay all ones); sets *flag when an entry's key
// or attribute needs more than 32 bits, the attribute is NULL, or the packed
// word would read as empty. `flag` must be zero before the launch.
hipError_t launch_pack_table(const uint64_t* ht, uint64_t cap, uint64_t* packed, uint32_t* flag, hipStream_t s);
// Adds the per-wave closing sums of a launch (LeanArgs::wave_sums,
// PipeArgs::wave_sums) of `nout` output columns into the counters' tuples,
// nnz and col_pages.
hipError_t launch_fold_sums(const unsigned long long* wave_sums, uint32_t stride, uint32_t nout, PipeCounters* c,
                            hipStream_t s);

// ---- join sampling (the planner's cardinality estimates) ----
// One scanned column of the sample gather: rows rowids[first .. first + n) of
// `col` are decoded into out[first ..]. The descriptors and row ids are in
// device memory.
struct SampleCol {
    ScanCol  col;
    uint32_t first;
    uint32_t n;
};
hipError_t launch_sample_gather(const SampleCol* cols, uint32_t ncols, const uint64_t* rowids, uint64_t* out,
                                hipStream_t s);
// A built hash table as the sample re-filter sees it: plain entries of
// `words` words (key first) or grouped two-word slots; linear probing wraps
// inside regions of mask + 1 slots.
// ---- scan restrictions (reference engine/storage/RestrictionLogic.cpp:40-447) ----
// A restriction keeps the rows whose value in one column passes it; NULL never
// passes. Kinds: equal to a, equal to a or b, a <= v <= b (unsigned 64-bit
// slots: the reference's Lt/Gt/Between over contiguous keys), membership in a
// direct-mapped table of up to 32 values (TinyTable; slot = value mod 2^bits,
// Fibonacci hash, or folded Fibonacci hash), a table's bloom filter (Join), or
// exact membership in a hash table (JoinPrecise).
constexpr int kMaxFilters = 4;
constexpr int kTinySlots  = 32;
enum : uint32_t { FLT_EQ = 0, FLT_EQ2 = 1, FLT_RANGE = 2, FLT_TINY = 3, FLT_BLOOM = 4, FLT_TABLE = 5 };
enum : uint32_t { TINY_MOD = 0, TINY_FIBO = 1, TINY_FIBO2 = 2 };
struct FilterSpec {
    ScanCol         col;
    uint32_t        kind;
    uint32_t        tiny_bits;  // FLT_TINY: 2^tiny_bits slots
    uint32_t        tiny_hash;  // FLT_TINY: TINY_*
    uint32_t        words;      // FLT_TABLE: words per slot
    uint64_t        a, b;       // FLT_EQ / FLT_EQ2 / FLT_RANGE
    const uint64_t* table;      // FLT_TABLE: slots; FLT_BLOOM: bloom words
    uint64_t        mask;       // FLT_TABLE: region slots - 1
    uint32_t        shift;      // FLT_TABLE: slot = hash_key(v) >> shift; FLT_BLOOM: log2(words)
    uint32_t        pad;
    uint64_t        tiny[kTinySlots]; // FLT_TINY: values, kNull in empty slots
};
struct FilterArgs {
    uint64_t   nrows;
    uint32_t   n;          // restrictions (<= kMaxFilters)
    uint32_t   accumulate; // AND into the existing selection words
    FilterSpec f[kMaxFilters];
};
// sel[b] = bit i set when row 64b + i passes every restriction (and, with
// `accumulate`, passed before).
hipError_t launch_filter_select(const FilterArgs& a, uint64_t* sel, hipStream_t s);
// Keys of the non-empty slots of a hash table (slot stride `stride` words):
// out[i] for i < min(*count, max), *count = keys found (zeroed by the caller).
hipError_t launch_table_keys(const uint64_t* ht, uint64_t cap, uint32_t stride, uint64_t* out, uint32_t max,
                             unsigned int* count, hipStream_t s);
// Slot of value v in a tiny table (reference RestrictionLogic.cpp:355-382).
__host__ __device__ inline uint32_t tiny_slot(uint64_t v, uint32_t hash, uint32_t bits) {
    switch (hash) {
    case TINY_FIBO: return uint32_t((v * 11400714819323198485ull) >> (64 - bits));
    case TINY_FIBO2: {
        const uint64_t m = v * 11400714819323198485ull;
        return uint32_t(m ^ (m >> 32)) >> (32 - bits);
    }
    default: return uint32_t(v & ((1ull << bits) - 1));
    }
}

struct SampleTable {
    const uint64_t* ht;
    uint64_t        mask;
    uint32_t        shift;
    uint32_t        words; // 2 for a grouped table
};
// found[i] = 1 when keys[i] is a key of the table (NULL never is), else 0.
hipError_t launch_sample_probe(SampleTable t, const uint64_t* keys, uint32_t n, uint32_t* found, hipStream_t s);

// Synthetic generators.
struct GenSpec {
    int32_t  type;
    int32_t  kind;
    uint64_t seed;
    uint64_t stream;
    uint64_t null_stream;
    double   null_frac;
    uint64_t domain;
    double   zipf_s;
    double   miss_frac;
    uint64_t lo, hi;
};
hipError_t launch_gen_fill(const GenSpec& g, const uint32_t* row_start, uint64_t npages, uint8_t* pages,
                           hipStream_t s);

} // namespace cgj
