// Host side of the join pipeline: device memory, device tables, plan import,
// join ordering and pipeline execution. Everything here is host C++; device
// work goes through kernels.hpp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <thread>
#include <functional>
#include <exception>
#include <condition_variable>
#include <atomic>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/gpujoin.h"
#include "kernels.hpp"
#include "planner.hpp"

namespace cgj {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define CGJ_HIP(expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ::cgj::HipError(std::string(#expr) + ": " + hipGetErrorString(e_));                  \
    } while (0)

// Size-bucketed caching allocator over hipMalloc: repeated queries of the same
// shape reuse their buffers instead of paying for hipMalloc/hipFree. Buffers
// share ownership of the heap, so tables and results may outlive their context.
class DeviceHeap : public std::enable_shared_from_this<DeviceHeap> {
public:
    DeviceHeap();
    ~DeviceHeap();
    void* alloc(size_t bytes);
    // `evict_first`: the block goes to the cold end of the cache (released
    // before any other). Result pools are freed in bulk when their results
    // go, and would otherwise push out the tables and staging areas that the
    // next queries reuse.
    void  free(void* p, bool evict_first = false);
    // Releases every cached block of this heap's device.
    void  trim();
    // The stream the owning context runs on: a released block becomes
    // reusable on another stream once this stream has passed the release.
    // nullptr (the context is gone, its stream drained): blocks are idle.
    void  set_stream(hipStream_t s);
    // Cached blocks released on `s` become idle (the stream is drained and
    // about to be destroyed).
    static void forget_stream(int device, hipStream_t s);
    // While worker contexts run queries on one device, releases (each waits
    // for the whole device) are put off until the cached total passes twice
    // the bound; settle() brings the device's cache back under the bound
    // once the workers are done.
    struct Parallel {
        Parallel();
        ~Parallel();
    };
    static void settle(int device);

private:
    static size_t round(size_t bytes);
    // Releases the cached blocks of `device`: an allocation that failed
    // retries after this.
    static void trim_device(int device);
    int                               device_ = 0;
    std::mutex                        mu_;
    hipStream_t                       stream_ = nullptr;
    std::unordered_map<void*, size_t> live_;
    // released blocks below the shared cache's minimum, for this heap alone
    static constexpr size_t           kOwnMax = size_t(256) << 20;
    std::multimap<size_t, void*>      own_;
    size_t                            own_bytes_ = 0;
};

// RAII device buffer from a DeviceHeap.
class DevBuf {
public:
    DevBuf() = default;
    DevBuf(DeviceHeap* h, size_t bytes)
        : heap_(h->shared_from_this()), bytes_(bytes), ptr_(bytes ? h->alloc(bytes) : nullptr) {}
    ~DevBuf() { reset(); }
    DevBuf(DevBuf&& o) noexcept
        : heap_(std::move(o.heap_)), bytes_(o.bytes_), ptr_(o.ptr_), evict_first_(o.evict_first_) {
        o.ptr_ = nullptr;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            reset();
            heap_        = std::move(o.heap_);
            bytes_       = o.bytes_;
            ptr_         = o.ptr_;
            evict_first_ = o.evict_first_;
            o.ptr_       = nullptr;
        }
        return *this;
    }
    // On release the block goes to the cold end of its heap's cache.
    void set_evict_first() { evict_first_ = true; }
    DevBuf(const DevBuf&)            = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    void reset() {
        if (ptr_) heap_->free(ptr_, evict_first_);
        ptr_ = nullptr;
    }
    template <class T = void>
    T*     get() const { return static_cast<T*>(ptr_); }
    size_t bytes() const { return bytes_; }

private:
    std::shared_ptr<DeviceHeap> heap_;
    size_t                      bytes_       = 0;
    void*                       ptr_         = nullptr;
    bool                        evict_first_ = false;
};

// One column of a device-resident table.
struct DevColumn {
    int32_t  type   = 0;
    uint64_t npages = 0;
    uint64_t nrows  = 0;
    DevBuf   pages;     // npages x 8 KiB
    DevBuf   row_start; // u32 [npages + 1]
    DevBuf   blocks;    // BlockIdx [ceil(nrows / 64)]
    uint32_t dense_rows = 0; // see ScanCol::dense_rows

    ScanCol view(uint32_t col_id) const;
};

// Process-wide distinct id of a device table (key of cached join statistics).
uint64_t next_table_uid();

// Key columns (table uid, column) of base tables on which a hash table had to
// be built grouped: a key repeated past the walk cap. A later build on such a
// column starts grouped instead of failing the plain layouts first. A context
// shares its hints with its worker contexts.
class BuildHints {
public:
    bool grouped(uint64_t uid, uint64_t col) const {
        std::lock_guard<std::mutex> lock(mu_);
        return grouped_.count({uid, col}) != 0;
    }
    void mark_grouped(uint64_t uid, uint64_t col) {
        std::lock_guard<std::mutex> lock(mu_);
        grouped_.insert({uid, col});
    }

private:
    mutable std::mutex                       mu_;
    std::set<std::pair<uint64_t, uint64_t>> grouped_;
};

struct DevTable {
    uint64_t               nrows = 0;
    std::vector<DevColumn> cols;
    uint64_t               uid = next_table_uid();
};

// Output pages of a query, still in HBM: `pool` holds `pool_pages` page slots
// followed by the page directory (column, wave stream, sequence, slot) of the
// `total_pages` pages written. The per-column page order is derived from the
// directory when the result is downloaded.
struct DevResult {
    uint64_t             nrows       = 0;
    std::vector<int32_t> types;
    // Output columns whose rows all hold one value (removed one-row
    // relations) are written on download: the pool holds the other columns,
    // pool column j being output column out_pos[j] of type pool_types[j].
    std::vector<ConstOut> consts;
    std::vector<uint32_t> out_pos;    // empty: every column is in the pool, in order
    std::vector<int32_t>  pool_types;
    uint64_t             total_pages    = 0;
    uint64_t             pool_pages     = 0;
    uint64_t             reserved_pages = 0; // directory entries to scan (valid flag in word 3)
    DevBuf               pool;

    const uint32_t* dir() const { return reinterpret_cast<const uint32_t*>(pool.get<uint8_t>() + pool_pages * 8192); }
};

// Output pages in a pool (a DevResult or one job of a batch): `pool` holds
// `pool_pages` page slots followed by their directory.
struct PoolView {
    uint64_t             nrows = 0;
    std::vector<int32_t> types;
    const uint8_t*       pool           = nullptr;
    uint64_t             pool_pages     = 0;
    uint64_t             reserved_pages = 0;
    uint64_t             total_pages    = 0;
};

// Host threads copying pages between the pinned staging buffers and the
// callers' pages: started once per context, each run split in grains of 64
// pages taken in turn by the workers and the calling thread.
class CopyPool {
public:
    explicit CopyPool(uint32_t threads);
    ~CopyPool();
    void run(uint64_t n, const std::function<void(uint64_t)>& f); // f(i) for i in [0, n)

private:
    static constexpr uint64_t                 kGrain = 64;
    void                                      work();
    void                                      drain();
    std::vector<std::thread>                  th_;
    std::mutex                                mu_;
    std::condition_variable                   cv_, done_;
    const std::function<void(uint64_t)>*      job_ = nullptr;
    uint64_t                                  n_   = 0;
    std::atomic<uint64_t>                     next_{0};
    uint32_t                                  active_ = 0;
    uint64_t                                  gen_    = 0;
    bool                                      stop_   = false;
    std::exception_ptr                        err_;
};

class Context;
class Estimator;

// Tunables of one hash join, chosen per (build rows, probe rows, key width).
struct JoinTuning {
    double   max_load    = 0.75; // capacity = next_pow2(build / max_load)
    uint32_t bloom_bits  = 0;    // bloom bits per key (0 = no bloom filter)
    bool     partitioned = true; // radix-partitioned LDS build (else global CAS inserts)
};

// Execution options (cgj_set_option); automatic choices unless overridden.
struct Options {
    // Probe-side radix partitioning: its scatter costs more than the cache
    // misses it saves on MI355X (HBM latency is hidden by occupancy), so it is
    // off unless a size threshold is set.
    uint64_t radix_min_table_bytes = ~uint64_t(0);       // partition the probe side above this table size
    uint64_t radix_min_rows        = uint64_t(4) << 20;  // ... and this many driving rows
    double   max_load              = 0;                  // hash-table load factor (0 = automatic)
    int32_t  bloom_bits            = -1;                 // bloom bits per key (-1 = automatic)
    int32_t  partitioned_build     = -1;                 // radix-partitioned build (-1 = automatic)
    bool     lean_kernel           = true;               // use the one-probe kernel where it applies
    uint32_t lean_waves_per_cu     = 24;                 // lean kernel grid: waves per CU (6 per SIMD)
    uint32_t build_regions_log     = 14;                 // partitioned build: at most 2^this regions
    bool     aligned_pairs         = true;               // lean probe: 2-slot groups aligned to 2 slots
    // Partitioned build: threads per workgroup (256/512/1024). Its LDS tables
    // allow one or two workgroups per CU, so wider ones keep more loads in
    // flight (c2 build 4.8 -> 3.9 ms at 1024).
    uint32_t build_threads         = 1024;
    // Grouped tables: an insert that passes this many slots holding its own
    // key rebuilds the table grouped by key (0 = never group);
    // `force_grouped` groups every table (tests).
    uint32_t group_walk            = 32;
    bool     force_grouped         = false;
    // Hash tables on a plain scan are partitioned straight from the pages
    // (no staged tuples).
    bool     scan_build            = true;
    bool     block_prefilter       = true; // early bloom filters tested eight 64-row blocks at a time first
    uint64_t scan_insert_max_bytes = uint64_t(8) << 20; // hash tables on a plain scan up to this size are
                                                        // inserted straight from the pages (0: never)
    uint32_t scan_workgroups       = 512; // workgroups of the scan-driven partition (16..4096)
    // Re-plan after every hash-table build with the tables built so far as
    // fixed build sides (the reference's incremental loop); 0: plan once.
    bool     incremental           = true;
    bool     restrictions          = true; // incremental loop: restrict key producers after each build
    bool     singletons            = true; // remove one-row relations before planning
    // The one-probe kernel reads two-word INT32 tables packed (key | attr << 32).
    bool     packed_tables         = true;
};

// Tunables of one batched join.
struct BatchTune {
    double   max_load        = 0.5;    // hash-table capacity = next_pow2(build / max_load)
    uint32_t bloom_bits      = 0;      // bloom bits per key (0 = none)
    uint32_t probe_task_rows = 65536;  // probe rows per task (multiple of 64)
    uint32_t build_task_rows = 32768;  // build rows per task (multiple of 64)
};

// Chooses BatchTune per (build rows, probe rows, key width) size class. Starts
// from rules (cache sizes, task counts) and keeps the settings that measured
// fastest per class in Batch::autotune (per-task wall-clock ticks, summed per
// job in device memory). The table lives in the context, so later batches of
// similar shapes reuse it.
class BatchTuner {
public:
    using Key = std::tuple<uint32_t, uint32_t, uint32_t>; // log2 build rows, log2 probe/build, key bytes
    static Key key(uint64_t build_rows, uint64_t probe_rows, uint32_t key_bytes);
    BatchTune  pick(uint64_t build_rows, uint64_t probe_rows, uint32_t key_bytes) const;
    void       store(const Key& k, const BatchTune& t) { tuned_[k] = t; }
    size_t     tuned_classes() const { return tuned_.size(); }

private:
    std::map<Key, BatchTune> tuned_;
};

class Context {
public:
    explicit Context(int device);
    ~Context();

    Options options;
    // Sets a named option; false for an unknown name.
    bool set_option(const std::string& name, double value);

    int         device() const { return device_; }
    hipStream_t stream() const { return stream_; }
    DeviceHeap& heap() { return *heap_; }

    // Uploads the table's pages; with `used`, only the flagged columns.
    std::unique_ptr<DevTable> upload(const cgj_host_table& t, const std::vector<bool>* used = nullptr);
    std::unique_ptr<DevTable> generate(uint64_t seed, uint64_t nrows, const cgj_gen_column* cols, uint32_t ncols);
    void prepare_column(DevColumn& c);
    // Prepares several columns with one host round trip.
    void prepare_columns(const std::vector<DevColumn*>& cols);

    std::unique_ptr<DevResult> execute(const cgj_plan& plan, const std::vector<const DevTable*>& inputs);
    // Runs independent plans concurrently: `workers` host threads, each with
    // its own worker context (HIP stream, device heap), take plans from a
    // shared counter, so one plan's host round trips overlap another's
    // kernels. Worker contexts persist across calls (warm heaps). Queries are
    // planned by this context's estimator (by the workers, with the
    // max-of-inputs rule, when that is the model).
    std::vector<std::unique_ptr<DevResult>> execute_many(const std::vector<const cgj_plan*>& plans,
                                                         const std::vector<std::vector<const DevTable*>>& inputs,
                                                         uint32_t workers);
    uint64_t count(const cgj_plan& plan, const std::vector<const DevTable*>& inputs);
    void     download(const DevResult& r, cgj_result& out);
    void     download_pool(const PoolView& v, cgj_result& out);
    // Host copy of a device column's pages, one operator-new block per page.
    void     download_column(const DevColumn& c, cgj_result_column& out);
    // Pinned host staging of page uploads and downloads: two buffers of
    // kStagingBytes, allocated on first use and kept for the context's
    // lifetime; staging(b) waits until the last copy through buffer b is done.
    static constexpr size_t kStagingBytes = size_t(64) << 20;
    uint8_t*                staging(int b);
    // Pinned host buffer of at least `bytes` for small transfers (the
    // estimator's features and outputs), grown on demand and kept for the
    // context's lifetime.
    uint8_t* small_pinned(size_t bytes);
    CopyPool& copy_pool();
    // Pinned host buffer of kSamplePinnedBytes for the sample re-filter of
    // the incremental loop (keys out, hit flags back), apart from
    // small_pinned so a re-filter queued ahead of a build's last
    // synchronisation is not overwritten by the build's own transfers.
    static constexpr size_t kSamplePinnedBytes = size_t(64) << 10; // 64 relations x 64 keys x 12 B
    uint8_t*                sample_pinned();
    void                    staging_done(int b); // records the copy just issued through buffer b
    double   copy_probe(uint64_t bytes, uint32_t reps);

    cgj_stats   stats{};
    std::string error;

    JoinTuning tune(uint64_t build_rows, uint64_t probe_rows, uint32_t key_bytes) const;
    BatchTuner batch_tuner; // per size class settings of batched joins (kept across batches)

    // Cardinality model of the join ordering (the max-of-inputs rule unless
    // configured otherwise) and its cached join statistics.
    Estimator& estimator() { return *est_; }
    BuildHints& build_hints() { return *hints_; }
    // Result rows of an imported query (any join order; no pages written).
    uint64_t count_query(const Query& q, const std::vector<const DevTable*>& inputs);
    // Result rows of the join of the connected relation set `relset` of a plan
    // (relations numbered in the plan's depth-first scan order, left first).
    uint64_t count_subset(const cgj_plan& plan, const std::vector<const DevTable*>& inputs, uint64_t relset);
    // The join tree the planner picks for a plan under the current estimator:
    // per join, the relation sets of its build side, its probe side, and the
    // estimated cardinality of its result.
    struct PlannedJoin {
        uint64_t build, probe;
        double   card;
    };
    std::vector<PlannedJoin> plan_tree(const cgj_plan& plan, const std::vector<const DevTable*>& inputs);
    uint32_t   cu_count() const { return cus_; }
    uint64_t   lds_bytes() const { return lds_; } // LDS per workgroup
    uint64_t   mem_bytes() const { return mem_; } // device memory

    // A query imported and ordered by this context's estimator, and its
    // execution (on any context of the device) as planned.
    PlannedQuery               plan_query(const cgj_plan& plan, const std::vector<const DevTable*>& inputs);
    // Planned with another context's estimator (its counts and inference run here).
    PlannedQuery plan_query(const cgj_plan& plan, const std::vector<const DevTable*>& inputs, Estimator& est);
    std::unique_ptr<DevResult> run_planned(const PlannedQuery& pq, const std::vector<const DevTable*>& inputs,
                                           bool count_only);
    // Up to 64 deterministic sample rows of every relation's join columns,
    // gathered on the device in one launch; NOT NULL applied.
    std::shared_ptr<Samples> take_samples(const Query& q, const std::vector<const DevTable*>& inputs);
    // Slot values of every attribute of the one-row relations (others: empty).
    std::vector<std::vector<uint64_t>> singleton_values(const Query& q, const std::vector<const DevTable*>& inputs);
    // The bytes of the string a VARCHAR slot (string handle) of column c names.
    std::string read_string(const DevColumn& c, uint64_t handle);

private:
    void check_types(const Query& q, const std::vector<const DevTable*>& inputs) const;
    std::unique_ptr<DevResult> run_query(const cgj_plan& plan, const std::vector<const DevTable*>& inputs,
                                         bool count_only);

    int         device_;
    uint32_t    cus_    = 256;
    uint64_t    lds_    = 65536;
    uint64_t    mem_    = uint64_t(64) << 30;
    hipStream_t stream_ = nullptr;
    std::shared_ptr<DeviceHeap> heap_ = std::make_shared<DeviceHeap>();
    std::unique_ptr<Estimator>  est_;
    std::shared_ptr<BuildHints> hints_ = std::make_shared<BuildHints>();
    void*                       pin_[2]    = {nullptr, nullptr};
    hipEvent_t                  pin_ev_[2] = {nullptr, nullptr};
    std::unique_ptr<CopyPool>   copy_pool_;
    void*                       spin_       = nullptr;
    void*                       sample_pin_ = nullptr;
    size_t                      spin_bytes_ = 0;
    std::vector<std::unique_ptr<Context>> workers_;
};

} // namespace cgj
