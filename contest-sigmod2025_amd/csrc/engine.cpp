// Host engine: device tables, plan import, join ordering and pipeline
// execution over the HIP kernels.
//
// Plan import follows the reference's semantics (engine/query/PlanImport.cpp:24-89):
// every scan-node output column is an attribute, join conditions merge
// attributes into equivalence classes, and the query result is the inner join
// of all scanned relations over those classes, projected on the root's
// output attributes. The given join order and build sides are advisory, as in
// the reference (QueryPlan.cpp:748-782 re-plans from scratch): the planner
// below orders joins itself with a DP over connected sub-graphs.
#include "engine.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <functional>
#include <limits>
#include <malloc.h>
#include <mutex>
#include <numeric>
#include <thread>

#include "estimator.hpp"
#include "gen.hpp"
#include "noalloc.hpp"
#include "planner.hpp"

namespace cgj {

// ---------------------------------------------------------------------------
// device heap
// ---------------------------------------------------------------------------
// Size classes: 4 KiB steps up to 1 MiB, then four classes per power of two
// (2^k x 1, 1.25, 1.5, 1.75; at most 25% over the request), so buffers of
// varying sizes (hash tables, staging areas, page pools of successive
// queries) map to few classes and are served from the cache after warm-up
// instead of by hipMalloc.
size_t DeviceHeap::round(size_t b) {
    if (b <= (1u << 20)) return (b + 4095) & ~size_t(4095);
    uint32_t k = 63 - uint32_t(__builtin_clzll(b)); // 2^k <= b < 2^(k+1)
    size_t   q = size_t(1) << (k - 2);              // class step: a quarter of 2^k
    return (b + q - 1) & ~(q - 1);
}

// Released blocks of at least kSharedMin bytes go to one cache per device,
// shared by the heaps of every context on it (execute_many's workers
// included), so a table or page pool one worker released serves the next
// query of any worker; smaller blocks stay in their heap's own cache. A
// cached block remembers the stream of the heap that released it, and a heap
// takes a block of its own stream when one fits (stream order makes its reuse
// safe). For another stream: while worker contexts run, the release recorded
// an event on the releasing stream and the new stream waits for it; outside
// that, if the old stream still has work queued, an event recorded on it at
// reuse (after the release, so after every kernel that could still use the
// block) is waited for instead.
//
// The cache stays below a quarter of the device's memory: caches that grew
// without bound would leave the runtime itself no memory (kernel launches
// allocate too). A release that takes it over the bound frees the oldest
// blocks until it is 1/8 below it, in one go, since a hipFree waits for the
// whole device. While worker contexts run, the bound is doubled and the cache
// settles under it when the workers are done.
namespace {
constexpr size_t kSharedMin = size_t(4) << 20;
struct CachedBlock {
    size_t      bytes;
    void*       ptr;
    hipStream_t stream; // stream of the releasing heap (nullptr: idle)
    hipEvent_t  ready;  // recorded on `stream` at release (nullptr: none)
    std::multimap<size_t, std::list<CachedBlock>::iterator>::iterator by_size;
};
struct DeviceCache {
    std::mutex                                              mu;
    std::list<CachedBlock>                                  order; // oldest (coldest) first
    std::multimap<size_t, std::list<CachedBlock>::iterator> by_size;
    uint64_t                                                bytes = 0;
    std::vector<hipEvent_t>                                 events; // idle events
    uint32_t                                                heaps = 0;
};
constexpr int         kCacheDevices = 64;
DeviceCache           g_cache[kCacheDevices];
std::atomic<uint32_t> g_parallel{0};
std::atomic<uint64_t> g_cache_cap{0};

DeviceCache& device_cache(int device) { return g_cache[device >= 0 && device < kCacheDevices ? device : 0]; }

uint64_t cache_cap() {
    uint64_t c = g_cache_cap.load();
    if (!c) {
        size_t freeb = 0, total = 0;
        c = hipMemGetInfo(&freeb, &total) == hipSuccess && total ? total / 4 : (uint64_t(16) << 30);
        g_cache_cap.store(c);
    }
    return c;
}

// Moves blocks from the cold end of `c` into `out` until at most `keep` bytes
// stay cached (the caller holds c.mu and frees `out` after unlocking).
void take_oldest(DeviceCache& c, uint64_t keep, std::vector<CachedBlock>& out) {
    while (c.bytes > keep && !c.order.empty()) {
        CachedBlock& b = c.order.front();
        c.bytes -= b.bytes;
        c.by_size.erase(b.by_size);
        out.push_back(b);
        c.order.pop_front();
    }
}

// hipFree waits for the whole device, so no stream needs waiting for.
void release(DeviceCache& c, std::vector<CachedBlock>& blocks) {
    if (blocks.empty()) return;
    for (auto& b : blocks) (void)hipFree(b.ptr);
    std::lock_guard<std::mutex> lock(c.mu);
    for (auto& b : blocks)
        if (b.ready) c.events.push_back(b.ready);
    blocks.clear();
}

// An idle event from the device's pool, or a new one (nullptr on failure).
hipEvent_t take_event(DeviceCache& c) {
    {
        std::lock_guard<std::mutex> lock(c.mu);
        if (!c.events.empty()) {
            hipEvent_t e = c.events.back();
            c.events.pop_back();
            return e;
        }
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return e;
}
} // namespace

DeviceHeap::DeviceHeap() {
    if (hipGetDevice(&device_) != hipSuccess) device_ = 0;
    DeviceCache&                c = device_cache(device_);
    std::lock_guard<std::mutex> lock(c.mu);
    ++c.heaps;
}

void DeviceHeap::set_stream(hipStream_t s) {
    std::lock_guard<std::mutex> lock(mu_);
    stream_ = s;
}

DeviceHeap::Parallel::Parallel() { ++g_parallel; }
DeviceHeap::Parallel::~Parallel() { --g_parallel; }

void DeviceHeap::settle(int device) {
    const uint64_t           cap = cache_cap();
    DeviceCache&             c   = device_cache(device);
    std::vector<CachedBlock> out;
    {
        std::lock_guard<std::mutex> lock(c.mu);
        if (c.bytes > cap) take_oldest(c, cap - cap / 8, out);
    }
    release(c, out);
}

void DeviceHeap::trim_device(int device) {
    DeviceCache&             c = device_cache(device);
    std::vector<CachedBlock> out;
    {
        std::lock_guard<std::mutex> lock(c.mu);
        take_oldest(c, 0, out);
    }
    release(c, out);
}

void DeviceHeap::trim() { trim_device(device_); }

void DeviceHeap::forget_stream(int device, hipStream_t s) {
    DeviceCache&                c = device_cache(device);
    std::lock_guard<std::mutex> lock(c.mu);
    for (auto& b : c.order)
        if (b.stream == s) b.stream = nullptr;
}

void* DeviceHeap::alloc(size_t bytes) {
    const size_t r = round(bytes);
    DeviceCache& c = device_cache(device_);
    hipStream_t  s;
    {
        std::lock_guard<std::mutex> lock(mu_);
        s = stream_;
        if (r < kSharedMin) {
            auto it = own_.lower_bound(r);
            if (it != own_.end() && it->first <= r + r / 2) {
                void* p  = it->second;
                live_[p] = it->first;
                own_bytes_ -= it->first;
                own_.erase(it);
                return p;
            }
        }
    }
    void*       p     = nullptr;
    size_t      got   = 0;
    hipStream_t prev  = nullptr;
    hipEvent_t  ready = nullptr;
    if (r >= kSharedMin) {
        std::lock_guard<std::mutex> lock(c.mu);
        // the first fitting block of this stream, else the first fitting one
        auto best = c.by_size.end();
        for (auto it = c.by_size.lower_bound(r); it != c.by_size.end() && it->first <= r + r / 2; ++it) {
            if (best == c.by_size.end()) best = it;
            if (it->second->stream == s) {
                best = it;
                break;
            }
        }
        if (best != c.by_size.end()) {
            const CachedBlock b = *best->second;
            c.order.erase(best->second);
            c.by_size.erase(best);
            c.bytes -= b.bytes;
            p     = b.ptr;
            got   = b.bytes;
            prev  = b.stream == s ? nullptr : b.stream; // the same stream orders its own reuse
            ready = b.ready;
            if (ready && !prev) {
                c.events.push_back(ready);
                ready = nullptr;
            }
        }
    }
    if (p && prev) {
        hipError_t e = hipSuccess;
        if (ready) {
            // recorded at the release
            e = s ? hipStreamWaitEvent(s, ready, 0) : hipEventSynchronize(ready);
        } else if (hipStreamQuery(prev) != hipSuccess) {
            (void)hipGetLastError(); // hipErrorNotReady: work still queued on the old stream
            ready = take_event(c);
            e     = ready ? hipEventRecord(ready, prev) : hipErrorOutOfMemory;
            if (e == hipSuccess) e = s ? hipStreamWaitEvent(s, ready, 0) : hipEventSynchronize(ready);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            e = hipStreamSynchronize(prev);
        }
        if (ready) {
            std::lock_guard<std::mutex> lock(c.mu);
            c.events.push_back(ready);
        }
        CGJ_HIP(e);
    }
    if (!p) {
        got          = r;
        hipError_t e = hipMalloc(&p, r);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            trim_device(device_);
            e = hipMalloc(&p, r);
        }
        if (e != hipSuccess) throw HipError("hipMalloc(" + std::to_string(r) + "): " + hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> lock(mu_);
    live_[p] = got;
    return p;
}

void DeviceHeap::free(void* p, bool evict_first) {
    size_t      bytes;
    hipStream_t s;
    {
        std::lock_guard<std::mutex> lock(mu_);
        auto                        it = live_.find(p);
        if (it == live_.end()) return;
        bytes = it->second;
        live_.erase(it);
        s = stream_;
        if (bytes < kSharedMin) {
            // small blocks stay with their heap, bounded by kOwnMax bytes
            own_.emplace(bytes, p);
            own_bytes_ += bytes;
            while (own_bytes_ > kOwnMax && !own_.empty()) {
                auto big = std::prev(own_.end());
                own_bytes_ -= big->first;
                (void)hipFree(big->second);
                own_.erase(big);
            }
            return;
        }
    }
    DeviceCache& c     = device_cache(device_);
    hipEvent_t   ready = nullptr;
    if (s && g_parallel) {
        // other workers may take the block while this stream still runs
        ready = take_event(c);
        if (ready && hipEventRecord(ready, s) != hipSuccess) {
            (void)hipGetLastError();
            std::lock_guard<std::mutex> lock(c.mu);
            c.events.push_back(ready);
            ready = nullptr; // the lazy path at reuse still orders it
        }
    }
    std::vector<CachedBlock> out;
    {
        std::lock_guard<std::mutex> lock(c.mu);
        auto blk =
            c.order.insert(evict_first ? c.order.begin() : c.order.end(), CachedBlock{bytes, p, s, ready, {}});
        blk->by_size = c.by_size.emplace(bytes, blk);
        c.bytes += bytes;
        const uint64_t cap = g_parallel ? 2 * cache_cap() : cache_cap();
        if (c.bytes > cap) take_oldest(c, cap - cap / 8, out);
    }
    release(c, out);
}

DeviceHeap::~DeviceHeap() {
    DeviceCache& c    = device_cache(device_);
    bool         last = false;
    {
        std::lock_guard<std::mutex> lock(c.mu);
        last = --c.heaps == 0;
    }
    // the last heap of the device leaves nothing cached behind
    if (last) trim_device(device_);
    for (auto& kv : own_) (void)hipFree(kv.second);
    own_.clear();
    for (auto& kv : live_) (void)hipFree(kv.first);
    live_.clear();
}

ScanCol DevColumn::view(uint32_t col_id) const {
    ScanCol c;
    c.pages     = pages.get<uint8_t>();
    c.row_start = row_start.get<uint32_t>();
    c.blocks    = blocks.get<BlockIdx>();
    c.type       = type;
    c.col_id     = col_id;
    c.dense_rows = dense_rows;
    return c;
}

// ---------------------------------------------------------------------------
// context, tables
// ---------------------------------------------------------------------------
uint64_t next_table_uid() {
    static std::atomic<uint64_t> next{1};
    return next.fetch_add(1);
}

namespace {
// Result pages are one operator-new block of 8 KiB each, freed by the caller
// after use. glibc returns freed memory at the top of its heaps to the
// kernel once it passes the trim threshold (128 KiB by default), so the next
// result's pages are faulted in again, page by page: on c2 at 1/10 scale
// (1 GB of result pages) that was 215 of 253 ms. With the threshold at its
// maximum and the heap grown 256 MiB at a time, freed pages stay mapped and
// the next result reuses them (45 ms). The reference keeps its pages for
// reuse through a global operator new/delete (engine/infra/PageMemory.cpp:138-201).
void keep_freed_pages() {
    static std::once_flag once;
    std::call_once(once, [] {
        (void)mallopt(M_TRIM_THRESHOLD, std::numeric_limits<int>::max());
        (void)mallopt(M_TOP_PAD, 256 << 20);
    });
}
} // namespace

Context::Context(int device) : device_(device), est_(std::make_unique<Estimator>()) {
    keep_freed_pages();
    CGJ_HIP(hipSetDevice(device));
    CGJ_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    heap_->set_stream(stream_);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        if (prop.multiProcessorCount > 0) cus_ = uint32_t(prop.multiProcessorCount);
        if (prop.sharedMemPerBlock > 0) lds_ = uint64_t(prop.sharedMemPerBlock);
        if (prop.totalGlobalMem > 0) mem_ = uint64_t(prop.totalGlobalMem);
    }
    (void)noalloc_scratch(device); // per-device address of the no-allocation entry point's array
}

Context::~Context() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (int b = 0; b < 2; ++b) {
        if (pin_ev_[b]) (void)hipEventDestroy(pin_ev_[b]);
        if (pin_[b]) (void)hipHostFree(pin_[b]);
    }
    if (spin_) (void)hipHostFree(spin_);
    if (sample_pin_) (void)hipHostFree(sample_pin_);
    workers_.clear();
    est_.reset(); // its network's buffers go back to the heap
    // blocks released from now on (tables and results outliving the
    // context) are idle: the stream is drained before it goes
    if (stream_) (void)hipStreamSynchronize(stream_);
    heap_->set_stream(nullptr);
    if (stream_) {
        DeviceHeap::forget_stream(device_, stream_);
        (void)hipStreamDestroy(stream_);
    }
}

void Context::prepare_column(DevColumn& c) { prepare_columns({&c}); }

void Context::prepare_columns(const std::vector<DevColumn*>& cols) {
    std::vector<DevColumn*> todo;
    for (DevColumn* c : cols) {
        c->row_start = DevBuf(heap_.get(), (c->npages + 1) * sizeof(uint32_t));
        c->blocks    = DevBuf(heap_.get(), ((c->nrows + 63) / 64 + 1) * sizeof(BlockIdx));
        if (c->npages == 0) {
            CGJ_HIP(hipMemsetAsync(c->row_start.get(), 0, sizeof(uint32_t), stream_));
            if (c->nrows) throw std::runtime_error("column has rows but no pages");
            continue;
        }
        todo.push_back(c);
    }
    if (todo.empty()) return;
    // per column: rows, then the two dense words
    DevBuf              total(heap_.get(), todo.size() * 2 * sizeof(uint64_t));
    std::vector<DevBuf> scratch;
    CGJ_HIP(hipMemsetAsync(total.get(), 0, todo.size() * 2 * sizeof(uint64_t), stream_));
    for (size_t i = 0; i < todo.size(); ++i) {
        DevColumn& c = *todo[i];
        scratch.emplace_back(heap_.get(), prepare_scratch_words(c.npages) * sizeof(uint32_t) + 16);
        uint64_t* t = total.get<uint64_t>() + 2 * i;
        CGJ_HIP(launch_prepare_column(c.pages.get<uint8_t>(), c.npages, c.type, c.nrows, c.row_start.get<uint32_t>(),
                                      c.blocks.get<BlockIdx>(), scratch.back().get<uint32_t>(), t,
                                      reinterpret_cast<uint32_t*>(t + 1), stream_));
    }
    std::vector<uint64_t> back(todo.size() * 2);
    CGJ_HIP(hipMemcpyAsync(back.data(), total.get(), back.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, stream_));
    CGJ_HIP(hipStreamSynchronize(stream_));
    for (size_t i = 0; i < todo.size(); ++i) {
        DevColumn&     c    = *todo[i];
        const uint64_t rows = back[2 * i];
        if (rows != c.nrows)
            throw std::runtime_error("column pages hold " + std::to_string(rows) + " rows, table has " +
                                     std::to_string(c.nrows));
        const uint32_t uneven = uint32_t(back[2 * i + 1]), n0 = uint32_t(back[2 * i + 1] >> 32);
        c.dense_rows          = (!uneven && n0 >= 64 && c.nrows < (uint64_t(1) << 32)) ? n0 : 0;
    }
}

uint8_t* Context::staging(int b) {
    if (!pin_[b]) {
        CGJ_HIP(hipHostMalloc(&pin_[b], kStagingBytes, hipHostMallocDefault));
        CGJ_HIP(hipEventCreateWithFlags(&pin_ev_[b], hipEventDisableTiming));
    } else {
        CGJ_HIP(hipEventSynchronize(pin_ev_[b]));
    }
    return static_cast<uint8_t*>(pin_[b]);
}

void Context::staging_done(int b) { CGJ_HIP(hipEventRecord(pin_ev_[b], stream_)); }

uint8_t* Context::small_pinned(size_t bytes) {
    if (bytes > spin_bytes_) {
        CGJ_HIP(hipStreamSynchronize(stream_)); // no copy through the old buffer is in flight
        if (spin_) CGJ_HIP(hipHostFree(spin_));
        spin_       = nullptr;
        spin_bytes_ = 0;
        const size_t want = std::max<size_t>(bytes, size_t(256) << 10);
        CGJ_HIP(hipHostMalloc(&spin_, want, hipHostMallocDefault));
        spin_bytes_ = want;
    }
    return static_cast<uint8_t*>(spin_);
}

uint8_t* Context::sample_pinned() {
    if (!sample_pin_) CGJ_HIP(hipHostMalloc(&sample_pin_, kSamplePinnedBytes, hipHostMallocDefault));
    return static_cast<uint8_t*>(sample_pin_);
}


CopyPool::CopyPool(uint32_t threads) {
    for (uint32_t t = 0; t + 1 < threads; ++t) th_.emplace_back([this] { work(); });
}

CopyPool::~CopyPool() {
    {
        std::lock_guard<std::mutex> lock(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
}

void CopyPool::drain() {
    for (;;) {
        const uint64_t i0 = next_.fetch_add(kGrain);
        if (i0 >= n_) return;
        const uint64_t i1 = std::min(n_, i0 + kGrain);
        try {
            for (uint64_t i = i0; i < i1; ++i) (*job_)(i);
        } catch (...) {
            std::lock_guard<std::mutex> lock(mu_);
            if (!err_) err_ = std::current_exception();
            next_ = n_;
        }
    }
}

void CopyPool::work() {
    uint64_t seen = 0;
    for (;;) {
        {
            std::unique_lock<std::mutex> lock(mu_);
            cv_.wait(lock, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            ++active_;
        }
        drain();
        {
            std::lock_guard<std::mutex> lock(mu_);
            if (--active_ == 0) done_.notify_all();
        }
    }
}

void CopyPool::run(uint64_t n, const std::function<void(uint64_t)>& f) {
    if (n < 2 * kGrain || th_.empty()) {
        for (uint64_t i = 0; i < n; ++i) f(i);
        return;
    }
    {
        std::lock_guard<std::mutex> lock(mu_);
        job_  = &f;
        n_    = n;
        next_ = 0;
        err_  = nullptr;
        ++gen_;
    }
    cv_.notify_all();
    drain(); // the caller works too
    std::unique_lock<std::mutex> lock(mu_);
    done_.wait(lock, [&] { return active_ == 0 && next_ >= n_; });
    job_ = nullptr;
    if (err_) std::rethrow_exception(err_);
}

CopyPool& Context::copy_pool() {
    if (!copy_pool_)
        copy_pool_ = std::make_unique<CopyPool>(std::min<uint32_t>(16, std::max(1u, std::thread::hardware_concurrency())));
    return *copy_pool_;
}

std::unique_ptr<DevTable> Context::upload(const cgj_host_table& t, const std::vector<bool>* used) {
    auto dt   = std::make_unique<DevTable>();
    dt->nrows = t.num_rows;
    // Chunks alternate between the two staging buffers: gathering chunk k+1
    // on the host overlaps the copy of chunk k.
    const uint64_t per   = kStagingBytes / kPageSize;
    int            chunk = 0;
    for (uint32_t ci = 0; ci < t.num_columns; ++ci) {
        const cgj_host_column& hc = t.columns[ci];
        DevColumn              c;
        c.type  = hc.type;
        c.nrows = t.num_rows;
        if (used && !(*used)[ci]) { // never scanned: keep an empty placeholder
            dt->cols.push_back(std::move(c));
            continue;
        }
        c.npages = hc.num_pages;
        c.pages  = DevBuf(heap_.get(), c.npages * kPageSize + 64);
        for (uint64_t p0 = 0; p0 < c.npages; p0 += per, ++chunk) {
            const uint64_t n   = std::min(per, c.npages - p0);
            const int      b   = chunk & 1;
            uint8_t*       buf = staging(b);
            copy_pool().run(n, [&](uint64_t i) { std::memcpy(buf + i * kPageSize, hc.pages[p0 + i], kPageSize); });
            CGJ_HIP(hipMemcpyAsync(c.pages.get<uint8_t>() + p0 * kPageSize, buf, n * kPageSize,
                                   hipMemcpyHostToDevice, stream_));
            staging_done(b);
        }
        dt->cols.push_back(std::move(c));
    }
    std::vector<DevColumn*> prep;
    for (auto& c : dt->cols)
        if (c.npages || c.nrows == 0) prep.push_back(&c); // placeholders of unscanned columns stay unprepared
    prepare_columns(prep);
    CGJ_HIP(hipStreamSynchronize(stream_));
    return dt;
}

// Page boundaries of a generated fixed-width column, by the contest writer's rule.
static std::vector<uint32_t> gen_row_starts(const cgj_gen_column& g, uint64_t seed, uint64_t nrows) {
    std::vector<uint32_t> rs;
    uint32_t              w     = g.type == CGJ_INT32 ? 4 : 8;
    uint32_t              first = w;
    rs.push_back(0);
    if (g.null_frac <= 0.0) {
        uint32_t cap = 0;
        {
            uint32_t rows = 0, end = first;
            while (!page_full(rows, end, true)) {
                ++rows;
                end += w;
            }
            cap = rows;
        }
        for (uint64_t r = cap; r < nrows; r += cap) rs.push_back(uint32_t(r));
        rs.push_back(uint32_t(nrows));
        return rs;
    }
    uint64_t base = stream_base(seed, g.null_stream);
    uint32_t rows = 0, end = first;
    for (uint64_t r = 0; r < nrows; ++r) {
        bool valid = !(unit(draw(base, r)) < g.null_frac);
        if (page_full(rows, end, valid)) {
            rs.push_back(uint32_t(r));
            rows = 0;
            end  = first;
        }
        ++rows;
        if (valid) end += w;
    }
    if (rows) rs.push_back(uint32_t(nrows));
    return rs;
}

static GenSpec to_spec(const cgj_gen_column& g, uint64_t seed) {
    GenSpec s;
    s.type        = g.type;
    s.kind        = g.kind;
    s.seed        = seed;
    s.stream      = g.stream;
    s.null_stream = g.null_stream;
    s.null_frac   = g.null_frac;
    s.domain      = g.domain ? g.domain : 1;
    s.zipf_s      = g.zipf_s;
    s.miss_frac   = g.miss_frac;
    s.lo          = g.lo;
    s.hi          = g.hi;
    return s;
}

std::unique_ptr<DevTable> Context::generate(uint64_t seed, uint64_t nrows, const cgj_gen_column* cols,
                                            uint32_t ncols) {
    auto dt   = std::make_unique<DevTable>();
    dt->nrows = nrows;
    for (uint32_t ci = 0; ci < ncols; ++ci) {
        const cgj_gen_column& g = cols[ci];
        if (g.type == CGJ_VARCHAR) throw std::runtime_error("generator: VARCHAR columns are not generated on device");
        std::vector<uint32_t> rs = gen_row_starts(g, seed, nrows);
        DevColumn             c;
        c.type   = g.type;
        c.nrows  = nrows;
        c.npages = nrows ? rs.size() - 1 : 0;
        c.pages  = DevBuf(heap_.get(), c.npages * kPageSize + 64);
        if (c.npages) {
            DevBuf drs(heap_.get(), rs.size() * sizeof(uint32_t));
            CGJ_HIP(hipMemcpyAsync(drs.get(), rs.data(), rs.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
            CGJ_HIP(hipMemsetAsync(c.pages.get(), 0, c.npages * kPageSize, stream_));
            CGJ_HIP(launch_gen_fill(to_spec(g, seed), drs.get<uint32_t>(), c.npages, c.pages.get<uint8_t>(), stream_));
            CGJ_HIP(hipStreamSynchronize(stream_));
        }
        prepare_column(c);
        dt->cols.push_back(std::move(c));
    }
    return dt;
}

JoinTuning Context::tune(uint64_t build_rows, uint64_t probe_rows, uint32_t key_bytes) const {
    JoinTuning t = autotune_join(build_rows, probe_rows, key_bytes);
    if (options.max_load > 0) t.max_load = options.max_load;
    if (options.bloom_bits >= 0) t.bloom_bits = uint32_t(options.bloom_bits);
    if (options.partitioned_build >= 0) t.partitioned = options.partitioned_build != 0;
    return t;
}

bool Context::set_option(const std::string& name, double v) {
    if (name == "radix_min_table_bytes") options.radix_min_table_bytes = v >= 1.8e19 ? ~uint64_t(0) : uint64_t(v);
    else if (name == "radix_min_rows") options.radix_min_rows = uint64_t(v);
    else if (name == "max_load") options.max_load = (v > 0 && v < 1) ? v : 0;
    else if (name == "bloom_bits") options.bloom_bits = int32_t(v);
    else if (name == "partitioned_build") options.partitioned_build = int32_t(v);
    else if (name == "lean_kernel") options.lean_kernel = v != 0;
    else if (name == "build_regions_log") options.build_regions_log = v >= 1 && v <= 14 ? uint32_t(v) : 14;
    else if (name == "lean_waves_per_cu") options.lean_waves_per_cu = v >= 4 && v <= 64 ? uint32_t(v) : 24;
    else if (name == "aligned_pairs") options.aligned_pairs = v != 0;
    else if (name == "build_threads") options.build_threads = (v == 256 || v == 512) ? uint32_t(v) : 1024u;
    else if (name == "group_walk") options.group_walk = v >= 0 && v < 1e6 ? uint32_t(v) : 32u;
    else if (name == "force_grouped") options.force_grouped = v != 0;
    else if (name == "scan_build") options.scan_build = v != 0;
    else if (name == "block_prefilter") options.block_prefilter = v != 0;
    else if (name == "scan_insert_max_bytes") options.scan_insert_max_bytes = v > 0 ? uint64_t(v) : 0;
    else if (name == "scan_workgroups") options.scan_workgroups = v >= 16 && v <= 4096 ? uint32_t(v) : 512u;
    else if (name == "incremental") options.incremental = v != 0;
    else if (name == "restrictions") options.restrictions = v != 0;
    else if (name == "singletons") options.singletons = v != 0;
    else if (name == "packed_tables") options.packed_tables = v != 0;
    else return false;
    return true;
}


// ---------------------------------------------------------------------------
// query execution
// ---------------------------------------------------------------------------
namespace {

uint32_t log2_ceil(uint64_t x) {
    uint32_t l = 0;
    while ((1ull << l) < x) ++l;
    return l;
}

// A hash table built on device: entries [key, attrs...], one word per class.
struct HashTable {
    DevBuf                slots;
    DevBuf                bloom;
    // Grouped layout (keys that repeat many times): `slots` holds 2-word
    // entries [key, start << 24 | count]; words 1.. of the key's tuples are
    // rows start.. of `payload` (words - 1 words each).
    bool                  grouped = false;
    DevBuf                payload;
    uint64_t              cap       = 0;
    uint64_t              rmask     = 0; // linear probing wraps inside regions of rmask + 1 slots
    uint32_t              shift     = 64;
    uint32_t              words     = 1;
    uint32_t              bloom_log = 0;
    bool                  has_bloom = false;
    bool                  unique    = true;
    uint64_t              tuples    = 0;
    std::vector<uint32_t> eqs; // class of each word (word 0 = key)
    // Packed copy for the one-probe kernel (two-word tables of INT32 classes):
    // one word per slot, key | attr << 32, made on the first probe that can
    // use it; `pack_flag` nonzero means an entry did not fit and the kernel
    // reads `slots`.
    mutable DevBuf        packed;
    mutable DevBuf        pack_flag;
};

// Materialised tuples: the driver of a pipeline that continues a long chain.
struct Tuples {
    DevBuf                buf;
    uint64_t              n     = 0;
    std::vector<uint32_t> eqs; // class of each word
};

struct Driver {
    int           rel    = -1; // base relation, or -1 when `tuples` drives
    const Tuples* tuples = nullptr;
    uint64_t      set    = 0;  // relations covered
};

struct ProbeStep {
    const HashTable* ht;
    uint32_t         key_eq;
    uint64_t         set; // relations of the build side
};

struct Target {
    uint32_t              kind = TARGET_STAGE;
    std::vector<uint32_t> eqs;       // TARGET_STAGE: class of each staged word
    double                est  = 0;  // upper bound of the tuples (at least the estimate)
    double                card = 0;  // the planner's estimate (0 = none)
    DevBuf                stage;     // TARGET_STAGE output
    uint64_t              count = 0; // tuples produced
    uint64_t              slots = 0; // TARGET_STAGE: staging slots used (unused ones have key kEmpty)
    DevResult*            result = nullptr;
};

class Executor {
public:
    Executor(Context& ctx, const Query& q, const PlanTree& t, const std::vector<const DevTable*>& inputs)
        : ctx_(ctx), q_(q), t_(&t), inputs_(inputs) {
        counters_ = DevBuf(&ctx.heap(), sizeof(PipeCounters));
        CGJ_HIP(hipEventCreate(&ev0_));
        CGJ_HIP(hipEventCreate(&ev1_));
        rfilt_.resize(q.rels.size());
        rtables_.resize(q.rels.size());
        sel_.resize(q.rels.size());
        sel_done_.assign(q.rels.size(), false);
        for (const QFilter& f : q.filters) add_filter(q.attrs[f.attr].rel, host_filter(f));
    }
    ~Executor() {
        (void)hipEventDestroy(ev0_);
        (void)hipEventDestroy(ev1_);
    }

    std::unique_ptr<DevResult> run(bool count_only = false);
    // The reference's incremental execution (engine/query/QueryPlan.cpp:748-782):
    // plan the remaining inputs (base relations and hash tables built so far),
    // run the first hash-table build of that plan, re-filter the other
    // relations' samples through the new table, plan again; until the plan's
    // next pipeline is the output pipeline. `card` prices vertex sets.
    using VertexCard = std::function<double(const std::vector<PlanVertex>&, uint64_t)>;
    std::unique_ptr<DevResult> run_incremental(bool count_only, const VertexCard& card,
                                               std::shared_ptr<Samples> samples);
    uint64_t group_rows(int g) const { return groups_[g].ht->tuples; }

    double   ms_build = 0, ms_probe = 0, ms_probe_kernel = 0, ms_partition = 0, ms_replan = 0;
    uint32_t launches = 0, retries = 0, retries_build = 0, probe_lean = 0, probe_waves = 0, replans = 0;
    uint32_t probe_packed = 0;
    uint32_t pools_compacted = 0;
    uint32_t blooms_skipped  = 0;
    uint32_t restrictions = 0, semijoins = 0;
    uint64_t build_tuples = 0, probe_rows = 0, build_bytes = 0, out_ref_pages = 0, probe_input_pages = 0;

private:
    double                                  last_ms_ = 0;
    Context&                                ctx_;
    const Query&                            q_;
    const PlanTree*                         t_;
    const std::vector<const DevTable*>&     inputs_;
    // Hash tables built by earlier iterations of the incremental loop: the
    // relations they cover, their key class and the table.
    struct Group {
        uint64_t         set;
        uint32_t         key;
        const HashTable* ht;
    };
    std::vector<Group>       groups_;
    std::shared_ptr<Samples> samples_;
    // The class a join uses as hash key: a built table's own key, else the
    // first class the build side shares with the probe side so far.
    uint32_t key_for(int build_node, uint64_t cur) const;
    // The first hash table the plan under `node` needs that is not built yet
    // (false: every build of node's probe chain is a built table).
    bool     next_build(int node, int& x, uint32_t& key, uint64_t& probe_rows) const;
    // Runs the probe chain of `node` into target t (builds first); false when a
    // build side is empty (the inner join is empty).
    bool     run_chain(int node, Target& t);
    // Sample re-filter through a table the incremental loop builds: the
    // loop arms it before the build, the build queues the probe of the
    // samples ahead of its last synchronisation (queue_refilter), and the
    // loop applies the hits afterwards (finish_refilter), so the re-filter
    // adds no round trip of its own.
    // A table that gets a bloom filter is probed with the samples as well
    // (`for_bloom`), even where nothing is left to plan: a pipeline probes it
    // without its filter when most of the driving relation's sampled keys
    // found a match (the filter would reject few rows).
    struct PendingRefilter {
        bool                                       armed   = false;
        bool                                       for_bloom = false;
        uint32_t                                   key     = 0;
        uint64_t                                   grouped = 0;       // relations covered by tables after the build
        const HashTable*                           table   = nullptr; // table the queued probe reads (null: none)
        std::vector<std::pair<uint32_t, uint32_t>> todo;              // relation, sample keys queued for it
        DevBuf                                     dev;
    };
    PendingRefilter rf_;
    // Restrictions of each relation's scans (reference Restriction.hpp,
    // RestrictionLogic.cpp): the query's own (singleton constants) and those
    // the incremental loop derives from built tables. Before a relation's
    // first scan its selection words are computed once (launch_filter_select)
    // and every kernel scanning it reads them.
    std::vector<std::vector<FilterSpec>> rfilt_;
    std::vector<DevBuf>                  sel_;
    std::vector<bool>                    sel_done_;
    // Semi-join tables restricting a relation (exact membership of its key):
    // a pipeline driven by the relation probes them as early, payload-free
    // probes (with their bloom filters); a scan building a hash table turns
    // them into selection words.
    std::vector<std::vector<std::pair<const HashTable*, uint32_t>>> rtables_;
    uint64_t                             done_ = 0; // relations a semi-join consumed (no longer planned)
    FilterSpec      host_filter(const QFilter& f) const;
    FilterSpec      table_filter(uint32_t attr, const HashTable* ht) const;
    void            add_filter(uint32_t rel, const FilterSpec& f);
    const uint64_t* sel_for(uint32_t rel);
    // After building `ht` on class `key` over relations `set`: restricts the
    // relations still unjoined that produce the key; returns true when the
    // table is consumed as a semi-join (no longer a join input).
    bool            restrict_after_build(const HashTable* ht, uint32_t key, uint64_t set);
    void            queue_refilter(const HashTable* ht);
    void            drop_refilter() { rf_.table = nullptr; }
    void            finish_refilter(const HashTable* ht);
    std::vector<std::unique_ptr<HashTable>> tables_;
    std::vector<std::unique_ptr<Tuples>>    tuples_;
    std::vector<const DevColumn*>           colreg_;
    std::map<std::pair<uint32_t, uint64_t>, uint32_t> colid_;
    DevBuf                                  counters_;
    DevBuf                                  wave_sums_; // one-probe kernel: per-wave closing sums
    DevBuf                                  col_base_, col_npages_;
    hipEvent_t                              ev0_ = nullptr, ev1_ = nullptr;

    const DevColumn& column(uint32_t rel, uint64_t col) const {
        return inputs_[q_.rels[rel].table]->cols[col];
    }
    // Rows to size a join result's staging area or page pool for: the
    // planner's estimate, but never less than the largest input of the set.
    // An estimate that steers the join order well can still fall short of the
    // true size, and a short buffer re-runs the whole pipeline; the largest
    // input bounds every foreign-key join.
    double size_bound(uint64_t set, double est) const {
        double b = est;
        for (uint32_t r = 0; r < q_.rels.size(); ++r)
            if ((set >> r) & 1) b = std::max(b, double(q_.rels[r].rows));
        return b;
    }
    uint32_t col_id(uint32_t rel, uint64_t col) {
        uint64_t table = q_.rels[rel].table;
        auto     key   = std::make_pair(uint32_t(table), col);
        auto     it    = colid_.find(key);
        if (it != colid_.end()) return it->second;
        uint32_t id = uint32_t(colreg_.size());
        if (id >= 4095) throw std::runtime_error("too many scanned columns");
        colreg_.push_back(&column(rel, col));
        colid_[key] = id;
        return id;
    }
    void register_all_columns();
    void walk(int node, std::vector<int>& joins, int& leaf) const;
    // Builds the hash table of join-tree node `node` on class `key_eq`;
    // `probe_rows` = rows expected to probe it (tunes load factor and bloom).
    const HashTable* build(int node, uint32_t key_eq, uint64_t probe_rows);
    uint64_t         probe_rows_of(const TreeNode& join) const {
        const TreeNode& p = t_->nodes[join.probe];
        return uint64_t(std::min(size_bound(p.set, p.card), 1.8e19));
    }
    // known_grouped: the key column is known to need the grouped layout.
    const HashTable* make_table(Target& staged, const std::vector<uint32_t>& eqs, uint64_t probe_est,
                                bool known_grouped);
    // Hash table on a plain scan of relation `rel`, partitioned straight from
    // its pages (no staged tuples). False when that path does not apply or the
    // table needs more than the plain partitioned build (the caller stages).
    bool scan_build(uint32_t rel, const std::vector<uint32_t>& eqs, uint64_t probe_est, const HashTable*& out);
    // Small hash table on a plain scan inserted straight from its pages (one
    // kernel, one round trip). False when the table is too large for it or a
    // key repeats past the walk cap (`walked`: the caller builds grouped).
    bool scan_insert(uint32_t rel, const std::vector<uint32_t>& eqs, uint64_t probe_est, const HashTable*& out,
                     bool& walked);
    // Tuple words of relation `rel`'s columns for classes `eqs`, NULL-checked join columns.
    ScanBuildArgs scan_args(uint32_t rel, const std::vector<uint32_t>& eqs);
    void pipeline(Driver d, std::vector<ProbeStep> probes, Target& t);
    uint64_t launch(PipeArgs& a, uint32_t nslots, uint64_t nrows, double& ms);
    uint64_t grid_waves(uint64_t nrows, bool xcd_order = false, uint32_t per_cu = 16) const;
    // Rows per morsel: up to kMorselRows, fewer when that leaves fewer
    // morsels than the grid's `per_cu` waves per CU (a mid-size input would
    // otherwise run on a handful of waves).
    uint32_t morsel_rows(uint64_t nrows, bool xcd_order, uint32_t per_cu) const {
        if (xcd_order) return kMorselRows;
        const uint64_t target = std::max<uint64_t>(1, uint64_t(ctx_.cu_count()) * per_cu);
        const uint64_t m      = ((nrows + target - 1) / target + 63) / 64 * 64;
        return uint32_t(std::min<uint64_t>(std::max<uint64_t>(m, 64), kMorselRows));
    }
    uint32_t per_cu_for(const PipeArgs& a, uint32_t nslots) const {
        if (a.grid_per_cu) return a.grid_per_cu;
        return use_lean(a, nslots) ? ctx_.options.lean_waves_per_cu : 16;
    }
    // Bytes of one build tuple's values: 4 per INT32 class, 8 per other class
    // (the algorithmic size of a compact tuple).
    uint64_t value_bytes(const std::vector<uint32_t>& eqs) const {
        uint64_t b = 0;
        for (uint32_t e : eqs) b += q_.attrs[q_.eq_attrs[e].front()].type == T_INT32 ? 4 : 8;
        return b;
    }
    // A two-word table whose key and attribute classes are INT32: the
    // one-probe kernel reads it packed, 8 bytes per slot instead of 16.
    bool packable(const HashTable& ht) const {
        if (!ctx_.options.packed_tables || !ctx_.options.lean_kernel || ht.grouped || ht.words != 2) return false;
        for (uint32_t w = 0; w < 2; ++w)
            if (q_.attrs[q_.eq_attrs[ht.eqs[w]].front()].type != T_INT32) return false;
        return true;
    }
    // Whether the pipeline runs on the lean one-probe kernel, and its grid.
    bool     use_lean(const PipeArgs& a, uint32_t nslots) const {
        return ctx_.options.lean_kernel && probe1_applicable(a, nslots);
    }
    uint64_t waves_for(const PipeArgs& a, uint32_t nslots, uint64_t nrows) const {
        return grid_waves(nrows, a.xcd_order != 0, per_cu_for(a, nslots));
    }
    const Tuples* partition_driver(uint32_t rel, uint32_t key_eq, uint64_t table_bytes);
    // Output pools are sized for an upper bound of the result (the largest
    // input of the set). A result whose pages fill little of its pool gets a
    // pool of exactly its pages, so results held together (execute_many over
    // thousands of plans) do not keep the unused slots: 1,850 such c5 results
    // had asked for 715 GB of pools.
    void compact_pool(DevResult& res);
    static constexpr uint32_t kMorselRows = 4096;
};

void Executor::walk(int node, std::vector<int>& joins, int& leaf) const {
    // node = Join(build, probe); the probe side continues the pipeline down to a leaf
    std::vector<int> stack;
    int              n = node;
    while (!t_->leaf(n)) {
        stack.push_back(n);
        n = t_->nodes[n].probe;
    }
    leaf = t_->nodes[n].rel;
    if (leaf < 0) throw std::runtime_error("a built hash table cannot drive a pipeline");
    joins.assign(stack.rbegin(), stack.rend()); // innermost join first
}

void Executor::register_all_columns() {
    for (uint32_t r = 0; r < q_.rels.size(); ++r)
        for (uint32_t e : q_.needed(1ull << r)) col_id(r, q_.attrs[q_.attr_of(r, e)].col);
    for (uint32_t a : q_.out) col_id(q_.attrs[a].rel, q_.attrs[a].col);
    std::vector<const uint8_t*> base;
    std::vector<uint32_t>       np;
    for (auto* c : colreg_) {
        base.push_back(c->pages.get<uint8_t>());
        np.push_back(uint32_t(c->npages));
    }
    if (base.empty()) {
        base.push_back(nullptr);
        np.push_back(0);
    }
    col_base_   = DevBuf(&ctx_.heap(), base.size() * sizeof(void*));
    col_npages_ = DevBuf(&ctx_.heap(), np.size() * sizeof(uint32_t));
    CGJ_HIP(hipMemcpyAsync(col_base_.get(), base.data(), base.size() * sizeof(void*), hipMemcpyHostToDevice,
                           ctx_.stream()));
    CGJ_HIP(hipMemcpyAsync(col_npages_.get(), np.data(), np.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                           ctx_.stream()));
}

uint64_t Executor::grid_waves(uint64_t nrows, bool xcd_order, uint32_t per_cu) const {
    const uint64_t mr      = morsel_rows(nrows, xcd_order, per_cu);
    uint64_t       morsels = (nrows + mr - 1) / mr;
    uint64_t waves   = std::min<uint64_t>(std::max<uint64_t>(morsels, 1), uint64_t(ctx_.cu_count()) * per_cu);
    uint32_t per     = pipeline_waves_per_block() * (xcd_order ? 8 : 1);
    return (waves + per - 1) / per * per;
}

void Executor::compact_pool(DevResult& res) {
    const uint64_t used = res.total_pages;
    // below 256 MiB unused the move costs more than the memory it frees
    // (c4's small joins: 0.13 ms each, 9 ms per 256 joins at a 32 MiB bar)
    if (res.pool_pages <= 2 * used + 64 || (res.pool_pages - used) * kPageSize < (uint64_t(256) << 20)) return;
    DevBuf pool(&ctx_.heap(), std::max<uint64_t>(used, 1) * (kPageSize + 16) + 64);
    // counters_ is free again once the pipeline's counters were read back
    PipeCounters* pc = counters_.get<PipeCounters>();
    CGJ_HIP(hipMemsetAsync(&pc->morsel, 0, sizeof(pc->morsel), ctx_.stream()));
    CGJ_HIP(hipMemsetAsync(&pc->fault, 0, sizeof(pc->fault), ctx_.stream()));
    CGJ_HIP(launch_compact_pages(res.pool.get<uint8_t>(), res.pool_pages, res.reserved_pages, pool.get<uint8_t>(), used,
                                 &pc->morsel, &pc->fault, ctx_.stream()));
    unsigned long long moved = 0, fault = 0;
    CGJ_HIP(hipMemcpyAsync(&moved, &pc->morsel, sizeof(moved), hipMemcpyDeviceToHost, ctx_.stream()));
    CGJ_HIP(hipMemcpyAsync(&fault, &pc->fault, sizeof(fault), hipMemcpyDeviceToHost, ctx_.stream()));
    CGJ_HIP(hipStreamSynchronize(ctx_.stream()));
    if (fault || moved != used) throw std::runtime_error("result page directory does not match the page count");
    res.pool           = std::move(pool);
    res.pool.set_evict_first();
    res.pool_pages     = used;
    res.reserved_pages = used;
    ++pools_compacted;
}

const Tuples* Executor::partition_driver(uint32_t rel, uint32_t key_eq, uint64_t table_bytes) {
    std::vector<uint32_t> eqs = q_.needed(1ull << rel);
    auto                  it  = std::find(eqs.begin(), eqs.end(), key_eq);
    if (it == eqs.end()) throw std::runtime_error("probe key is not a column of the driving table");
    std::iter_swap(eqs.begin(), it);
    if (eqs.size() > size_t(kMaxScan)) throw std::runtime_error("too many scanned columns in a pipeline");
    RadixArgs ra;
    std::memset(&ra, 0, sizeof(ra));
    ra.nrows = q_.rels[rel].rows;
    for (uint32_t e : eqs) {
        uint64_t col = q_.attrs[q_.attr_of(rel, e)].col;
        if (q_.is_join_eq(e)) ra.not_null |= 1u << ra.nscan;
        ra.scan[ra.nscan++] = column(rel, col).view(col_id(rel, col));
    }
    // partitions of about 2 MiB of hash table each: half an XCD's L2
    uint64_t nparts = 8;
    while (nparts < 4096 && table_bytes / nparts > (uint64_t(2) << 20)) nparts <<= 1;
    ra.nparts        = uint32_t(nparts);
    ra.rshift        = 64 - log2_ceil(nparts);
    uint64_t nb      = (ra.nrows + 63) / 64;
    uint32_t wgs     = uint32_t(std::min<uint64_t>(1024, std::max<uint64_t>(1, (nb + 15) / 16)));
    ra.blocks_per_wg = (nb + wgs - 1) / wgs;
    DevBuf bh(&ctx_.heap(), nparts * wgs * sizeof(uint32_t));
    DevBuf scratch(&ctx_.heap(), scan_scratch_words(nparts * wgs) * sizeof(uint32_t));
    DevBuf total(&ctx_.heap(), sizeof(uint64_t));
    auto   tp = std::make_unique<Tuples>();
    tp->buf   = DevBuf(&ctx_.heap(), std::max<uint64_t>(ra.nrows, 1) * ra.nscan * sizeof(uint64_t));
    tp->eqs   = eqs;
    ra.bh     = bh.get<uint32_t>();
    ra.out    = tp->buf.get<uint64_t>();
    CGJ_HIP(hipEventRecord(ev0_, ctx_.stream()));
    CGJ_HIP(launch_radix_partition(ra, wgs, scratch.get<uint32_t>(), total.get<uint64_t>(), ctx_.stream()));
    CGJ_HIP(hipEventRecord(ev1_, ctx_.stream()));
    uint64_t n = 0;
    CGJ_HIP(hipMemcpyAsync(&n, total.get(), sizeof(n), hipMemcpyDeviceToHost, ctx_.stream()));
    CGJ_HIP(hipStreamSynchronize(ctx_.stream()));
    float f = 0;
    CGJ_HIP(hipEventElapsedTime(&f, ev0_, ev1_));
    ms_probe += f;
    ms_partition += f;
    tp->n = n;
    for (uint32_t c = 0; c < ra.nscan; ++c) probe_input_pages += column(rel, q_.attrs[q_.attr_of(rel, eqs[c])].col).npages;
    tuples_.push_back(std::move(tp));
    return tuples_.back().get();
}

uint64_t Executor::launch(PipeArgs& a, uint32_t nslots, uint64_t nrows, double& ms) {
    a.counters    = counters_.get<PipeCounters>();
    a.morsel_rows = morsel_rows(nrows, a.xcd_order != 0, per_cu_for(a, nslots));
    a.nrows       = nrows;
    a.nmorsels    = (nrows + a.morsel_rows - 1) / a.morsel_rows;
    a.col_base    = col_base_.get<const uint8_t* const>();
    a.col_npages  = col_npages_.get<uint32_t>();
    CGJ_HIP(hipMemsetAsync(a.counters, 0, sizeof(PipeCounters), ctx_.stream()));
    uint64_t waves = waves_for(a, nslots, nrows);
    CGJ_HIP(hipEventRecord(ev0_, ctx_.stream()));
    if (a.nmorsels) {
        // Closing sums as per-wave stores, added up by a second small kernel,
        // and each wave's first pages or staging slots given out statically:
        // no wave queues an atomic behind the others' at its start or end.
        const uint64_t stride = (waves + 63) / 64 * 64;
        if (wave_sums_.bytes() < stride * kSumFields * sizeof(unsigned long long))
            wave_sums_ = DevBuf(&ctx_.heap(), stride * kSumFields * sizeof(unsigned long long));
        a.wave_sums        = wave_sums_.get<unsigned long long>();
        a.wave_sums_stride = uint32_t(stride);
        a.static_chunk     = 1;
        if (use_lean(a, nslots))
            CGJ_HIP(launch_probe1(a, nslots, uint32_t(waves), ctx_.options.aligned_pairs, ctx_.stream()));
        else
            CGJ_HIP(launch_pipeline(a, nslots, uint32_t(waves), ctx_.stream()));
        CGJ_HIP(hipEventRecord(ev1_, ctx_.stream()));
        const uint32_t nout = a.target == TARGET_PAGES ? a.nout : 0;
        CGJ_HIP(launch_fold_sums(a.wave_sums, a.wave_sums_stride, nout, a.counters, ctx_.stream()));
    } else {
        CGJ_HIP(hipEventRecord(ev1_, ctx_.stream()));
    }
    CGJ_HIP(hipStreamSynchronize(ctx_.stream())); // the kernel and the sums' fold
    float f = 0;
    CGJ_HIP(hipEventElapsedTime(&f, ev0_, ev1_));
    ms += f;
    last_ms_ = f;
    ++launches;
    return waves;
}

const HashTable* Executor::make_table(Target& staged, const std::vector<uint32_t>& eqs, uint64_t probe_est,
                                      bool known_grouped) {
    auto ht    = std::make_unique<HashTable>();
    ht->eqs    = eqs;
    ht->words  = uint32_t(eqs.size());
    ht->tuples = staged.count;
    const uint64_t n  = staged.count;
    JoinTuning     tn = ctx_.tune(n, probe_est, 8);
    uint32_t       lg = log2_ceil(uint64_t(std::ceil(double(std::max<uint64_t>(n, 1)) / tn.max_load)));
    if (lg < 6) lg = 6;
    ht->cap   = 1ull << lg;
    ht->shift = 64 - lg;
    ht->rmask = ht->cap - 1;
    ht->slots = DevBuf(&ctx_.heap(), ht->cap * ht->words * sizeof(uint64_t));
    DevBuf flags(&ctx_.heap(), 4 * sizeof(uint32_t));
    CGJ_HIP(hipMemsetAsync(flags.get(), 0, 4 * sizeof(uint32_t), ctx_.stream()));
    if (tn.bloom_bits) {
        uint64_t words = std::max<uint64_t>(1, n * tn.bloom_bits / 64);
        ht->bloom_log  = log2_ceil(words);
        ht->bloom      = DevBuf(&ctx_.heap(), (1ull << ht->bloom_log) * sizeof(uint64_t));
        ht->has_bloom  = true;
    }
    uint64_t* bloom = ht->has_bloom ? ht->bloom.get<uint64_t>() : nullptr;
    CGJ_HIP(hipEventRecord(ev0_, ctx_.stream()));
    if (bloom) CGJ_HIP(launch_fill_u64(bloom, 1ull << ht->bloom_log, 0, ctx_.stream()));

    // Region size: the largest power of two whose entries fit a 64 KiB LDS
    // table, raised when the table would need more than 16K regions (the
    // partition histogram keeps one u32 per region in 64 KiB of LDS), as long
    // as the region still fits the workgroup's LDS (160 KiB on gfx950).
    uint32_t rs_log = 12;
    while (rs_log > 4 && (uint64_t(1) << rs_log) * ht->words * 8 > 65536) --rs_log;
    const uint32_t max_rlog = ctx_.options.build_regions_log;
    if (lg > rs_log + max_rlog) rs_log = lg - max_rlog;
    bool partitioned = tn.partitioned && n >= 65536 && lg > rs_log &&
                       (uint64_t(1) << rs_log) * ht->words * 8 <= ctx_.lds_bytes();
    // With grouping on, an insert that passes `walk_cap` slots holding its own
    // key (a key repeated that often) abandons the plain layout: the table is
    // rebuilt grouped instead of finishing a quadratic insert.
    const bool     can_group = ctx_.options.group_walk > 0;
    const uint32_t walk_cap  = can_group ? ctx_.options.group_walk : 0;
    bool           done = false, group = ctx_.options.force_grouped || (known_grouped && can_group);
    auto           read_flags = [&](uint32_t* f) {
        CGJ_HIP(hipMemcpyAsync(f, flags.get(), 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx_.stream()));
        CGJ_HIP(hipStreamSynchronize(ctx_.stream()));
        CGJ_HIP(hipMemsetAsync(flags.get(), 0, 4 * sizeof(uint32_t), ctx_.stream()));
    };
    if (partitioned && !group) {
        uint32_t log_regions = lg - rs_log;
        uint64_t nregions    = 1ull << log_regions;
        uint64_t slots       = staged.slots;
        uint32_t blocks      = uint32_t(std::min<uint64_t>(512, (slots + 16383) / 16384));
        uint64_t chunk       = (slots + blocks - 1) / blocks;
        DevBuf   bh(&ctx_.heap(), nregions * blocks * sizeof(uint32_t));
        DevBuf   scratch(&ctx_.heap(), scan_scratch_words(nregions * blocks) * sizeof(uint32_t));
        DevBuf   total(&ctx_.heap(), sizeof(uint64_t));
        DevBuf   parts(&ctx_.heap(), n * ht->words * sizeof(uint64_t));
        CGJ_HIP(launch_part_build(staged.stage.get<uint64_t>(), slots, n, ht->words, parts.get<uint64_t>(), bh.get<uint32_t>(),
                                  blocks, chunk, scratch.get<uint32_t>(), total.get<uint64_t>(),
                                  ht->slots.get<uint64_t>(), lg, log_regions, bloom, ht->bloom_log,
                                  flags.get<uint32_t>(), ctx_.options.build_threads, walk_cap, ctx_.stream()));
        uint32_t f[4];
        read_flags(f);
        if (!f[1] && !f[2]) {
            done       = true;
            ht->rmask  = (1ull << rs_log) - 1;
            ht->unique = f[0] == 0;
        } else {
            ++retries;
            ++retries_build; // a region overflowed or an insert walked too far (heavy duplicates)
            group = can_group;
        }
    }
    if (!done && !group) {
        CGJ_HIP(launch_fill_u64(ht->slots.get<uint64_t>(), ht->cap * ht->words, kEmpty, ctx_.stream()));
        CGJ_HIP(launch_ht_insert(staged.stage.get<uint64_t>(), staged.slots, ht->words, ht->slots.get<uint64_t>(), ht->shift,
                                 ht->cap - 1, bloom, ht->bloom_log, flags.get<uint32_t>(), walk_cap, ctx_.stream()));
        uint32_t f[4];
        read_flags(f);
        if (!f[2]) {
            done       = true;
            ht->rmask  = ht->cap - 1;
            ht->unique = f[0] == 0;
        } else {
            ++retries;
            ++retries_build;
            group = true;
        }
    }
    if (!done) {
        // Grouped build: regions of up to 8K two-word slots built in LDS (keys,
        // counts, cursors: 16 bytes per slot). Capacity follows the tuple
        // count, an upper bound on the distinct keys; a region whose distinct
        // keys still overflow doubles the capacity.
        const uint32_t threads = ctx_.options.build_threads;
        for (int attempt = 0;; ++attempt) {
            uint32_t glog = 13;
            while (glog > 6 && group_build_lds(1u << glog, threads) > ctx_.lds_bytes()) --glog;
            if (glog > lg) glog = lg;
            if (lg > glog + ctx_.options.build_regions_log) glog = lg - ctx_.options.build_regions_log;
            if (group_build_lds(1u << glog, threads) > ctx_.lds_bytes())
                throw std::runtime_error("grouped hash table: region does not fit the LDS");
            const uint32_t log_regions = lg - glog;
            const uint64_t nregions    = 1ull << log_regions;
            const uint64_t slots       = staged.slots;
            const uint32_t blocks      = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(512, (slots + 16383) / 16384)));
            const uint64_t chunk       = (slots + blocks - 1) / blocks;
            ht->cap     = 1ull << lg;
            ht->shift   = 64 - lg;
            ht->rmask   = (1ull << glog) - 1;
            ht->slots   = DevBuf(&ctx_.heap(), ht->cap * 2 * sizeof(uint64_t));
            ht->payload = DevBuf(&ctx_.heap(), std::max<uint64_t>(1, n * (ht->words - 1)) * sizeof(uint64_t));
            DevBuf bh(&ctx_.heap(), nregions * blocks * sizeof(uint32_t));
            DevBuf scratch(&ctx_.heap(), scan_scratch_words(nregions * blocks) * sizeof(uint32_t));
            DevBuf total(&ctx_.heap(), sizeof(uint64_t));
            DevBuf parts(&ctx_.heap(), n * ht->words * sizeof(uint64_t));
            if (bloom) CGJ_HIP(launch_fill_u64(bloom, 1ull << ht->bloom_log, 0, ctx_.stream()));
            CGJ_HIP(launch_group_build(staged.stage.get<uint64_t>(), slots, n, ht->words, parts.get<uint64_t>(),
                                       bh.get<uint32_t>(), blocks, chunk, scratch.get<uint32_t>(), total.get<uint64_t>(),
                                       ht->slots.get<uint64_t>(), ht->payload.get<uint64_t>(), lg, log_regions, bloom,
                                       ht->bloom_log, flags.get<uint32_t>(), threads, ctx_.stream()));
            uint32_t f[4];
            read_flags(f);
            if (f[3]) throw std::runtime_error("grouped hash table: a key repeats 2^24 or more times");
            if (!f[1]) break;
            if (attempt >= 3) throw std::runtime_error("grouped hash table: regions overflow");
            ++retries;
            ++retries_build;
            ++lg;
        }
        ht->grouped = true;
        ht->unique  = false;
    }
    queue_refilter(ht.get());
    CGJ_HIP(hipEventRecord(ev1_, ctx_.stream()));
    CGJ_HIP(hipEventSynchronize(ev1_));
    float fms = 0;
    CGJ_HIP(hipEventElapsedTime(&fms, ev0_, ev1_));
    ms_build += fms;
    build_tuples += n;
    build_bytes += n * value_bytes(ht->eqs);
    staged.stage.reset();
    tables_.push_back(std::move(ht));
    return tables_.back().get();
}

uint32_t Executor::key_for(int build_node, uint64_t cur) const {
    const TreeNode& bn = t_->nodes[build_node];
    const auto      sh = q_.shared(bn.set, cur);
    if (sh.empty()) throw std::runtime_error("join without a predicate");
    if (bn.group >= 0) {
        const uint32_t k = groups_[bn.group].key;
        if (std::find(sh.begin(), sh.end(), k) == sh.end())
            throw std::runtime_error("a built hash table is joined off its key");
        return k;
    }
    return sh.front();
}

const HashTable* Executor::build(int node, uint32_t key_eq, uint64_t probe_rows) {
    if (t_->nodes[node].group >= 0) return groups_[t_->nodes[node].group].ht;
    const TreeNode&       tn = t_->nodes[node];
    std::vector<uint32_t> eqs{key_eq};
    for (uint32_t e : q_.needed(tn.set))
        if (e != key_eq) eqs.push_back(e);
    if (eqs.size() > kMaxWords) throw std::runtime_error("hash-table entry wider than supported");
    Target t;
    t.kind = TARGET_STAGE;
    t.eqs  = eqs;
    t.est  = size_bound(tn.set, tn.card);
    t.card = tn.card;
    std::vector<int> joins;
    int              leaf;
    walk(node, joins, leaf);
    std::vector<ProbeStep> probes;
    uint64_t               cur = 1ull << leaf;
    for (int j : joins) {
        const TreeNode& jn  = t_->nodes[j];
        const TreeNode& bn  = t_->nodes[jn.build];
        const uint32_t  key = key_for(jn.build, cur);
        const HashTable* ht = build(jn.build, key, probe_rows_of(jn));
        if (!ht) return nullptr;
        probes.push_back(ProbeStep{ht, key, bn.set});
        cur |= bn.set;
    }
    // The key column may be known to need the grouped layout: for a plain
    // scan its own column, for a join result the base column its key values
    // come from (the first relation of the set in the key's class).
    const bool plain   = probes.empty();
    bool       has_key = false; // a relation of the set produces the key class (always, for a join key)
    uint32_t   key_rel = 0;
    for (uint32_t r = 0; r < q_.rels.size() && !has_key; ++r)
        if (((tn.set >> r) & 1u) && std::binary_search(q_.rels[r].eqs.begin(), q_.rels[r].eqs.end(), key_eq)) {
            key_rel = r;
            has_key = true;
        }
    const uint64_t key_uid = has_key ? inputs_[q_.rels[key_rel].table]->uid : 0;
    const uint64_t key_col = has_key ? q_.attrs[q_.attr_of(key_rel, key_eq)].col : 0;
    bool           known   = has_key && ctx_.build_hints().grouped(key_uid, key_col);
    if (plain && ctx_.options.scan_build && !known) {
        const HashTable* ht     = nullptr;
        bool             walked = false;
        if (scan_insert(uint32_t(leaf), eqs, probe_rows, ht, walked)) return ht; // nullptr: empty build side
        if (walked && has_key) { // a key repeats past the walk cap: build grouped at once
            ctx_.build_hints().mark_grouped(key_uid, key_col);
            known = true;
        } else if (scan_build(uint32_t(leaf), eqs, probe_rows, ht)) {
            return ht;
        }
    }
    Driver d;
    d.rel = leaf;
    d.set = 1ull << leaf;
    pipeline(d, probes, t);
    if (t.count == 0) return nullptr; // empty build side: the inner join is empty
    const HashTable* ht = make_table(t, eqs, probe_rows, known);
    if (has_key && !known && ht->grouped && !ctx_.options.force_grouped)
        ctx_.build_hints().mark_grouped(key_uid, key_col);
    return ht;
}

ScanBuildArgs Executor::scan_args(uint32_t rel, const std::vector<uint32_t>& eqs) {
    ScanBuildArgs sa{};
    sa.words = uint32_t(eqs.size());
    for (uint32_t w = 0; w < sa.words; ++w) {
        const uint32_t a = q_.attr_of(rel, eqs[w]);
        const uint64_t c = q_.attrs[a].col;
        sa.col[w]        = column(rel, c).view(col_id(rel, c));
        if (q_.is_join_eq(eqs[w])) sa.notnull |= 1u << w;
    }
    sa.notnull |= 1u; // the key
    sa.nrows = q_.rels[rel].rows;
    for (auto [ht, key] : rtables_[rel]) add_filter(rel, table_filter(q_.attr_of(rel, key), ht));
    rtables_[rel].clear();
    sa.sel = sel_for(rel);
    return sa;
}

FilterSpec Executor::host_filter(const QFilter& f) const {
    const QAttr& at = q_.attrs[f.attr];
    FilterSpec   s{};
    s.col       = column(at.rel, at.col).view(0);
    s.kind      = f.kind;
    s.a         = f.a;
    s.b         = f.b;
    s.tiny_bits = f.tiny_bits;
    s.tiny_hash = f.tiny_hash;
    for (int i = 0; i < kTinySlots; ++i) s.tiny[i] = size_t(i) < f.tiny.size() ? f.tiny[size_t(i)] : kNull;
    return s;
}

FilterSpec Executor::table_filter(uint32_t attr, const HashTable* ht) const {
    const QAttr& at = q_.attrs[attr];
    FilterSpec   s{};
    s.col   = column(at.rel, at.col).view(0);
    s.kind  = FLT_TABLE;
    s.table = ht->slots.get<uint64_t>();
    s.mask  = ht->rmask;
    s.shift = ht->shift;
    s.words = ht->grouped ? 2 : ht->words;
    return s;
}

void Executor::add_filter(uint32_t rel, const FilterSpec& f) {
    if (rel >= rfilt_.size() || q_.rels[rel].rows == 0) return;
    ++restrictions;
    if (!sel_done_[rel]) {
        rfilt_[rel].push_back(f);
        return;
    }
    // the selection words exist already: narrow them
    FilterArgs fa{};
    fa.nrows      = q_.rels[rel].rows;
    fa.n          = 1;
    fa.accumulate = 1;
    fa.f[0]       = f;
    CGJ_HIP(launch_filter_select(fa, sel_[rel].get<uint64_t>(), ctx_.stream()));
}

const uint64_t* Executor::sel_for(uint32_t rel) {
    if (rel >= rfilt_.size()) return nullptr;
    if (sel_done_[rel]) return sel_[rel].get<uint64_t>();
    if (rfilt_[rel].empty()) return nullptr;
    const uint64_t nrows = q_.rels[rel].rows;
    sel_[rel]            = DevBuf(&ctx_.heap(), std::max<uint64_t>(1, (nrows + 63) / 64) * sizeof(uint64_t));
    const auto& fs       = rfilt_[rel];
    for (size_t i = 0; i < fs.size(); i += kMaxFilters) {
        FilterArgs fa{};
        fa.nrows      = nrows;
        fa.accumulate = i > 0 ? 1 : 0;
        fa.n          = uint32_t(std::min<size_t>(kMaxFilters, fs.size() - i));
        for (uint32_t k = 0; k < fa.n; ++k) fa.f[k] = fs[i + k];
        CGJ_HIP(launch_filter_select(fa, sel_[rel].get<uint64_t>(), ctx_.stream()));
    }
    sel_done_[rel] = true;
    return sel_[rel].get<uint64_t>();
}

bool Executor::restrict_after_build(const HashTable* ht, uint32_t key, uint64_t set) {
    if (!ctx_.options.restrictions) return false;
    // relations still planned on their own that produce the key class
    uint64_t grouped = done_ | set;
    bool     other   = false; // another built table produces the key
    for (const Group& g : groups_) {
        if ((g.set & set) == g.set) continue; // consumed by the new table
        grouped |= g.set;
        if (q_.produced(g.set).count(key)) other = true;
    }
    std::vector<uint32_t> targets;
    for (uint32_t r = 0; r < q_.rels.size(); ++r)
        if (!((grouped >> r) & 1) && std::binary_search(q_.rels[r].eqs.begin(), q_.rels[r].eqs.end(), key))
            targets.push_back(r);
    if (targets.empty()) return false;
    // The join is a semi-join when the table is duplicate-free and carries
    // only its key, and no other built table produces the key (whose rows
    // could not be restricted any more): the table then only restricts the
    // key's producers exactly (reference QueryPlan.cpp:602-625, JoinPrecise).
    const bool semijoin = ht->unique && ht->words == 1 && !other && q_.eq_attrs[key].size() > 1;
    // A table of at most 32 tuples restricts by its keys (Eq, Eq2, a range or
    // a tiny table, RestrictionLogic.cpp:397-447); else a semi-join keeps the
    // exact table membership, and a join leaves its probe to filter.
    QFilter small;
    bool    have_small = false;
    if (ht->tuples <= 32) {
        DevBuf         buf(&ctx_.heap(), 32 * sizeof(uint64_t) + sizeof(uint32_t));
        auto*          cnt = reinterpret_cast<unsigned int*>(buf.get<uint8_t>() + 32 * sizeof(uint64_t));
        hipStream_t    st  = ctx_.stream();
        uint8_t*       pin = ctx_.sample_pinned() + Context::kSamplePinnedBytes - 512;
        CGJ_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t), st));
        CGJ_HIP(launch_table_keys(ht->slots.get<uint64_t>(), ht->cap, ht->grouped ? 2 : ht->words, buf.get<uint64_t>(),
                                  32, cnt, st));
        CGJ_HIP(hipMemcpyAsync(pin, buf.get(), 32 * sizeof(uint64_t) + sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        CGJ_HIP(hipStreamSynchronize(st));
        const uint32_t n = std::min<uint32_t>(*reinterpret_cast<const uint32_t*>(pin + 32 * sizeof(uint64_t)), 32);
        std::vector<uint64_t> keys(reinterpret_cast<const uint64_t*>(pin), reinterpret_cast<const uint64_t*>(pin) + n);
        have_small = small_restriction(keys, 0, small);
    }
    if (!have_small && !semijoin) return false;
    for (uint32_t r : targets) {
        const uint32_t attr = q_.attr_of(r, key);
        if (have_small) {
            QFilter f = small;
            f.attr    = attr;
            add_filter(r, host_filter(f));
            if (samples_ && samples_->size[r]) { // the sample rows that pass
                auto it = samples_->keys[r].find(key);
                if (it != samples_->keys[r].end())
                    for (size_t i = 0; i < it->second.size(); ++i)
                        if (!f.pass(it->second[i])) samples_->matches[r] &= ~(1ull << i);
            }
        } else {
            rtables_[r].emplace_back(ht, key);
            ++restrictions;
        }
    }
    if (semijoin) {
        done_ |= set;
        ++semijoins;
    }
    return semijoin;
}

bool Executor::scan_insert(uint32_t rel, const std::vector<uint32_t>& eqs, uint64_t probe_est, const HashTable*& out,
                           bool& walked) {
    out               = nullptr;
    walked            = false;
    const uint64_t n0 = q_.rels[rel].rows; // upper bound on the tuples (NULL keys are dropped)
    if (n0 == 0 || eqs.size() > size_t(kMaxWords) || ctx_.options.force_grouped) return false;
    const uint32_t words = uint32_t(eqs.size());
    JoinTuning     tn    = ctx_.tune(n0, probe_est, 8);
    uint32_t       lg    = log2_ceil(uint64_t(std::ceil(double(n0) / tn.max_load)));
    if (lg < 6) lg = 6;
    if ((uint64_t(1) << lg) * words * sizeof(uint64_t) > ctx_.options.scan_insert_max_bytes) return false;
    ScanBuildArgs sa = scan_args(rel, eqs);
    auto          ht = std::make_unique<HashTable>();
    ht->eqs          = eqs;
    ht->words        = words;
    ht->cap          = 1ull << lg;
    ht->shift        = 64 - lg;
    ht->rmask        = ht->cap - 1;
    ht->slots        = DevBuf(&ctx_.heap(), ht->cap * words * sizeof(uint64_t));
    DevBuf fc(&ctx_.heap(), 4 * sizeof(uint32_t) + sizeof(uint64_t)); // flags, then the count
    if (tn.bloom_bits) {
        uint64_t bw   = std::max<uint64_t>(1, n0 * tn.bloom_bits / 64);
        ht->bloom_log = log2_ceil(bw);
        ht->bloom     = DevBuf(&ctx_.heap(), (1ull << ht->bloom_log) * sizeof(uint64_t));
        ht->has_bloom = true;
    }
    uint64_t*   bloom = ht->has_bloom ? ht->bloom.get<uint64_t>() : nullptr;
    hipStream_t s     = ctx_.stream();
    CGJ_HIP(hipEventRecord(ev0_, s));
    CGJ_HIP(hipMemsetAsync(fc.get(), 0, 4 * sizeof(uint32_t) + sizeof(uint64_t), s));
    CGJ_HIP(launch_fill_u64(ht->slots.get<uint64_t>(), ht->cap * words, kEmpty, s));
    if (bloom) CGJ_HIP(launch_fill_u64(bloom, 1ull << ht->bloom_log, 0, s));
    auto* count = reinterpret_cast<unsigned long long*>(fc.get<uint32_t>() + 4);
    CGJ_HIP(launch_scan_insert(sa, ht->slots.get<uint64_t>(), lg, bloom, ht->bloom_log, fc.get<uint32_t>(),
                               ctx_.options.group_walk, count, s));
    uint32_t back[6];
    CGJ_HIP(hipMemcpyAsync(back, fc.get(), sizeof(back), hipMemcpyDeviceToHost, s));
    queue_refilter(ht.get());
    CGJ_HIP(hipEventRecord(ev1_, s));
    CGJ_HIP(hipEventSynchronize(ev1_));
    float fms = 0;
    CGJ_HIP(hipEventElapsedTime(&fms, ev0_, ev1_));
    ms_build += fms;
    const uint64_t n = uint64_t(back[4]) | (uint64_t(back[5]) << 32);
    if (back[2]) { // an insert passed the walk cap: the table is incomplete
        drop_refilter();
        ++retries;
        ++retries_build;
        walked = true;
        return false;
    }
    if (n == 0) {
        drop_refilter();
        return true; // empty build side
    }
    ht->tuples = n;
    ht->unique = back[0] == 0;
    build_tuples += n;
    build_bytes += n * value_bytes(ht->eqs);
    tables_.push_back(std::move(ht));
    out = tables_.back().get();
    return true;
}

bool Executor::scan_build(uint32_t rel, const std::vector<uint32_t>& eqs, uint64_t probe_est, const HashTable*& out) {
    out               = nullptr;
    const uint64_t n0 = q_.rels[rel].rows; // upper bound on the tuples (NULL keys are dropped)
    if (eqs.size() > size_t(kMaxWords) || n0 < 65536) return false;
    const uint32_t words = uint32_t(eqs.size());
    JoinTuning     tn    = ctx_.tune(n0, probe_est, 8);
    uint32_t       lg    = log2_ceil(uint64_t(std::ceil(double(n0) / tn.max_load)));
    if (lg < 6) lg = 6;
    uint32_t rs_log = 12;
    while (rs_log > 4 && (uint64_t(1) << rs_log) * words * 8 > 65536) --rs_log;
    if (lg > rs_log + ctx_.options.build_regions_log) rs_log = lg - ctx_.options.build_regions_log;
    if (!tn.partitioned || lg <= rs_log || (uint64_t(1) << rs_log) * words * 8 > ctx_.lds_bytes() ||
        ctx_.options.force_grouped)
        return false;
    ScanBuildArgs sa = scan_args(rel, eqs);
    const uint32_t log_regions  = lg - rs_log;
    sa.nregions                 = 1u << log_regions;
    sa.rshift                   = 64 - log_regions;
    const uint64_t nb           = (n0 + 63) / 64;
    const uint32_t blocks       = uint32_t(std::min<uint64_t>(ctx_.options.scan_workgroups, std::max<uint64_t>(1, nb / 64)));
    sa.blocks_per_wg            = (nb + blocks - 1) / blocks;
    auto ht    = std::make_unique<HashTable>();
    ht->eqs    = eqs;
    ht->words  = words;
    ht->cap    = 1ull << lg;
    ht->shift  = 64 - lg;
    ht->rmask  = (1ull << rs_log) - 1;
    ht->slots  = DevBuf(&ctx_.heap(), ht->cap * words * sizeof(uint64_t));
    DevBuf flags(&ctx_.heap(), 4 * sizeof(uint32_t));
    DevBuf bh(&ctx_.heap(), uint64_t(sa.nregions) * blocks * sizeof(uint32_t));
    DevBuf scratch(&ctx_.heap(), scan_scratch_words(uint64_t(sa.nregions) * blocks) * sizeof(uint32_t));
    DevBuf total(&ctx_.heap(), sizeof(uint64_t));
    DevBuf parts(&ctx_.heap(), n0 * words * sizeof(uint64_t));
    if (tn.bloom_bits) {
        uint64_t bw   = std::max<uint64_t>(1, n0 * tn.bloom_bits / 64);
        ht->bloom_log = log2_ceil(bw);
        ht->bloom     = DevBuf(&ctx_.heap(), (1ull << ht->bloom_log) * sizeof(uint64_t));
        ht->has_bloom = true;
    }
    uint64_t*   bloom = ht->has_bloom ? ht->bloom.get<uint64_t>() : nullptr;
    hipStream_t s     = ctx_.stream();
    CGJ_HIP(hipEventRecord(ev0_, s));
    CGJ_HIP(hipMemsetAsync(flags.get(), 0, 4 * sizeof(uint32_t), s));
    if (bloom) CGJ_HIP(launch_fill_u64(bloom, 1ull << ht->bloom_log, 0, s));
    CGJ_HIP(launch_scan_partition(sa, blocks, bh.get<uint32_t>(), scratch.get<uint32_t>(), total.get<uint64_t>(),
                                  parts.get<uint64_t>(), ctx_.options.build_threads, s));
    // the region build reads the tuple count on the device; one round trip
    // brings back the count and the flags
    const uint32_t walk_cap = ctx_.options.group_walk;
    CGJ_HIP(launch_region_build(parts.get<uint64_t>(), bh.get<uint32_t>(), blocks, total.get<uint64_t>(), words,
                                ht->slots.get<uint64_t>(), lg, log_regions, bloom, ht->bloom_log, flags.get<uint32_t>(),
                                ctx_.options.build_threads, walk_cap, s));
    uint64_t n = 0;
    uint32_t f[4];
    CGJ_HIP(hipMemcpyAsync(&n, total.get(), sizeof(n), hipMemcpyDeviceToHost, s));
    CGJ_HIP(hipMemcpyAsync(f, flags.get(), sizeof(f), hipMemcpyDeviceToHost, s));
    queue_refilter(ht.get());
    CGJ_HIP(hipEventRecord(ev1_, s));
    CGJ_HIP(hipEventSynchronize(ev1_));
    float fms = 0;
    CGJ_HIP(hipEventElapsedTime(&fms, ev0_, ev1_));
    ms_build += fms;
    if (n == 0 || f[1] || f[2]) drop_refilter();
    if (n == 0) return true; // empty build side
    if (f[1] || f[2]) { // a region overflowed or a key repeats heavily: the staged path handles it
        ++retries;
        ++retries_build;
        return false;
    }
    ht->tuples = n;
    ht->unique = f[0] == 0;
    build_tuples += n;
    build_bytes += n * value_bytes(ht->eqs);
    tables_.push_back(std::move(ht));
    out = tables_.back().get();
    return true;
}

void Executor::pipeline(Driver d, std::vector<ProbeStep> probes, Target& t) {
    // semi-join tables restricting the driving relation: early probes, first
    if (d.rel >= 0 && !rtables_[uint32_t(d.rel)].empty()) {
        std::vector<ProbeStep> semi;
        for (auto [ht, key] : rtables_[uint32_t(d.rel)]) semi.push_back(ProbeStep{ht, key, 0});
        rtables_[uint32_t(d.rel)].clear();
        probes.insert(probes.begin(), semi.begin(), semi.end());
    }
    // Probe order: where the driving relation's samples were probed into the
    // tables, the most selective table goes first (fewest sample rows found),
    // among those whose key the pipeline already carries; the reference
    // orders a scan's filters by selectivity the same way
    // (engine/op/TableScan.cpp:461-474).
    if (samples_ && d.rel >= 0 && probes.size() > 1 && samples_->size[d.rel] > 0) {
        const auto&            pass = samples_->pass[d.rel];
        std::vector<ProbeStep> ordered;
        std::vector<bool>      used(probes.size(), false);
        uint64_t               have = d.set;
        auto                   sel  = [&](const ProbeStep& p) {
            auto it = pass.find(p.ht);
            return it == pass.end() ? 2.0 : double(__builtin_popcountll(it->second)) / samples_->size[d.rel];
        };
        while (ordered.size() < probes.size()) {
            int best = -1;
            for (size_t i = 0; i < probes.size(); ++i) {
                if (used[i]) continue;
                const auto pr = q_.produced(have);
                if (!pr.count(probes[i].key_eq)) continue;
                if (best < 0 || sel(probes[i]) < sel(probes[size_t(best)])) best = int(i);
            }
            if (best < 0) { // keep the remaining ones in their order (keys come from them)
                for (size_t i = 0; i < probes.size(); ++i)
                    if (!used[i]) ordered.push_back(probes[i]);
                break;
            }
            used[size_t(best)] = true;
            have |= probes[size_t(best)].set;
            ordered.push_back(probes[size_t(best)]);
        }
        probes = std::move(ordered);
    }
    // Chains longer than the kernel's probe limit are cut: the first part is
    // materialised as tuples that drive the rest.
    while (probes.size() > size_t(kMaxProbes)) {
        std::vector<ProbeStep> head(probes.begin(), probes.begin() + kMaxProbes);
        uint64_t               set = d.set;
        for (auto& p : head) set |= p.set;
        Target mid;
        mid.kind = TARGET_STAGE;
        mid.eqs  = q_.needed(set);
        mid.est  = t.est;
        // word 0 must be a join class (never NULL): unused slots are marked by an empty word 0
        auto je = std::find_if(mid.eqs.begin(), mid.eqs.end(), [&](uint32_t e) { return q_.is_join_eq(e); });
        if (je == mid.eqs.end()) throw std::runtime_error("materialised chain carries no join column");
        std::iter_swap(mid.eqs.begin(), je);
        pipeline(d, head, mid);
        auto tp = std::make_unique<Tuples>();
        tp->buf = std::move(mid.stage);
        tp->n   = mid.slots;
        tp->eqs = mid.eqs;
        tuples_.push_back(std::move(tp));
        d.rel    = -1;
        d.tuples = tuples_.back().get();
        d.set    = set;
        probes.erase(probes.begin(), probes.begin() + kMaxProbes);
    }

    // A table probing a hash table far larger than an XCD's L2 is first
    // radix-partitioned by the probe key's hash; the pipeline then walks the
    // partitions XCD by XCD so each probes a cache-resident range of the table.
    bool xcd_order = false;
    if (d.rel >= 0 && !probes.empty() && rfilt_[uint32_t(d.rel)].empty()) {
        const HashTable* h0 = probes[0].ht;
        uint64_t         hb = h0->cap * h0->words * sizeof(uint64_t);
        if (hb >= ctx_.options.radix_min_table_bytes && q_.rels[d.rel].rows >= ctx_.options.radix_min_rows) {
            d.tuples  = partition_driver(uint32_t(d.rel), probes[0].key_eq, hb);
            d.rel     = -1;
            xcd_order = true;
        }
    }

    PipeArgs a;
    std::memset(&a, 0, sizeof(a));
    a.abort_pages = ~uint64_t(0);
    a.xcd_order = xcd_order ? 1 : 0;
    a.prefilter = ctx_.options.block_prefilter ? 1 : 0;
    std::map<uint32_t, uint32_t> slot; // class -> tuple slot
    uint32_t                     nslots = 0;
    uint64_t                     nrows  = 0;
    std::vector<uint32_t>        scan_eqs;
    std::vector<uint64_t>        scan_cols;
    if (d.rel >= 0) {
        a.driver = DRIVER_PAGES;
        a.sel    = sel_for(uint32_t(d.rel));
        scan_eqs = q_.needed(1ull << d.rel);
        nrows    = q_.rels[d.rel].rows;
        if (scan_eqs.size() > size_t(kMaxScan)) throw std::runtime_error("too many scanned columns in a pipeline");
        for (uint32_t e : scan_eqs) {
            uint32_t attr = q_.attr_of(uint32_t(d.rel), e);
            uint64_t col  = q_.attrs[attr].col;
            a.scan[a.nscan++] = column(uint32_t(d.rel), col).view(col_id(uint32_t(d.rel), col));
            scan_cols.push_back(col);
            slot[e]           = nslots++;
        }
    } else {
        a.driver      = DRIVER_TUPLES;
        a.tuples      = d.tuples->buf.get<uint64_t>();
        a.tuple_words = uint32_t(d.tuples->eqs.size());
        nrows         = d.tuples->n;
        if (d.tuples->eqs.size() > size_t(kMaxScan)) throw std::runtime_error("too many columns in a pipeline");
        for (uint32_t w = 0; w < d.tuples->eqs.size(); ++w) {
            a.tuple_word[a.nscan++] = w;
            slot[d.tuples->eqs[w]]  = nslots++;
        }
    }
    uint32_t scanned = nslots;
    a.nprobes        = uint32_t(probes.size());
    for (uint32_t i = 0; i < probes.size(); ++i) {
        const HashTable* ht = probes[i].ht;
        ProbeSpec&       P  = a.probes[i];
        auto             ks = slot.find(probes[i].key_eq);
        if (ks == slot.end()) throw std::runtime_error("probe key is not available in the pipeline");
        P.ht        = ht->slots.get<uint64_t>();
        P.shift     = ht->shift;
        P.words     = ht->grouped ? 2 : ht->words;
        P.mask      = ht->rmask;
        P.key_slot  = ks->second;
        P.unique    = ht->unique ? 1 : 0;
        P.grouped   = ht->grouped ? 1 : 0;
        P.pwords    = ht->words - 1;
        P.payload   = ht->grouped ? ht->payload.get<uint64_t>() : nullptr;
        P.early     = ks->second < scanned ? 1 : 0;
        // A bloom filter pays only when most probe keys miss: a table that
        // found at least half of the driving relation's sampled non-NULL keys
        // is probed without it (c1 and the c4 joins: foreign keys into
        // unfiltered build sides; the pre-filtered c3 dimensions keep theirs).
        bool bloom = ht->has_bloom;
        if (bloom && P.early && d.rel >= 0 && samples_) {
            const auto it = samples_->pass[uint32_t(d.rel)].find(ht);
            const auto kv = samples_->keys[uint32_t(d.rel)].find(probes[i].key_eq);
            if (it != samples_->pass[uint32_t(d.rel)].end() && kv != samples_->keys[uint32_t(d.rel)].end()) {
                uint32_t keys = 0, hits = 0;
                for (size_t j = 0; j < kv->second.size() && j < 64; ++j)
                    if (kv->second[j] != kNull) {
                        ++keys;
                        hits += uint32_t((it->second >> j) & 1u);
                    }
                if (keys >= 16 && 2 * hits >= keys) {
                    bloom = false;
                    ++blooms_skipped;
                }
            }
        }
        P.bloom     = bloom ? ht->bloom.get<uint64_t>() : nullptr;
        P.bloom_log = ht->bloom_log;
        if (probes.size() == 1 && (t.kind == TARGET_PAGES || t.kind == TARGET_COUNT) && packable(*ht)) {
            if (!ht->packed.get()) {
                ht->packed    = DevBuf(&ctx_.heap(), ht->cap * sizeof(uint64_t));
                ht->pack_flag = DevBuf(&ctx_.heap(), sizeof(uint32_t));
                CGJ_HIP(hipMemsetAsync(ht->pack_flag.get(), 0, sizeof(uint32_t), ctx_.stream()));
                CGJ_HIP(launch_pack_table(ht->slots.get<uint64_t>(), ht->cap, ht->packed.get<uint64_t>(),
                                          ht->pack_flag.get<uint32_t>(), ctx_.stream()));
            }
            P.ht_packed = ht->packed.get<uint64_t>();
            P.pack_flag = ht->pack_flag.get<uint32_t>();
        }
        for (uint32_t w = 1; w < ht->words; ++w) {
            uint32_t e  = ht->eqs[w];
            auto     it = slot.find(e);
            if (it != slot.end()) {
                if (P.nchecks >= uint32_t(kMaxChecks)) throw std::runtime_error("too many join predicates");
                P.check_word[P.nchecks] = w;
                P.check_slot[P.nchecks] = it->second;
                ++P.nchecks;
            } else {
                P.attr_word[P.nattrs]   = w;
                P.attr_slot[P.nattrs++] = nslots;
                slot[e]                 = nslots++;
            }
        }
    }
    if (nslots > 24) throw std::runtime_error("pipeline carries too many columns");
    // Scanned join columns must be non-NULL (NULL never matches); slots filled
    // from hash-table entries need no check, their build pipelines filtered them.
    for (auto& kv : slot)
        if (kv.second < scanned && q_.is_join_eq(kv.first)) a.not_null |= 1u << kv.second;

    a.target = t.kind;
    DevResult* res = t.result;
    double*    ms  = t.kind == TARGET_PAGES ? &ms_probe : &ms_build;
    if (t.kind == TARGET_STAGE) {
        a.stage_words = uint32_t(t.eqs.size());
        for (uint32_t w = 0; w < t.eqs.size(); ++w) {
            auto it = slot.find(t.eqs[w]);
            if (it == slot.end()) throw std::runtime_error("staged class is not available in the pipeline");
            a.stage_slot[w] = it->second;
        }
        // Each wave reserves `chunk` slots per atomic; the last reservation of a
        // wave can stay partly unused (those slots carry key kEmpty). The area
        // holds four times the planner's estimate (at least 64K tuples), never
        // more than the bound (a larger result re-runs at its exact size), as
        // the output pool does.
        const double est = t.card > 0 ? std::min(t.est, std::max(4.0 * t.card, 65536.0)) : t.est;
        uint64_t waves = grid_waves(nrows, xcd_order);
        double   per   = (probes.empty() ? double(nrows) : std::max(est, 1.0)) / double(waves);
        uint64_t chunk = std::min<uint64_t>(4096, std::max<uint64_t>(64, uint64_t(per)));
        uint64_t cap   = probes.empty() ? nrows : uint64_t(std::min(std::max(est * 1.25, 1024.0), 4e9));
        cap            = std::max<uint64_t>(cap, 1) + waves * chunk;
        // staged tuples beyond what device memory holds cannot be built: the
        // pipeline stops reserving slots there instead of enumerating them all
        a.abort_pages  = ctx_.mem_bytes() / (uint64_t(a.stage_words) * sizeof(uint64_t));
        for (int attempt = 0;; ++attempt) {
            t.stage       = DevBuf(&ctx_.heap(), cap * a.stage_words * sizeof(uint64_t));
            a.stage       = t.stage.get<uint64_t>();
            a.stage_cap   = cap;
            a.stage_chunk = chunk;
            launch(a, nslots, nrows, *ms);
            PipeCounters c;
            CGJ_HIP(hipMemcpy(&c, counters_.get(), sizeof(c), hipMemcpyDeviceToHost));
            if (c.fault) throw std::runtime_error("hash-table probe met no empty slot in its region");
            if (!c.overflow) {
                t.count = c.tuples;
                t.slots = std::min<uint64_t>(c.reserved(), cap);
                break;
            }
            if (c.reserved() >= a.abort_pages)
                throw std::runtime_error("join result too large: more than " + std::to_string(a.abort_pages) +
                                         " staged tuples, beyond device memory");
            if (attempt > 2) throw std::runtime_error("staging overflow persists");
            ++retries;
            cap = c.reserved() + waves * chunk + 64;
            if (cap * a.stage_words * sizeof(uint64_t) > ctx_.mem_bytes())
                throw std::runtime_error("join result too large: " + std::to_string(c.pages) +
                                         " staged tuples need more than the device memory");
        }
        return;
    }
    if (t.kind == TARGET_COUNT) {
        launch(a, nslots, nrows, *ms);
        PipeCounters c;
        CGJ_HIP(hipMemcpy(&c, counters_.get(), sizeof(c), hipMemcpyDeviceToHost));
        if (c.fault) throw std::runtime_error("hash-table probe met no empty slot in its region");
        t.count = c.tuples;
        return;
    }
    // TARGET_PAGES
    a.nout = uint32_t(q_.out.size());
    if (a.nout > uint32_t(kMaxOut)) throw std::runtime_error("too many output columns");
    double bytes_per_row = 0;
    for (uint32_t i = 0; i < a.nout; ++i) {
        uint32_t attr = q_.out[i];
        auto     it   = slot.find(q_.attrs[attr].eq);
        if (it == slot.end()) throw std::runtime_error("output class is not available in the pipeline");
        a.out[i].slot = it->second;
        a.out[i].type = q_.attrs[attr].type;
        a.out[i].vc   = a.out[i].type == CGJ_VARCHAR ? a.nvarchar++ : 0;
        bytes_per_row += a.out[i].type == CGJ_INT32 ? 4.2 : (a.out[i].type == CGJ_VARCHAR ? 48.0 : 8.2);
    }
    // Expected data pages: for four times the planner's estimate (at least
    // 64K rows), never more than the bound; a larger result overflows the pool
    // and the pipeline re-runs at its exact size. A pool sized for the bound
    // alone (the largest input of the set) is mostly unused when the result
    // is small, and thousands of such results held together do not fit the
    // device. Every wave also ends with a partly filled page per column and
    // may hold a partly used reservation. So that those stay small beside the
    // data, the grid gets about as many waves per CU as there are pages per
    // column per CU (4..the default), and a reservation holds about one wave's
    // share of the pages (1..kPageChunk).
    const double   pool_rows = t.card > 0 ? std::min(t.est, std::max(4.0 * t.card, 65536.0)) : t.est;
    const double   dp      = std::min(pool_rows, 4e9) * bytes_per_row / 7800.0;
    const uint32_t def_pcu = per_cu_for(a, nslots);
    a.grid_per_cu          = uint32_t(std::min<double>(
        def_pcu, std::max<double>(4.0, dp / (std::max<uint32_t>(a.nout, 1) * double(ctx_.cu_count())))));
    uint64_t waves   = waves_for(a, nslots, nrows);
    a.page_chunk     = uint32_t(std::min<double>(kPageChunk, std::max<double>(1.0, std::ceil(dp / double(waves)))));
    DevBuf   stage_s = DevBuf(&ctx_.heap(), std::max<uint64_t>(waves * a.nvarchar * kPageSize, 64));
    a.str_stage      = stage_s.get<uint8_t>();
    uint64_t slack   = waves * a.page_chunk + 64;
    uint64_t pages   = uint64_t(dp) + waves * a.nout + slack;
    // more output pages than device memory holds: the pipeline stops there
    a.abort_pages    = ctx_.mem_bytes() / kPageSize;
    for (int attempt = 0;; ++attempt) {
        res->pool    = DevBuf(&ctx_.heap(), pages * kPageSize + pages * 16 + 64);
        res->pool.set_evict_first();
        a.pool       = res->pool.get<uint8_t>();
        a.pool_pages = pages;
        a.dir        = reinterpret_cast<uint32_t*>(a.pool + pages * kPageSize);
        a.dir_cap    = pages;
        launch(a, nslots, nrows, *ms);
        PipeCounters c;
        CGJ_HIP(hipMemcpy(&c, counters_.get(), sizeof(c), hipMemcpyDeviceToHost));
        if (c.fault) throw std::runtime_error("hash-table probe met no empty slot in its region");
        if (!c.overflow) {
            res->nrows          = c.tuples;
            res->total_pages    = 0;
            for (uint32_t i = 0; i < a.nout; ++i) res->total_pages += c.col_pages[i];
            res->pool_pages     = pages;
            res->reserved_pages = std::min<uint64_t>(c.reserved(), pages);
            compact_pool(*res);
            ms_probe_kernel     = last_ms_;
            probe_lean          = use_lean(a, nslots) ? 1 : 0;
            probe_packed        = c.packed ? 1 : 0;
            probe_waves         = uint32_t(waves_for(a, nslots, nrows));
            // output bytes at the reference's packing (TableTarget.cpp:111-182)
            for (uint32_t i = 0; i < a.nout; ++i) {
                int32_t ty = a.out[i].type;
                if (ty == CGJ_VARCHAR) {
                    out_ref_pages += c.col_pages[i];
                    continue;
                }
                uint64_t w      = ty == CGJ_INT32 ? 32 : 64;
                uint64_t budget = 8 * (kPageSize - 4 - (ty == CGJ_INT32 ? 0 : 4));
                out_ref_pages += (c.tuples + w * c.nnz[i] + budget - 1) / budget;
            }
            break;
        }
        if (c.reserved() >= a.abort_pages)
            throw std::runtime_error("query result too large: more than " + std::to_string(a.abort_pages) +
                                     " output pages (" + std::to_string(a.abort_pages * kPageSize >> 30) +
                                     " GiB), beyond device memory");
        if (attempt > 2) throw std::runtime_error("output page pool overflow persists");
        ++retries;
        pages = c.reserved() + slack;
        if (pages * (kPageSize + 16) > ctx_.mem_bytes())
            throw std::runtime_error("query result too large: " + std::to_string(c.pages) +
                                     " output pages need more than the device memory");
    }
    probe_rows += nrows;
    if (d.rel >= 0)
        for (uint32_t i = 0; i < a.nscan; ++i) probe_input_pages += column(uint32_t(d.rel), scan_cols[i]).npages;
}

bool Executor::run_chain(int node, Target& t) {
    std::vector<int> joins;
    int              leaf;
    walk(node, joins, leaf);
    std::vector<ProbeStep> probes;
    uint64_t               cur = 1ull << leaf;
    for (int j : joins) {
        const TreeNode&  jn  = t_->nodes[j];
        const TreeNode&  bn  = t_->nodes[jn.build];
        const uint32_t   key = key_for(jn.build, cur);
        const HashTable* ht  = build(jn.build, key, probe_rows_of(jn));
        if (!ht) return false; // empty build side
        probes.push_back(ProbeStep{ht, key, bn.set});
        cur |= bn.set;
    }
    Driver d;
    d.rel = leaf;
    d.set = 1ull << leaf;
    pipeline(d, probes, t);
    return true;
}

std::unique_ptr<DevResult> Executor::run(bool count_only) {
    auto res = std::make_unique<DevResult>();
    for (uint32_t a : q_.out) res->types.push_back(q_.attrs[a].type);
    for (auto& r : q_.rels)
        if (r.rows == 0) return res; // an empty input empties the inner join
    register_all_columns();
    Target t;
    t.kind   = count_only ? TARGET_COUNT : TARGET_PAGES;
    t.est    = size_bound(t_->nodes[t_->root].set, t_->nodes[t_->root].card);
    t.card   = t_->nodes[t_->root].card;
    t.result = res.get();
    if (!run_chain(t_->root, t)) return res;
    if (count_only) res->nrows = t.count;
    return res;
}

bool Executor::next_build(int node, int& x, uint32_t& key, uint64_t& probe_rows) const {
    std::vector<int> joins;
    int              leaf;
    walk(node, joins, leaf);
    uint64_t cur = 1ull << leaf;
    for (int j : joins) {
        const TreeNode& jn = t_->nodes[j];
        const int       b  = jn.build;
        const uint32_t  k  = key_for(b, cur);
        cur |= t_->nodes[b].set;
        if (t_->nodes[b].group >= 0) continue; // built already
        if (t_->leaf(b) || !next_build(b, x, key, probe_rows)) {
            // a plain scan, or a chain whose builds are all built tables
            x          = b;
            key        = k;
            probe_rows = probe_rows_of(jn);
        }
        return true;
    }
    return false;
}

void Executor::queue_refilter(const HashTable* ht) {
    rf_.table = nullptr;
    rf_.todo.clear();
    if (!samples_ || !(rf_.armed || (rf_.for_bloom && ht->has_bloom))) return;
    // relations still planned on their own that produce the key class
    size_t n = 0;
    for (uint32_t r = 0; r < q_.rels.size(); ++r) {
        if (((rf_.grouped >> r) & 1) || samples_->size[r] == 0) continue;
        auto it = samples_->keys[r].find(rf_.key);
        if (it == samples_->keys[r].end()) continue;
        rf_.todo.emplace_back(r, uint32_t(it->second.size()));
        n += it->second.size();
    }
    rf_.table = ht;
    if (!n) return;
    if (n * (sizeof(uint64_t) + sizeof(uint32_t)) > Context::kSamplePinnedBytes)
        throw std::runtime_error("sample re-filter: more sample keys than its buffer holds");
    uint8_t*  pin  = ctx_.sample_pinned();
    uint64_t* keys = reinterpret_cast<uint64_t*>(pin);
    size_t    k    = 0;
    for (auto& [r, cnt] : rf_.todo)
        for (uint64_t x : samples_->keys[r].find(rf_.key)->second) keys[k++] = x;
    rf_.dev       = DevBuf(&ctx_.heap(), n * (sizeof(uint64_t) + sizeof(uint32_t)));
    hipStream_t s = ctx_.stream();
    CGJ_HIP(hipMemcpyAsync(rf_.dev.get(), keys, n * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    SampleTable st;
    st.ht    = ht->slots.get<uint64_t>();
    st.mask  = ht->rmask;
    st.shift = ht->shift;
    st.words = ht->grouped ? 2 : ht->words;
    uint32_t* dfound = reinterpret_cast<uint32_t*>(rf_.dev.get<uint8_t>() + n * sizeof(uint64_t));
    CGJ_HIP(launch_sample_probe(st, rf_.dev.get<uint64_t>(), uint32_t(n), dfound, s));
    CGJ_HIP(hipMemcpyAsync(pin + n * sizeof(uint64_t), dfound, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
}

void Executor::finish_refilter(const HashTable* ht) {
    if (!samples_) return;
    if (rf_.table != ht) { // the build's path did not queue it: probe now
        queue_refilter(ht);
        CGJ_HIP(hipStreamSynchronize(ctx_.stream()));
    }
    // the build synchronised after the queued copies: the hits are here
    size_t n = 0;
    for (auto& [r, cnt] : rf_.todo) n += cnt;
    const uint32_t* found = reinterpret_cast<const uint32_t*>(ctx_.sample_pinned() + n * sizeof(uint64_t));
    size_t          k     = 0;
    for (auto& [r, cnt] : rf_.todo) {
        uint64_t m = 0;
        for (uint32_t i = 0; i < cnt; ++i, ++k)
            if (found[k]) m |= 1ull << i;
        samples_->matches[r] &= m;
        samples_->pass[r][ht] = m;
    }
    rf_.table = nullptr;
    rf_.todo.clear();
    rf_.dev.reset();
}

std::unique_ptr<DevResult> Executor::run_incremental(bool count_only, const VertexCard& card,
                                                     std::shared_ptr<Samples> samples) {
    samples_ = std::move(samples);
    auto res = std::make_unique<DevResult>();
    for (uint32_t a : q_.out) res->types.push_back(q_.attrs[a].type);
    for (auto& r : q_.rels)
        if (r.rows == 0) return res; // an empty input empties the inner join
    register_all_columns();
    PlanTree tree;
    for (;;) {
        // the join graph now: built tables plus the relations none of them covers
        const auto              t0 = std::chrono::steady_clock::now();
        std::vector<PlanVertex> vx;
        uint64_t                grouped = 0;
        for (size_t g = 0; g < groups_.size(); ++g) {
            PlanVertex v;
            v.set   = groups_[g].set;
            v.group = int(g);
            v.key   = int32_t(groups_[g].key);
            vx.push_back(v);
            grouped |= groups_[g].set;
        }
        for (uint32_t r = 0; r < q_.rels.size(); ++r)
            if (!(((grouped | done_) >> r) & 1)) {
                PlanVertex v;
                v.set = 1ull << r;
                v.rel = int(r);
                vx.push_back(v);
            }
        tree = plan_joins_vertices(q_, vx, [&](uint64_t vmask) { return card(vx, vmask); });
        t_   = &tree;
        ms_replan += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        ++replans;
        int      x   = -1;
        uint32_t key = 0;
        uint64_t prows = 0;
        if (!next_build(tree.root, x, key, prows)) {
            // every build of the root's chain is done: the output pipeline
            Target t;
            t.kind   = count_only ? TARGET_COUNT : TARGET_PAGES;
            t.est    = size_bound(tree.nodes[tree.root].set, tree.nodes[tree.root].card);
            t.card   = tree.nodes[tree.root].card;
            t.result = res.get();
            if (!run_chain(tree.root, t)) return res;
            if (count_only) res->nrows = t.count;
            return res;
        }
        // The samples steer the next join order and the probe order of a
        // pipeline with several probes: with one relation and one table left
        // after this build there is nothing for them to decide.
        const uint64_t set     = tree.nodes[x].set;
        uint64_t       covered = set | done_;
        size_t         ntables = 1;
        for (const Group& g : groups_)
            if ((g.set & set) != g.set) { // tables the new one consumes drop out
                covered |= g.set;
                ++ntables;
            }
        const uint32_t left = uint32_t(q_.rels.size()) - uint32_t(__builtin_popcountll(covered));
        rf_.armed     = samples_ && (left >= 2 || ntables >= 2);
        // (a filter's cost grows with the probe side: below 1M probe rows
        // the extra sample probe costs more than skipping the filter saves)
        rf_.for_bloom = samples_ != nullptr && prows >= (uint64_t(1) << 20);
        rf_.key       = key;
        rf_.grouped = covered;
        rf_.table   = nullptr;
        const HashTable* ht = build(x, key, prows);
        if (!ht) return res; // an empty build side empties the inner join
        const bool         semi = restrict_after_build(ht, key, set);
        std::vector<Group> keep;
        for (const Group& g : groups_)
            if ((g.set & set) != g.set) keep.push_back(g);
        if (!semi) keep.push_back(Group{set, key, ht});
        groups_ = std::move(keep);
        if (rf_.armed || (rf_.for_bloom && ht->has_bloom)) finish_refilter(ht);
        rf_.for_bloom = false;
        rf_.armed = false;
    }
}

} // namespace

std::unique_ptr<DevResult> Context::execute(const cgj_plan& plan, const std::vector<const DevTable*>& inputs) {
    return run_query(plan, inputs, false);
}

std::vector<std::unique_ptr<DevResult>> Context::execute_many(const std::vector<const cgj_plan*>& plans,
                                                              const std::vector<std::vector<const DevTable*>>& inputs,
                                                              uint32_t workers) {
    if (plans.size() != inputs.size()) throw std::runtime_error("one input list per plan");
    workers = std::max<uint32_t>(1, std::min<uint32_t>({workers, 16u, uint32_t(std::max<size_t>(plans.size(), 1))}));
    while (workers_.size() < workers) {
        workers_.push_back(std::make_unique<Context>(device_));
        workers_.back()->options = options;
        workers_.back()->hints_  = hints_;
    }
    for (uint32_t w = 0; w < workers; ++w) workers_[w]->options = options;
    // Each worker plans its query with this context's estimator (statistics
    // catalog shared under a lock, the network's inference on the worker's own
    // stream) and executes it, so planning runs on as many threads as
    // execution does.
    if (est_->mode == EST_LEARNED && est_->mlp) est_->mlp->refresh_bf16(); // before any thread reads the weights
    CGJ_HIP(hipStreamSynchronize(stream_)); // inputs uploaded on this stream are complete
    std::vector<std::unique_ptr<DevResult>> out(plans.size());
    std::vector<std::string>                errors(workers);
    std::atomic<size_t>                     next{0};
    auto                                    run = [&](uint32_t w) {
        try {
            CGJ_HIP(hipSetDevice(device_));
            Context& wc = *workers_[w];
            for (size_t i; (i = next.fetch_add(1)) < plans.size();)
                out[i] = wc.run_planned(wc.plan_query(*plans[i], inputs[i], *est_), inputs[i], false);
        } catch (const std::exception& e) {
            errors[w] = e.what();
            next      = plans.size(); // stop the other workers early
        }
    };
    {
        DeviceHeap::Parallel     parallel;
        std::vector<std::thread> threads;
        for (uint32_t w = 0; w < workers; ++w) threads.emplace_back(run, w);
        for (auto& t : threads) t.join();
    }
    DeviceHeap::settle(device_);
    for (const auto& e : errors)
        if (!e.empty()) throw std::runtime_error("execute_many: " + e);
    return out;
}

uint64_t Context::count(const cgj_plan& plan, const std::vector<const DevTable*>& inputs) {
    return run_query(plan, inputs, true)->nrows;
}

double Context::copy_probe(uint64_t bytes, uint32_t reps) {
    uint64_t   words = std::max<uint64_t>(bytes / 16, 1);
    DevBuf     a(heap_.get(), words * 16), b(heap_.get(), words * 16);
    hipEvent_t e0, e1;
    CGJ_HIP(hipEventCreate(&e0));
    CGJ_HIP(hipEventCreate(&e1));
    CGJ_HIP(launch_copy(b.get(), a.get(), words, stream_)); // warm-up
    CGJ_HIP(hipEventRecord(e0, stream_));
    for (uint32_t i = 0; i < std::max<uint32_t>(reps, 1); ++i) CGJ_HIP(launch_copy(b.get(), a.get(), words, stream_));
    CGJ_HIP(hipEventRecord(e1, stream_));
    CGJ_HIP(hipEventSynchronize(e1));
    float ms = 0;
    CGJ_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return double(ms) / std::max<uint32_t>(reps, 1);
}

void Context::check_types(const Query& q, const std::vector<const DevTable*>& inputs) const {
    for (auto& r : q.rels)
        for (uint32_t a : r.attrs)
            if (r.rows && inputs[r.table]->cols[q.attrs[a].col].type != q.attrs[a].type)
                throw std::runtime_error("plan column type differs from the input column type");
}

static Query import_inputs(const cgj_plan& plan, const std::vector<const DevTable*>& inputs) {
    std::vector<uint64_t> rows;
    std::vector<uint32_t> cols;
    for (auto* t : inputs) {
        rows.push_back(t->nrows);
        cols.push_back(uint32_t(t->cols.size()));
    }
    return import_plan(plan, rows, cols);
}

uint64_t Context::count_query(const Query& q, const std::vector<const DevTable*>& inputs) {
    if (q.rels.size() == 1) return q.rels[0].rows;
    std::vector<double> base;
    for (auto& r : q.rels) base.push_back(double(r.rows));
    PlanTree tree = plan_joins(q, max_card_model(q, base));
    Executor ex(*this, q, tree, inputs);
    return ex.run(true)->nrows;
}

uint64_t Context::count_subset(const cgj_plan& plan, const std::vector<const DevTable*>& inputs, uint64_t relset) {
    Query q = import_inputs(plan, inputs);
    check_types(q, inputs);
    return count_query(restrict_query(q, relset), inputs);
}

std::vector<Context::PlannedJoin> Context::plan_tree(const cgj_plan& plan, const std::vector<const DevTable*>& inputs) {
    Query q = import_inputs(plan, inputs);
    check_types(q, inputs);
    EstimateReport           rep;
    std::shared_ptr<Samples> smp;
    bool                     empty = false;
    for (const auto& r : q.rels) empty = empty || r.rows == 0;
    if ((est_->mode == EST_SAMPLE || est_->mode == EST_LEARNED) && !empty) smp = take_samples(q, inputs);
    PlanTree                 tree = plan_joins(q, est_->model(*this, q, inputs, rep, smp));
    std::vector<PlannedJoin> out;
    std::function<void(int)> rec = [&](int n) {
        if (tree.leaf(n)) return;
        const TreeNode& t = tree.nodes[n];
        rec(t.build);
        rec(t.probe);
        out.push_back(PlannedJoin{tree.nodes[t.build].set, tree.nodes[t.probe].set, t.card});
    };
    rec(tree.root);
    return out;
}

PlannedQuery Context::plan_query(const cgj_plan& plan, const std::vector<const DevTable*>& inputs) {
    return plan_query(plan, inputs, *est_);
}

PlannedQuery Context::plan_query(const cgj_plan& plan, const std::vector<const DevTable*>& inputs, Estimator& est) {
    PlannedQuery pq;
    pq.q = import_inputs(plan, inputs);
    check_types(pq.q, inputs);
    for (uint32_t a : pq.q.out) pq.types.push_back(pq.q.attrs[a].type);
    const auto t0 = std::chrono::steady_clock::now();
    pq.mode       = est.mode;
    bool empty    = false;
    for (const auto& r : pq.q.rels) empty = empty || r.rows == 0;
    bool single = false;
    for (const auto& r : pq.q.rels) single = single || r.rows == 1;
    if (options.singletons && single && !empty && pq.q.rels.size() > 1) {
        SingletonReduction red = eliminate_singletons(pq.q, singleton_values(pq.q, inputs));
        if (red.eliminated) {
            for (ConstOut& c : red.consts)
                if (c.type == T_VARCHAR && c.value != kNull) {
                    const QAttr& at = pq.q.attrs[pq.q.rels[c.rel].attrs[c.slot]];
                    c.str           = read_string(inputs[pq.q.rels[c.rel].table]->cols[at.col], c.value);
                }
            pq.q          = std::move(red.q);
            pq.consts     = std::move(red.consts);
            pq.out_pos    = std::move(red.out_pos);
            pq.empty      = red.empty;
            pq.singletons = red.eliminated;
        }
    }
    if (pq.empty) { // no row can result: nothing to plan
        pq.ms_plan = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return pq;
    }
    if ((est.mode == EST_SAMPLE || est.mode == EST_LEARNED) && !empty) pq.samples = take_samples(pq.q, inputs);
    pq.card       = est.model(*this, pq.q, inputs, pq.rep, pq.samples);
    pq.tree       = plan_joins(pq.q, pq.card);
    pq.ms_plan =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() - pq.rep.ms_stats;
    return pq;
}

std::vector<std::vector<uint64_t>> Context::singleton_values(const Query& q,
                                                             const std::vector<const DevTable*>& inputs) {
    std::vector<std::vector<uint64_t>>          vals(q.rels.size());
    std::vector<SampleCol>                      cols;
    std::vector<std::pair<uint32_t, uint32_t>> where; // (relation, attribute index)
    for (uint32_t r = 0; r < q.rels.size(); ++r) {
        if (q.rels[r].rows != 1) continue;
        for (uint32_t i = 0; i < q.rels[r].attrs.size(); ++i) {
            const DevColumn& c = inputs[q.rels[r].table]->cols[q.attrs[q.rels[r].attrs[i]].col];
            // column k reads row id k of the list and writes output k
            cols.push_back(SampleCol{c.view(0), uint32_t(cols.size()), 1});
            where.emplace_back(r, i);
        }
        vals[r].assign(q.rels[r].attrs.size(), kNull);
    }
    if (cols.empty()) return vals;
    const size_t cb = cols.size() * sizeof(SampleCol), ob = cols.size() * sizeof(uint64_t);
    uint8_t*     pin = small_pinned(cb + 2 * ob);
    std::memcpy(pin, cols.data(), cb);
    std::memset(pin + cb, 0, ob); // every row id is 0: the relation's one row
    DevBuf dev(heap_.get(), cb + 2 * ob);
    CGJ_HIP(hipMemcpyAsync(dev.get(), pin, cb + ob, hipMemcpyHostToDevice, stream_));
    uint8_t* d = dev.get<uint8_t>();
    CGJ_HIP(launch_sample_gather(reinterpret_cast<const SampleCol*>(d), uint32_t(cols.size()),
                                 reinterpret_cast<const uint64_t*>(d + cb), reinterpret_cast<uint64_t*>(d + cb + ob),
                                 stream_));
    CGJ_HIP(hipMemcpyAsync(pin + cb + ob, d + cb + ob, ob, hipMemcpyDeviceToHost, stream_));
    CGJ_HIP(hipStreamSynchronize(stream_));
    const uint64_t* out = reinterpret_cast<const uint64_t*>(pin + cb + ob);
    for (size_t k = 0; k < where.size(); ++k) vals[where[k].first][where[k].second] = out[k];
    return vals;
}

std::string Context::read_string(const DevColumn& c, uint64_t handle) {
    uint64_t       p   = handle_page(handle);
    const uint32_t idx = handle_idx(handle);
    if (p >= c.npages) throw std::runtime_error("string handle past the column's pages");
    uint8_t* pin  = small_pinned(kPageSize);
    auto     read = [&](uint64_t page) {
        CGJ_HIP(hipMemcpyAsync(pin, c.pages.get<uint8_t>() + page * kPageSize, kPageSize, hipMemcpyDeviceToHost,
                               stream_));
        CGJ_HIP(hipStreamSynchronize(stream_));
    };
    auto u16 = [&](uint32_t off) { return uint32_t(pin[off]) | (uint32_t(pin[off + 1]) << 8); };
    read(p);
    std::string s;
    if (idx == 0xffff) { // a long string: the first page, then its continuations
        for (;;) {
            const uint32_t len = std::min<uint32_t>(u16(2), kPageSize - 4);
            s.append(reinterpret_cast<const char*>(pin + 4), len);
            if (++p >= c.npages) break;
            read(p);
            if (u16(0) != 0xfffe) break;
        }
        return s;
    }
    const uint32_t nv = u16(2);
    if (idx >= nv) throw std::runtime_error("string handle past the page's values");
    const uint32_t base = 4 + 2 * nv;
    const uint32_t st = idx ? u16(4 + 2 * (idx - 1)) : 0, en = u16(4 + 2 * idx);
    if (en < st || base + en > kPageSize) throw std::runtime_error("corrupt string page");
    return std::string(reinterpret_cast<const char*>(pin + base + st), en - st);
}

std::shared_ptr<Samples> Context::take_samples(const Query& q, const std::vector<const DevTable*>& inputs) {
    auto           sm = std::make_shared<Samples>();
    const uint32_t n  = uint32_t(q.rels.size());
    sm->rows.resize(n);
    sm->size.resize(n);
    sm->keys.resize(n);
    sm->matches.resize(n);
    sm->pass.resize(n);
    std::vector<SampleCol>                     cols;
    std::vector<uint64_t>                      rowids;
    std::vector<std::pair<uint32_t, uint32_t>> where; // (relation, class) of each gathered column
    for (uint32_t r = 0; r < n; ++r) {
        const uint64_t rows = q.rels[r].rows;
        const uint32_t size = uint32_t(std::min<uint64_t>(rows, 64));
        sm->rows[r]         = rows;
        sm->size[r]         = size;
        sm->matches[r]      = size >= 64 ? ~0ull : ((1ull << size) - 1);
        if (!size) continue;
        // sample rows drawn with replacement from a stream fixed by the
        // relation's size and position: the same plan samples the same rows
        std::vector<uint64_t> ids(size);
        const uint64_t        seed = sm_mix(rows * 0x9e3779b97f4a7c15ull + r + 1);
        for (uint32_t i = 0; i < size; ++i) ids[i] = sm_mix(seed + (i + 1) * kGolden) % rows;
        std::sort(ids.begin(), ids.end());
        for (uint32_t e : q.rels[r].eqs) {
            bool restricted = false;
            for (const QFilter& f : q.filters) restricted = restricted || (q.attrs[f.attr].rel == r && q.attrs[f.attr].eq == e);
            if (!q.is_join_eq(e) && !restricted) continue;
            const DevColumn& c = inputs[q.rels[r].table]->cols[q.attrs[q.attr_of(r, e)].col];
            cols.push_back(SampleCol{c.view(0), uint32_t(rowids.size()), size});
            rowids.insert(rowids.end(), ids.begin(), ids.end());
            where.emplace_back(r, e);
        }
    }
    if (!cols.empty()) {
        const size_t cb = cols.size() * sizeof(SampleCol), rb = rowids.size() * sizeof(uint64_t);
        uint8_t*     pin = small_pinned(cb + 2 * rb);
        std::memcpy(pin, cols.data(), cb);
        std::memcpy(pin + cb, rowids.data(), rb);
        DevBuf dev(heap_.get(), cb + 2 * rb);
        CGJ_HIP(hipMemcpyAsync(dev.get(), pin, cb + rb, hipMemcpyHostToDevice, stream_));
        uint8_t* d = dev.get<uint8_t>();
        CGJ_HIP(launch_sample_gather(reinterpret_cast<const SampleCol*>(d), uint32_t(cols.size()),
                                     reinterpret_cast<const uint64_t*>(d + cb), reinterpret_cast<uint64_t*>(d + cb + rb),
                                     stream_));
        CGJ_HIP(hipMemcpyAsync(pin + cb + rb, d + cb + rb, rb, hipMemcpyDeviceToHost, stream_));
        CGJ_HIP(hipStreamSynchronize(stream_));
        const uint64_t* out = reinterpret_cast<const uint64_t*>(pin + cb + rb);
        for (size_t k = 0; k < cols.size(); ++k) {
            const auto [r, e] = where[k];
            std::vector<uint64_t> v(out + cols[k].first, out + cols[k].first + cols[k].n);
            for (uint32_t i = 0; i < v.size(); ++i)
                if (v[i] == kNull) sm->matches[r] &= ~(1ull << i); // NULL never joins or passes a restriction
            for (const QFilter& f : q.filters) // the scan's restrictions
                if (q.attrs[f.attr].rel == r && q.attrs[f.attr].eq == e)
                    for (uint32_t i = 0; i < v.size(); ++i)
                        if (!f.pass(v[i])) sm->matches[r] &= ~(1ull << i);
            sm->keys[r][e] = std::move(v);
        }
    }
    return sm;
}

std::unique_ptr<DevResult> Context::run_query(const cgj_plan& plan, const std::vector<const DevTable*>& inputs,
                                              bool count_only) {
    PlannedQuery pq = plan_query(plan, inputs);
    return run_planned(pq, inputs, count_only);
}

std::unique_ptr<DevResult> Context::run_planned(const PlannedQuery& pq, const std::vector<const DevTable*>& inputs,
                                                bool count_only) {
    const Query&          q   = pq.q;
    const EstimateReport& rep = pq.rep;
    Executor              ex(*this, q, pq.tree, inputs);
    std::unique_ptr<DevResult> res;
    if (pq.empty) {
        res = std::make_unique<DevResult>();
        for (uint32_t a : q.out) res->types.push_back(q.attrs[a].type);
    } else if (!options.incremental) {
        res = ex.run(count_only);
    } else {
        // Vertex cardinalities for re-planning. The max-of-inputs rule (sampled
        // or exact relation sizes) takes a built table's exact tuple count as
        // its cardinality; the other models scale their estimate of the
        // covered relations by each built table's measured / estimated size.
        std::vector<uint64_t> group_tuples; // filled through the executor
        const CardFn          base  = pq.card;
        const uint32_t        mode  = pq.mode;
        auto                  smp   = pq.samples;
        Executor::VertexCard  vcard = [&, base, mode, smp](const std::vector<PlanVertex>& vx, uint64_t vmask) {
            const bool max_rule = mode == EST_MAX || mode == EST_SAMPLE;
            double     c = 0, corr = 1;
            uint64_t   set = 0;
            for (uint64_t t = vmask; t; t &= t - 1) {
                const PlanVertex& v = vx[__builtin_ctzll(t)];
                set |= v.set;
                double vc;
                if (v.group >= 0)
                    vc = double(ex.group_rows(v.group));
                else if (mode == EST_SAMPLE && smp)
                    vc = smp->card(uint32_t(v.rel));
                else
                    vc = double(q.rels[v.rel].rows);
                if (max_rule) c = std::max(c, vc);
                else if (v.group >= 0) corr *= vc / std::max(base(v.set), 1.0);
            }
            if (max_rule) return c;
            if (!(vmask & (vmask - 1))) {
                const PlanVertex& v = vx[__builtin_ctzll(vmask)];
                return v.group >= 0 ? double(ex.group_rows(v.group)) : base(v.set);
            }
            return base(set) * corr;
        };
        res = ex.run_incremental(count_only, vcard, smp);
    }
    if (pq.singletons) { // constant columns and the output order
        res->pool_types = res->types;
        res->types      = pq.types;
        res->out_pos    = pq.out_pos;
        res->consts     = pq.consts;
    }
    stats                   = cgj_stats{};
    stats.singletons        = pq.singletons;
    stats.restrictions      = ex.restrictions;
    stats.semijoins         = ex.semijoins;
    stats.ms_plan           = pq.ms_plan + ex.ms_replan;
    stats.ms_stats          = rep.ms_stats;
    stats.est_sets          = rep.sets;
    stats.est_fallbacks     = rep.fallbacks;
    stats.est_stat_joins    = rep.stat_joins;
    stats.ms_build          = ex.ms_build;
    stats.ms_probe          = ex.ms_probe;
    stats.ms_total          = ex.ms_build + ex.ms_probe;
    stats.ms_probe_kernel   = ex.ms_probe_kernel;
    stats.ms_probe_partition = ex.ms_partition;
    stats.probe_kernel      = ex.probe_lean;
    stats.probe_packed      = ex.probe_packed;
    stats.pools_compacted   = ex.pools_compacted;
    stats.blooms_skipped    = ex.blooms_skipped;
    stats.probe_waves       = ex.probe_waves;
    stats.build_tuples      = ex.build_tuples;
    stats.build_bytes       = ex.build_bytes;
    stats.probe_rows        = ex.probe_rows;
    stats.probe_input_pages = ex.probe_input_pages;
    stats.output_rows       = res->nrows;
    stats.output_pages      = res->total_pages;
    stats.output_ref_pages  = ex.out_ref_pages;
    stats.pipelines         = ex.launches;
    stats.retries           = ex.retries;
    stats.retries_build     = ex.retries_build;
    stats.replans           = ex.replans;
    uint64_t in_pages       = 0;
    for (uint32_t ri = 0; ri < q.rels.size(); ++ri) {
        const QRel& r = q.rels[ri];
        if (!r.rows) continue;
        for (uint32_t e : q.needed(1ull << ri)) in_pages += inputs[r.table]->cols[q.attrs[q.attr_of(ri, e)].col].npages;
    }
    stats.input_pages = in_pages;
    return res;
}

namespace {
// `n` rows of one value as contest pages (the library's page writer, in
// chunks of bounded size).
void encode_constant(const ConstOut& c, uint64_t n, cgj_result_column& out) {
    out.type      = c.type;
    out.num_pages = 0;
    std::vector<void*> pages;
    const bool         null = c.value == kNull;
    const uint64_t     step = c.type == T_VARCHAR ? std::max<uint64_t>(1, (uint64_t(32) << 20) / (c.str.size() + 1))
                                                  : (uint64_t(1) << 20);
    for (uint64_t done = 0; done < n; done += step) {
        const uint64_t       m = std::min(step, n - done);
        std::vector<uint8_t> valid(m, null ? 0 : 1);
        cgj_result_column    part{};
        int                  rc;
        if (c.type == T_VARCHAR) {
            const uint64_t        len = null ? 0 : c.str.size();
            std::vector<uint64_t> offs(m + 1);
            std::string           chars;
            chars.reserve(m * len);
            for (uint64_t i = 0; i < m; ++i) {
                offs[i] = chars.size();
                if (!null) chars += c.str;
            }
            offs[m] = chars.size();
            rc      = cgj_encode_column(c.type, offs.data(), chars.data(), valid.data(), m, &part);
        } else if (c.type == T_INT32) {
            std::vector<int32_t> v(m, int32_t(uint32_t(c.value)));
            rc = cgj_encode_column(c.type, v.data(), nullptr, valid.data(), m, &part);
        } else {
            std::vector<uint64_t> v(m, null ? 0 : c.value);
            rc = cgj_encode_column(c.type, v.data(), nullptr, valid.data(), m, &part);
        }
        if (rc != 0) throw std::runtime_error("constant column: page encoding failed");
        for (uint64_t k = 0; k < part.num_pages; ++k) pages.push_back(part.pages[k]);
        delete[] part.pages;
    }
    out.num_pages = pages.size();
    out.pages     = new void*[pages.empty() ? 1 : pages.size()];
    for (size_t k = 0; k < pages.size(); ++k) out.pages[k] = pages[k];
}
} // namespace

void Context::download(const DevResult& r, cgj_result& out) {
    PoolView v;
    v.nrows          = r.nrows;
    v.types          = r.out_pos.empty() && r.consts.empty() ? r.types : r.pool_types;
    v.pool           = r.pool.get<uint8_t>();
    v.pool_pages     = r.pool_pages;
    v.reserved_pages = r.reserved_pages;
    v.total_pages    = r.total_pages;
    if (r.out_pos.empty() && r.consts.empty()) {
        download_pool(v, out);
        return;
    }
    cgj_result pool{};
    download_pool(v, pool);
    out.num_rows    = r.nrows;
    out.num_columns = uint32_t(r.types.size());
    out.columns     = new cgj_result_column[r.types.size()]();
    for (uint32_t j = 0; j < pool.num_columns; ++j) out.columns[r.out_pos[j]] = pool.columns[j];
    delete[] pool.columns;
    for (const ConstOut& c : r.consts) encode_constant(c, r.nrows, out.columns[c.pos]);
}

void Context::download_column(const DevColumn& c, cgj_result_column& out) {
    out.type      = c.type;
    out.num_pages = c.npages; // entries stay null until their page is copied
    out.pages     = new void*[c.npages ? c.npages : 1]();
    const uint64_t per     = kStagingBytes / kPageSize;
    const uint64_t nchunks = (c.npages + per - 1) / per;
    auto           issue   = [&](uint64_t k) {
        const uint64_t p0 = k * per, n = std::min(per, c.npages - p0);
        CGJ_HIP(hipMemcpyAsync(staging(int(k & 1)), c.pages.get<uint8_t>() + p0 * kPageSize, n * kPageSize,
                               hipMemcpyDeviceToHost, stream_));
        staging_done(int(k & 1));
    };
    if (nchunks) issue(0);
    for (uint64_t k = 0; k < nchunks; ++k) {
        if (k + 1 < nchunks) issue(k + 1);
        const uint8_t* buf = staging(int(k & 1)); // waits for chunk k's copy
        const uint64_t p0 = k * per, n = std::min(per, c.npages - p0);
        copy_pool().run(n, [&](uint64_t i) {
            void* pg = ::operator new(kPageSize);
            std::memcpy(pg, buf + i * kPageSize, kPageSize);
            out.pages[p0 + i] = pg;
        });
    }
}

void Context::download_pool(const PoolView& view, cgj_result& out) {
    PoolView r      = view;
    out.num_rows    = r.nrows;
    out.num_columns = uint32_t(r.types.size());
    out.columns     = new cgj_result_column[r.types.size()];
    for (size_t c = 0; c < r.types.size(); ++c) {
        out.columns[c].type      = r.types[c];
        out.columns[c].num_pages = 0;
        out.columns[c].pages     = new void*[1];
    }
    if (r.total_pages == 0) return;
    // Reserved pool pages that hold nothing (the unused part of each wave's
    // last page reservation) would cross the link with the rest: a pool with
    // more than 1/8 of them is first moved into a dense one on the device,
    // at device bandwidth, so only written pages are copied to the host.
    DevBuf                            dense, moved;
    std::array<unsigned long long, 2> mv{0, 0};
    if (r.reserved_pages > r.total_pages + r.total_pages / 8 + 256) {
        dense = DevBuf(heap_.get(), r.total_pages * (kPageSize + 16) + 64);
        moved = DevBuf(heap_.get(), 2 * sizeof(unsigned long long));
        auto* m = moved.get<unsigned long long>();
        CGJ_HIP(hipMemsetAsync(m, 0, 2 * sizeof(unsigned long long), stream_));
        CGJ_HIP(launch_compact_pages(r.pool, r.pool_pages, r.reserved_pages, dense.get<uint8_t>(), r.total_pages, m,
                                     m + 1, stream_));
        CGJ_HIP(hipMemcpyAsync(mv.data(), m, sizeof(mv), hipMemcpyDeviceToHost, stream_));
        r.pool           = dense.get<uint8_t>();
        r.pool_pages     = r.total_pages;
        r.reserved_pages = r.total_pages;
    }
    // The directory has one entry per reserved pool page: (column, wave
    // stream, sequence, valid). Page order per column: stream, then sequence.
    std::vector<uint32_t> dir(r.reserved_pages * 4);
    CGJ_HIP(hipMemcpyAsync(dir.data(), r.pool + r.pool_pages * kPageSize, dir.size() * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, stream_));
    CGJ_HIP(hipStreamSynchronize(stream_));
    if (dense.get() && (mv[1] || mv[0] != r.total_pages))
        throw std::runtime_error("page directory does not match the page count");
    std::vector<uint64_t> order;
    order.reserve(r.total_pages);
    for (uint64_t i = 0; i < r.reserved_pages; ++i)
        if (dir[4 * i + 3] == 1) order.push_back(i);
    if (order.size() != r.total_pages) throw std::runtime_error("page directory does not match the page count");
    std::sort(order.begin(), order.end(), [&](uint64_t x, uint64_t y) {
        const uint32_t* a0 = &dir[4 * x];
        const uint32_t* b0 = &dir[4 * y];
        if (a0[0] != b0[0]) return a0[0] < b0[0];
        if (a0[1] != b0[1]) return a0[1] < b0[1];
        return a0[2] < b0[2];
    });
    // Every (column, stream, sequence) key names one page: a repeated key
    // would make the order of that column's pages ambiguous.
    for (size_t k = 1; k < order.size(); ++k) {
        const uint32_t* a0 = &dir[4 * order[k - 1]];
        const uint32_t* b0 = &dir[4 * order[k]];
        if (a0[0] == b0[0] && a0[1] == b0[1] && a0[2] == b0[2])
            throw std::runtime_error("page directory repeats a (column, stream, sequence) key");
    }
    std::vector<std::vector<uint32_t>> pages(r.types.size());
    for (uint64_t i : order) {
        if (dir[4 * i] >= pages.size()) throw std::runtime_error("corrupt page directory");
        pages[dir[4 * i]].push_back(uint32_t(i));
    }
    uint64_t used = 0;
    for (auto& col : pages)
        for (uint32_t p : col) used = std::max<uint64_t>(used, uint64_t(p) + 1);
    const uint64_t kChunkPages = kStagingBytes / kPageSize;
    for (size_t c = 0; c < r.types.size(); ++c) {
        delete[] out.columns[c].pages;
        out.columns[c].pages     = new void*[pages[c].size() ? pages[c].size() : 1];
        out.columns[c].num_pages = pages[c].size();
        for (size_t i = 0; i < pages[c].size(); ++i) out.columns[c].pages[i] = nullptr;
    }
    // Copy the pool in chunks, alternating between the two staging buffers,
    // and scatter each chunk's pages to their columns while the next chunk's
    // copy runs.
    std::vector<std::pair<uint32_t, uint32_t>> where(used, {~0u, ~0u});
    for (size_t c = 0; c < r.types.size(); ++c)
        for (size_t i = 0; i < pages[c].size(); ++i) where[pages[c][i]] = {uint32_t(c), uint32_t(i)};
    const uint64_t nchunks = (used + kChunkPages - 1) / kChunkPages;
    auto issue = [&](uint64_t k) {
        const uint64_t p0 = k * kChunkPages, n = std::min(kChunkPages, used - p0);
        CGJ_HIP(hipMemcpyAsync(staging(int(k & 1)), r.pool + p0 * kPageSize, n * kPageSize, hipMemcpyDeviceToHost,
                               stream_));
        staging_done(int(k & 1));
    };
    if (nchunks) issue(0);
    for (uint64_t k = 0; k < nchunks; ++k) {
        if (k + 1 < nchunks) issue(k + 1);
        const uint8_t* buf = staging(int(k & 1)); // waits for chunk k's copy
        const uint64_t p0 = k * kChunkPages, n = std::min(kChunkPages, used - p0);
        copy_pool().run(n, [&](uint64_t i) {
            const auto w = where[p0 + i];
            if (w.first == ~0u) return;
            void* pg = ::operator new(kPageSize);
            std::memcpy(pg, buf + i * kPageSize, kPageSize);
            out.columns[w.first].pages[w.second] = pg;
        });
    }
}

} // namespace cgj
