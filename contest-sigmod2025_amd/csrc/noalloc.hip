// Count of a two-scan equi-join that allocates nothing: the hash table, its
// flags and the pipeline counters live in one fixed device array per GPU.
//
// The array's address is looked up once per device (hipGetSymbolAddress, at
// context creation); host code only ever passes that address on, never the
// array's name. Every call on one GPU shares the array and leaves it dirty:
// each call clears what must start empty (the table) or at zero (flags and
// counters) with fill kernels, and calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "engine.hpp"
#include "noalloc.hpp"

namespace cgj {

namespace {

// Layout: [0, kNoAllocHeader) flags, inserted-row count and pipeline
// counters; the hash table after it.
constexpr uint64_t kNoAllocHeader = 4096;

static __device__ __align__(256) uint8_t scratch[kNoAllocHeader + kNoAllocTableBytes];

constexpr int kMaxDevices = 64;
std::mutex    g_mu;
void*         g_ptr[kMaxDevices] = {};

} // namespace

void* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("no-allocation scratch: bad device index");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        int cur = -1;
        CGJ_HIP(hipGetDevice(&cur));
        if (cur != device) CGJ_HIP(hipSetDevice(device));
        void*      p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess)
            throw HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[device] = p;
    }
    return g_ptr[device];
}

uint64_t count_join_noalloc(Context& ctx, const cgj_plan& plan, const std::vector<const DevTable*>& inputs) {
    std::vector<uint64_t> rows;
    std::vector<uint32_t> cols;
    for (const DevTable* t : inputs) {
        rows.push_back(t->nrows);
        cols.push_back(uint32_t(t->cols.size()));
    }
    const Query q = import_plan(plan, rows, cols);
    if (q.rels.size() != 2) throw std::runtime_error("count_noalloc: the plan must join exactly two scans");
    const auto shared = q.shared(1, 2);
    if (shared.size() != 1) throw std::runtime_error("count_noalloc: the two scans must join on one key");
    for (const QRel& r : q.rels)
        for (uint32_t a : r.attrs)
            if (r.rows && inputs[r.table]->cols[q.attrs[a].col].type != q.attrs[a].type)
                throw std::runtime_error("plan column type differs from the input column type");
    if (q.rels[0].rows == 0 || q.rels[1].rows == 0) return 0;
    const uint32_t  key = shared[0];
    const uint32_t  b = q.rels[0].rows <= q.rels[1].rows ? 0 : 1, p = 1 - b;
    const uint64_t  nb = q.rels[b].rows, np = q.rels[p].rows;
    const QAttr&    bk = q.attrs[q.attr_of(b, key)];
    const QAttr&    pk = q.attrs[q.attr_of(p, key)];
    const DevColumn& bc = inputs[q.rels[b].table]->cols[bk.col];
    const DevColumn& pc = inputs[q.rels[p].table]->cols[pk.col];
    if (bc.npages == 0 || pc.npages == 0) throw std::runtime_error("count_noalloc: key column not on the device");

    // key-only table at load <= 0.5
    uint32_t lg = 6;
    while ((uint64_t(1) << lg) < 2 * nb) ++lg;
    const uint64_t cap = uint64_t(1) << lg;
    if (cap * sizeof(uint64_t) > kNoAllocTableBytes)
        throw std::runtime_error("count_noalloc: a build side of " + std::to_string(nb) + " rows needs " +
                                 std::to_string(cap * sizeof(uint64_t)) + " bytes of hash table, more than the " +
                                 std::to_string(kNoAllocTableBytes) + "-byte scratch");

    uint8_t*      base     = static_cast<uint8_t*>(noalloc_scratch(ctx.device()));
    uint32_t*     flags    = reinterpret_cast<uint32_t*>(base);                      // 4 x u32
    auto*         inserted = reinterpret_cast<unsigned long long*>(base + 16);       // u64
    PipeCounters* counters = reinterpret_cast<PipeCounters*>(base + 256);
    uint64_t*     ht       = reinterpret_cast<uint64_t*>(base + kNoAllocHeader);
    static_assert(256 + sizeof(PipeCounters) <= kNoAllocHeader, "scratch header too small");
    hipStream_t s = ctx.stream();

    // clear the header (flags, count, counters) and empty the table
    CGJ_HIP(launch_fill_u64(reinterpret_cast<uint64_t*>(base), kNoAllocHeader / sizeof(uint64_t), 0, s));
    CGJ_HIP(launch_fill_u64(ht, cap, kEmpty, s));

    ScanBuildArgs sa{};
    sa.words   = 1;
    sa.notnull = 1;
    sa.col[0]  = bc.view(0);
    sa.nrows   = nb;
    CGJ_HIP(launch_scan_insert(sa, ht, lg, nullptr, 0, flags, 0, inserted, s));

    PipeArgs a;
    std::memset(&a, 0, sizeof(a));
    a.abort_pages = ~uint64_t(0);
    a.driver      = DRIVER_PAGES;
    a.nscan       = 1;
    a.scan[0]     = pc.view(0);
    a.not_null    = 1;
    a.nprobes     = 1;
    ProbeSpec& P  = a.probes[0];
    P.ht          = ht;
    P.shift       = 64 - lg;
    P.words       = 1;
    P.mask        = cap - 1;
    P.key_slot    = 0;
    P.unique      = 0; // duplicates are counted by walking to the empty slot
    P.early       = 1;
    a.target      = TARGET_COUNT;
    a.counters    = counters;
    a.page_chunk  = 1;
    // morsels as the engine sizes them: up to 4096 rows, at least 32 waves per CU
    const uint64_t target = std::max<uint64_t>(1, uint64_t(ctx.cu_count()) * 32);
    a.morsel_rows = uint32_t(std::min<uint64_t>(std::max<uint64_t>(((np + target - 1) / target + 63) / 64 * 64, 64), 4096));
    a.nrows       = np;
    a.nmorsels    = (np + a.morsel_rows - 1) / a.morsel_rows;
    const uint64_t waves = std::min<uint64_t>(a.nmorsels, target);
    CGJ_HIP(launch_probe1(a, 1, uint32_t((waves + 3) / 4 * 4), false, s));

    PipeCounters c;
    CGJ_HIP(hipMemcpyAsync(&c, counters, sizeof(c), hipMemcpyDeviceToHost, s));
    CGJ_HIP(hipStreamSynchronize(s));
    if (c.fault) throw std::runtime_error("count_noalloc: hash-table probe met no empty slot");
    return c.tuples;
}

} // namespace cgj
