// Device-side building blocks shared by the kernels: constants, hashing,
// wave64 collectives and the page reader.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cgj {

constexpr uint32_t kPageSize   = 8192;
constexpr uint64_t kNull       = ~0ull; // NULL slot value (reference RuntimeValue.hpp:10)
constexpr uint64_t kEmpty      = ~0ull; // empty hash-table slot; NULL keys are never inserted
constexpr uint32_t kWave       = 64;
constexpr uint32_t kMaxPageRows = 4096; // rows per output page: 64 lanes x 64 bitmap bits

enum : int32_t { T_INT32 = 0, T_INT64 = 1, T_FP64 = 2, T_VARCHAR = 3 };

// 64-bit finaliser (xor-shift-multiply); good avalanche on sequential and
// random keys alike. Slot positions use its top bits.
__host__ __device__ inline uint64_t hash_key(uint64_t k) {
    k ^= k >> 31;
    k *= 0x7fb5d329728ea185ull;
    k ^= k >> 27;
    k *= 0x81dadef4bc2dd44dull;
    k ^= k >> 33;
    return k;
}

// Second, independent-looking hash for the blocked bloom filter.
__host__ __device__ inline uint64_t hash_bloom(uint64_t h) { return h * 0x9e3779b97f4a7c15ull; }

// Bloom block: one 64-bit word per key with 4 bits set.
__host__ __device__ inline uint64_t bloom_bits(uint64_t h2) {
    return (1ull << ((h2 >> 0) & 63)) | (1ull << ((h2 >> 6) & 63)) | (1ull << ((h2 >> 12) & 63)) |
           (1ull << ((h2 >> 18) & 63));
}
__host__ __device__ inline uint64_t bloom_word(uint64_t h2, uint32_t log_words) {
    return log_words ? (h2 >> (64 - log_words)) : 0;
}

// String handle: [column id : 12][page : 36][value index : 16]; a value index
// of 0xffff marks a long string whose first page is `page`.
__host__ __device__ inline uint64_t make_handle(uint32_t col, uint64_t page, uint32_t idx) {
    return (uint64_t(col) << 52) | (page << 16) | idx;
}
__host__ __device__ inline uint32_t handle_col(uint64_t h) { return uint32_t(h >> 52); }
__host__ __device__ inline uint64_t handle_page(uint64_t h) { return (h >> 16) & ((1ull << 36) - 1); }
__host__ __device__ inline uint32_t handle_idx(uint64_t h) { return uint32_t(h & 0xffff); }

// ---- wave64 collectives ----
__device__ inline uint32_t lane_id() { return __lane_id(); }
// Index of the wave in its workgroup. readfirstlane tells the compiler the
// value is wave-uniform, so loop control derived from it stays in scalar
// registers and index loads become scalar loads.
__device__ inline uint32_t wave_in_block() { return __builtin_amdgcn_readfirstlane(threadIdx.x / 64); }
__device__ inline uint64_t mask_lt() {
    uint32_t l = lane_id();
    return l ? (~0ull >> (64 - l)) : 0ull;
}
__device__ inline uint64_t mask_le() {
    uint32_t l = lane_id();
    return ~0ull >> (63 - l);
}
__device__ inline uint32_t wave_incl_scan(uint32_t v) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane_id() >= uint32_t(d)) v += t;
    }
    return v;
}
__device__ inline uint64_t wave_or(uint64_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v |= __shfl_xor(v, d, 64);
    return v;
}
__device__ inline uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
__device__ inline uint32_t highest_lane(uint64_t m) { return 63 - __clzll(m); }

__device__ inline uint32_t uniform_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline uint64_t uniform_u64(uint64_t v) {
    return uint64_t(uniform_u32(uint32_t(v))) | (uint64_t(uniform_u32(uint32_t(v >> 32))) << 32);
}

// Lane 0 test for one-lane side effects (header stores, counter updates).
// The empty asm makes every call an independent value: two such tests in a
// loop cannot be correlated by the compiler. Correlated `lane == 0` branches
// at the end of one iteration and the start of the next let jump threading
// route lanes 1..63 around the loop head on their own, after which a
// readfirstlane in the loop reads lane 1 instead of lane 0 (a task loop built
// that way re-ran its first task on lanes 1..63 without end).
__device__ inline bool lead_lane() {
    uint32_t l = lane_id();
    asm volatile("" : "+v"(l));
    return l == 0;
}

// Adds `n` to *p once for the whole wave and returns the previous value to
// every lane: the first active lane does the atomic and its result is
// broadcast. The lane test is an opaque value per call (as in lead_lane), so
// it cannot be correlated with another one, and the broadcast sits right
// after it. Work queues and reservations use this.
__device__ inline unsigned long long wave_fetch_add(unsigned long long* p, unsigned long long n) {
    uint32_t l = lane_id();
    asm volatile("" : "+v"(l));
    unsigned long long old = 0;
    if (l == uniform_u32(l)) old = atomicAdd(p, n);
    return uniform_u64(old);
}
__device__ inline uint32_t wave_fetch_or(uint32_t* p, uint32_t v) {
    uint32_t l = lane_id();
    asm volatile("" : "+v"(l));
    uint32_t old = 0;
    if (l == uniform_u32(l)) old = atomicOr(p, v);
    return uniform_u32(old);
}

// Load of data that no kernel writes while the reading kernel runs (column
// indexes, hash tables built by an earlier launch). Through the constant
// address space a wave-uniform address becomes a scalar load; otherwise the
// compiler has to assume the kernel's own stores may alias it and issues a
// vector load plus readfirstlane. The cast goes through an integer: a direct
// pointer cast is folded back into a global-memory load. Adjacent dword loads
// merge into one s_load_dwordx4.
__device__ inline uint4 load_invariant_u4(const void* p) {
    const __attribute__((address_space(4))) uint32_t* q = (const __attribute__((address_space(4))) uint32_t*)(uintptr_t)p;
    return make_uint4(q[0], q[1], q[2], q[3]);
}

__device__ inline uint32_t load_invariant_u32(const uint32_t* p) {
    return *(const __attribute__((address_space(4))) uint32_t*)(uintptr_t)p;
}

__device__ inline uint64_t load_invariant_u64(const uint64_t* p) {
    const __attribute__((address_space(4))) uint32_t* q = (const __attribute__((address_space(4))) uint32_t*)(uintptr_t)p;
    return uint64_t(q[0]) | (uint64_t(q[1]) << 32);
}

// ---- page reader ----

// Decode index of one 64-row block of a column (rows 64b .. 64b+63): the page
// holding the block's first row, the number of non-NULL values before that row
// in its page, the validity of the 64 rows, and the first lane that lies on
// the next page (64 = none). A block that touches more than two pages, or a
// long-string page, is marked `complex` and decoded row by row.
struct BlockIdx {
    uint32_t page;
    uint16_t nn;
    uint8_t  split;
    uint8_t  complex;
    uint64_t valid;
};

// One scanned column as the kernels see it.
struct ScanCol {
    const uint8_t*  pages;      // npages x 8 KiB
    const uint32_t* row_start;  // npages + 1 prefix of rows per page
    const BlockIdx* blocks;     // ceil(nrows / 64) block entries
    int32_t         type;
    uint32_t        col_id;     // id used in string handles
    uint32_t        dense_rows; // > 0: fixed width, no NULLs, every page but the last holds this many
                                // rows (>= 64) and the column fewer than 2^32: row r is value r % d of
                                // page r / d, found without the block index
};

// Value slot of non-NULL value `nn` of page `p`.
__device__ inline uint64_t load_value(const ScanCol& c, uint32_t p, uint32_t nn, bool is_long) {
    const uint8_t* page = c.pages + uint64_t(p) * kPageSize;
    switch (c.type) {
    case T_INT32: return *reinterpret_cast<const uint32_t*>(page + 4 + 4 * nn);
    case T_INT64:
    case T_FP64: return *reinterpret_cast<const uint64_t*>(page + 8 + 8 * uint64_t(nn));
    default: return make_handle(c.col_id, p, is_long ? 0xffffu : nn);
    }
}

// Row-by-row decode (complex blocks): the page by search over the row prefix,
// validity from the bitmap, value index by counting bitmap bits.
__device__ inline uint64_t decode_row_slow(const ScanCol& c, uint32_t p, uint32_t row) {
    while (c.row_start[p + 1] <= row) ++p;
    uint32_t       j    = row - c.row_start[p];
    const uint8_t* page = c.pages + uint64_t(p) * kPageSize;
    uint32_t       nr   = *reinterpret_cast<const uint16_t*>(page);
    uint32_t       nv   = *reinterpret_cast<const uint16_t*>(page + 2);
    if (c.type == T_VARCHAR && nr == 0xffff) return load_value(c, p, 0, true);
    if (nv == nr) return load_value(c, p, j, false);
    const uint8_t* bm = page + kPageSize - (nr + 7) / 8;
    if (!((bm[j >> 3] >> (j & 7)) & 1)) return kNull;
    uint32_t nn = 0;
    for (uint32_t b = 0; b < (j >> 3); ++b) nn += __popc(bm[b]);
    nn += __popc(bm[j >> 3] & ((1u << (j & 7)) - 1));
    return load_value(c, p, nn, false);
}

// Decodes row 64*block + lane for every lane (lanes with !active yield NULL).
// Replaces the reference's skipTo/step64 (TableScan.cpp:135-324): validity and
// the value index come from the block's 64-bit mask by popcount, so a step
// costs one uniform index load plus one coalesced value load.
// The entry is read as four dwords (a scalar load when `block` is
// wave-uniform; its u16/u8 fields alone would force byte-wise vector loads).
__device__ inline BlockIdx load_block_idx(const BlockIdx* blocks, uint64_t block) {
    const uint4 w = load_invariant_u4(blocks + block);
    BlockIdx    bi;
    bi.page    = w.x;
    bi.nn      = uint16_t(w.y);
    bi.split   = uint8_t(w.y >> 16);
    bi.complex = uint8_t(w.y >> 24);
    bi.valid   = uint64_t(w.z) | (uint64_t(w.w) << 32);
    return bi;
}

// The vector part of decode_block, given the block's index entry.
__device__ inline uint64_t decode_with(const ScanCol& c, const BlockIdx& bi, uint64_t block, bool active) {
    const uint32_t lane = lane_id();
    if (bi.complex) return active ? decode_row_slow(c, bi.page, uint32_t(block * 64 + lane)) : kNull;
    bool second = lane >= bi.split;
    bool valid  = active && ((bi.valid >> lane) & 1ull);
    if (!valid) return kNull;
    uint64_t before = bi.valid & mask_lt();
    uint32_t nn     = second ? __popcll(before & ~((bi.split >= 64 ? ~0ull : (1ull << bi.split)) - 1ull))
                             : bi.nn + __popcll(before);
    return load_value(c, bi.page + (second ? 1 : 0), nn, false);
}

// decode_block of a dense column: a block spans at most two pages (d >= 64).
__device__ inline uint64_t decode_dense(const ScanCol& c, uint64_t block, bool active) {
    if (!active) return kNull;
    const uint32_t d  = c.dense_rows;
    const uint32_t r0 = uint32_t(block) * 64;
    const uint32_t p0 = r0 / d;
    const uint32_t j  = r0 - p0 * d + lane_id();
    const uint32_t s  = j >= d ? 1u : 0u;
    return load_value(c, p0 + s, j - s * d, false);
}

__device__ inline uint64_t decode_block(const ScanCol& c, uint64_t block, bool active) {
    if (c.dense_rows) return decode_dense(c, block, active);
    return decode_with(c, load_block_idx(c.blocks, block), block, active);
}

} // namespace cgj
