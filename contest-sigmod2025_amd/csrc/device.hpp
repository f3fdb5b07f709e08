// Device-side building blocks shared by the kernels: constants, hashing,
// wave64 collectives and the page reader.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cgj {

constexpr uint32_t kPageSize   = 8192;
constexpr uint64_t kNull       = ~0ull; // NULL slot value (reference RuntimeValue.hpp:10)
constexpr uint64_t kEmpty      = ~0ull; // empty hash-table slot; NULL keys are never inserted
constexpr uint32_t kIndexRows  = 1024;  // row granularity of the per-column morsel index
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

// ---- page reader ----

// One scanned column as the kernels see it.
struct ScanCol {
    const uint8_t*  pages;     // npages x 8 KiB
    const uint32_t* row_start; // npages + 1 prefix of rows per page
    const uint32_t* idx_page;  // page holding row k * kIndexRows
    const uint32_t* idx_nn;    // non-NULL values before that row within its page
    int32_t         type;
    uint32_t        col_id;    // id used in string handles
};

// Per-wave decode position: the page of the wave's next row and the number of
// non-NULL values that precede that row in the page.
struct Cursor {
    uint32_t page;
    uint32_t nn;
};

__device__ inline Cursor cursor_at(const ScanCol& c, uint64_t row) {
    uint64_t k = row / kIndexRows;
    return Cursor{c.idx_page[k], c.idx_nn[k]};
}

// Decodes row `row` of column `c` for every lane; the active lanes must form
// a prefix of the wave and hold consecutive rows. Replaces the reference's
// skipTo/step64 (TableScan.cpp:135-324): the validity of each row comes from
// its own bitmap bit, and its value index is the ballot-prefix of valid rows
// of the same page, so no pext/pdep or offset tables are needed.
__device__ inline uint64_t decode_step(const ScanCol& c, Cursor& cur, uint32_t row, bool active) {
    uint32_t p = cur.page;
    if (active) {
        while (c.row_start[p + 1] <= row) ++p;
    }
    uint32_t       j    = active ? row - c.row_start[p] : 1;
    const uint8_t* page = c.pages + uint64_t(p) * kPageSize;
    uint32_t       hdr  = active ? *reinterpret_cast<const uint32_t*>(page) : 0;
    uint32_t       nr   = hdr & 0xffff;
    uint32_t       nv   = hdr >> 16;
    bool           is_long = c.type == T_VARCHAR && nr == 0xffff;
    bool           valid   = false;
    if (active) {
        if (is_long || nv == nr) {
            valid = true;
        } else {
            uint32_t nb = (nr + 7) >> 3;
            valid       = (page[kPageSize - nb + (j >> 3)] >> (j & 7)) & 1;
        }
    }
    uint64_t jz     = __ballot(active && j == 0);
    uint64_t starts = jz | 1ull;
    uint32_t seg    = highest_lane(starts & mask_le());
    uint64_t vmask  = __ballot(valid);
    uint32_t base   = (seg == 0 && !(jz & 1ull)) ? cur.nn : 0;
    uint64_t segm   = mask_lt() & ~((1ull << seg) - 1ull);
    uint32_t nn     = base + __popcll(vmask & segm);

    uint64_t slot = kNull;
    if (valid) {
        switch (c.type) {
        case T_INT32: slot = *reinterpret_cast<const uint32_t*>(page + 4 + 4 * nn); break;
        case T_INT64:
        case T_FP64: slot = *reinterpret_cast<const uint64_t*>(page + 8 + 8 * uint64_t(nn)); break;
        default: slot = make_handle(c.col_id, p, is_long ? 0xffffu : nn); break;
        }
    }
    uint64_t act = __ballot(active);
    if (act) {
        uint32_t last = highest_lane(act);
        cur.page      = __shfl(p, last, 64);
        cur.nn        = __shfl(nn + (valid ? 1u : 0u), last, 64);
    }
    return slot;
}

} // namespace cgj
