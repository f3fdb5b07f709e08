// HIP kernels of the join pipeline for gfx950 (wave64).
//
//  * column preparation: rows per page -> prefix (reference TableScan.cpp:546-582
//    makeSearchVector) plus a block index (page, non-NULL offset and validity
//    mask of every 64-row block) so that any wave decodes any block with one
//    index load and one coalesced value load;
//  * page decode into 64-bit slots (operator entry point and tests);
//  * hash-table insert: open addressing with linear probing over
//    [key, attrs...] entries, plus a blocked bloom filter;
//  * the fused pipeline: scan pages -> up to four probes -> target, where the
//    target is a staged tuple array (hash-table build side), the contest output
//    pages, or a tuple count.
#include <hip/hip_runtime.h>

#include <atomic>

#include "gen.hpp"
#include "kernels.hpp"

namespace cgj {

// ---------------------------------------------------------------------------
// exclusive scan of u32 counts (block-local scan, block sums, add-back)
// ---------------------------------------------------------------------------
constexpr int kScanBlock = 1024;

// Rows per page; dense[0] becomes 1 unless every page holds no NULL and as
// many rows as page 0 (the last page at most as many), dense[1] = rows of page 0.
__global__ void k_page_counts(const uint8_t* pages, uint64_t npages, int32_t type, uint32_t* counts,
                              uint32_t* dense) {
    const uint64_t p0 = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool     in = p0 < npages;
    const uint64_t p  = in ? p0 : npages - 1; // lanes past the end repeat the last page and write nothing
    uint32_t nr = *reinterpret_cast<const uint16_t*>(pages + p * kPageSize);
    uint32_t c  = nr;
    if (type == T_VARCHAR && nr >= 0xfffe) c = nr - 0xfffe; // long string: 1 row on its first page
    if (in) counts[p] = c;
    const uint32_t nv = *reinterpret_cast<const uint16_t*>(pages + p * kPageSize + 2);
    const uint32_t n0 = *reinterpret_cast<const uint16_t*>(pages);
    if (p0 == 0) dense[1] = n0;
    const bool mixed = type == T_VARCHAR || nv != nr || (p + 1 < npages ? nr != n0 : nr > n0);
    if (__ballot(mixed) && lead_lane()) atomicOr(dense, 1u); // one atomic per wave, not per page
}

__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
    // blockDim.x == kScanBlock, 16 waves
    uint32_t lane = lane_id(), w = threadIdx.x / 64;
    uint32_t inc  = wave_incl_scan(v);
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    if (w == 0) {
        uint32_t x = lane < kScanBlock / 64 ? sh[lane] : 0;
        uint32_t y = wave_incl_scan(x);
        if (lane < kScanBlock / 64) sh[lane] = y - x;
        if (lane == kScanBlock / 64 - 1) sh[16] = y;
    }
    __syncthreads();
    uint32_t r = sh[w] + inc - v;
    *total     = sh[16];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_local(uint32_t* data, uint64_t n, uint32_t* sums) {
    __shared__ uint32_t sh[17];
    uint64_t i = uint64_t(blockIdx.x) * kScanBlock + threadIdx.x;
    uint32_t v = i < n ? data[i] : 0;
    uint32_t tot;
    uint32_t e = block_excl_scan(v, sh, &tot);
    if (i < n) data[i] = e;
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// One block scans all block sums in chunks; writes the grand total.
__global__ __launch_bounds__(kScanBlock) void k_scan_sums(uint32_t* sums, uint64_t nblocks, uint64_t* total) {
    __shared__ uint32_t sh[17];
    uint32_t carry = 0;
    for (uint64_t base = 0; base < nblocks; base += kScanBlock) {
        uint64_t i = base + threadIdx.x;
        uint32_t v = i < nblocks ? sums[i] : 0;
        uint32_t tot;
        uint32_t e = block_excl_scan(v, sh, &tot);
        if (i < nblocks) sums[i] = e + carry;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_add(uint32_t* data, uint64_t n, const uint32_t* sums, const uint64_t* total,
                           uint32_t* out, uint64_t out_n) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = data[i] + sums[i / kScanBlock];
    if (i == 0) out[out_n] = uint32_t(*total);
}


// Validity of rows [bit, bit + 64) of a page (nr rows, nv non-NULL values),
// bit i = row bit + i; rows at or past nr read as 0. A page without NULLs is
// all valid and its bitmap is not read; otherwise the 9 bitmap bytes from
// byte bit / 8 on come from two aligned 8-byte loads (the page end is
// 8-aligned, so neither reaches past it; bytes past the page end read as 0).
__device__ inline uint64_t page_bits64(const uint8_t* page, uint32_t nr, uint32_t nv, uint32_t bit) {
    if (bit >= nr) return 0;
    const uint32_t n = nr - bit < 64 ? nr - bit : 64;
    const uint64_t m = n >= 64 ? ~0ull : ((1ull << n) - 1ull);
    if (nv == nr) return m;
    const uint32_t off = kPageSize - (nr + 7) / 8 + (bit >> 3); // first bitmap byte wanted
    const uint32_t al  = off & ~7u, o = off & 7u, sh = bit & 7u;
    const uint64_t w0  = *reinterpret_cast<const uint64_t*>(page + al);
    const uint64_t w1  = al + 8 < kPageSize ? *reinterpret_cast<const uint64_t*>(page + al + 8) : 0ull;
    const uint64_t lo  = o ? ((w0 >> (8 * o)) | (w1 << (64 - 8 * o))) : w0; // bytes off .. off + 7
    const uint64_t hi  = (w1 >> (8 * o)) & 0xffull;                         // byte off + 8
    const uint64_t w   = sh ? ((lo >> sh) | (hi << (64 - sh))) : lo;
    return w & m;
}

// Block index, four pages per wave, 16 lanes per page: the blocks whose first
// row lies in page p (16 at a time, one per lane) get their entries from the
// page's bitmap; a block's non-NULL offset is the page's valid rows before it,
// from an exclusive scan of the blocks' in-page valid counts over the page's
// 16 lanes. A block that runs past the page continues on page p + 1 when that
// page holds all its remaining rows (and is not a long-string page), else it
// is marked complex, as are blocks on long-string pages. Reads only headers
// and bitmaps.
__global__ __launch_bounds__(256) void k_block_index_pages(const uint8_t* pages, uint64_t npages, int32_t type,
                                                          const uint32_t* row_start, uint64_t nrows,
                                                          BlockIdx* blocks) {
    constexpr uint32_t kLanesPerPage = 16;
    const uint32_t     lane          = lane_id();
    const uint32_t     gl            = lane & (kLanesPerPage - 1);
    const uint32_t     glast         = lane | (kLanesPerPage - 1); // last lane of this page's group
    const uint64_t     wave          = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
    const uint64_t     p             = wave * (64 / kLanesPerPage) + lane / kLanesPerPage;
    uint32_t           s = 0, e = 0;
    if (p < npages) {
        s = row_start[p];
        e = row_start[p + 1];
    }
    // an empty page, or one whose rows all belong to a block started earlier, starts no block
    const uint64_t b_first = (uint64_t(s) + 63) / 64;
    const uint64_t b_last  = e > s ? (uint64_t(e) - 1) / 64 : 0;
    const bool     has     = e > s && b_first <= b_last;
    const uint8_t* page    = pages + (has ? p : 0) * kPageSize;
    uint32_t       nr_raw = 0, nv = 0;
    if (has) {
        nr_raw = *reinterpret_cast<const uint16_t*>(page);
        nv     = *reinterpret_cast<const uint16_t*>(page + 2);
    }
    const bool     is_long = type == T_VARCHAR && nr_raw >= 0xfffe;
    const uint32_t nr      = e - s;
    // valid rows of the page before its first block
    const uint32_t j00   = has ? uint32_t(b_first * 64 - s) : 0;
    uint32_t       carry = (has && !is_long && j00)
                               ? uint32_t(__popcll(page_bits64(page, nr, nv, 0) & ((1ull << j00) - 1ull)))
                               : 0;
    for (uint64_t k = 0;; k += kLanesPerPage) {
        const uint64_t b    = b_first + k + gl;
        const bool     have = has && b <= b_last;
        BlockIdx       x;
        x.page    = uint32_t(p);
        x.nn      = 0;
        x.split   = 64;
        x.complex = is_long ? 1 : 0;
        x.valid   = 0;
        uint32_t cnt = 0;
        if (have && !is_long) {
            const uint64_t r0 = b * 64;
            const uint64_t r1 = r0 + 64 < nrows ? r0 + 64 : nrows;
            const uint32_t j0 = uint32_t(r0 - s);
            x.valid           = page_bits64(page, nr, nv, j0);
            cnt               = uint32_t(__popcll(x.valid));
            if (r1 > e) {
                const uint32_t n1 = uint32_t(e - r0);
                const uint64_t q  = p + 1;
                bool           ok = q < npages && row_start[q + 1] >= r1;
                if (ok) {
                    const uint8_t* pq  = pages + q * kPageSize;
                    const uint32_t nrq = *reinterpret_cast<const uint16_t*>(pq);
                    const uint32_t nvq = *reinterpret_cast<const uint16_t*>(pq + 2);
                    ok                 = !(type == T_VARCHAR && nrq >= 0xfffe);
                    if (ok) {
                        const uint32_t rows_q = row_start[q + 1] - row_start[q];
                        const uint32_t n2     = uint32_t(r1 - e);
                        x.split               = uint8_t(n1);
                        x.valid |= (page_bits64(pq, rows_q, nvq, 0) & ((1ull << n2) - 1ull)) << n1;
                    }
                }
                if (!ok) x.complex = 1;
            }
        }
        // inclusive scan of cnt over the page's lanes
        uint32_t incl = cnt;
#pragma unroll
        for (uint32_t d = 1; d < kLanesPerPage; d <<= 1) {
            const uint32_t t = __shfl_up(incl, d, kLanesPerPage);
            if (gl >= d) incl += t;
        }
        x.nn = uint16_t(carry + incl - cnt);
        carry += __shfl(incl, glast, 64);
        if (have) blocks[b] = x;
        if (!__ballot(has && b_first + k + kLanesPerPage <= b_last)) break;
    }
}

size_t prepare_scratch_words(uint64_t npages) { return npages + (npages + kScanBlock - 1) / kScanBlock + 2; }

// In-place exclusive scan of n u32 values; `sums` holds ceil(n/1024) words.
__global__ void k_scan_finish(uint32_t* data, uint64_t n, const uint32_t* sums) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) data[i] += sums[i / kScanBlock];
}

size_t scan_scratch_words(uint64_t n) { return (n + kScanBlock - 1) / kScanBlock + 2; }

hipError_t launch_exclusive_scan(uint32_t* data, uint64_t n, uint32_t* scratch, uint64_t* total, hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint64_t nblocks = (n + kScanBlock - 1) / kScanBlock;
    k_scan_local<<<dim3(nblocks), dim3(kScanBlock), 0, s>>>(data, n, scratch);
    k_scan_sums<<<dim3(1), dim3(kScanBlock), 0, s>>>(scratch, nblocks, total);
    k_scan_finish<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(data, n, scratch);
    return hipGetLastError();
}

hipError_t launch_prepare_column(const uint8_t* pages, uint64_t npages, int32_t type, uint64_t nrows,
                                 uint32_t* row_start, BlockIdx* blocks, uint32_t* scratch, uint64_t* total_rows_dev,
                                 uint32_t* dense, hipStream_t s) {
    if (npages == 0) return hipSuccess;
    uint32_t* counts  = scratch;
    uint32_t* sums    = scratch + npages;
    uint64_t  nblocks = (npages + kScanBlock - 1) / kScanBlock;
    k_page_counts<<<dim3((npages + 255) / 256), dim3(256), 0, s>>>(pages, npages, type, counts, dense);
    k_scan_local<<<dim3(nblocks), dim3(kScanBlock), 0, s>>>(counts, npages, sums);
    k_scan_sums<<<dim3(1), dim3(kScanBlock), 0, s>>>(sums, nblocks, total_rows_dev);
    k_scan_add<<<dim3((npages + 255) / 256), dim3(256), 0, s>>>(counts, npages, sums, total_rows_dev, row_start,
                                                                npages);
    uint64_t nb = (nrows + 63) / 64;
    if (nb)
        k_block_index_pages<<<dim3((npages + 15) / 16), dim3(256), 0, s>>>(pages, npages, type, row_start, nrows,
                                                                           blocks);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// standalone decode: one wave per 1024-row morsel
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_decode(ScanCol c, uint64_t nrows, uint64_t* out) {
    uint64_t wave   = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) / 64;
    uint64_t nwaves = uint64_t(gridDim.x) * blockDim.x / 64;
    uint64_t nb     = (nrows + 63) / 64;
    for (uint64_t b = wave; b < nb; b += nwaves) {
        uint64_t row    = b * 64 + lane_id();
        bool     active = row < nrows;
        uint64_t v      = decode_block(c, b, active);
        if (active) out[row] = v;
    }
}

hipError_t launch_decode(const ScanCol& c, uint64_t nrows, uint64_t* out, hipStream_t s) {
    if (nrows == 0) return hipSuccess;
    uint64_t nb     = (nrows + 63) / 64;
    uint64_t blocks = (nb + 3) / 4;
    if (blocks > 8192) blocks = 8192;
    k_decode<<<dim3(blocks), dim3(256), 0, s>>>(c, nrows, out);
    return hipGetLastError();
}

// Plain streaming copy (16-byte vectors, grid-stride): the HBM bandwidth probe.
__global__ __launch_bounds__(256) void k_copy(uint4* dst, const uint4* src, uint64_t n) {
    uint64_t i      = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy(void* dst, const void* src, uint64_t n16, hipStream_t s) {
    k_copy<<<dim3(4096), dim3(256), 0, s>>>(static_cast<uint4*>(dst), static_cast<const uint4*>(src), n16);
    return hipGetLastError();
}

// Moves the valid pages of a result pool (directory word 3 == 1) into a pool
// of exactly `dst_pages` slots, each with its directory entry, so a result
// stops holding the slots its pool was sized for but did not use. One
// workgroup per source page; destination slots come from a counter (the page
// order is in the directory keys, not in the slot positions). A page beyond
// `dst_pages` (a count that does not match the directory) sets *fault.
__global__ __launch_bounds__(256) void k_compact_pages(const uint8_t* src, uint64_t src_pages, uint64_t reserved,
                                                       uint8_t* dst, uint64_t dst_pages,
                                                       unsigned long long* counter, unsigned long long* fault) {
    __shared__ uint64_t slot;
    const uint32_t*     sdir = reinterpret_cast<const uint32_t*>(src + src_pages * kPageSize);
    uint32_t*           ddir = reinterpret_cast<uint32_t*>(dst + dst_pages * kPageSize);
    for (uint64_t i = blockIdx.x; i < reserved; i += gridDim.x) {
        const uint4 e = *reinterpret_cast<const uint4*>(sdir + 4 * i);
        if (e.w != 1) continue; // uniform over the workgroup
        if (threadIdx.x == 0) slot = atomicAdd(counter, 1ull);
        __syncthreads();
        const uint64_t d = slot;
        __syncthreads(); // every thread has read `slot` before the next page's atomic
        if (d >= dst_pages) {
            if (threadIdx.x == 0) atomicExch(fault, 1ull);
            continue;
        }
        const uint4* s = reinterpret_cast<const uint4*>(src + i * kPageSize);
        uint4*       o = reinterpret_cast<uint4*>(dst + d * kPageSize);
        for (uint32_t k = threadIdx.x; k < kPageSize / 16; k += blockDim.x) o[k] = s[k];
        if (threadIdx.x == 0) *reinterpret_cast<uint4*>(ddir + 4 * d) = e;
    }
}

hipError_t launch_compact_pages(const uint8_t* src, uint64_t src_pages, uint64_t reserved, uint8_t* dst,
                                uint64_t dst_pages, unsigned long long* counter, unsigned long long* fault,
                                hipStream_t s) {
    if (reserved == 0) return hipSuccess;
    const uint64_t wgs = reserved < 8192 ? reserved : 8192;
    k_compact_pages<<<dim3(uint32_t(wgs)), dim3(256), 0, s>>>(src, src_pages, reserved, dst, dst_pages, counter,
                                                             fault);
    return hipGetLastError();
}

__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t v) {
    uint64_t i      = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

hipError_t launch_fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    k_fill_u64<<<dim3(blocks), dim3(256), 0, s>>>(p, n, v);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// hash-table insert (open addressing, linear probing, entries [key, attrs...])
// ---------------------------------------------------------------------------
// flags[0] = duplicate keys seen, flags[2] = an insert passed `walk_cap`
// slots holding its own key (0 = no cap) and stopped: the table is then
// incomplete and gets rebuilt with grouped duplicates.
__global__ void k_ht_insert(const uint64_t* stage, uint64_t n, uint32_t words, uint64_t* ht, uint32_t shift,
                            uint64_t mask, uint64_t* bloom, uint32_t bloom_log, uint32_t* flags, uint32_t walk_cap) {
    uint64_t i      = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) {
        const uint64_t* t   = stage + i * words;
        uint64_t        key = t[0];
        if (key == kEmpty) continue; // unused staging slot
        uint64_t        h   = hash_key(key);
        uint64_t        pos = h >> shift;
        bool            dup = false, stop = false;
        for (uint32_t same = 0;;) {
            unsigned long long prev =
                atomicCAS(reinterpret_cast<unsigned long long*>(ht + pos * words), kEmpty, key);
            if (prev == kEmpty) break;
            if (prev == key) {
                dup = true;
                if (walk_cap && ++same >= walk_cap) {
                    stop = true;
                    break;
                }
            }
            pos = (pos + 1) & mask;
        }
        if (dup) flags[0] = 1;
        if (stop) {
            flags[2] = 1;
            continue;
        }
        for (uint32_t w = 1; w < words; ++w) ht[pos * words + w] = t[w];
        if (bloom) {
            uint64_t h2 = hash_bloom(h);
            atomicOr(reinterpret_cast<unsigned long long*>(bloom + bloom_word(h2, bloom_log)), bloom_bits(h2));
        }
    }
}

// ---------------------------------------------------------------------------
// radix-partitioned build: the table is split into F regions of S slots (the
// top bits of a key's slot position select its region, and linear probing
// wraps inside the region), so each region can be built in LDS by one block
// and written out with full-line stores.
// ---------------------------------------------------------------------------

// Per-block histogram of region ids over a contiguous chunk of staged tuples;
// stored region-major (bh[region * blocks + block]) so that one exclusive scan
// yields every block's write offset in every region.
__global__ __launch_bounds__(1024) void k_part_hist(const uint64_t* stage, uint64_t n, uint32_t words,
                                                   uint32_t rshift, uint32_t nregions, uint64_t chunk,
                                                   uint32_t* bh) {
    extern __shared__ uint32_t hist[];
    for (uint32_t i = threadIdx.x; i < nregions; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    uint64_t begin = uint64_t(blockIdx.x) * chunk;
    uint64_t end   = begin + chunk < n ? begin + chunk : n;
    for (uint64_t i = begin + threadIdx.x; i < end; i += blockDim.x) {
        uint64_t key = stage[i * words];
        if (key == kEmpty) continue; // unused staging slot
        atomicAdd(&hist[rshift >= 64 ? 0u : uint32_t(hash_key(key) >> rshift)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nregions; i += blockDim.x) bh[uint64_t(i) * gridDim.x + blockIdx.x] = hist[i];
}

// Scatters the same chunk into region order using the scanned offsets.
__global__ __launch_bounds__(1024) void k_part_scatter(const uint64_t* stage, uint64_t n, uint32_t words,
                                                      uint32_t rshift, uint32_t nregions, uint64_t chunk,
                                                      const uint32_t* offs, uint64_t* out) {
    extern __shared__ uint32_t cur[];
    for (uint32_t i = threadIdx.x; i < nregions; i += blockDim.x) cur[i] = offs[uint64_t(i) * gridDim.x + blockIdx.x];
    __syncthreads();
    uint64_t begin = uint64_t(blockIdx.x) * chunk;
    uint64_t end   = begin + chunk < n ? begin + chunk : n;
    if (words == 2) {
        // two-word tuples: one 16-byte load and store each, four per thread in
        // flight before their position atomics
        constexpr int   U  = 4;
        const ulonglong2* src = reinterpret_cast<const ulonglong2*>(stage);
        ulonglong2*       dst = reinterpret_cast<ulonglong2*>(out);
        for (uint64_t i0 = begin + threadIdx.x; i0 < end; i0 += U * uint64_t(blockDim.x)) {
            ulonglong2 t[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t i = i0 + uint64_t(u) * blockDim.x;
                t[u]             = i < end ? src[i] : make_ulonglong2(kEmpty, kEmpty);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (t[u].x == kEmpty) continue; // unused staging slot or past the chunk
                const uint32_t r   = rshift >= 64 ? 0u : uint32_t(hash_key(t[u].x) >> rshift);
                const uint32_t pos = atomicAdd(&cur[r], 1u);
                dst[pos]           = t[u];
            }
        }
        return;
    }
    for (uint64_t i = begin + threadIdx.x; i < end; i += blockDim.x) {
        const uint64_t* t = stage + i * words;
        if (t[0] == kEmpty) continue; // unused staging slot
        uint32_t r   = rshift >= 64 ? 0u : uint32_t(hash_key(t[0]) >> rshift);
        uint32_t pos = atomicAdd(&cur[r], 1u);
        uint64_t*       d   = out + uint64_t(pos) * words;
        for (uint32_t w = 0; w < words; ++w) d[w] = t[w];
    }
}

// One block per region: inserts the region's tuples into an LDS table of
// `rslots` slots (linear probing inside the region), then stores the region.
// flags[0] = duplicate keys seen, flags[1] = a region was too full,
// flags[2] = an insert passed `walk_cap` slots holding its own key (0 = no
// cap) and the region was abandoned (heavy duplicates: rebuilt grouped).
__global__ __launch_bounds__(1024) void k_part_build(const uint64_t* parts, const uint32_t* offs, uint32_t blocks,
                                                    uint64_t n, uint32_t nregions, uint32_t words,
                                                    uint32_t rslots, uint32_t shift, uint64_t* ht,
                                                    uint64_t* bloom, uint32_t bloom_log, uint32_t* flags,
                                                    uint32_t walk_cap, const uint64_t* n_dev) {
    extern __shared__ __attribute__((aligned(16))) uint64_t tab[];
    uint32_t r     = blockIdx.x;
    uint64_t start = offs[uint64_t(r) * blocks];
    uint64_t end   = r + 1 < nregions ? offs[uint64_t(r + 1) * blocks] : (n_dev ? *n_dev : n);
    uint64_t cnt   = end - start;
    uint64_t* dst  = ht + uint64_t(r) * rslots * words;
    if (cnt >= rslots) { // a full region would leave probes without an empty slot
        if (threadIdx.x == 0) flags[1] = 1;
        return;
    }
    // 16-byte LDS and global accesses: rslots * words is even (rslots >= 16)
    const uint32_t n2 = rslots * words / 2;
    for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x)
        reinterpret_cast<ulonglong2*>(tab)[i] = make_ulonglong2(kEmpty, kEmpty);
    __syncthreads();
    bool     dup   = false, stop = false;
    uint32_t rmask = rslots - 1;
    for (uint64_t i = start + threadIdx.x; i < end && !stop; i += blockDim.x) {
        const uint64_t* t   = parts + i * words;
        uint64_t        key = t[0];
        uint64_t        h   = hash_key(key);
        uint32_t        pos = uint32_t(h >> shift) & rmask;
        for (uint32_t same = 0;;) {
            unsigned long long prev =
                atomicCAS(reinterpret_cast<unsigned long long*>(tab + uint64_t(pos) * words), kEmpty, key);
            if (prev == kEmpty) break;
            if (prev == key) {
                dup = true;
                if (walk_cap && ++same >= walk_cap) {
                    stop = true;
                    break;
                }
            }
            pos = (pos + 1) & rmask;
        }
        if (stop) break;
        for (uint32_t w = 1; w < words; ++w) tab[uint64_t(pos) * words + w] = t[w];
        if (bloom) {
            uint64_t h2 = hash_bloom(h);
            atomicOr(reinterpret_cast<unsigned long long*>(bloom + bloom_word(h2, bloom_log)), bloom_bits(h2));
        }
    }
    if (dup) flags[0] = 1;
    if (stop) flags[2] = 1;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n2; i += blockDim.x)
        reinterpret_cast<ulonglong2*>(dst)[i] = reinterpret_cast<const ulonglong2*>(tab)[i];
}

hipError_t launch_part_build(const uint64_t* stage, uint64_t n, uint64_t n_real, uint32_t words, uint64_t* parts,
                             uint32_t* bh, uint32_t blocks, uint64_t chunk, uint32_t* scan_scratch, uint64_t* total,
                             uint64_t* ht, uint32_t log_cap, uint32_t log_regions, uint64_t* bloom,
                             uint32_t bloom_log, uint32_t* flags, uint32_t threads, uint32_t walk_cap,
                             hipStream_t s) {
    if (n == 0) return hipSuccess;
    // The LDS histograms and region tables allow one or two workgroups per CU:
    // wider workgroups keep more waves (memory requests) in flight.
    if (threads != 256 && threads != 512 && threads != 1024) return hipErrorInvalidValue;
    uint32_t nregions = 1u << log_regions;
    uint32_t rslots   = 1u << (log_cap - log_regions);
    uint32_t rshift   = 64 - log_regions;
    size_t   hist_lds = size_t(nregions) * sizeof(uint32_t);
    k_part_hist<<<dim3(blocks), dim3(threads), hist_lds, s>>>(stage, n, words, rshift, nregions, chunk, bh);
    hipError_t e = launch_exclusive_scan(bh, uint64_t(nregions) * blocks, scan_scratch, total, s);
    if (e != hipSuccess) return e;
    k_part_scatter<<<dim3(blocks), dim3(threads), hist_lds, s>>>(stage, n, words, rshift, nregions, chunk, bh,
                                                                 parts);
    size_t tab_lds = size_t(rslots) * words * sizeof(uint64_t);
    if (tab_lds > 65536) {
        // raise the kernel's dynamic LDS limit once per device, to the most any region needs
        constexpr int             kMaxDev = 16;
        static std::atomic<size_t> lds_set[kMaxDev];
        int                        dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev ||
            lds_set[dev].load(std::memory_order_relaxed) < tab_lds) {
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_part_build),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(tab_lds));
            if (e != hipSuccess) return e;
            if (dev >= 0 && dev < kMaxDev) lds_set[dev].store(tab_lds, std::memory_order_relaxed);
        }
    }
    k_part_build<<<dim3(nregions), dim3(threads), tab_lds, s>>>(parts, bh, blocks, n_real, nregions, words, rslots,
                                                                64 - log_cap, ht, bloom, bloom_log, flags, walk_cap,
                                                                nullptr);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// partitioning straight from pages (hash table on a plain scan)
// ---------------------------------------------------------------------------
// Decodes the tuple words of a 64-row block; returns active and no NULL in a
// `notnull` word (the key, word 0, is always one).
__device__ inline bool scan_build_row(const ScanBuildArgs& a, uint64_t blk, bool active, uint64_t (&v)[kMaxWords]) {
    bool valid = active && (!a.sel || ((load_invariant_u64(a.sel + blk) >> lane_id()) & 1));
#pragma unroll
    for (int w = 0; w < kMaxWords; ++w) {
        v[w] = kNull;
        if (uint32_t(w) >= a.words) continue;
        v[w] = decode_block(a.col[w], blk, active);
        if ((a.notnull >> w) & 1u) valid = valid && v[w] != kNull;
    }
    return valid;
}

__global__ __launch_bounds__(1024) void k_scan_hist(ScanBuildArgs a, uint32_t* bh) {
    extern __shared__ uint32_t hist[];
    for (uint32_t i = threadIdx.x; i < a.nregions; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const uint64_t nb     = (a.nrows + 63) / 64;
    const uint64_t b0     = uint64_t(blockIdx.x) * a.blocks_per_wg;
    const uint64_t b1     = b0 + a.blocks_per_wg < nb ? b0 + a.blocks_per_wg : nb;
    const uint32_t nwaves = blockDim.x / 64;
    for (uint64_t b = b0 + wave_in_block(); b < b1; b += nwaves) {
        const bool     active = b * 64 + lane_id() < a.nrows;
        const uint64_t key    = decode_block(a.col[0], b, active);
        bool           valid  = active && key != kNull;
        if (a.sel) valid = valid && ((load_invariant_u64(a.sel + b) >> lane_id()) & 1);
        for (uint32_t w = 1; w < a.words; ++w) // NULL checks of the other join columns
            if ((a.notnull >> w) & 1u) valid = valid && decode_block(a.col[w], b, active) != kNull;
        if (valid) atomicAdd(&hist[a.rshift >= 64 ? 0u : uint32_t(hash_key(key) >> a.rshift)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < a.nregions; i += blockDim.x) bh[uint64_t(i) * gridDim.x + blockIdx.x] = hist[i];
}

__global__ __launch_bounds__(1024) void k_scan_scatter(ScanBuildArgs a, const uint32_t* offs, uint64_t* out) {
    extern __shared__ uint32_t cur[];
    for (uint32_t i = threadIdx.x; i < a.nregions; i += blockDim.x) cur[i] = offs[uint64_t(i) * gridDim.x + blockIdx.x];
    __syncthreads();
    const uint64_t nb     = (a.nrows + 63) / 64;
    const uint64_t b0     = uint64_t(blockIdx.x) * a.blocks_per_wg;
    const uint64_t b1     = b0 + a.blocks_per_wg < nb ? b0 + a.blocks_per_wg : nb;
    const uint32_t nwaves = blockDim.x / 64;
    for (uint64_t b = b0 + wave_in_block(); b < b1; b += nwaves) {
        uint64_t   v[kMaxWords];
        const bool valid = scan_build_row(a, b, b * 64 + lane_id() < a.nrows, v);
        if (!valid) continue;
        const uint32_t r   = a.rshift >= 64 ? 0u : uint32_t(hash_key(v[0]) >> a.rshift);
        const uint32_t pos = atomicAdd(&cur[r], 1u);
        if (a.words == 2) {
            reinterpret_cast<ulonglong2*>(out)[pos] = make_ulonglong2(v[0], v[1]);
        } else {
            uint64_t* d = out + uint64_t(pos) * a.words;
#pragma unroll
            for (int w = 0; w < kMaxWords; ++w)
                if (uint32_t(w) < a.words) d[w] = v[w];
        }
    }
}

// Hash table on a plain scan inserted straight from its pages, for tables
// small enough that random inserts stay in the caches (larger ones take the
// partitioned build): each lane decodes its row's tuple words (a NULL join
// column drops the row) and inserts them as k_ht_insert does, with the same
// flags; `count` receives the rows inserted. One 64-row block per wave step.
__global__ __launch_bounds__(256) void k_scan_insert(ScanBuildArgs a, uint64_t* ht, uint32_t shift, uint64_t mask,
                                                     uint64_t* bloom, uint32_t bloom_log, uint32_t* flags,
                                                     uint32_t walk_cap, unsigned long long* count) {
    const uint64_t nb     = (a.nrows + 63) / 64;
    const uint32_t wpb    = blockDim.x / 64;
    const uint64_t nwaves = uint64_t(gridDim.x) * wpb;
    const uint32_t words  = a.words;
    uint32_t       mine   = 0;
    for (uint64_t b = uint64_t(blockIdx.x) * wpb + wave_in_block(); b < nb; b += nwaves) {
        uint64_t   v[kMaxWords];
        const bool valid = scan_build_row(a, b, b * 64 + lane_id() < a.nrows, v);
        if (!valid) continue;
        const uint64_t key = v[0];
        const uint64_t h   = hash_key(key);
        uint64_t       pos = h >> shift;
        bool           dup = false, stop = false;
        for (uint32_t same = 0;;) {
            const unsigned long long prev =
                atomicCAS(reinterpret_cast<unsigned long long*>(ht + pos * words), kEmpty, key);
            if (prev == kEmpty) break;
            if (prev == key) {
                dup = true;
                if (walk_cap && ++same >= walk_cap) {
                    stop = true;
                    break;
                }
            }
            pos = (pos + 1) & mask;
        }
        if (dup) flags[0] = 1;
        if (stop) {
            flags[2] = 1;
            continue;
        }
#pragma unroll
        for (int w = 1; w < kMaxWords; ++w)
            if (uint32_t(w) < words) ht[pos * words + w] = v[w];
        if (bloom) {
            const uint64_t h2 = hash_bloom(h);
            atomicOr(reinterpret_cast<unsigned long long*>(bloom + bloom_word(h2, bloom_log)), bloom_bits(h2));
        }
        ++mine;
    }
    const uint32_t total = wave_sum(mine);
    if (lead_lane() && total) atomicAdd(count, (unsigned long long)total);
}

hipError_t launch_scan_insert(const ScanBuildArgs& a, uint64_t* ht, uint32_t log_cap, uint64_t* bloom,
                              uint32_t bloom_log, uint32_t* flags, uint32_t walk_cap, unsigned long long* count,
                              hipStream_t s) {
    if (a.nrows == 0) return hipSuccess;
    if (a.words == 0 || a.words > uint32_t(kMaxWords) || log_cap < 1 || log_cap > 40) return hipErrorInvalidValue;
    const uint64_t nb     = (a.nrows + 63) / 64;
    uint64_t       blocks = (nb + 3) / 4;
    if (blocks > 4096) blocks = 4096;
    k_scan_insert<<<dim3(blocks), dim3(256), 0, s>>>(a, ht, 64 - log_cap, (uint64_t(1) << log_cap) - 1, bloom,
                                                     bloom_log, flags, walk_cap, count);
    return hipGetLastError();
}

hipError_t launch_scan_partition(const ScanBuildArgs& a, uint32_t blocks, uint32_t* bh, uint32_t* scan_scratch,
                                 uint64_t* total, uint64_t* parts, uint32_t threads, hipStream_t s) {
    if (threads != 256 && threads != 512 && threads != 1024) return hipErrorInvalidValue;
    if (a.words == 0 || a.words > uint32_t(kMaxWords)) return hipErrorInvalidValue;
    if (uint64_t(blocks) * a.blocks_per_wg * 64 < a.nrows) return hipErrorInvalidValue; // every row covered
    const size_t hist_lds = size_t(a.nregions) * sizeof(uint32_t);
    k_scan_hist<<<dim3(blocks), dim3(threads), hist_lds, s>>>(a, bh);
    hipError_t e = launch_exclusive_scan(bh, uint64_t(a.nregions) * blocks, scan_scratch, total, s);
    if (e != hipSuccess) return e;
    k_scan_scatter<<<dim3(blocks), dim3(threads), hist_lds, s>>>(a, bh, parts);
    return hipGetLastError();
}

hipError_t launch_region_build(const uint64_t* parts, const uint32_t* bh, uint32_t blocks, const uint64_t* n_dev,
                               uint32_t words, uint64_t* ht, uint32_t log_cap, uint32_t log_regions,
                               uint64_t* bloom, uint32_t bloom_log, uint32_t* flags, uint32_t threads,
                               uint32_t walk_cap, hipStream_t s) {
    const uint32_t nregions = 1u << log_regions;
    const uint32_t rslots   = 1u << (log_cap - log_regions);
    const size_t   tab_lds  = size_t(rslots) * words * sizeof(uint64_t);
    if (tab_lds > 65536) {
        constexpr int              kMaxDev = 16;
        static std::atomic<size_t> lds_set[kMaxDev];
        int                        dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev ||
            lds_set[dev].load(std::memory_order_relaxed) < tab_lds) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_part_build),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, int(tab_lds));
            if (e != hipSuccess) return e;
            if (dev >= 0 && dev < kMaxDev) lds_set[dev].store(tab_lds, std::memory_order_relaxed);
        }
    }
    k_part_build<<<dim3(nregions), dim3(threads), tab_lds, s>>>(parts, bh, blocks, 0, nregions, words, rslots,
                                                                64 - log_cap, ht, bloom, bloom_log, flags, walk_cap,
                                                                n_dev);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// grouped build: a table whose keys repeat many times keeps each distinct key
// once, as [key, start << 24 | count], and the key's tuples (words 1.. of the
// staged tuple) contiguously in a payload array. Probe sequences then stay
// as short as with unique keys, and a probe that matches walks its group.
// ---------------------------------------------------------------------------
constexpr uint32_t kGroupCountBits = 24;
constexpr uint32_t kGroupPosBits   = 64 - kGroupCountBits; // payload row index bits

// Exclusive scan of cnt[0..n) into out[0..n) by one workgroup; `tmp` holds
// blockDim.x words of LDS.
__device__ inline void block_exclusive_scan(const uint32_t* cnt, uint32_t* out, uint32_t n, uint32_t* tmp) {
    const uint32_t T = blockDim.x, t = threadIdx.x;
    const uint32_t per = (n + T - 1) / T;
    const uint32_t b0 = t * per, b1 = b0 + per < n ? b0 + per : n;
    uint32_t       local = 0;
    for (uint32_t j = b0; j < b1; ++j) local += cnt[j];
    tmp[t] = local;
    __syncthreads();
    for (uint32_t d = 1; d < T; d <<= 1) { // Hillis-Steele inclusive scan
        const uint32_t v = t >= d ? tmp[t - d] : 0u;
        __syncthreads();
        tmp[t] += v;
        __syncthreads();
    }
    uint32_t run = tmp[t] - local;
    for (uint32_t j = b0; j < b1; ++j) {
        const uint32_t c = cnt[j];
        out[j]           = run;
        run += c;
    }
    __syncthreads();
}

// One workgroup per region of `rslots` slots (tuples already in region order
// in `parts`, region r starting at offs[r * blocks]). LDS: keys, counts and
// cursors of the region's slots, plus the scan's scratch.
// flags[1] = a region's distinct keys did not fit, flags[3] = a group
// exceeded the count field.
__global__ __launch_bounds__(1024) void k_group_build(const uint64_t* parts, const uint32_t* offs, uint32_t blocks,
                                                     uint64_t n, uint32_t nregions, uint32_t words, uint32_t rslots,
                                                     uint32_t shift, uint64_t* ht, uint64_t* payload,
                                                     uint64_t* bloom, uint32_t bloom_log, uint32_t* flags) {
    extern __shared__ uint64_t lds[];
    uint64_t*      keys  = lds;
    uint32_t*      cnt   = reinterpret_cast<uint32_t*>(lds + rslots);
    uint32_t*      cur   = cnt + rslots;
    uint32_t*      tmp   = cur + rslots;
    const uint32_t r     = blockIdx.x;
    const uint64_t start = offs[uint64_t(r) * blocks];
    const uint64_t end   = r + 1 < nregions ? offs[uint64_t(r + 1) * blocks] : n;
    const uint32_t rmask = rslots - 1;
    const uint32_t pw    = words - 1;
    for (uint32_t i = threadIdx.x; i < rslots; i += blockDim.x) {
        keys[i] = kEmpty;
        cnt[i]  = 0;
    }
    __syncthreads();
    bool over = false;
    for (uint64_t i = start + threadIdx.x; i < end; i += blockDim.x) {
        const uint64_t key = parts[i * words];
        const uint64_t h   = hash_key(key);
        uint32_t       pos = uint32_t(h >> shift) & rmask;
        uint32_t       steps = 0;
        for (;; ++steps) {
            if (steps >= rslots) {
                over = true;
                break;
            }
            const unsigned long long prev =
                atomicCAS(reinterpret_cast<unsigned long long*>(keys + pos), kEmpty, key);
            if (prev == kEmpty || prev == key) {
                atomicAdd(cnt + pos, 1u);
                break;
            }
            pos = (pos + 1) & rmask;
        }
        if (bloom && steps < rslots) {
            const uint64_t h2 = hash_bloom(h);
            atomicOr(reinterpret_cast<unsigned long long*>(bloom + bloom_word(h2, bloom_log)), bloom_bits(h2));
        }
    }
    if (over) flags[1] = 1;
    __syncthreads();
    block_exclusive_scan(cnt, cur, rslots, tmp);
    uint64_t* dst = ht + uint64_t(r) * rslots * 2;
    bool      big = false;
    for (uint32_t i = threadIdx.x; i < rslots; i += blockDim.x) {
        const uint64_t k = keys[i];
        big              = big || cnt[i] >= (1u << kGroupCountBits);
        dst[2 * uint64_t(i)]     = k;
        dst[2 * uint64_t(i) + 1] = k == kEmpty ? 0ull : ((start + cur[i]) << kGroupCountBits) | cnt[i];
    }
    if (big) flags[3] = 1;
    __syncthreads();
    if (pw == 0 || over) return; // key-only tuples: the counts are the table
    for (uint64_t i = start + threadIdx.x; i < end; i += blockDim.x) {
        const uint64_t* t   = parts + i * words;
        const uint64_t  key = t[0];
        uint32_t        pos = uint32_t(hash_key(key) >> shift) & rmask;
        while (keys[pos] != key) pos = (pos + 1) & rmask; // inserted above: always found
        const uint32_t q = atomicAdd(cur + pos, 1u);
        uint64_t*      d = payload + (start + q) * pw;
        for (uint32_t w = 0; w < pw; ++w) d[w] = t[1 + w];
    }
}

size_t group_build_lds(uint32_t rslots, uint32_t threads) {
    return size_t(rslots) * 16 + size_t(threads) * 4;
}

hipError_t launch_group_build(const uint64_t* stage, uint64_t n, uint64_t n_real, uint32_t words, uint64_t* parts,
                              uint32_t* bh, uint32_t blocks, uint64_t chunk, uint32_t* scan_scratch,
                              uint64_t* total, uint64_t* ht, uint64_t* payload, uint32_t log_cap,
                              uint32_t log_regions, uint64_t* bloom, uint32_t bloom_log, uint32_t* flags,
                              uint32_t threads, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (threads != 256 && threads != 512 && threads != 1024) return hipErrorInvalidValue;
    const uint32_t nregions = 1u << log_regions;
    const uint32_t rslots   = 1u << (log_cap - log_regions);
    const uint32_t rshift   = 64 - log_regions; // 64: one region (index 0)
    const size_t   hist_lds = size_t(nregions) * sizeof(uint32_t);
    k_part_hist<<<dim3(blocks), dim3(threads), hist_lds, s>>>(stage, n, words, rshift, nregions, chunk, bh);
    hipError_t e = launch_exclusive_scan(bh, uint64_t(nregions) * blocks, scan_scratch, total, s);
    if (e != hipSuccess) return e;
    k_part_scatter<<<dim3(blocks), dim3(threads), hist_lds, s>>>(stage, n, words, rshift, nregions, chunk, bh,
                                                                 parts);
    const size_t lds = group_build_lds(rslots, threads);
    if (lds > 65536) {
        constexpr int              kMaxDev = 16;
        static std::atomic<size_t> lds_set[kMaxDev];
        int                        dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev ||
            lds_set[dev].load(std::memory_order_relaxed) < lds) {
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_group_build),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
            if (e != hipSuccess) return e;
            if (dev >= 0 && dev < kMaxDev) lds_set[dev].store(lds, std::memory_order_relaxed);
        }
    }
    k_group_build<<<dim3(nregions), dim3(threads), lds, s>>>(parts, bh, blocks, n_real, nregions, words, rslots,
                                                             64 - log_cap, ht, payload, bloom, bloom_log, flags);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// radix-partitioned probe side: the scanned columns of the driving table are
// decoded once and scattered by the top bits of the probe key's hash, so that
// each partition probes one contiguous, cache-sized range of the hash table.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_radix_hist(RadixArgs a) {
    extern __shared__ uint32_t hist[];
    for (uint32_t i = threadIdx.x; i < a.nparts; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    uint64_t nb    = (a.nrows + 63) / 64;
    uint64_t b0    = uint64_t(blockIdx.x) * a.blocks_per_wg;
    uint64_t b1    = b0 + a.blocks_per_wg < nb ? b0 + a.blocks_per_wg : nb;
    uint32_t wave  = wave_in_block();
    for (uint64_t b = b0 + wave; b < b1; b += blockDim.x / 64) {
        uint64_t row    = b * 64 + lane_id();
        bool     active = row < a.nrows;
        uint64_t key    = decode_block(a.scan[0], b, active);
        bool     ok     = active && key != kNull;
        for (uint32_t c = 1; c < a.nscan && ok; ++c)
            if ((a.not_null >> c) & 1u) ok = decode_block(a.scan[c], b, active) != kNull;
        if (ok) atomicAdd(&hist[uint32_t(hash_key(key) >> a.rshift)], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < a.nparts; i += blockDim.x) a.bh[uint64_t(i) * gridDim.x + blockIdx.x] = hist[i];
}

__global__ __launch_bounds__(256) void k_radix_scatter(RadixArgs a) {
    extern __shared__ uint32_t cur[];
    for (uint32_t i = threadIdx.x; i < a.nparts; i += blockDim.x) cur[i] = a.bh[uint64_t(i) * gridDim.x + blockIdx.x];
    __syncthreads();
    uint64_t nb   = (a.nrows + 63) / 64;
    uint64_t b0   = uint64_t(blockIdx.x) * a.blocks_per_wg;
    uint64_t b1   = b0 + a.blocks_per_wg < nb ? b0 + a.blocks_per_wg : nb;
    uint32_t wave = wave_in_block();
    for (uint64_t b = b0 + wave; b < b1; b += blockDim.x / 64) {
        uint64_t row    = b * 64 + lane_id();
        bool     active = row < a.nrows;
        uint64_t v[kMaxScan];
        bool     ok = active;
        for (uint32_t c = 0; c < a.nscan; ++c) {
            v[c] = decode_block(a.scan[c], b, active);
            if ((a.not_null >> c) & 1u) ok = ok && v[c] != kNull;
        }
        if (ok) {
            uint32_t  pos = atomicAdd(&cur[uint32_t(hash_key(v[0]) >> a.rshift)], 1u);
            uint64_t* d   = a.out + uint64_t(pos) * a.nscan;
            for (uint32_t c = 0; c < a.nscan; ++c) d[c] = v[c];
        }
    }
}

hipError_t launch_radix_partition(const RadixArgs& a, uint32_t wgs, uint32_t* scan_scratch, uint64_t* total,
                                  hipStream_t s) {
    size_t lds = size_t(a.nparts) * sizeof(uint32_t);
    k_radix_hist<<<dim3(wgs), dim3(256), lds, s>>>(a);
    hipError_t e = launch_exclusive_scan(a.bh, uint64_t(a.nparts) * wgs, scan_scratch, total, s);
    if (e != hipSuccess) return e;
    k_radix_scatter<<<dim3(wgs), dim3(256), lds, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_ht_insert(const uint64_t* stage, uint64_t n, uint32_t words, uint64_t* ht, uint32_t shift,
                            uint64_t mask, uint64_t* bloom, uint32_t bloom_log, uint32_t* flags, uint32_t walk_cap,
                            hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 32768) blocks = 32768;
    k_ht_insert<<<dim3(blocks), dim3(256), 0, s>>>(stage, n, words, ht, shift, mask, bloom, bloom_log, flags, walk_cap);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// fused pipeline
// ---------------------------------------------------------------------------
constexpr uint32_t kWavesPerBlock = 4;
uint32_t           pipeline_waves_per_block() { return kWavesPerBlock; }

template <int NA>
__device__ inline uint64_t sget(const uint64_t (&s)[NA], uint32_t i) {
    uint64_t r = 0;
#pragma unroll
    for (int k = 0; k < NA; ++k) r = (uint32_t(k) == i) ? s[k] : r;
    return r;
}
template <int NA>
__device__ inline void sset(uint64_t (&s)[NA], uint32_t i, uint64_t v) {
#pragma unroll
    for (int k = 0; k < NA; ++k) s[k] = (uint32_t(k) == i) ? v : s[k];
}

// Per-wave state shared by all output columns (wave-uniform values in LDS).
// Pool pages and staging slots are reserved in chunks with one atomic per
// chunk, so no counter sees more than one atomic per chunk of work.
struct WaveState {
    uint64_t rows;       // rows emitted (output) / tuples staged
    uint64_t next;       // next reserved page index / staging slot
    uint64_t end;        // end of the reserved range
    uint64_t stop;       // a reservation passed the abort bound: the result cannot fit, the wave stops
};

// Per-wave state of one output column's open page (wave-uniform values).
struct OutState {
    uint32_t page;   // pool index of the open page, kNoPage if none
    uint32_t nr;     // rows in the page
    uint32_t nv;     // non-NULL values in the page
    int32_t  avail;  // remaining bit budget
    uint32_t chars;  // VARCHAR: characters staged for the page
    uint32_t seq;    // pages emitted by this wave for this column
    uint32_t bad;    // pool exhausted: keep counting, write nothing
    uint32_t npages; // pages written (flushed to the counters at the end)
    uint64_t nnz;    // non-NULL values written (flushed at the end)
    uint64_t pad;
};
constexpr uint32_t kNoPage    = 0xffffffffu;

__device__ inline int32_t fixed_budget(int32_t type) { return 8 * (int32_t(kPageSize) - 4 - (type == T_INT32 ? 0 : 4)); }

// Wave-uniform state: every lane holds (or writes) the same values.
__device__ inline void page_open(const PipeArgs& a, WaveState& ws, OutState& st, uint64_t* bits, int32_t budget) {
    uint64_t next = ws.next, end = ws.end;
    if (next == end) {
        const uint64_t           sbase = a.static_chunk ? uint64_t(gridDim.x) * kWavesPerBlock * a.page_chunk : 0;
        const unsigned long long base  = sbase + wave_fetch_add(&a.counters->pages, (unsigned long long)a.page_chunk);
        next   = base;
        ws.end = base + a.page_chunk;
        if (base >= a.abort_pages) ws.stop = 1;
    }
    uint64_t idx = next;
    ws.next      = next + 1;
    st.bad = idx >= a.pool_pages ? 1u : 0u;
    if (st.bad && lead_lane()) atomicExch(&a.counters->overflow, 1ull);
    st.page  = st.bad ? 0u : uint32_t(idx);
    st.nr    = 0;
    st.nv    = 0;
    st.chars = 0;
    st.avail = budget;
    bits[lane_id()] = 0;
}

// Records page `page` in the directory slot of the same index.
__device__ inline void dir_put(const PipeArgs& a, uint32_t page, uint32_t col, uint32_t stream, uint32_t seq) {
    uint4 e = make_uint4(col, stream, seq, 1u);
    *reinterpret_cast<uint4*>(a.dir + 4 * uint64_t(page)) = e;
}

// Writes header, (staged characters) and bitmap of the open page and records it.
__device__ inline void page_close(const PipeArgs& a, OutState& st, uint64_t* bits, uint32_t col, uint32_t stream,
                                  const uint8_t* chars) {
    if (!st.bad) {
        uint8_t* page = a.pool + uint64_t(st.page) * kPageSize;
        if (chars) {
            uint8_t* dst = page + 4 + 2 * st.nv;
            for (uint32_t i = lane_id(); i < st.chars; i += 64) dst[i] = chars[i];
        }
        uint32_t nb   = (st.nr + 7) / 8;
        uint8_t* bm   = page + kPageSize - nb;
        uint64_t word = bits[lane_id()];
        uint32_t b0   = lane_id() * 8;
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b)
            if (b0 + b < nb) bm[b0 + b] = uint8_t(word >> (8 * b));
        if (lead_lane()) {
            *reinterpret_cast<uint32_t*>(page) = st.nr | (st.nv << 16);
            dir_put(a, st.page, col, stream, st.seq);
        }
        st.npages++;
        st.nnz += st.nv;
    }
    st.seq++;
    st.page = kNoPage;
}

// Appends the rows of lanes with `have` (rank = position in the batch) to a
// fixed-width column. Page packing follows the reference's bit budget
// (TableTarget.cpp:111-182): each row costs 1 bit, a non-NULL row the value
// width on top; the values are written straight to the page, the bitmap is
// held by the lanes (lane i holds rows 64i..64i+63) until the page closes.
__device__ inline void append_fixed(const PipeArgs& a, WaveState& ws, OutState& st, uint64_t* bits, int32_t type,
                                    uint64_t v, bool have, uint32_t rank, uint32_t cnt, uint32_t col,
                                    uint32_t stream) {
    const uint32_t w      = type == T_INT32 ? 4 : 8;
    const uint32_t off    = type == T_INT32 ? 4 : 8;
    const int32_t  budget = fixed_budget(type);
    const bool     valid  = have && v != kNull;
    const uint64_t vall   = __ballot(valid);
    const uint64_t le     = mask_le();
    uint32_t       start  = 0;
    while (start < cnt) {
        if (st.page == kNoPage) page_open(a, ws, st, bits, budget);
        // bit cost of the batch up to this lane: 1 per row + width per value
        bool     part = have && rank >= start;
        uint64_t pm   = __ballot(part);
        uint32_t cum  = __popcll(pm & le) + 8 * w * __popcll(pm & vall & le);
        bool     fits = part && int32_t(cum) <= st.avail && (st.nr + (rank - start)) < kMaxPageRows;
        uint64_t fm   = __ballot(fits);
        uint32_t k    = __popcll(fm);
        uint64_t vm   = fm & vall;
        if (fits && valid && !st.bad) {
            uint32_t vr  = __popcll(vm & mask_lt());
            uint8_t* dst = a.pool + uint64_t(st.page) * kPageSize + off + uint64_t(st.nv + vr) * w;
            if (w == 4)
                *reinterpret_cast<uint32_t*>(dst) = uint32_t(v);
            else
                *reinterpret_cast<uint64_t*>(dst) = v;
            uint32_t bit = st.nr + (rank - start);
            atomicOr(reinterpret_cast<unsigned long long*>(&bits[bit >> 6]), 1ull << (bit & 63));
        }
        uint32_t nvk = __popcll(vm);
        st.nr += k;
        st.nv += nvk;
        st.avail -= int32_t(k + 8 * w * nvk);
        start += k;
        if (start < cnt || st.nr >= kMaxPageRows) page_close(a, st, bits, col, stream, nullptr);
    }
}

// Resolves a string handle to its bytes (short strings) or flags a long string.
__device__ inline const uint8_t* resolve_string(const PipeArgs& a, uint64_t h, uint32_t& len, bool& is_long) {
    const uint8_t* page = a.col_base[handle_col(h)] + handle_page(h) * kPageSize;
    uint32_t       idx  = handle_idx(h);
    if (idx == 0xffff) {
        is_long = true;
        len     = 0;
        return page;
    }
    is_long               = false;
    uint32_t        nv    = *reinterpret_cast<const uint16_t*>(page + 2);
    const uint16_t* offs  = reinterpret_cast<const uint16_t*>(page + 4);
    uint32_t        end   = offs[idx];
    uint32_t        begin = idx ? offs[idx - 1] : 0;
    len                   = end - begin;
    return page + 4 + 2 * nv + begin;
}

// VARCHAR counterpart of append_fixed (reference TableTarget.cpp:183-391):
// NULL costs 1 bit, a short string 17 + 8*len bits; the end offsets go
// straight to the page, the characters to the wave's 8 KiB staging area until
// the page closes; a long string closes the page and is copied page by page.
__device__ void append_varchar(const PipeArgs& a, WaveState& ws, OutState& st, uint64_t* bits, uint8_t* stage,
                               uint64_t v, bool have, uint32_t rank, uint32_t cnt, uint32_t col, uint32_t stream) {
    const int32_t  budget  = 8 * (int32_t(kPageSize) - 4);
    const bool     valid   = have && v != kNull;
    uint32_t       len     = 0;
    bool           is_long = false;
    const uint8_t* src     = nullptr;
    if (valid) src = resolve_string(a, v, len, is_long);
    uint32_t start = 0;
    while (start < cnt) {
        uint64_t lm      = __ballot(have && is_long && rank >= start);
        uint32_t seg_end = cnt;
        if (lm) seg_end = __shfl(rank, __ffsll((unsigned long long)lm) - 1, 64);
        while (start < seg_end) {
            if (st.page == kNoPage) page_open(a, ws, st, bits, budget);
            bool     part = have && rank >= start && rank < seg_end;
            uint32_t cost = part ? (valid ? 17u + 8u * len : 1u) : 0u;
            uint32_t cum  = wave_incl_scan(cost);
            bool     fits = part && int32_t(cum) <= st.avail && (st.nr + (rank - start)) < kMaxPageRows;
            uint64_t fm   = __ballot(fits);
            uint32_t k    = __popcll(fm);
            uint64_t vm   = __ballot(fits && valid);
            uint32_t clen = (fits && valid) ? len : 0u;
            uint32_t cinc = wave_incl_scan(clen);
            if (fits && valid && !st.bad) {
                uint32_t vr     = __popcll(vm & mask_lt());
                uint8_t* pg     = a.pool + uint64_t(st.page) * kPageSize;
                uint32_t endoff = st.chars + cinc;
                *reinterpret_cast<uint16_t*>(pg + 4 + 2 * (st.nv + vr)) = uint16_t(endoff);
                uint8_t* dst = stage + (endoff - len);
                for (uint32_t i = 0; i < len; ++i) dst[i] = src[i];
            }
            uint64_t pat = wave_or((fits && valid) ? (1ull << (rank - start)) : 0ull);
            uint32_t wi = st.nr >> 6, sh = st.nr & 63;
            if (lane_id() == wi) bits[wi] |= pat << sh;
            if (sh && lane_id() == wi + 1) bits[wi + 1] |= pat >> (64 - sh);
            uint32_t used  = k ? __shfl(cum, highest_lane(fm), 64) : 0;
            uint32_t added = k ? __shfl(cinc, highest_lane(fm), 64) : 0;
            st.nr += k;
            st.nv += __popcll(vm);
            st.avail -= int32_t(used);
            st.chars += added;
            start += k;
            if (start < seg_end || st.nr >= kMaxPageRows) page_close(a, st, bits, col, stream, stage);
        }
        if (start < cnt) {
            // the row of rank `start` is a long string: copy its pages
            if (st.page != kNoPage) page_close(a, st, bits, col, stream, stage);
            uint32_t       src_lane = __ffsll((unsigned long long)lm) - 1;
            uint64_t       h        = __shfl(v, src_lane, 64);
            uint32_t       cid      = handle_col(h);
            uint64_t       p        = handle_page(h);
            const uint8_t* base     = a.col_base[cid];
            uint32_t       np       = a.col_npages[cid];
            do {
                page_open(a, ws, st, bits, budget);
                if (!st.bad) {
                    const uint64_t* s64 = reinterpret_cast<const uint64_t*>(base + p * kPageSize);
                    uint64_t*       d64 = reinterpret_cast<uint64_t*>(a.pool + uint64_t(st.page) * kPageSize);
                    for (uint32_t i = lane_id(); i < kPageSize / 8; i += 64) d64[i] = s64[i];
                    if (lead_lane()) dir_put(a, st.page, col, stream, st.seq);
                    st.npages++;
                }
                st.seq++;
                st.page = kNoPage;
                ++p;
            } while (p < np && *reinterpret_cast<const uint16_t*>(base + p * kPageSize) == 0xfffe);
            st.nnz += 1;
            start += 1;
        }
    }
}

// Stages the tuples of lanes with `have` into the wave's reserved slots;
// a batch that does not fit the rest of the reservation spills into a fresh one.
__device__ inline void stage_put(const PipeArgs& a, WaveState& ws, bool have, uint32_t rank, uint32_t cnt,
                                 const uint64_t* words_src, uint32_t nwords) {
    uint64_t next = ws.next, end = ws.end;
    uint64_t rem  = end - next;
    uint64_t base2 = 0;
    if (cnt > rem) {
        const uint64_t sbase = a.static_chunk ? uint64_t(gridDim.x) * kWavesPerBlock * a.stage_chunk : 0;
        base2                = sbase + wave_fetch_add(&a.counters->pages, (unsigned long long)a.stage_chunk);
        if (base2 >= a.abort_pages) ws.stop = 1;
        if (base2 + a.stage_chunk > a.stage_cap && lead_lane()) atomicExch(&a.counters->overflow, 1ull);
    }
    if (have) {
        uint64_t pos = rank < rem ? next + rank : base2 + (rank - rem);
        if (pos < a.stage_cap) {
            uint64_t* dst = a.stage + pos * a.stage_words;
            // fixed trip count: the caller's word array stays in registers
#pragma unroll
            for (int w = 0; w < kMaxWords; ++w)
                if (uint32_t(w) < nwords) dst[w] = words_src[w];
        }
    }
    if (cnt > rem) {
        ws.next = base2 + (cnt - rem);
        ws.end  = base2 + a.stage_chunk;
    } else {
        ws.next = next + cnt;
    }
}

// Closes a column's open page at the end of a wave and flushes its counters.
__device__ inline void finish_column(const PipeArgs& a, OutState& st, uint64_t* bits, uint32_t c, uint32_t stream,
                                     const uint8_t* stg) {
    if (st.page != kNoPage && st.nr > 0)
        page_close(a, st, bits, c, stream, stg);
    else if (st.page != kNoPage && !st.bad && lead_lane())
        a.dir[4 * uint64_t(st.page) + 3] = 0; // opened, never filled
    if (lead_lane()) {
        if (a.wave_sums) {
            a.wave_sums[uint64_t(1 + c) * a.wave_sums_stride + stream]           = st.nnz;
            a.wave_sums[uint64_t(1 + kMaxOut + c) * a.wave_sums_stride + stream] = st.npages;
        } else {
            if (st.nnz) atomicAdd(&a.counters->nnz[c], (unsigned long long)st.nnz);
            if (st.npages) atomicAdd(&a.counters->col_pages[c], (unsigned long long)st.npages);
        }
    }
}

// Flushes the wave's tuple count and retires the unused part of its last
// reservation (staging slots become holes, pool pages invalid directory entries).
__device__ inline void finish_wave(const PipeArgs& a, const WaveState& ws, uint64_t local, uint32_t stream) {
    uint32_t lane = lane_id();
    if (lead_lane()) {
        if (a.wave_sums)
            a.wave_sums[stream] = local;
        else if (local)
            atomicAdd(&a.counters->tuples, (unsigned long long)local);
    }
    if (a.target == TARGET_STAGE) {
        for (uint64_t i = ws.next + lane; i < ws.end && i < a.stage_cap; i += 64) a.stage[i * a.stage_words] = kEmpty;
    } else if (a.target == TARGET_PAGES) {
        for (uint64_t i = ws.next + lane; i < ws.end && i < a.pool_pages; i += 64) a.dir[4 * i + 3] = 0;
    }
}

// Early bloom filters of a pages-driven pipeline, tested before the full
// per-block work. Returns the blocks (bit j = block `first + j`, j < nblk <=
// 64) that hold a row passing every filter tested. Blocks go through eight at
// a time, so a wave has eight blocks' index, key and filter loads in flight
// where the full body has one. Column c (a bit of `bmask`) is tested against
// the filter of probe (bprobe >> 4c) & 15, the first probe keyed on it; the
// full cascade tests every filter again, so no block it would keep is
// dropped. Only k_pipeline<8> uses it: in the other variants its registers
// would cost a wave per SIMD (<4> would pass 128 VGPRs, <16> 168). A later
// column's keys are decoded only for the rows that passed the earlier
// filters: loading them ahead for every row costs more than the round trip
// it saves.
template <int NA>
__device__ inline uint64_t prefilter_blocks(const PipeArgs& a, uint64_t first, uint32_t nblk, uint64_t end,
                                            uint32_t bmask, uint64_t bprobe) {
    constexpr int  U    = 8;
    const uint32_t lane = lane_id();
    uint64_t       live = 0;
    for (uint32_t g = 0; g < nblk; g += U) {
        bool alive[U];
#pragma unroll
        for (int j = 0; j < U; ++j) alive[j] = g + j < nblk && (first + g + j) * 64 + lane < end;
#pragma unroll
        for (int c = 0; c < kMaxScan; ++c) {
            if (c >= NA || !((bmask >> c) & 1u)) continue;
            uint64_t any = 0;
#pragma unroll
            for (int j = 0; j < U; ++j) any |= __ballot(alive[j]);
            if (!any) break;
            const ProbeSpec& P     = a.probes[(bprobe >> (4 * c)) & 15u];
            const bool       dense = a.scan[c].dense_rows != 0;
            BlockIdx         bi[U] = {};
            if (!dense) // index entries of all U blocks first, then their values
#pragma unroll
                for (int j = 0; j < U; ++j) bi[j] = load_block_idx(a.scan[c].blocks, first + g + (g + j < nblk ? j : 0));
            uint64_t w[U], bb[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const uint64_t v  = dense ? decode_dense(a.scan[c], first + g + j, alive[j])
                                          : decode_with(a.scan[c], bi[j], first + g + j, alive[j]);
                alive[j]          = alive[j] && v != kNull;
                const uint64_t h2 = hash_bloom(hash_key(v));
                bb[j]             = bloom_bits(h2);
                w[j]              = alive[j] ? P.bloom[bloom_word(h2, P.bloom_log)] : 0;
            }
#pragma unroll
            for (int j = 0; j < U; ++j) alive[j] = alive[j] && (w[j] & bb[j]) == bb[j];
        }
#pragma unroll
        for (int j = 0; j < U; ++j)
            if (__ballot(alive[j])) live |= 1ull << (g + j);
    }
    return live;
}

// The narrow variant runs at 3 waves per SIMD (134 VGPRs). Held to 4 (128
// VGPRs) it spilled 20 bytes per lane to scratch memory: every launch then
// needs a scratch allocation per queue, which can fail when many worker
// streams run with device memory nearly full, and the c5 plan set ran 2-3%
// slower.
template <int NA>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NA <= 4 ? 3 : 1, 8))) void k_pipeline(PipeArgs a) {
    extern __shared__ uint64_t lds[];
    const uint32_t wib    = wave_in_block();
    const uint32_t stream = blockIdx.x * kWavesPerBlock + wib;
    const uint32_t nwaves = gridDim.x * kWavesPerBlock;
    const uint32_t lane   = lane_id();

    // per-wave LDS: kMaxOut OutState (6 words each), nout x 64 bitmap words
    const uint32_t per_wave = kMaxOut * 6 + a.nout * 64;
    OutState*      ost      = reinterpret_cast<OutState*>(lds + wib * per_wave);
    uint64_t*      obits    = lds + wib * per_wave + kMaxOut * 6;
    WaveState      ws{0, 0, 0, 0};
    if (a.target == TARGET_PAGES && lane < a.nout) {
        OutState s{kNoPage, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        ost[lane] = s;
    }
    if (a.wave_sums && stream == 0 && lead_lane()) a.counters->waves = nwaves;
    if (a.static_chunk && (a.target == TARGET_PAGES || a.target == TARGET_STAGE)) {
        // the wave's first pages (or staging slots) need no atomic
        const uint64_t chunk = a.target == TARGET_PAGES ? a.page_chunk : a.stage_chunk;
        ws.next              = uint64_t(stream) * chunk;
        ws.end               = ws.next + chunk;
        if (stream == 0 && lead_lane()) a.counters->static_pages = uint64_t(nwaves) * chunk;
        if (a.target == TARGET_STAGE && ws.end > a.stage_cap && lead_lane()) atomicExch(&a.counters->overflow, 1ull);
    }

    uint64_t local_count = 0;

    // Morsel schedule. With xcd_order, workgroups are taken to be dispatched
    // round-robin over the 8 XCDs (block b on XCD b % 8): XCD x walks the x-th
    // eighth of the morsels in order, so its concurrently running waves touch
    // one narrow, L2-resident range of a partition-ordered probe.
    uint64_t m_first = stream, m_end = a.nmorsels, m_step = nwaves;
    if (a.xcd_order) {
        uint32_t x    = blockIdx.x % 8;
        uint32_t bpx  = gridDim.x / 8;
        uint64_t lo   = a.nmorsels * x / 8;
        uint64_t hi   = a.nmorsels * (x + 1) / 8;
        m_first       = lo + uint64_t(blockIdx.x / 8) * kWavesPerBlock + wib;
        m_end         = hi;
        m_step        = uint64_t(bpx) * kWavesPerBlock;
    }
    // Scanned columns keying an early bloom filter, and per column the first
    // probe keyed on it (4 bits each).
    uint32_t bmask  = 0;
    uint64_t bprobe = 0;
    if (a.driver == DRIVER_PAGES)
        for (uint32_t p = a.nprobes; p-- > 0;)
            if (a.probes[p].bloom && a.probes[p].early && a.probes[p].key_slot < uint32_t(NA)) {
                const uint32_t c = a.probes[p].key_slot;
                bmask |= 1u << c;
                bprobe = (bprobe & ~(15ull << (4 * c))) | (uint64_t(p) << (4 * c));
            }
    // A wave tests the filters ahead of the full per-block work while at most
    // half of the blocks it has tested survive them.
    const bool try_prefilter = NA == 8 && bmask != 0 && a.prefilter;
    uint32_t   pf_seen = 0, pf_live = 0;

    for (uint64_t m = m_first; m < m_end && !ws.stop; m += m_step) {
        uint64_t r0  = m * a.morsel_rows;
        uint64_t end = r0 + a.morsel_rows < a.nrows ? r0 + a.morsel_rows : a.nrows;
        // runs of 64 blocks (a morsel: <= 4096 rows); past the abort bound a
        // block ends after one round and the morsel loop stops (checks here
        // cost registers)
        for (uint64_t cb = r0; cb < end; cb += 64 * 64) {
        const uint64_t cend = cb + 64 * 64 < end ? cb + 64 * 64 : end;
        const uint32_t nblk = uint32_t((cend - cb + 63) / 64);
        uint64_t       live = nblk >= 64 ? ~0ull : ((1ull << nblk) - 1);
        if (try_prefilter && (pf_seen < 256 || 2 * pf_live <= pf_seen)) {
            live = prefilter_blocks<NA>(a, cb / 64, nblk, end, bmask, bprobe);
            pf_seen += nblk;
            pf_live += uint32_t(__popcll(live));
        }
        for (; live; live &= live - 1) {
            const uint64_t rb = cb + 64 * uint64_t(__builtin_ctzll(live));
            uint64_t row    = rb + lane;
            bool     active = row < end;
            uint64_t s[NA];
#pragma unroll
            for (int k = 0; k < NA; ++k) s[k] = kNull;
            bool alive = active;
            if (a.driver == DRIVER_PAGES && a.sel) { // the scan's restrictions
                const uint64_t w = load_invariant_u64(a.sel + rb / 64);
                if (!w) continue;
                alive = alive && ((w >> lane) & 1);
            }
            if (a.driver == DRIVER_PAGES) {
                // Bloom cascade on keys that come straight from the scan, each
                // key column decoded just before its filter and only while a
                // lane of the block is alive; the other columns are decoded
                // once the block still has a row after the cascade.
                // Filters commute, so the cascade runs in column order; every
                // column index is a compile-time constant (no runtime-indexed
                // register array).
                uint32_t decoded = 0; // wave-uniform bitmask of decoded columns
#pragma unroll
                for (int c = 0; c < kMaxScan; ++c) {
                    if (c >= NA || !((bmask >> c) & 1u) || !__ballot(alive)) continue;
                    s[c] = decode_block(a.scan[c], rb / 64, active);
                    decoded |= 1u << c;
                    if (s[c] == kNull) alive = false; // NULL keys never match
                    for (uint32_t p = 0; p < a.nprobes; ++p) {
                        const ProbeSpec& P = a.probes[p];
                        if (P.bloom && P.early && P.key_slot == uint32_t(c) && alive) {
                            const uint64_t h2 = hash_bloom(hash_key(s[c]));
                            const uint64_t bb = bloom_bits(h2);
                            alive             = (P.bloom[bloom_word(h2, P.bloom_log)] & bb) == bb;
                        }
                    }
                }
                if (__ballot(alive)) {
#pragma unroll
                    for (int c = 0; c < kMaxScan; ++c)
                        if (uint32_t(c) < a.nscan && c < NA && !((decoded >> c) & 1u))
                            s[c] = decode_block(a.scan[c], rb / 64, active);
                }
            } else if (active) {
                // materialised tuples: word 0 holds a join class, never NULL except in unused slots
                alive = a.tuples[row * a.tuple_words] != kEmpty;
#pragma unroll
                for (int c = 0; c < kMaxScan; ++c)
                    if (uint32_t(c) < a.nscan && c < NA) s[c] = a.tuples[row * a.tuple_words + a.tuple_word[c]];
                // bloom cascade on keys that come straight from the tuples
                for (uint32_t p = 0; p < a.nprobes; ++p) {
                    const ProbeSpec& P = a.probes[p];
                    if (P.bloom && P.early && alive) {
                        const uint64_t key = sget(s, P.key_slot);
                        if (key == kNull) {
                            alive = false;
                        } else {
                            uint64_t h2 = hash_bloom(hash_key(key));
                            uint64_t bb = bloom_bits(h2);
                            alive       = (P.bloom[bloom_word(h2, P.bloom_log)] & bb) == bb;
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < NA; ++k)
                if ((a.not_null >> k) & 1u) alive = alive && s[k] != kNull;

            // depth-first enumeration of probe matches, one tuple per lane per
            // round. Per level, pos = ~0 (no more matches), the next slot to
            // inspect, or (grouped table) the next payload row of the matched
            // key's group in bits 0..39 with the rows left in bits 40..63.
            uint32_t level = 0;
            uint64_t pos[kMaxProbes];
            uint64_t pkey[kMaxProbes];
            uint32_t mult = 1; // tuples this lane's match stands for (count target)
#pragma unroll
            for (int k = 0; k < kMaxProbes; ++k) {
                pos[k]  = ~0ull;
                pkey[k] = kNull;
            }
            if (alive && a.nprobes > 0) {
                pkey[0] = sget(s, a.probes[0].key_slot);
                pos[0]  = hash_key(pkey[0]) >> a.probes[0].shift;
            }
            for (;;) {
                bool have = false;
                if (a.nprobes == 0) {
                    have  = alive;
                    alive = false;
                } else {
                    while (alive && !have) {
#pragma unroll
                        for (int k = 0; k < kMaxProbes; ++k) {
                            if (uint32_t(k) != level || uint32_t(k) >= a.nprobes) continue;
                            const ProbeSpec& P     = a.probes[k];
                            bool             found = false;
                            uint64_t         p     = pos[k];
                            uint32_t         rem   = 0; // grouped: rows left in the group, from row p
                            if (P.grouped && p != ~0ull) {
                                rem = uint32_t(p >> kGroupPosBits);
                                p &= (1ull << kGroupPosBits) - 1;
                                if (rem == 0) {
                                    // the key's slot (keys are unique in a grouped table)
                                    for (;;) {
                                        const uint64_t* e  = P.ht + 2 * p;
                                        const uint64_t  kw = e[0];
                                        if (kw == kEmpty) break;
                                        if (kw == pkey[k]) {
                                            const uint64_t g = e[1];
                                            p                = g >> kGroupCountBits;
                                            rem              = uint32_t(g & ((1u << kGroupCountBits) - 1));
                                            break;
                                        }
                                        p = (p & ~P.mask) | ((p + 1) & P.mask);
                                    }
                                }
                                // next row of the group that passes the residual equalities
                                while (rem > 0) {
                                    const uint64_t* t  = P.payload + p * P.pwords;
                                    bool            ok = true;
                                    for (uint32_t c = 0; c < P.nchecks; ++c)
                                        ok = ok && t[P.check_word[c] - 1] == sget(s, P.check_slot[c]);
                                    if (ok) {
                                        found = true;
                                        break;
                                    }
                                    ++p;
                                    --rem;
                                }
                            } else if (p != ~0ull) {
                                for (;;) {
                                    const uint64_t* e  = P.ht + p * P.words;
                                    uint64_t        kw = e[0];
                                    if (kw == kEmpty) break;
                                    if (kw == pkey[k]) {
                                        bool ok = true;
                                        for (uint32_t c = 0; c < P.nchecks; ++c)
                                            ok = ok && e[P.check_word[c]] == sget(s, P.check_slot[c]);
                                        if (ok) {
                                            found = true;
                                            break;
                                        }
                                    }
                                    p = (p & ~P.mask) | ((p + 1) & P.mask);
                                }
                            }
                            if (found && P.grouped) {
                                const uint64_t* t = P.payload + p * P.pwords;
                                for (uint32_t u = 0; u < P.nattrs; ++u) sset(s, P.attr_slot[u], t[P.attr_word[u] - 1]);
                                if (a.target == TARGET_COUNT && P.nchecks == 0 && uint32_t(k) + 1 == a.nprobes) {
                                    mult = rem; // the rest of the group, counted at once
                                    rem  = 0;
                                } else {
                                    --rem;
                                }
                                pos[k] = rem ? ((p + 1) | (uint64_t(rem) << kGroupPosBits)) : ~0ull;
                            } else if (found) {
                                const uint64_t* e = P.ht + p * P.words;
                                for (uint32_t t = 0; t < P.nattrs; ++t) sset(s, P.attr_slot[t], e[P.attr_word[t]]);
                                pos[k] = P.unique ? ~0ull : ((p & ~P.mask) | ((p + 1) & P.mask));
                            }
                            if (found) {
                                if (uint32_t(k) + 1 == a.nprobes) {
                                    have = true;
                                } else if (k + 1 < kMaxProbes) {
                                    const ProbeSpec& Q    = a.probes[k + 1];
                                    uint64_t         nk   = sget(s, Q.key_slot);
                                    bool             pass = nk != kNull;
                                    if (pass && Q.bloom) {
                                        uint64_t h2 = hash_bloom(hash_key(nk));
                                        uint64_t bb = bloom_bits(h2);
                                        pass        = (Q.bloom[bloom_word(h2, Q.bloom_log)] & bb) == bb;
                                    }
                                    pkey[k + 1] = nk;
                                    pos[k + 1]  = pass ? (hash_key(nk) >> Q.shift) : ~0ull;
                                    level       = k + 1;
                                }
                            } else {
                                pos[k] = ~0ull;
                                if (k == 0)
                                    alive = false;
                                else
                                    level = k - 1;
                            }
                        }
                    }
                }
                uint64_t hm = __ballot(have);
                if (hm) {
                    uint32_t cnt  = __popcll(hm);
                    uint32_t rank = __popcll(hm & mask_lt());
                    if (a.target == TARGET_STAGE) {
                        uint64_t words[kMaxWords];
#pragma unroll
                        for (int w = 0; w < kMaxWords; ++w)
                            words[w] = uint32_t(w) < a.stage_words ? sget(s, a.stage_slot[w]) : 0;
                        stage_put(a, ws, have, rank, cnt, words, a.stage_words);
                        local_count += cnt;
                    } else if (a.target == TARGET_COUNT) {
                        uint64_t m = have ? uint64_t(mult) : 0;
#pragma unroll
                        for (int d = 32; d >= 1; d >>= 1) m += __shfl_xor(m, d, 64);
                        local_count += m;
                    } else {
                        local_count += cnt;
                        for (uint32_t c = 0; c < a.nout; ++c) {
                            uint64_t v = sget(s, a.out[c].slot);
                            OutState st = ost[c];
                            if (a.out[c].type == T_VARCHAR) {
                                uint8_t* stg = a.str_stage + (uint64_t(stream) * a.nvarchar + a.out[c].vc) * kPageSize;
                                append_varchar(a, ws, st, obits + c * 64, stg, v, have, rank, cnt, c, stream);
                            } else {
                                append_fixed(a, ws, st, obits + c * 64, a.out[c].type, v, have, rank, cnt, c, stream);
                            }
                            ost[c] = st;
                        }
                    }
                }
                mult = 1;
                if (!__ballot(alive) || ws.stop) break;
            }
        }
        } // runs of 64 blocks
    }

    if (a.target == TARGET_PAGES) {
        for (uint32_t c = 0; c < a.nout; ++c) {
            OutState st = ost[c];
            const uint8_t* stg = a.out[c].type == T_VARCHAR
                                     ? a.str_stage + (uint64_t(stream) * a.nvarchar + a.out[c].vc) * kPageSize
                                     : nullptr;
            finish_column(a, st, obits + c * 64, c, stream, stg);
        }
    }
    finish_wave(a, ws, local_count, stream);
}

// ---------------------------------------------------------------------------
// lean pipeline: one probe without residual predicates whose only scanned
// join column is the probe key, and up to four fixed-width output columns
// (or a count)
// ---------------------------------------------------------------------------
// Each output column names where its value comes from: the probe key, a
// column of the driving table (decoded straight into the output register), or
// a word of the matching hash-table entry. Without tuple slots there are no
// runtime-indexed register selects. The page-writer state is wave-uniform
// (scalar registers) and each output column's validity bitmap lives in one
// 64-bit register per lane (lane i holds rows 64i..64i+63 of the open page):
// a 64-row batch is appended to a column with one ballot, one masked store and
// a register OR; the batch's validity in row order comes from a single forward
// lane permute shared by all columns.
// (LeanArgs, LeanOutCol and the SRC_* codes are declared in kernels.hpp.)

// Open page of one output column (wave-uniform).
struct LeanPage {
    uint32_t page;   // pool index, kNoPage if none
    uint32_t nr;     // rows in the page
    uint32_t nv;     // non-NULL values in the page
    int32_t  avail;  // remaining bit budget
    uint32_t seq;    // pages emitted by this wave for this column
    uint32_t npages; // pages written
    uint32_t bad;    // pool exhausted: keep counting, write nothing
    uint32_t pad;
    uint64_t nnz;    // non-NULL values written
};

// The kernel's by-value argument struct, read from the kernarg segment (the
// first explicit argument sits at offset 0). The empty asm hides where the
// pointer comes from, so argument loads inside a loop are not hoisted and kept
// live in scalar registers for the whole kernel (which runs out of scalar
// registers and spills them into vector lanes); each iteration re-reads the
// fields it needs with scalar loads that hit the scalar cache.
template <class T>
__device__ inline const T& kernel_args() {
    typedef const __attribute__((address_space(4))) T* KPtr;
    KPtr p = (KPtr)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return *(const T*)p;
}
// Same for a read-only descriptor at device address `addr` (written before the
// launch): scalar loads through the constant address space, re-read per use.
template <class T>
__device__ inline const T& args_at(uintptr_t addr) {
    typedef const __attribute__((address_space(4))) T* KPtr;
    KPtr p = (KPtr)addr;
    asm volatile("" : "+s"(p));
    return *(const T*)p;
}
__device__ inline uintptr_t kernarg_addr() { return (uintptr_t)__builtin_amdgcn_kernarg_segment_ptr(); }

__device__ inline uint64_t low_mask(uint32_t n) { return n >= 64 ? ~0ull : ((1ull << n) - 1ull); }

__device__ inline void lean_open(const LeanArgs& a, WaveState& ws, LeanPage& st, uint64_t& bits, int32_t budget) {
    uint64_t next = ws.next;
    if (next == ws.end) {
        const uint32_t chunk = a.page_chunk;
        const uint64_t base  = a.static_chunk ? uint64_t(gridDim.x) * kWavesPerBlock * chunk : 0;
        next   = base + wave_fetch_add(&a.counters->pages, (unsigned long long)chunk);
        ws.end = next + chunk;
        if (next >= a.abort_pages) ws.stop = 1;
    }
    ws.next  = next + 1;
    // readfirstlane keeps these scalar: the compiler otherwise folds the
    // uniform test into the lane-0 branch below and the page index comes out
    // of a divergent join
    st.bad   = uniform_u32(next >= a.pool_pages ? 1u : 0u);
    st.page  = uniform_u32(st.bad ? 0u : uint32_t(next));
    if (st.bad && lead_lane()) atomicExch(&a.counters->overflow, 1ull);
    st.nr    = 0;
    st.nv    = 0;
    st.avail = budget;
    bits     = 0;
}

// Same page layout as page_close: header, bitmap bytes at the page end, directory entry.
__device__ inline void lean_close(const LeanArgs& a, LeanPage& st, uint64_t bits, uint32_t col, uint32_t stream) {
    if (!st.bad) {
        uint8_t* page = a.pool + uint64_t(st.page) * kPageSize;
        uint32_t nb   = (st.nr + 7) / 8;
        uint8_t* bm   = page + kPageSize - nb;
        uint32_t b0   = lane_id() * 8;
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b)
            if (b0 + b < nb) bm[b0 + b] = uint8_t(bits >> (8 * b));
        if (lead_lane()) {
            *reinterpret_cast<uint32_t*>(page) = st.nr | (st.nv << 16);
            uint4 e = make_uint4(col, stream, st.seq, 1u);
            *reinterpret_cast<uint4*>(a.dir + 4 * uint64_t(st.page)) = e;
        }
        st.npages++;
        st.nnz += st.nv;
    }
    st.seq++;
    st.page = kNoPage;
}

// Writes batch rows [lo, hi) (row order) to the open page. `vm` = lanes with a
// non-NULL value, `pat` = validity of the batch rows in row order.
__device__ inline void lean_segment(const LeanArgs& a, LeanPage& st, uint64_t& bits, uint32_t w, uint64_t v,
                                    bool valid, uint32_t rank, uint64_t vm, uint64_t pat, uint32_t lo, uint32_t hi) {
    const uint32_t n    = hi - lo;
    const uint64_t sp   = (pat >> lo) & low_mask(n);
    const uint32_t nvs  = __popcll(sp);
    const uint32_t skip = __popcll(pat & low_mask(lo)); // valid rows before the segment
    if (valid && rank >= lo && rank < hi && !st.bad) {
        uint32_t vi  = st.nv + __popcll(vm & mask_lt()) - skip;
        uint8_t* dst = a.pool + uint64_t(st.page) * kPageSize;
        if (w == 4)
            *reinterpret_cast<uint32_t*>(dst + 4 + 4 * vi) = uint32_t(v);
        else
            *reinterpret_cast<uint64_t*>(dst + 8 + 8 * uint64_t(vi)) = v;
    }
    const uint32_t wi = st.nr >> 6, sh = st.nr & 63, lane = lane_id();
    if (lane == wi) bits |= sp << sh;
    if (sh && lane == wi + 1) bits |= sp >> (64 - sh);
    st.nr += n;
    st.nv += nvs;
    st.avail -= int32_t(n + 8 * w * nvs);
}

// Appends a batch of `cnt` rows to a fixed-width column following the
// reference's bit budget (TableTarget.cpp:111-182): 1 bit per row plus the
// value width per non-NULL row. A fresh page holds more than 64 rows of any
// fixed-width type, so a batch spans at most two pages.
__device__ inline void lean_append(const LeanArgs& a, WaveState& ws, LeanPage& st, uint64_t& bits, uint32_t w,
                                   uint64_t v, bool valid, uint32_t rank, uint32_t cnt, uint64_t vm, uint64_t pat,
                                   uint32_t col, uint32_t stream) {
    const int32_t budget = 8 * (int32_t(kPageSize) - 4 - (w == 4 ? 0 : 4));
    const int32_t vbits  = int32_t(8 * w);
    if (st.page == kNoPage) lean_open(a, ws, st, bits, budget);
    uint32_t k = cnt;
    if (int32_t(cnt) + vbits * int32_t(__popcll(vm)) > st.avail || st.nr + cnt > kMaxPageRows) {
        // the leading rows that still fit: lane r tests rows 0..r (bitwise &,
        // not &&: a short-circuit branch here makes the compiler treat the
        // page state after it as divergent and move it to vector registers)
        uint32_t r    = lane_id();
        int32_t  cost = int32_t(r + 1) + vbits * int32_t(__popcll(pat & mask_le()));
        bool     fits = (r < cnt) & (cost <= st.avail) & (st.nr + r + 1 <= kMaxPageRows);
        k             = uniform_u32(__popcll(__ballot(fits)));
    }
    if (k) lean_segment(a, st, bits, w, v, valid, rank, vm, pat, 0, k);
    if (k < cnt) {
        lean_close(a, st, bits, col, stream);
        lean_open(a, ws, st, bits, budget);
        lean_segment(a, st, bits, w, v, valid, rank, vm, pat, k, cnt);
    }
    if (st.nr >= kMaxPageRows) lean_close(a, st, bits, col, stream);
}

__device__ inline void lean_finish(const LeanArgs& a, LeanPage& st, uint64_t bits, uint32_t c, uint32_t stream) {
    if (st.page != kNoPage && st.nr > 0)
        lean_close(a, st, bits, c, stream);
    else if (st.page != kNoPage && !st.bad && lead_lane())
        a.dir[4 * uint64_t(st.page) + 3] = 0; // opened, never filled
    if (lead_lane()) {
        if (a.wave_sums) {
            a.wave_sums[uint64_t(1 + c) * a.wave_sums_stride + stream]           = st.nnz;
            a.wave_sums[uint64_t(1 + kMaxOut + c) * a.wave_sums_stride + stream] = st.npages;
        } else {
            if (st.nnz) atomicAdd(&a.counters->nnz[c], (unsigned long long)st.nnz);
            if (st.npages) atomicAdd(&a.counters->col_pages[c], (unsigned long long)st.npages);
        }
    }
}

// Key and first attribute word of hash-table entry `slot`; two-word entries
// (16-byte aligned) take one 16-byte load. Every slot of a region exists, so
// inspecting a neighbour slot never reads out of bounds.
__device__ inline void ht_entry2(const uint64_t* ht, uint32_t words, uint64_t slot, uint64_t& k, uint64_t& w1) {
    const uint64_t* e = ht + slot * words;
    if (words == 2) {
        const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(e);
        k                  = v.x;
        w1                 = v.y;
    } else {
        k  = e[0];
        w1 = words > 1 ? e[1] : 0;
    }
}

// Per-wave state of the lean pipeline: page reservations, one open page and
// its bitmap register per output column, and the row count.
template <int NS>
struct LeanWave {
    WaveState ws;
    LeanPage  st[NS];
    uint64_t  bits[NS];
    uint64_t  local;
};

template <int NS>
__device__ inline void lean_wave_init(LeanWave<NS>& w) {
    w.ws = WaveState{0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < NS; ++c) {
        w.st[c]   = LeanPage{kNoPage, 0, 0, 0, 0, 0, 0, 0, 0};
        w.bits[c] = 0;
    }
    w.local = 0;
}

// Launch shapes of the one-probe kernel whose descriptor flags are known at
// compile time. Every test of a flag costs scalar instructions in each block
// or probe step, and the scalar unit (one per CU, shared by its four SIMDs)
// is the kernel's busiest issue port: with the c2 shape compiled in, the
// probe kernel went from 11.48 to 10.56 ms.
enum : int {
    LEAN_ANY    = 0, // flags read from the descriptor
    LEAN_TWO    = 1, // page driver, no bloom filter, no selection words, 2-word entries read as aligned pairs
    LEAN_PACKED = 2, // page driver, no bloom filter, no selection words, packed table (its flag still read)
};

// Probe state of the lanes that one probe step left searching, per wave in
// LDS: key, next slot and the scanned output values. Once 64 have gathered
// they are probed as a block of their own. Without it a wave steps as long as
// the longest probe sequence of its 64 lanes (at load factor 0.35 and two
// slots per step: 1.15 steps per lane on average, 3.0 for the longest of 64);
// with it a block of fresh rows takes one step, and the 12% of lanes that
// need more finish in full blocks of deferred lanes (about 1.45 steps per 64
// rows in all). Unique tables only: a lane then stops at its first match.
template <int NS>
struct LeanQueue {
    static constexpr uint32_t kCap = 128; // fewer than 64 held plus up to 64 added
    uint64_t key[kCap];
    uint64_t v[NS][kCap];
    uint32_t slot[kCap];
};

// Runs the one-probe pipeline over rows [r0, end) (multiples of 64 except at
// the end) of the descriptor at `args_addr` (re-read per 64-row block).
// `unique` >= 0 overrides the descriptor's duplicate-free flag. `Q` (DQ only)
// is the wave's deferral queue; it is empty again when the call returns.
template <int NO, int NS, int VAR, bool DQ = false>
__device__ inline void lean_rows(uintptr_t args_addr, uint64_t r0, uint64_t end, uint32_t stream, LeanWave<NS>& W,
                                 int32_t unique, LeanQueue<NS>* Q = nullptr) {
    constexpr bool kPages = VAR != LEAN_ANY; // the driver is a page-stored table
    constexpr bool kPlain = VAR != LEAN_ANY; // no bloom filter, no selection words
    constexpr bool kTwo   = VAR == LEAN_TWO; // 2-word entries, never packed, aligned pairs
    const uint32_t lane   = lane_id();
    // the packed table is used unless its build flagged an entry that does not fit
    bool packed = false;
    if (!kTwo) {
        const LeanArgs& a0 = args_at<LeanArgs>(args_addr);
        packed             = a0.ht_packed && load_invariant_u32(a0.pack_flag) == 0;
    }
    // Walks the searching lanes until each has decided (a match or an empty
    // slot), at most `max_steps` steps. Each step inspects two consecutive
    // slots of every searching lane, which halves the dependent hash-table
    // round trips. While the region holds an empty slot no lane takes more
    // than rmask + 2 steps over all its rounds, so a block takes at most 64
    // times that; past it the table is corrupt or full, and the block stops
    // with the fault flag set instead of walking on.
    auto walk_lanes = [&](const LeanArgs& a, uint64_t key, uint64_t& q, bool& live, bool& have, uint64_t& hslot,
                          uint64_t& k1, uint64_t& walk, uint32_t max_steps) {
        const uint64_t* ht         = packed ? a.ht_packed : a.ht;
        const uint32_t  words      = kTwo ? 2u : a.words;
        const uint64_t  rmask      = a.mask;
        const uint64_t  walk_limit = 64 * (rmask + 2);
        const uint64_t  odd        = (kTwo || a.aligned_pairs || packed) ? 1 : 0; // uniform
        bool            searching  = live;
        have                       = false;
        hslot                      = 0;
        k1                         = 0;
        uint32_t steps             = 0;
        while (__ballot(searching) && steps < max_steps) {
            ++steps;
            if (++walk > walk_limit) {
                if (lead_lane()) atomicExch(&a.counters->fault, 1ull);
                live = false;
                break;
            }
            if (searching) {
                // slots qa (ignored when `skip`) and qb, then continue at qn
                const uint64_t qa   = q & ~odd;
                const bool     skip = qa != q;
                const uint64_t qb   = (qa & ~rmask) | ((qa + 1) & rmask);
                const uint64_t qn   = (qb & ~rmask) | ((qb + 1) & rmask);
                uint64_t       ka, wa, kb, wb;
                if (!kTwo && packed) {
                    // one 8-byte word per slot, always aligned pairs: both
                    // slots come from one 16-byte load (two loads of one
                    // line can each miss the L2 and fetch it twice)
                    const ulonglong2 e2 = *reinterpret_cast<const ulonglong2*>(ht + qa);
                    const uint64_t   ea = e2.x, eb = e2.y;
                    ka = ea == kEmpty ? kEmpty : (ea & 0xffffffffull);
                    wa = ea >> 32;
                    kb = eb == kEmpty ? kEmpty : (eb & 0xffffffffull);
                    wb = eb >> 32;
                } else {
                    ht_entry2(ht, words, qa, ka, wa);
                    ht_entry2(ht, words, qb, kb, wb);
                }
                // the first deciding slot wins: a match, or an empty
                // slot that ends the walk. Selects, not an if-chain:
                // the chain's exec-mask bookkeeping cost more scalar
                // instructions than the whole step.
                const bool ma = !skip & (ka == key);
                const bool ea = !skip & (ka == kEmpty);
                const bool mb = !ma & !ea & (kb == key);
                const bool eb = !ma & !ea & !mb & (kb == kEmpty);
                have          = ma | mb;
                hslot         = ma ? qa : qb;
                k1            = ma ? wa : wb;
                q             = ma ? qb : qn;
                live          = live & !(ea | eb);
                searching     = live & !have;
            }
        }
    };
    // Appends the matching lanes' output rows to the wave's open pages.
    auto emit = [&](const LeanArgs& a, uint64_t key, const uint64_t (&v)[NS], bool have, uint64_t hslot,
                    uint64_t k1) {
        const uint64_t  hm = __ballot(have);
        if (!hm) return;
        const uint64_t* ht  = packed ? a.ht_packed : a.ht;
        const uint64_t* he  = ht + hslot * (kTwo ? 2u : a.words);
        const uint32_t  cnt = __popcll(hm);
        W.local += cnt;
        if (NO > 0) {
            const uint32_t rank = __popcll(hm & mask_lt());
            // forward permute: have-lanes to their rank, the others to cnt.. (a permutation)
            const uint32_t dst = have ? rank : cnt + __popcll(~hm & mask_lt());
            uint32_t       vbm = 0;
            uint64_t       o[NS];
#pragma unroll
            for (int c = 0; c < NO; ++c) {
                const uint32_t src = a.out[c].src, w = a.out[c].word;
                o[c] = src == SRC_KEY ? key : src == SRC_SCAN ? v[c] : (!have ? kNull : w == 1 ? k1 : he[w]);
                if (have && o[c] != kNull) vbm |= 1u << c;
            }
            const uint32_t recv = uint32_t(__builtin_amdgcn_ds_permute(int(dst << 2), int(vbm)));
#pragma unroll
            for (int c = 0; c < NO; ++c) {
                const bool     valid = (vbm >> c) & 1u;
                const uint64_t vm    = __ballot(valid);
                const uint64_t pat   = __ballot((recv >> c) & 1u);
                lean_append(a, W.ws, W.st[c], W.bits[c], a.out[c].type == T_INT32 ? 4u : 8u, o[c], valid, rank,
                            cnt, vm, pat, uint32_t(c), stream);
            }
        }
    };
    // Probes the top `n` (1..64) entries of the queue to the end as one block.
    uint32_t nq    = 0; // entries in the queue (wave-uniform, < 64 between blocks)
    auto     drain = [&](uint32_t n) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const LeanArgs& a   = args_at<LeanArgs>(args_addr);
        const bool      act = lane < n;
        const uint32_t  e   = nq - n + lane; // in [nq - n, nq) for the active lanes
        uint64_t        key = kNull, q = 0;
        uint64_t        v[NS];
#pragma unroll
        for (int c = 0; c < NS; ++c) v[c] = kNull;
        if (act) {
            key = Q->key[e];
            q   = Q->slot[e];
#pragma unroll
            for (int c = 0; c < NO; ++c)
                if (a.out[c].src == SRC_SCAN) v[c] = Q->v[c][e];
        }
        bool     live = act, have;
        uint64_t hslot, k1, walk = 0;
        walk_lanes(a, key, q, live, have, hslot, k1, walk, ~0u);
        emit(a, key, v, have, hslot, k1);
        nq -= n;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    };
    for (uint64_t rb = r0; rb < end && !W.ws.stop; rb += 64) {
        const LeanArgs& a      = args_at<LeanArgs>(args_addr); // re-read per block (scalar cache)
        const uint64_t  row    = rb + lane;
        const bool      active = row < end;
        const uint64_t  blk    = rb / 64;
        // probe key and the scanned output columns: all loads issued up front
        uint64_t key  = kNull;
        bool     hole = false;
        uint64_t v[NS];
#pragma unroll
        for (int c = 0; c < NS; ++c) v[c] = kNull;
        if (kPages || a.driver == DRIVER_PAGES) {
            key = decode_block(a.key, blk, active);
#pragma unroll
            for (int c = 0; c < NO; ++c) v[c] = a.out[c].src == SRC_SCAN ? decode_block(a.out[c].col, blk, active) : kNull;
        } else {
            const uint64_t* t = a.tuples + row * a.tuple_words;
            if (active) {
                hole = t[0] == kEmpty;
                key  = t[a.key_word];
            }
#pragma unroll
            for (int c = 0; c < NO; ++c) v[c] = (a.out[c].src == SRC_SCAN && active) ? t[a.out[c].word] : kNull;
        }
        bool alive = active && !hole && key != kNull;
        if (!kPlain && a.driver == DRIVER_PAGES && a.sel) alive = alive && ((load_invariant_u64(a.sel + blk) >> lane) & 1);
        const uint64_t h = hash_key(key);
        if (!kPlain && alive && a.bloom) {
            uint64_t h2 = hash_bloom(h);
            uint64_t bb = bloom_bits(h2);
            alive       = (a.bloom[bloom_word(h2, a.bloom_log)] & bb) == bb;
        }
        const bool uniq = unique > 0 || (unique < 0 && a.unique);
        // deferral keeps slots as 32-bit values: tables of at most 2^32 slots
        const bool defer = DQ && uniq && a.shift >= 32;
        uint64_t   q     = h >> a.shift; // next slot to inspect
        bool       live  = alive;
        uint64_t   walk  = 0;
        for (;;) {
            bool     have;
            uint64_t hslot, k1;
            walk_lanes(a, key, q, live, have, hslot, k1, walk, defer ? 1u : ~0u);
            if (DQ && defer) {
                // lanes still searching after one step join the queue
                const bool     left = live & !have;
                const uint64_t lm   = __ballot(left);
                if (lm) {
                    const uint32_t at = nq + __popcll(lm & mask_lt()); // nq < 64: at < 128
                    if (left) {
                        Q->key[at]  = key;
                        Q->slot[at] = uint32_t(q);
#pragma unroll
                        for (int c = 0; c < NO; ++c)
                            if (a.out[c].src == SRC_SCAN) Q->v[c][at] = v[c];
                    }
                    nq += __popcll(lm);
                }
            }
            if (uniq) live = false; // a lane without a match has stopped
            emit(a, key, v, have, hslot, k1);
            if (!__ballot(live) || W.ws.stop) break;
        }
        if (DQ && nq >= 64) drain(64);
    }
    if (DQ)
        while (nq > 0 && !W.ws.stop) drain(nq < 64 ? nq : 64u);
}

// Closes the wave's open pages, adds its counts to the counters and retires
// the unused part of its page reservation; the wave state starts over.
template <int NO, int NS>
__device__ inline void lean_flush(const LeanArgs& a, LeanWave<NS>& W, uint32_t stream) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int c = 0; c < NO; ++c) lean_finish(a, W.st[c], W.bits[c], uint32_t(c), stream);
    if (lead_lane()) {
        if (a.wave_sums)
            a.wave_sums[stream] = W.local;
        else if (W.local)
            atomicAdd(&a.counters->tuples, (unsigned long long)W.local);
    }
    if (NO > 0)
        for (uint64_t i = W.ws.next + lane; i < W.ws.end && i < a.pool_pages; i += 64) a.dir[4 * i + 3] = 0;
    lean_wave_init(W);
}

// NO = output columns (0: count only); VAR = launch shape (LEAN_*).
template <int NO, int VAR>
__global__ __launch_bounds__(256) void k_lean(LeanArgs karg) {
    (void)karg; // read through kernel_args()
    const LeanArgs& a     = kernel_args<LeanArgs>();
    constexpr int  NS     = NO > 0 ? NO : 1;
    const uint32_t wib    = wave_in_block();
    const uint32_t stream = blockIdx.x * kWavesPerBlock + wib;
    const uint32_t nwaves = gridDim.x * kWavesPerBlock;
    LeanWave<NS>   W;
    lean_wave_init(W);
    // the compiled-in shapes defer long probe walks (LeanQueue)
    constexpr bool DQ = VAR != LEAN_ANY;
    LeanQueue<NS>* Q  = nullptr;
    if constexpr (DQ) {
        __shared__ LeanQueue<NS> queues[kWavesPerBlock];
        Q = &queues[wib];
    }
    if (stream == 0 && a.ht_packed && load_invariant_u32(a.pack_flag) == 0 && lead_lane()) a.counters->packed = 1;
    if (a.wave_sums && stream == 0 && lead_lane()) a.counters->waves = nwaves;
    if (NO > 0 && a.static_chunk) {
        // the wave's first pages need no atomic
        W.ws.next = uint64_t(stream) * a.page_chunk;
        W.ws.end  = W.ws.next + a.page_chunk;
        if (stream == 0 && lead_lane()) a.counters->static_pages = uint64_t(nwaves) * a.page_chunk;
    }

    // Morsel schedule: each wave starts on morsel `stream`, then takes the next
    // unclaimed one from a device counter (one atomic per 4096-row morsel),
    // which balances skewed morsels (hot keys, misses) across the waves. The
    // XCD-ordered walk over a partitioned probe side stays static.
    uint64_t m_first = stream, m_end = a.nmorsels, m_step = nwaves;
    if (a.xcd_order) {
        uint32_t x   = blockIdx.x % 8;
        uint32_t bpx = gridDim.x / 8;
        m_first      = a.nmorsels * x / 8 + uint64_t(blockIdx.x / 8) * kWavesPerBlock + wib;
        m_end        = a.nmorsels * (x + 1) / 8;
        m_step       = uint64_t(bpx) * kWavesPerBlock;
    }
    for (uint64_t m = m_first; m < m_end && !W.ws.stop;) {
        uint64_t m_next = m + m_step;
        if (!a.xcd_order) {
            // no atomic when every morsel was some wave's first
            m_next = a.nmorsels <= nwaves ? m_end : nwaves + wave_fetch_add(&a.counters->morsel, 1ull);
        }
        const uint64_t r0  = m * a.morsel_rows;
        const uint64_t end = r0 + a.morsel_rows < a.nrows ? r0 + a.morsel_rows : a.nrows;
        lean_rows<NO, NS, VAR, DQ>(kernarg_addr(), r0, end, stream, W, -1, Q);
        m = m_next;
    }
    lean_flush<NO>(a, W, stream);
}

// ---------------------------------------------------------------------------
// batched joins: persistent kernels over device queues of build and probe tasks
// ---------------------------------------------------------------------------
// One launch takes the build tasks of every job of the batch, the next launch
// the probe tasks: the kernel boundary orders every probe after every build
// (and makes the tables visible to all XCDs), so no wave ever waits on another.
//
// Build task: inserts rows [index * morsel_rows, ...) of the job's build table
// (NULL keys skipped) with compare-and-swap on the key word, then stores the
// attribute words, and flags duplicate keys.
__device__ inline void batch_build_task(uintptr_t job_addr, uint64_t index) {
    const uint32_t  lane = lane_id();
    const BatchJob& J0   = args_at<BatchJob>(job_addr);
    const uint64_t  r0   = index * J0.build.morsel_rows;
    const uint64_t  end  = r0 + J0.build.morsel_rows < J0.build.rows ? r0 + J0.build.morsel_rows : J0.build.rows;
    bool            dup  = false, lost = false;
    for (uint64_t rb = r0; rb < end; rb += 64) {
        const BatchBuild& B      = args_at<BatchJob>(job_addr).build;
        const uint64_t    row    = rb + lane;
        const bool        active = row < end;
        const uint64_t    blk    = rb / 64;
        const uint64_t    key    = decode_block(B.key, blk, active);
        uint64_t          at[kBatchAttrs];
#pragma unroll
        for (int t = 0; t < kBatchAttrs; ++t) at[t] = uint32_t(t) < B.nattr ? decode_block(B.attr[t], blk, active) : 0;
        if (active && key != kNull) {
            const uint64_t h     = hash_key(key);
            const uint32_t words = B.words;
            uint64_t       pos   = h >> B.shift;
            bool           full  = false;
            for (uint64_t step = 0;; ++step) {
                if (step > B.mask) { // no empty slot left: report instead of walking on
                    full = true;
                    break;
                }
                const unsigned long long prev =
                    atomicCAS(reinterpret_cast<unsigned long long*>(B.ht + pos * words), kEmpty, key);
                if (prev == kEmpty) break;
                dup = dup || prev == key;
                pos = (pos + 1) & B.mask;
            }
            lost = lost || full;
            if (!full) {
#pragma unroll
                for (int t = 0; t < kBatchAttrs; ++t)
                    if (uint32_t(t) < B.nattr) B.ht[pos * words + 1 + t] = at[t];
                if (B.bloom) {
                    const uint64_t h2 = hash_bloom(h);
                    atomicOr(reinterpret_cast<unsigned long long*>(B.bloom + bloom_word(h2, B.bloom_log)),
                             bloom_bits(h2));
                }
            }
        }
    }
    BatchSync* sync = args_at<BatchJob>(job_addr).sync;
    if (__ballot(dup)) wave_fetch_or(&sync->dup, 1u);
    if (__ballot(lost)) wave_fetch_or(&sync->dup, 2u);
}

// Probe task (a later launch than every build task): runs the one-probe
// pipeline over rows [index * morsel_rows, ...) of the probe table and flushes
// its pages; a table without duplicate keys stops at the first match. The
// task's pages carry the task index as their stream id: a wave that runs two
// tasks of one job starts its page sequence at 0 for each, and the job's
// directory keys (column, task, sequence) stay unique; a download orders the
// pages by task, i.e. by probe row range.
template <int NO, int NS>
__device__ inline void batch_probe_task(uintptr_t job_addr, uint64_t index, LeanWave<NS>& W, LeanQueue<NS>* Q) {
    const uint32_t  stream = uint32_t(index);
    const BatchJob& J    = args_at<BatchJob>(job_addr);
    BatchSync*      sync = J.sync;
    const uint32_t  dup  = uniform_u32(__hip_atomic_load(&sync->dup, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    const uintptr_t pa    = job_addr + offsetof(BatchJob, probe);
    const LeanArgs& a     = args_at<LeanArgs>(pa);
    const uint64_t  r0    = index * a.morsel_rows;
    const uint64_t  end   = r0 + a.morsel_rows < a.nrows ? r0 + a.morsel_rows : a.nrows;
    // jobs without a bloom filter on 2-word tables take the compiled-in shape
    if (a.driver == DRIVER_PAGES && !a.bloom && !a.sel && !a.ht_packed && a.words == 2 && a.aligned_pairs)
        lean_rows<NO, NS, LEAN_TWO, true>(pa, r0, end, stream, W, dup ? 0 : 1, Q);
    else
        lean_rows<NO, NS, LEAN_ANY>(pa, r0, end, stream, W, dup ? 0 : 1);
    lean_flush<NO>(args_at<LeanArgs>(pa), W, stream);
}

// Takes tasks from `queue` until `ntasks` are handed out; each wave's task
// time (wall clock) is added to its job's counters.
__global__ __launch_bounds__(256) void k_batch_build(const BatchJob* jobs, const BatchTask* tasks, uint64_t ntasks,
                                                     unsigned long long* queue) {
    for (;;) {
        const unsigned long long t = wave_fetch_add(queue, 1ull);
        if (t >= ntasks) break;
        const uint4     tk  = load_invariant_u4(tasks + t);
        const uintptr_t ja  = uintptr_t(jobs + tk.x);
        const uint64_t  idx = uint64_t(tk.z) | (uint64_t(tk.w) << 32);
        const uint64_t  t0  = wall_clock64();
        batch_build_task(ja, idx);
        const uint64_t dt   = wall_clock64() - t0;
        BatchSync*     sync = args_at<BatchJob>(ja).sync;
        wave_fetch_add(&sync->build_clocks, (unsigned long long)dt);
    }
}

template <int NO>
__global__ __launch_bounds__(256) void k_batch_probe(const BatchJob* jobs, const BatchTask* tasks, uint64_t ntasks,
                                                     unsigned long long* queue) {
    constexpr int NS = NO > 0 ? NO : 1;
    LeanWave<NS>  W;
    lean_wave_init(W);
    // deferral queue of the compiled-in shape (empty between tasks)
    __shared__ LeanQueue<NS> queues[kWavesPerBlock];
    LeanQueue<NS>* Q = &queues[wave_in_block()];
    for (;;) {
        const unsigned long long t = wave_fetch_add(queue, 1ull);
        if (t >= ntasks) break;
        const uint4     tk  = load_invariant_u4(tasks + t);
        const uintptr_t ja  = uintptr_t(jobs + tk.x);
        const uint64_t  idx = uint64_t(tk.z) | (uint64_t(tk.w) << 32);
        const uint64_t  t0  = wall_clock64();
        batch_probe_task<NO>(ja, idx, W, Q);
        const uint64_t dt   = wall_clock64() - t0;
        BatchSync*     sync = args_at<BatchJob>(ja).sync;
        wave_fetch_add(&sync->probe_clocks, (unsigned long long)dt);
    }
}

uint32_t batch_resident_waves(uint32_t nout) {
    static const void* const kernels[kLeanOut + 2] = {
        reinterpret_cast<const void*>(&k_batch_probe<0>), reinterpret_cast<const void*>(&k_batch_probe<1>),
        reinterpret_cast<const void*>(&k_batch_probe<2>), reinterpret_cast<const void*>(&k_batch_probe<3>),
        reinterpret_cast<const void*>(&k_batch_probe<4>), reinterpret_cast<const void*>(&k_batch_build)};
    const uint32_t k = nout == kBatchBuildKernel ? uint32_t(kLeanOut) + 1
                                                 : (nout < uint32_t(kLeanOut) ? nout : uint32_t(kLeanOut));
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernels[k], 256, 0) != hipSuccess || per_cu <= 0 ||
        cus <= 0)
        return 256 * kWavesPerBlock;
    return uint32_t(per_cu) * uint32_t(cus) * kWavesPerBlock;
}

double batch_clock_hz() {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
        return double(khz) * 1000.0;
    return 100e6;
}

hipError_t launch_batch(const BatchJob* jobs, const BatchTask* tasks, uint64_t ntasks, unsigned long long* queue,
                        uint32_t nout, uint32_t max_waves, hipStream_t s) {
    if (nout > uint32_t(kLeanOut) && nout != kBatchBuildKernel) return hipErrorInvalidValue;
    if (ntasks == 0) return hipSuccess;
    // A persistent grid: no more waves than fit the GPU at once (more would
    // only queue behind the first ones), nor more than there are tasks. The
    // caller passes the resident count when it launches under stream capture.
    uint32_t waves = max_waves ? max_waves : batch_resident_waves(nout);
    if (uint64_t(waves) > ntasks) waves = uint32_t(ntasks);
    uint32_t blocks = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
    if (blocks == 0) blocks = 1;
    switch (nout) {
    case 0: k_batch_probe<0><<<dim3(blocks), dim3(256), 0, s>>>(jobs, tasks, ntasks, queue); break;
    case 1: k_batch_probe<1><<<dim3(blocks), dim3(256), 0, s>>>(jobs, tasks, ntasks, queue); break;
    case 2: k_batch_probe<2><<<dim3(blocks), dim3(256), 0, s>>>(jobs, tasks, ntasks, queue); break;
    case 3: k_batch_probe<3><<<dim3(blocks), dim3(256), 0, s>>>(jobs, tasks, ntasks, queue); break;
    case 4: k_batch_probe<4><<<dim3(blocks), dim3(256), 0, s>>>(jobs, tasks, ntasks, queue); break;
    default: k_batch_build<<<dim3(blocks), dim3(256), 0, s>>>(jobs, tasks, ntasks, queue); break;
    }
    return hipGetLastError();
}

// The lean kernel applies when the pipeline has one probe without residual
// equalities, every scanned join column is the probe key (so NULL filtering is
// the key's), and the target is a count or at most four fixed-width columns.
bool probe1_applicable(const PipeArgs& a, uint32_t slots) {
    (void)slots;
    if (a.nprobes != 1 || a.probes[0].nchecks != 0 || a.probes[0].grouped || a.nscan > uint32_t(kMaxScan))
        return false;
    const ProbeSpec& P = a.probes[0];
    if (P.key_slot >= a.nscan || (a.not_null & ~(1u << P.key_slot)) != 0) return false;
    if (a.target == TARGET_COUNT) return true;
    if (a.target != TARGET_PAGES || a.nvarchar != 0 || a.nout > uint32_t(kLeanOut)) return false;
    for (uint32_t c = 0; c < a.nout; ++c) {
        if (a.out[c].type == T_VARCHAR) return false;
        const uint32_t slot = a.out[c].slot;
        if (slot == P.key_slot || slot < a.nscan) continue;
        bool found = false;
        for (uint32_t t = 0; t < P.nattrs; ++t) found = found || P.attr_slot[t] == slot;
        if (!found) return false;
    }
    return true;
}

// One thread per slot of a two-word table: key | attr << 32, or a flag when
// the entry does not fit (see launch_pack_table).
__global__ __launch_bounds__(256) void k_pack_table(const uint64_t* ht, uint64_t cap, uint64_t* packed,
                                                    uint32_t* flag) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    bool           bad = false;
    if (i < cap) {
        const ulonglong2 e = *reinterpret_cast<const ulonglong2*>(ht + 2 * i);
        uint64_t         w = kEmpty;
        if (e.x != kEmpty) {
            bad = (e.x >> 32) != 0 || (e.y >> 32) != 0 || (e.x == 0xffffffffull && e.y == 0xffffffffull);
            w   = e.x | (e.y << 32);
        }
        packed[i] = w;
    }
    if (__ballot(bad) && lead_lane()) atomicOr(flag, 1u); // one atomic per wave
}

// One workgroup per field: the sum over the grid's waves, added to the field.
__global__ __launch_bounds__(256) void k_fold_sums(const unsigned long long* ws, uint32_t stride, uint32_t nout,
                                                   PipeCounters* c) {
    __shared__ unsigned long long part[4];
    const uint32_t     f     = blockIdx.x; // 0: rows; 1..: nnz, then pages, of the used columns
    const uint32_t     field = f == 0 ? 0u : f <= nout ? f : 1u + kMaxOut + (f - 1 - nout);
    const uint64_t     n     = c->waves < stride ? c->waves : stride;
    unsigned long long s     = 0;
    for (uint64_t w = threadIdx.x; w < n; w += blockDim.x) s += ws[uint64_t(field) * stride + w];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if (lane_id() == 0) part[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = part[0] + part[1] + part[2] + part[3];
        if (f == 0)
            c->tuples += t;
        else if (f <= nout)
            c->nnz[f - 1] += t;
        else
            c->col_pages[f - 1 - nout] += t;
    }
}

hipError_t launch_fold_sums(const unsigned long long* wave_sums, uint32_t stride, uint32_t nout, PipeCounters* c,
                            hipStream_t s) {
    if (nout > uint32_t(kMaxOut)) return hipErrorInvalidValue;
    k_fold_sums<<<dim3(1 + 2 * nout), dim3(256), 0, s>>>(wave_sums, stride, nout, c);
    return hipGetLastError();
}

hipError_t launch_pack_table(const uint64_t* ht, uint64_t cap, uint64_t* packed, uint32_t* flag, hipStream_t s) {
    if (cap == 0) return hipSuccess;
    k_pack_table<<<dim3(uint32_t((cap + 255) / 256)), dim3(256), 0, s>>>(ht, cap, packed, flag);
    return hipGetLastError();
}

hipError_t launch_probe1(const PipeArgs& a, uint32_t slots, uint32_t grid_waves, bool aligned_pairs, hipStream_t s) {
    uint32_t blocks = (grid_waves + kWavesPerBlock - 1) / kWavesPerBlock;
    if (blocks == 0) blocks = 1;
    if (!probe1_applicable(a, slots)) return hipErrorInvalidValue;
    const ProbeSpec& P = a.probes[0];
    LeanArgs         l{};
    l.key         = a.scan[P.key_slot];
    l.tuples      = a.tuples;
    l.tuple_words = a.tuple_words;
    l.key_word    = a.tuple_word[P.key_slot];
    l.driver      = a.driver;
    l.morsel_rows = a.morsel_rows;
    l.xcd_order   = a.xcd_order;
    l.unique      = P.unique;
    l.nrows       = a.nrows;
    l.nmorsels    = a.nmorsels;
    l.ht          = P.ht;
    l.mask        = P.mask;
    l.shift       = P.shift;
    l.words       = P.words;
    l.bloom       = P.bloom;
    l.bloom_log   = P.bloom_log;
    l.nout        = a.target == TARGET_PAGES ? a.nout : 0;
    for (uint32_t c = 0; c < l.nout; ++c) {
        LeanOutCol&    o    = l.out[c];
        const uint32_t slot = a.out[c].slot;
        o.type              = a.out[c].type;
        if (slot == P.key_slot) {
            o.src = SRC_KEY;
        } else if (slot < a.nscan) {
            o.src  = SRC_SCAN;
            o.col  = a.scan[slot];
            o.word = a.tuple_word[slot];
        } else {
            o.src = SRC_HT;
            for (uint32_t t = 0; t < P.nattrs; ++t)
                if (P.attr_slot[t] == slot) o.word = P.attr_word[t];
        }
    }
    l.pool       = a.pool;
    l.pool_pages = a.pool_pages;
    l.dir        = a.dir;
    l.counters   = a.counters;
    l.page_chunk = a.page_chunk ? a.page_chunk : kPageChunk;
    l.aligned_pairs = aligned_pairs && P.words <= 2 ? 1u : 0u;
    l.abort_pages   = a.abort_pages;
    l.sel           = a.sel;
    l.static_chunk     = 1;
    l.wave_sums        = a.wave_sums;
    l.wave_sums_stride = a.wave_sums_stride;
    if (P.ht_packed && P.pack_flag && P.words == 2) {
        l.ht_packed = P.ht_packed;
        l.pack_flag = P.pack_flag;
    }
    // the launch shape whose flags can be compiled in
    int var = LEAN_ANY;
    if (l.driver == DRIVER_PAGES && !l.bloom && !l.sel && !l.xcd_order) {
        if (l.ht_packed)
            var = LEAN_PACKED;
        else if (l.words == 2 && l.aligned_pairs)
            var = LEAN_TWO;
    }
    // With the dynamic morsel schedule a grid larger than one resident wave
    // set only adds a tail: cap it at what the kernel's occupancy fits.
#define CGJ_LEAN_KERNELS(V)                                                                                  \
    {reinterpret_cast<const void*>(&k_lean<0, V>), reinterpret_cast<const void*>(&k_lean<1, V>),            \
     reinterpret_cast<const void*>(&k_lean<2, V>), reinterpret_cast<const void*>(&k_lean<3, V>),            \
     reinterpret_cast<const void*>(&k_lean<4, V>)}
    static const void* const kernels[3][kLeanOut + 1] = {CGJ_LEAN_KERNELS(LEAN_ANY), CGJ_LEAN_KERNELS(LEAN_TWO),
                                                          CGJ_LEAN_KERNELS(LEAN_PACKED)};
#undef CGJ_LEAN_KERNELS
    const uint32_t no = l.nout < uint32_t(kLeanOut) ? l.nout : uint32_t(kLeanOut);
    if (!a.xcd_order) {
        // resident workgroups per device and kernel, queried once (the
        // occupancy query costs more host time than a small launch)
        constexpr int              kMaxDev = 16;
        static std::atomic<uint32_t> resident_cache[kMaxDev][3][kLeanOut + 1];
        int                          dev = 0;
        if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < kMaxDev) {
            uint32_t resident = resident_cache[dev][var][no].load(std::memory_order_relaxed);
            if (resident == 0) {
                int cus = 0, per_cu = 0;
                if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
                    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernels[var][no], 256, 0) == hipSuccess &&
                    per_cu > 0 && cus > 0) {
                    resident = uint32_t(per_cu) * uint32_t(cus);
                    resident_cache[dev][var][no].store(resident, std::memory_order_relaxed);
                }
            }
            if (resident && blocks > resident) blocks = resident;
        }
    }
    // per-wave sums are stored at index `wave` < stride
    if (l.wave_sums && uint64_t(blocks) * kWavesPerBlock > l.wave_sums_stride) return hipErrorInvalidValue;
    const dim3 g(blocks), b(256);
    switch (var * 8 + int(no)) {
    case LEAN_ANY * 8 + 0: k_lean<0, LEAN_ANY><<<g, b, 0, s>>>(l); break;
    case LEAN_ANY * 8 + 1: k_lean<1, LEAN_ANY><<<g, b, 0, s>>>(l); break;
    case LEAN_ANY * 8 + 2: k_lean<2, LEAN_ANY><<<g, b, 0, s>>>(l); break;
    case LEAN_ANY * 8 + 3: k_lean<3, LEAN_ANY><<<g, b, 0, s>>>(l); break;
    case LEAN_ANY * 8 + 4: k_lean<4, LEAN_ANY><<<g, b, 0, s>>>(l); break;
    case LEAN_TWO * 8 + 0: k_lean<0, LEAN_TWO><<<g, b, 0, s>>>(l); break;
    case LEAN_TWO * 8 + 1: k_lean<1, LEAN_TWO><<<g, b, 0, s>>>(l); break;
    case LEAN_TWO * 8 + 2: k_lean<2, LEAN_TWO><<<g, b, 0, s>>>(l); break;
    case LEAN_TWO * 8 + 3: k_lean<3, LEAN_TWO><<<g, b, 0, s>>>(l); break;
    case LEAN_TWO * 8 + 4: k_lean<4, LEAN_TWO><<<g, b, 0, s>>>(l); break;
    case LEAN_PACKED * 8 + 0: k_lean<0, LEAN_PACKED><<<g, b, 0, s>>>(l); break;
    case LEAN_PACKED * 8 + 1: k_lean<1, LEAN_PACKED><<<g, b, 0, s>>>(l); break;
    case LEAN_PACKED * 8 + 2: k_lean<2, LEAN_PACKED><<<g, b, 0, s>>>(l); break;
    case LEAN_PACKED * 8 + 3: k_lean<3, LEAN_PACKED><<<g, b, 0, s>>>(l); break;
    default: k_lean<4, LEAN_PACKED><<<g, b, 0, s>>>(l); break;
    }
    return hipGetLastError();
}

hipError_t launch_pipeline(const PipeArgs& a, uint32_t slots, uint32_t grid_waves, hipStream_t s) {
    uint32_t blocks = (grid_waves + kWavesPerBlock - 1) / kWavesPerBlock;
    if (blocks == 0) blocks = 1;
    // per-wave sums are stored at index `wave` < stride
    if (a.wave_sums && uint64_t(blocks) * kWavesPerBlock > a.wave_sums_stride) return hipErrorInvalidValue;
    size_t lds = size_t(kWavesPerBlock) * (kMaxOut * 6 + a.nout * 64) * sizeof(uint64_t);
    if (slots <= 4)
        k_pipeline<4><<<dim3(blocks), dim3(256), lds, s>>>(a);
    else if (slots <= 8)
        k_pipeline<8><<<dim3(blocks), dim3(256), lds, s>>>(a);
    else if (slots <= 16)
        k_pipeline<16><<<dim3(blocks), dim3(256), lds, s>>>(a);
    else if (slots <= 24)
        k_pipeline<24><<<dim3(blocks), dim3(256), lds, s>>>(a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// join sampling: single-row decode of sample rows, and their membership in a
// built hash table
// ---------------------------------------------------------------------------
// Row `row` of a column, decoded by one thread: the block index gives the
// page, the validity bit and the non-NULL offset (dense columns need none).
__device__ inline uint64_t decode_one(const ScanCol& c, uint64_t row) {
    if (c.dense_rows) {
        const uint32_t d = c.dense_rows;
        return load_value(c, uint32_t(row / d), uint32_t(row % d), false);
    }
    const BlockIdx bi   = c.blocks[row / 64];
    const uint32_t lane = uint32_t(row % 64);
    if (bi.complex) return decode_row_slow(c, bi.page, uint32_t(row));
    if (!((bi.valid >> lane) & 1ull)) return kNull;
    const uint64_t before = bi.valid & ((1ull << lane) - 1ull);
    if (lane < bi.split) return load_value(c, bi.page, bi.nn + uint32_t(__popcll(before)), false);
    const uint64_t second = before & ~((bi.split >= 64 ? ~0ull : (1ull << bi.split)) - 1ull);
    return load_value(c, bi.page + 1, uint32_t(__popcll(second)), false);
}

__global__ __launch_bounds__(64) void k_sample_gather(const SampleCol* cols, const uint64_t* rowids, uint64_t* out) {
    const SampleCol& sc = cols[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < sc.n; i += blockDim.x) out[sc.first + i] = decode_one(sc.col, rowids[sc.first + i]);
}

hipError_t launch_sample_gather(const SampleCol* cols, uint32_t ncols, const uint64_t* rowids, uint64_t* out,
                                hipStream_t s) {
    if (ncols == 0) return hipSuccess;
    k_sample_gather<<<dim3(ncols), dim3(64), 0, s>>>(cols, rowids, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// scan restrictions: one wave per 64-row block, the ballot of the rows that
// pass is the block's selection word
// ---------------------------------------------------------------------------
__device__ inline bool filter_pass(const FilterSpec& f, uint64_t v) {
    if (v == kNull) return false;
    switch (f.kind) {
    case FLT_EQ: return v == f.a;
    case FLT_EQ2: return v == f.a || v == f.b;
    case FLT_RANGE: return v >= f.a && v <= f.b;
    case FLT_TINY: return f.tiny[tiny_slot(v, f.tiny_hash, f.tiny_bits)] == v;
    case FLT_BLOOM: {
        const uint64_t h2 = hash_bloom(hash_key(v));
        const uint64_t bb = bloom_bits(h2);
        return (f.table[bloom_word(h2, f.shift)] & bb) == bb;
    }
    default: { // FLT_TABLE: walk to the key or an empty slot (tables hold an empty slot per region)
        uint64_t p = hash_key(v) >> f.shift;
        for (uint64_t step = 0; step <= f.mask; ++step) {
            const uint64_t k = f.table[p * f.words];
            if (k == v) return true;
            if (k == kEmpty) return false;
            p = (p & ~f.mask) | ((p + 1) & f.mask);
        }
        return false;
    }
    }
}

__global__ __launch_bounds__(256) void k_filter_select(FilterArgs a, uint64_t* sel) {
    const uint64_t nb     = (a.nrows + 63) / 64;
    const uint64_t nwaves = uint64_t(gridDim.x) * (blockDim.x / 64);
    for (uint64_t b = uint64_t(blockIdx.x) * (blockDim.x / 64) + wave_in_block(); b < nb; b += nwaves) {
        const bool active = b * 64 + lane_id() < a.nrows;
        bool       ok     = active;
        for (uint32_t i = 0; i < a.n; ++i) {
            if (!__ballot(ok)) break;
            const uint64_t v = decode_block(a.f[i].col, b, active);
            ok               = ok && filter_pass(a.f[i], v);
        }
        const uint64_t m = __ballot(ok);
        if (lead_lane()) sel[b] = a.accumulate ? (sel[b] & m) : m;
    }
}

hipError_t launch_filter_select(const FilterArgs& a, uint64_t* sel, hipStream_t s) {
    const uint64_t nb = (a.nrows + 63) / 64;
    if (nb == 0) return hipSuccess;
    const uint64_t blocks = (nb + 3) / 4 < 8192 ? (nb + 3) / 4 : 8192;
    k_filter_select<<<dim3(uint32_t(blocks)), dim3(256), 0, s>>>(a, sel);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_table_keys(const uint64_t* ht, uint64_t cap, uint32_t stride, uint64_t* out,
                                                    uint32_t max, unsigned int* count) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < cap; i += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t k = ht[i * stride];
        if (k == kEmpty) continue;
        const unsigned int at = atomicAdd(count, 1u);
        if (at < max) out[at] = k;
    }
}

hipError_t launch_table_keys(const uint64_t* ht, uint64_t cap, uint32_t stride, uint64_t* out, uint32_t max,
                             unsigned int* count, hipStream_t s) {
    const uint64_t blocks = (cap + 255) / 256 < 4096 ? (cap + 255) / 256 : 4096;
    k_table_keys<<<dim3(uint32_t(blocks ? blocks : 1)), dim3(256), 0, s>>>(ht, cap, stride, out, max, count);
    return hipGetLastError();
}

// Linear probing from the key's home slot to a match or an empty slot, at
// most one region's worth of slots.
__global__ __launch_bounds__(256) void k_sample_probe(SampleTable t, const uint64_t* keys, uint32_t n,
                                                      uint32_t* found) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t key = keys[i];
    uint32_t       hit = 0;
    if (key != kNull) {
        uint64_t p = hash_key(key) >> t.shift;
        for (uint64_t step = 0; step <= t.mask; ++step) {
            const uint64_t k = t.ht[p * t.words];
            if (k == key) {
                hit = 1;
                break;
            }
            if (k == kEmpty) break;
            p = (p & ~t.mask) | ((p + 1) & t.mask);
        }
    }
    found[i] = hit;
}

hipError_t launch_sample_probe(SampleTable t, const uint64_t* keys, uint32_t n, uint32_t* found, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_sample_probe<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(t, keys, n, found);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// synthetic page generation: one block per page
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gen_fill(GenSpec g, const uint32_t* row_start, uint8_t* pages) {
    __shared__ uint32_t sh[17];
    __shared__ uint64_t bm[1024]; // bitmap of up to 65536 rows
    uint64_t p   = blockIdx.x;
    uint32_t r0  = row_start[p], r1 = row_start[p + 1];
    uint32_t n   = r1 - r0;
    uint8_t* page = pages + p * kPageSize;
    uint32_t w    = g.type == T_INT32 ? 4 : 8;
    uint32_t off  = g.type == T_INT32 ? 4 : 8;
    for (uint32_t i = threadIdx.x; i < 1024; i += 256) bm[i] = 0;
    __syncthreads();
    uint32_t nv = 0; // running non-NULL count (uniform)
    for (uint32_t c0 = 0; c0 < n; c0 += 256) {
        uint32_t j     = c0 + threadIdx.x;
        bool     in    = j < n;
        uint64_t r     = uint64_t(r0) + j;
        bool     valid = in && !gen_is_null(g.seed, g.null_stream, g.null_frac, r);
        // block prefix of valid flags (4 waves)
        uint32_t wv = threadIdx.x / 64;
        uint64_t b    = __ballot(valid);
        if (lead_lane()) {
            sh[wv] = __popcll(b);
            if (in) bm[j / 64] = b; // j is a multiple of 64 for lane 0
        }
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (uint32_t k = 0; k < 4; ++k) {
            if (k < wv) before += sh[k];
            total += sh[k];
        }
        uint32_t idx = nv + before + __popcll(b & mask_lt());
        if (valid) {
            uint64_t v = gen_value(g.type, g.kind, g.seed, g.stream, g.domain, g.zipf_s, g.miss_frac, g.lo, g.hi, r);
            if (w == 4)
                *reinterpret_cast<uint32_t*>(page + off + 4 * idx) = uint32_t(v);
            else
                *reinterpret_cast<uint64_t*>(page + off + 8 * uint64_t(idx)) = v;
        }
        nv += total;
        __syncthreads();
    }
    __syncthreads();
    __threadfence_block();
    // bitmap last (the contest writer copies it over the page tail after the values)
    uint32_t nb = (n + 7) / 8;
    uint8_t* dst = page + kPageSize - nb;
    for (uint32_t i = threadIdx.x; i < nb; i += 256) dst[i] = uint8_t(bm[i / 8] >> (8 * (i % 8)));
    if (threadIdx.x == 0) {
        *reinterpret_cast<uint16_t*>(page)     = uint16_t(n);
        *reinterpret_cast<uint16_t*>(page + 2) = uint16_t(nv);
    }
}

hipError_t launch_gen_fill(const GenSpec& g, const uint32_t* row_start, uint64_t npages, uint8_t* pages,
                           hipStream_t s) {
    if (npages == 0) return hipSuccess;
    k_gen_fill<<<dim3(npages), dim3(256), 0, s>>>(g, row_start, pages);
    return hipGetLastError();
}

} // namespace cgj
