// Counter-based synthetic data recipes, evaluated identically on host and
// device. Row r of a column draws from SplitMix64 streams (Steele, Lea and
// Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014)
// addressed by (seed, stream, r), so any row can be generated independently.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace cgj {

__host__ __device__ inline uint64_t sm_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ull;

__host__ __device__ inline uint64_t stream_base(uint64_t seed, uint64_t stream) {
    return sm_mix(seed ^ sm_mix(stream * 0xd1342543de82ef95ull + 1));
}
// r-th output of the stream (SplitMix64 state after r+1 increments).
__host__ __device__ inline uint64_t draw(uint64_t base, uint64_t r) { return sm_mix(base + (r + 1) * kGolden); }
__host__ __device__ inline double unit(uint64_t x) { return double(x >> 11) * (1.0 / 9007199254740992.0); }

__host__ __device__ inline bool gen_is_null(uint64_t seed, uint64_t null_stream, double frac, uint64_t r) {
    if (frac <= 0.0) return false;
    return unit(draw(stream_base(seed, null_stream), r)) < frac;
}

// Pseudo-random permutation of [0, n): 4-round Feistel network over the
// smallest even bit width covering n, with cycle walking.
__host__ __device__ inline uint64_t perm_index(uint64_t x, uint64_t n, uint64_t key) {
    uint32_t bits = 2;
    while ((1ull << bits) < n) bits += 2;
    uint32_t half = bits / 2;
    uint64_t hm   = (1ull << half) - 1;
    do {
        uint64_t l = x >> half, r = x & hm;
        for (uint32_t i = 0; i < 4; ++i) {
            uint64_t f  = sm_mix(r ^ (key + i * kGolden)) & hm;
            uint64_t nl = r;
            r           = l ^ f;
            l           = nl;
        }
        x = (l << half) | r;
    } while (x >= n);
    return x;
}

// Unique 64-bit key of build rank r: a bijection of r (never equal for
// different r, so keys of ranks >= domain never match a build key).
__host__ __device__ inline uint64_t unique_key64(uint64_t base, uint64_t r) { return sm_mix(r + base); }

// Rank drawn from a bounded Zipf(s) over [0, n) by inverting the continuous
// CDF of the density (x+1)^-s on [0, n): rank 0 is the most frequent.
__host__ __device__ inline uint64_t zipf_rank(double u, uint64_t n, double s) {
    double x;
    if (s <= 0.0) {
        x = u * double(n);
    } else if (fabs(s - 1.0) < 1e-9) {
        x = exp(u * log(double(n) + 1.0)) - 1.0;
    } else {
        double a = 1.0 - s;
        double t = pow(double(n) + 1.0, a);
        x        = pow(u * (t - 1.0) + 1.0, 1.0 / a) - 1.0;
    }
    uint64_t k = x <= 0.0 ? 0 : uint64_t(x);
    return k >= n ? n - 1 : k;
}

enum : int32_t { G_UNIQUE = 0, G_FK_ZIPF = 1, G_UNIFORM = 2, G_UNIT = 3, G_PERM = 4, G_FK_UNIFORM = 5 };

// Value bits (64-bit slot form) of row r; INT32 values use the low 32 bits.
__host__ __device__ inline uint64_t gen_value(int32_t type, int32_t kind, uint64_t seed, uint64_t stream,
                                              uint64_t domain, double zipf_s, double miss_frac, uint64_t lo,
                                              uint64_t hi, uint64_t r) {
    uint64_t base = stream_base(seed, stream);
    uint64_t x    = draw(base, r);
    switch (kind) {
    case G_UNIQUE: return type == 0 ? perm_index(r % domain, domain, base) : unique_key64(base, r);
    case G_PERM: return perm_index(r % domain, domain, base);
    case G_FK_ZIPF:
    case G_FK_UNIFORM: {
        bool     miss = unit(sm_mix(x ^ 0x5851f42d4c957f2dull)) < miss_frac;
        uint64_t rank = kind == G_FK_ZIPF ? zipf_rank(unit(x), domain, zipf_s) : (x % domain);
        if (type == 0) {
            // INT32 keys: build keys are a permutation of [0, domain); misses lie above it.
            if (miss) return domain + (sm_mix(x) % (0x7fffffffull - domain));
            return perm_index(rank, domain, base);
        }
        if (miss) return unique_key64(base, domain + (sm_mix(x) & ((1ull << 40) - 1)));
        return unique_key64(base, rank);
    }
    case G_UNIFORM: return lo + (hi > lo ? x % (hi - lo) : 0);
    default: {
        double d = unit(x);
        uint64_t b;
        memcpy(&b, &d, 8);
        return b;
    }
    }
}

// The contest page writer's rule for fixed-width columns: before appending a
// row to a page holding `rows` rows whose values end at `value_end`, the page
// is closed when the row would not fit (reference include/plan.h:204-221).
__host__ __device__ inline bool page_full(uint32_t rows, uint32_t value_end, bool valid) {
    return valid ? (value_end + 4 + rows / 8 + 1 > 8192) : (value_end + rows / 8 + 1 > 8192);
}

} // namespace cgj
