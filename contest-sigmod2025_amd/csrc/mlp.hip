// Kernels of the cardinality estimator's MLP (x -> 256 -> 256 -> 2, ReLU):
// FP32 training (tiled GEMM with fused bias/ReLU, backward helpers,
// heteroscedastic loss, Adam) and batched inference with BF16 inputs and FP32
// accumulation on the MFMA matrix cores (gfx950's v_mfma_f32_32x32x16_bf16).
#include <hip/hip_runtime.h>

#include "mlp.hpp"

namespace cgj {

// ---------------------------------------------------------------------------
// FP32 GEMM: C[M x N] = op(A)[M x K] op(B)[K x N] (+ bias[N]) (ReLU)
// op(A) = A (M x K, row-major, lda) or A^T (A stored K x M); same for B.
// 64 x 64 output tile per 256-thread workgroup, 4 x 4 per thread, K in steps
// of 16 through LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gemm(uint32_t M, uint32_t N, uint32_t K, const float* A, uint32_t lda,
                                              uint32_t ta, const float* B, uint32_t ldb, uint32_t tb, float* C,
                                              uint32_t ldc, const float* bias, uint32_t relu) {
    __shared__ float As[16][65];
    __shared__ float Bs[16][65];
    const uint32_t tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const uint32_t m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    float          acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
    for (uint32_t k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint32_t idx = threadIdx.x + e * 256;
            // A tile 64 (m) x 16 (k)
            {
                const uint32_t mm = ta ? idx % 64 : idx / 16, kk = ta ? idx / 64 : idx % 16;
                const uint32_t gm = m0 + mm, gk = k0 + kk;
                float          v  = 0.f;
                if (gm < M && gk < K) v = ta ? A[uint64_t(gk) * lda + gm] : A[uint64_t(gm) * lda + gk];
                As[kk][mm] = v;
            }
            // B tile 16 (k) x 64 (n)
            {
                const uint32_t nn = tb ? idx / 16 : idx % 64, kk = tb ? idx % 16 : idx / 64;
                const uint32_t gn = n0 + nn, gk = k0 + kk;
                float          v  = 0.f;
                if (gn < N && gk < K) v = tb ? B[uint64_t(gn) * ldb + gk] : B[uint64_t(gk) * ldb + gn];
                Bs[kk][nn] = v;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t gm = m0 + ty * 4 + i;
        if (gm >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t gn = n0 + tx * 4 + j;
            if (gn >= N) continue;
            float v = acc[i][j] + (bias ? bias[gn] : 0.f);
            if (relu) v = v > 0.f ? v : 0.f;
            C[uint64_t(gm) * ldc + gn] = v;
        }
    }
}

hipError_t launch_gemm(uint32_t M, uint32_t N, uint32_t K, const float* A, uint32_t lda, bool ta, const float* B,
                       uint32_t ldb, bool tb, float* C, uint32_t ldc, const float* bias, bool relu, hipStream_t s) {
    if (M == 0 || N == 0) return hipSuccess;
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    k_gemm<<<grid, dim3(256), 0, s>>>(M, N, K, A, lda, ta ? 1u : 0u, B, ldb, tb ? 1u : 0u, C, ldc, bias, relu ? 1u : 0u);
    return hipGetLastError();
}

// dZ = dA * (act > 0): the ReLU's gradient, from its output.
__global__ void k_relu_grad(float* d, const float* act, uint64_t n) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && !(act[i] > 0.f)) d[i] = 0.f;
}

// out[j] = sum_i d[i][j] (bias gradients); one thread per column.
__global__ void k_colsum(const float* d, uint32_t M, uint32_t N, float* out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.f;
    for (uint32_t i = 0; i < M; ++i) s += d[uint64_t(i) * N + j];
    out[j] = s;
}

// Heteroscedastic Gaussian loss per row of out = (mu, log_sigma) against y:
//   l = 0.5 (y - mu)^2 exp(-2 s) + s, averaged over the M rows.
// Writes d out / d (mu, s) and adds the summed loss to *loss_sum.
__global__ void k_nll(const float* out, const float* y, uint32_t M, float* dout, float* loss_sum) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float          l = 0.f;
    if (i < M) {
        const float mu = out[2 * i], s = out[2 * i + 1];
        const float r  = y[i] - mu;
        const float w  = expf(-2.f * s);
        l              = 0.5f * r * r * w + s;
        dout[2 * i]     = -r * w / float(M);
        dout[2 * i + 1] = (1.f - r * r * w) / float(M);
    }
    // wave sum, then one atomic per wave
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) l += __shfl_xor(l, d, 64);
    if (lead_lane()) atomicAdd(loss_sum, l);
}

// Adam step (Kingma & Ba, 2015) with L2 weight decay added to the gradient.
__global__ void k_adam(float* p, const float* g, float* m, float* v, uint64_t n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, float wd) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * p[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i]           = mi;
    v[i]           = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// Rows idx[0..M) of X (n x in) and Y into Xb / Yb.
__global__ void k_gather(const float* X, const float* Y, const uint32_t* idx, uint32_t M, uint32_t in, float* Xb,
                         float* Yb) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= uint64_t(M) * in) return;
    const uint32_t r = uint32_t(t / in), c = uint32_t(t % in);
    const uint32_t src = idx[r];
    Xb[t]              = X[uint64_t(src) * in + c];
    if (c == 0) Yb[r] = Y[src];
}

// ---------------------------------------------------------------------------
// BF16 inference
// ---------------------------------------------------------------------------
__device__ inline uint16_t f32_to_bf16(float f) {
    // round to nearest even on the upper 16 bits (inputs are finite)
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

__global__ void k_to_bf16(const float* src, uint16_t* dst, uint64_t n) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = f32_to_bf16(src[i]);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float  f32x16 __attribute__((ext_vector_type(16)));

// One 32-row block of activations (LDS, bf16, row stride `lds_stride`, K
// columns, zero past the real width) times a Kw x 256 bf16 weight matrix
// (global, row-major; rows past Kw read as zero) for the 32-column output
// tiles t of this wave; bias + ReLU, result to LDS as bf16.
// gfx950's double-K MFMA 32x32x16: lane l holds A[l % 32][8 (l / 32) + 0..7]
// and B[8 (l / 32) + 0..7][l % 32]; accumulator item r is
// D[(r % 4) + 8 (r / 4) + 4 (l / 32)][l % 32].
__device__ inline void mfma_layer(const uint16_t* act, uint32_t lds_stride, uint32_t K, uint32_t Kw,
                                  const uint16_t* W, const float* bias, uint16_t* out, uint32_t out_stride,
                                  uint32_t tile_first, uint32_t tiles) {
    const uint32_t lane = lane_id(), row = lane % 32, kq = 8 * (lane / 32);
    for (uint32_t t = tile_first; t < tile_first + tiles; ++t) {
        const uint32_t col = t * 32 + row;
        f32x16         acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = 0.f;
        for (uint32_t k0 = 0; k0 < K; k0 += 16) {
            bf16x8 a, b;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const uint32_t k = k0 + kq + q;
                a[q]             = __builtin_bit_cast(__bf16, act[row * lds_stride + k]);
                b[q]             = __builtin_bit_cast(__bf16, uint16_t(k < Kw ? W[uint64_t(k) * kMlpHidden + col] : 0));
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const uint32_t i = (r % 4) + 8 * (r / 4) + 4 * (lane / 32);
            float          v = acc[r] + bias[col];
            out[i * out_stride + col] = f32_to_bf16(v > 0.f ? v : 0.f);
        }
    }
}

// X: n x in (FP32, converted to BF16 on load), in_pad = in rounded up to 16.
// Layers 1-2 on MFMA, layer 3 (256 -> 2) in FP32 from the BF16 activations.
__global__ __launch_bounds__(256) void k_mlp_infer(const float* X, uint32_t n, uint32_t in, uint32_t in_pad,
                                                   const uint16_t* W1, const float* b1, const uint16_t* W2,
                                                   const float* b2, const float* W3, const float* b3, float* out) {
    __shared__ uint16_t xs[32 * kMlpMaxIn];
    __shared__ uint16_t h1[32 * (kMlpHidden + 8)];
    __shared__ uint16_t h2[32 * (kMlpHidden + 8)];
    constexpr uint32_t  hs = kMlpHidden + 8; // row stride (skews LDS banks)
    const uint32_t      r0 = blockIdx.x * 32;
    for (uint32_t t = threadIdx.x; t < 32 * in_pad; t += blockDim.x) {
        const uint32_t r = t / in_pad, c = t % in_pad;
        const float    v = (r0 + r < n && c < in) ? X[uint64_t(r0 + r) * in + c] : 0.f;
        xs[r * in_pad + c] = f32_to_bf16(v);
    }
    __syncthreads();
    const uint32_t wave = wave_in_block(); // 4 waves x 2 tiles of 32 columns = 256
    mfma_layer(xs, in_pad, in_pad, in, W1, b1, h1, hs, wave * 2, 2);
    __syncthreads();
    mfma_layer(h1, hs, kMlpHidden, kMlpHidden, W2, b2, h2, hs, wave * 2, 2);
    __syncthreads();
    if (wave == 0) {
        const uint32_t lane = lane_id(), r = lane % 32, o = lane / 32;
        float          s    = b3[o];
        for (uint32_t k = 0; k < kMlpHidden; ++k) s = fmaf(bf16_to_f32(h2[r * hs + k]), W3[k * kMlpOut + o], s);
        if (r0 + r < n) out[uint64_t(r0 + r) * kMlpOut + o] = s;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static uint32_t blocks_for(uint64_t n, uint32_t threads) { return uint32_t((n + threads - 1) / threads); }

hipError_t launch_relu_grad(float* d, const float* act, uint64_t n, hipStream_t s) {
    if (!n) return hipSuccess;
    k_relu_grad<<<blocks_for(n, 256), 256, 0, s>>>(d, act, n);
    return hipGetLastError();
}
hipError_t launch_colsum(const float* d, uint32_t M, uint32_t N, float* out, hipStream_t s) {
    k_colsum<<<blocks_for(N, 256), 256, 0, s>>>(d, M, N, out);
    return hipGetLastError();
}
hipError_t launch_nll(const float* out, const float* y, uint32_t M, float* dout, float* loss_sum, hipStream_t s) {
    k_nll<<<blocks_for(M, 256), 256, 0, s>>>(out, y, M, dout, loss_sum);
    return hipGetLastError();
}
hipError_t launch_adam(float* p, const float* g, float* m, float* v, uint64_t n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, float wd, hipStream_t s) {
    k_adam<<<blocks_for(n, 256), 256, 0, s>>>(p, g, m, v, n, lr, b1, b2, eps, bc1, bc2, wd);
    return hipGetLastError();
}
hipError_t launch_gather(const float* X, const float* Y, const uint32_t* idx, uint32_t M, uint32_t in, float* Xb,
                         float* Yb, hipStream_t s) {
    k_gather<<<blocks_for(uint64_t(M) * in, 256), 256, 0, s>>>(X, Y, idx, M, in, Xb, Yb);
    return hipGetLastError();
}
hipError_t launch_to_bf16(const float* src, uint16_t* dst, uint64_t n, hipStream_t s) {
    k_to_bf16<<<blocks_for(n, 256), 256, 0, s>>>(src, dst, n);
    return hipGetLastError();
}
hipError_t launch_mlp_infer(const float* X, uint32_t n, uint32_t in, const uint16_t* W1, const float* b1,
                            const uint16_t* W2, const float* b2, const float* W3, const float* b3, float* out,
                            hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (in > kMlpMaxIn) return hipErrorInvalidValue;
    const uint32_t in_pad = (in + 15) / 16 * 16;
    k_mlp_infer<<<blocks_for(n, 32), 256, 0, s>>>(X, n, in, in_pad, W1, b1, W2, b2, W3, b3, out);
    return hipGetLastError();
}

} // namespace cgj
