// The cardinality estimator's network: x (in features) -> 256 -> 256 -> 2,
// ReLU between layers; the two outputs are the predicted log10 cardinality
// and the log of its standard deviation (the estimator's confidence).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.hpp"

namespace cgj {

constexpr uint32_t kMlpHidden = 256;
constexpr uint32_t kMlpOut    = 2;
constexpr uint32_t kMlpMaxIn  = 64;

hipError_t launch_gemm(uint32_t M, uint32_t N, uint32_t K, const float* A, uint32_t lda, bool ta, const float* B,
                       uint32_t ldb, bool tb, float* C, uint32_t ldc, const float* bias, bool relu, hipStream_t s);
hipError_t launch_relu_grad(float* d, const float* act, uint64_t n, hipStream_t s);
hipError_t launch_colsum(const float* d, uint32_t M, uint32_t N, float* out, hipStream_t s);
hipError_t launch_nll(const float* out, const float* y, uint32_t M, float* dout, float* loss_sum, hipStream_t s);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, uint64_t n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, float wd, hipStream_t s);
hipError_t launch_gather(const float* X, const float* Y, const uint32_t* idx, uint32_t M, uint32_t in, float* Xb,
                         float* Yb, hipStream_t s);
hipError_t launch_to_bf16(const float* src, uint16_t* dst, uint64_t n, hipStream_t s);
// W1: in_pad x 256 (rows beyond `in` zero), W2: 256 x 256, both BF16 row-major.
hipError_t launch_mlp_infer(const float* X, uint32_t n, uint32_t in, const uint16_t* W1, const float* b1,
                            const uint16_t* W2, const float* b2, const float* W3, const float* b3, float* out,
                            hipStream_t s);

} // namespace cgj
