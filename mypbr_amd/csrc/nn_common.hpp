// nn_common.hpp - BF16 matrix-core building blocks shared by the learned
// parts (learn.hip: the kernel-predicting denoiser; nrc.hip: the neural
// radiance cache): BF16 conversion, the MFMA GEMM, ReLU and bias
// gradients, Adam. Each translation unit that includes this header gets
// its own copy of the kernels (internal linkage).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/mpt.h"

namespace mpt {

int set_error(int code, const std::string& m);

namespace nn {
namespace {

typedef uint16_t bf16;
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (NaN kept quiet)
__host__ __device__ __forceinline__ bf16 to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return bf16((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return bf16(u >> 16);
}
__host__ __device__ __forceinline__ float from_bf16(bf16 h) {
    uint32_t u = uint32_t(h) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// ------------------------------------------------------------------ GEMM
// The matrix-core instruction: v_mfma_f32_16x16x32_bf16 (gfx950), a 16 x 16
// FP32 tile += A (16 x 32, BF16) x B (32 x 16, BF16). Lane l supplies row
// (l % 16) of A and column (l % 16) of B, eight BF16 each, and receives
// column (l % 16), rows 4 * (l / 16) .. + 3 of the result. The eight values a
// lane supplies are the k of its slot group l / 16; the instruction pairs A's
// and B's slots position by position, so any assignment of k to the slots
// that A and B share gives the same sum. Here slot group g holds the eight
// consecutive k = 8 g .. 8 g + 7 of the 32, one 16-byte load.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float4_t mfma32(const bf16x8_t& a, const bf16x8_t& b, const float4_t& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8_t as_bf16x8(const uint4& u) {
    bf16x8_t v;
    __builtin_memcpy(&v, &u, 16);
    return v;
}

// C[m][n] = (accumulate ? C : 0) + sum_k A(m, k) B(n, k), FP32 accumulation.
// A(m, k) is A[m * lda + k] (AT = false) or A[k * lda + m] (AT = true); the
// same for B. A block of 256 threads computes a 64 x 64 tile of C, each wave
// 32 x 32 as 2 x 2 MFMA tiles, 32 k per step.
// Epilogue (EPI): 0 store FP32; 1 add bias[n], store BF16; 2 add bias[n],
// ReLU, store BF16; 3 add bias[n], store FP32; 4 atomically add FP32 (the
// K range is split over gridDim.z blocks of kchunk each).
// Fragment loads: 8 consecutive k of one row are one 16-byte load when the
// operand is k-contiguous and K and its leading dimension are multiples of
// 32 and 8; otherwise 8 element loads with bounds checks.
template <bool T>
__device__ __forceinline__ bf16x8_t frag(const bf16* __restrict__ X, int64_t ld, int64_t row, int64_t nrows, int k,
                                         int K, bool fast) {
    if (!T && fast) {
        uint4 u{0, 0, 0, 0};
        if (row < nrows) u = *reinterpret_cast<const uint4*>(X + row * ld + k);
        return as_bf16x8(u);
    }
    uint16_t v[8];
#pragma unroll
    for (int t = 0; t < 8; t++) {
        bf16 x = 0;
        if (row < nrows && k + t < K) x = T ? X[int64_t(k + t) * ld + row] : X[row * ld + k + t];
        v[t] = x;
    }
    bf16x8_t f;
    __builtin_memcpy(&f, v, 16);
    return f;
}

template <bool AT, bool BT, int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const bf16* __restrict__ A, int64_t lda,
                                              const bf16* __restrict__ B, int64_t ldb, void* __restrict__ Cv,
                                              int64_t ldc, const float* __restrict__ bias, int accumulate,
                                              int kchunk) {
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int64_t m0 = int64_t(blockIdx.x) * 64 + (wv & 1) * 32;
    const int64_t n0 = int64_t(blockIdx.y) * 64 + (wv >> 1) * 32;
    const int r = lane & 15;
    const int kq = (lane >> 4) * 8;
    const int kb = int(blockIdx.z) * kchunk;
    const int ke = K - kb < kchunk ? K : kb + kchunk;
    const bool fa = (K % 32) == 0 && (lda % 8) == 0, fb = (K % 32) == 0 && (ldb % 8) == 0;
    float4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += 32) {
        bf16x8_t a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) a[i] = frag<AT>(A, lda, m0 + 16 * i + r, M, k0 + kq, ke, fa);
#pragma unroll
        for (int j = 0; j < 2; j++) b[j] = frag<BT>(B, ldb, n0 + 16 * j + r, N, k0 + kq, ke, fb);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = mfma32(a[i], b[j], acc[i][j]);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int64_t n = n0 + 16 * j + r;
            if (n >= N) continue;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int64_t m = m0 + 16 * i + (lane >> 4) * 4 + t;
                if (m >= M) continue;
                float v = acc[i][j][t];
                if constexpr (EPI == 0) {
                    float* C = static_cast<float*>(Cv);
                    C[m * ldc + n] = accumulate ? C[m * ldc + n] + v : v;
                } else if constexpr (EPI == 4) {
                    atomicAdd(static_cast<float*>(Cv) + m * ldc + n, v);
                } else if constexpr (EPI == 3) {
                    static_cast<float*>(Cv)[m * ldc + n] = v + bias[n];
                } else {
                    v += bias[n];
                    if (EPI == 2 && !(v > 0.f)) v = 0.f;
                    static_cast<bf16*>(Cv)[m * ldc + n] = to_bf16(v);
                }
            }
        }
}

// G = bf16(dY * (Y > 0)) for a ReLU layer (relu = 1), or bf16(dY), over
// n = npix * C elements [pixel][channel]; with db, also the bias gradient
// db[c] += sum_p G[p][c] (FP32 over the BF16 values; db zeroed by the
// caller). The block holds (256 / C) * C threads and the grid stride is a
// multiple of C, so each thread keeps one channel: its partial sum stays in
// a register, the block adds them per channel in LDS, and one atomic per
// channel and block goes to db.
__global__ __launch_bounds__(256) void k_relu_grad(int64_t n, int C, const float* __restrict__ dY,
                                                   const bf16* __restrict__ Y, int relu, bf16* __restrict__ G,
                                                   float* __restrict__ db) {
    __shared__ float part[256];
    float s = 0.f;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += stride) {
        float g = dY[e];
        if (relu && !(from_bf16(Y[e]) > 0.f)) g = 0.f;
        const bf16 b = to_bf16(g);
        G[e] = b;
        s += from_bf16(b);
    }
    if (!db) return;
    part[threadIdx.x] = s;
    __syncthreads();
    if (int(threadIdx.x) < C) {
        float t = 0.f;
        for (int k = int(threadIdx.x); k < int(blockDim.x); k += C) t += part[k];
        atomicAdd(db + threadIdx.x, t);
    }
}

__global__ void k_to_bf16(int64_t n, const float* __restrict__ x, bf16* __restrict__ y) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < n) y[e] = to_bf16(x[e]);
}

// Adam (Kingma & Ba 2015) on FP32 master weights, refreshing the BF16 copy.
__global__ void k_adam(int64_t n, float* __restrict__ w, bf16* __restrict__ wb, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    const float wi = w[e] - lr * (mi / c1) / (std::sqrt(vi / c2) + eps);
    w[e] = wi;
    if (wb) wb[e] = to_bf16(wi);
}

}  // namespace
}  // namespace nn
}  // namespace mpt

namespace mpt {
namespace nn {
namespace {

#define NN_TRY(expr)                                                                                         \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return set_error(MPT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline unsigned blocks(int64_t n, int bs = 256) { return unsigned((n + bs - 1) / bs); }

// C = A * B^T in the layouts k_gemm takes.
// With epi 4 the K range is split over enough blocks to fill the GPU and
// the partial products are added atomically into C (which the caller zeroes).
int gemm(bool at, bool bt, int epi, int M, int N, int K, const bf16* A, int64_t lda, const bf16* B, int64_t ldb,
         void* C, int64_t ldc, const float* bias, bool accumulate, hipStream_t s) {
    const unsigned gm = unsigned((M + 63) / 64), gn = unsigned((N + 63) / 64);
    int kchunk = K, splits = 1;
    if (epi == 4) {
        const int64_t want = std::max<int64_t>(1, 4096 / (int64_t(gm) * gn));
        kchunk = int(std::max<int64_t>(256, ((int64_t(K) + want - 1) / want + 31) / 32 * 32));
        splits = int((int64_t(K) + kchunk - 1) / kchunk);
    }
    dim3 g(gm, gn, unsigned(splits));
#define GEMM_CASE(AT, BT, E)                                                                                      \
    if (at == AT && bt == BT && epi == E) {                                                                       \
        hipLaunchKernelGGL((k_gemm<AT, BT, E>), g, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias,        \
                           int(accumulate), kchunk);                                                              \
        return MPT_OK;                                                                                            \
    }
    GEMM_CASE(false, false, 0)
    GEMM_CASE(false, false, 2)
    GEMM_CASE(false, false, 3)
    GEMM_CASE(true, true, 0)
    GEMM_CASE(true, true, 4)
    GEMM_CASE(false, true, 0)
    GEMM_CASE(true, false, 0)
#undef GEMM_CASE
    return set_error(MPT_E_INVALID, "gemm: unsupported layout/epilogue");
}

}  // namespace
}  // namespace nn
}  // namespace mpt
