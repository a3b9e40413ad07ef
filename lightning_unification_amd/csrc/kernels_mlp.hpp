// Kernels of the learned surrogate: a small MLP trained and evaluated with
// BF16 matrix cores and FP32 accumulation.
//
// k_gemm_bf16 computes C[M x N] = op(A)[M x K] op(B)[K x N] with one
// v_mfma_f32_32x32x8bf16_1k chain per wave: a workgroup of four waves covers a
// 64 x 64 block of C, each wave a 32 x 32 tile, stepping K by 8.  Lane l
// supplies A[i = l % 32][k0 + 4 (l / 32) + 0..3] and B[k0 + 4 (l / 32) + 0..3]
// [j = l % 32]; the result tile holds C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
// in accumulator r.  Operands are read as FP32 or BF16 and rounded to BF16
// (nearest even) on load; everything out of range reads as zero, so any M, N,
// K is valid.  Epilogues fuse the bias, GELU (storing the pre-activation for
// the backward pass) and the GELU derivative.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels_reduce.hpp"

namespace plamd {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float mlp_floatx16 __attribute__((ext_vector_type(16)));

__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) { // inf / nan: keep, quiet the nan
        return static_cast<uint16_t>((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    }
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t b) {
    const uint32_t u = static_cast<uint32_t>(b) << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

__device__ __forceinline__ float mlp_load(const float *p, size_t i) { return p[i]; }
__device__ __forceinline__ float mlp_load(const uint16_t *p, size_t i) { return bf16_to_f32(p[i]); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum MlpEpilogue : int {
    kEpiStore = 0,    // C = acc (+ bias)
    kEpiGelu = 1,     // Z = acc + bias (FP32), act = GELU(Z) (BF16)
    kEpiGeluGrad = 2, // C = acc * GELU'(Zin)
};

struct GemmArgs {
    const void *A;
    int lda;
    const void *B;
    int ldb;
    float *C;
    int ldc;
    int M, N, K;
    const float *bias;  // may be null
    const float *Zin;   // kEpiGeluGrad: pre-activation, ld = ldc
    float *Zout;        // kEpiGelu: pre-activation out, ld = ldc
    uint16_t *act;      // kEpiGelu: BF16 activation out
    int ldact;
    int epi;
};

// TRA: A is stored K x M (A(i,k) = A[k lda + i]); TRB: B is stored N x K.
template <class TA, class TB, bool TRA, bool TRB>
__global__ __launch_bounds__(256) void k_gemm_bf16(const GemmArgs g) {
    const TA *A = static_cast<const TA *>(g.A);
    const TB *B = static_cast<const TB *>(g.B);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const int c0 = blockIdx.x * 64 + (wave & 1) * 32;
    const int i = r0 + (lane & 31), j = c0 + (lane & 31), kq = 4 * (lane >> 5);
    mlp_floatx16 acc = {};
    for (int k0 = 0; k0 < g.K; k0 += 8) {
        bf16x4 a, b;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int k = k0 + kq + q;
            float av = 0.0f, bv = 0.0f;
            if (i < g.M && k < g.K) {
                av = mlp_load(A, TRA ? static_cast<size_t>(k) * g.lda + i : static_cast<size_t>(i) * g.lda + k);
            }
            if (j < g.N && k < g.K) {
                bv = mlp_load(B, TRB ? static_cast<size_t>(j) * g.ldb + k : static_cast<size_t>(k) * g.ldb + j);
            }
            a[q] = static_cast<short>(f32_to_bf16(av));
            b[q] = static_cast<short>(f32_to_bf16(bv));
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int row = r0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (row >= g.M || j >= g.N) {
            continue;
        }
        float z = acc[r];
        if (g.bias) {
            z += g.bias[j];
        }
        const size_t o = static_cast<size_t>(row) * g.ldc + j;
        if (g.epi == kEpiGelu) {
            g.Zout[o] = z;
            g.act[static_cast<size_t>(row) * g.ldact + j] = f32_to_bf16(gelu(z));
        } else if (g.epi == kEpiGeluGrad) {
            g.C[o] = z * gelu_grad(g.Zin[o]);
        } else {
            g.C[o] = z;
        }
    }
}

// out[j] = sum_i X[i ld + j], i = 0..M-1 in order (bias gradients)
__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ X, int M, int N, int ld,
                                                float *__restrict__ out) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < N) {
        float s = 0.0f;
        for (int i = 0; i < M; i++) {
            s += X[static_cast<size_t>(i) * ld + j];
        }
        out[j] = s;
    }
}

// Weighted mean-squared error: loss = sum_ij w_j (Y - T)^2 / M;
// dY = 2 w_j (Y - T) / M; partial[block] = the block's share of M * loss.
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ Y, const float *__restrict__ T,
                                             int M, int N, const float *__restrict__ w,
                                             float *__restrict__ dY, double *__restrict__ partial) {
    double acc = 0;
    const size_t total = static_cast<size_t>(M) * N;
    for (size_t e = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; e < total;
         e += static_cast<size_t>(gridDim.x) * 256) {
        const int j = static_cast<int>(e % N);
        const float d = Y[e] - T[e];
        acc += static_cast<double>(w[j]) * d * d;
        dY[e] = 2.0f * w[j] * d / static_cast<float>(M);
    }
    const double2 s = block_sum2(double2{acc, 0});
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s.x;
    }
}

// Adam on FP32 master weights, refreshing the BF16 copy (wbf may be null).
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ w, uint16_t *__restrict__ wbf,
                                              float *__restrict__ m, float *__restrict__ v,
                                              const float *__restrict__ g, int64_t n, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (e < n) {
        const float gi = g[e];
        const float mi = b1 * m[e] + (1.0f - b1) * gi;
        const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
        m[e] = mi;
        v[e] = vi;
        const float wi = w[e] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[e] = wi;
        if (wbf) {
            wbf[e] = f32_to_bf16(wi);
        }
    }
}

__global__ __launch_bounds__(256) void k_to_bf16(const float *__restrict__ x, uint16_t *__restrict__ y,
                                                 int64_t n) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (e < n) {
        y[e] = f32_to_bf16(x[e]);
    }
}

} // namespace plamd
