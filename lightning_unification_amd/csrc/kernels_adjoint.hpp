// Fused adjoint-Jacobian step (one operation of the backward sweep of
// AdjointJacobianLQubit.hpp:282-329).
//
// Reference order per op i:  mu <- lambda;  lambda <- U_i^dag lambda;
// mu <- G_i mu;  J[t][k] = -2 s Im<H_k|mu>;  H_k <- U_i^dag H_k.
// Here, per amplitude group of op i (U_i and G_i lowered on one geometry):
// load lambda's group, form mu = G lambda in registers, store U^dag lambda;
// then for each H_k: load its group, accumulate Im(conj(H_k) . mu), store
// U^dag H_k.  mu never touches HBM and every state is read and written once
// per op: (2 + 2K) state passes against the reference's (6 + 4K).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels_reduce.hpp"

namespace plamd {

template <class T, int M> struct AdjOp {
    Geom geo;
    uint64_t off[M];
    dcx<T> ud[M * M];
    dcx<T> g[M * M];
    int ud_diag;
};

constexpr int kAdjMaxObs = 16; // observables per launch
// bytes of states one adjoint sweep keeps in flight when the states do not
// all fit on die: MI355X has 256 MB of Infinity Cache in front of HBM
constexpr size_t kAdjCacheBytes = size_t{192} << 20;
constexpr int kAdjChunk = 4;    // observables whose loads are in flight together

template <class T, int M>
__device__ __forceinline__ void matvec(const dcx<T> *m, int diag, const v2<T> *x, v2<T> *y) {
    if (diag) {
#pragma unroll
        for (int r = 0; r < M; r++) {
            y[r] = cmul(m[r * M + r], x[r]);
        }
    } else {
#pragma unroll
        for (int r = 0; r < M; r++) {
            v2<T> acc = cmul(m[r * M], x[0]);
#pragma unroll
            for (int c = 1; c < M; c++) {
                cfma(acc, m[r * M + c], x[c]);
            }
            y[r] = acc;
        }
    }
}

// Block totals of the per-thread accumulators acc[0 .. nobs): each wave sums
// with shuffles into an LDS table, one barrier, then thread k adds the waves'
// entries in order (nobs is uniform across the block).
__device__ __forceinline__ void adj_block_totals(const double *acc, int nobs, double *__restrict__ out) {
    __shared__ double wpart[kBlock / 64][kAdjMaxObs];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kAdjMaxObs; k++) {
        if (k < nobs) {
            const double s = wave_sum(acc[k]);
            if (lane == 0) {
                wpart[wave][k] = s;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < nobs) {
        double t = 0;
        for (int w = 0; w < kBlock / 64; w++) {
            t += wpart[w][threadIdx.x];
        }
        out[threadIdx.x] = t;
    }
}

// states[0] = lambda, states[1 .. nobs] = H_k of this launch.
// TRAIN: accumulate Im<H_k|G lambda>; partial[block * nobs + k].
// update_lambda: write U^dag lambda (only the last launch of an op does).
template <class T, int M, bool TRAIN>
__global__ __launch_bounds__(kBlock) void k_adjoint_step(v2<T> *const *__restrict__ states,
                                                         int nobs, uint64_t nwork,
                                                         const AdjOp<T, M> op,
                                                         int update_lambda,
                                                         double *__restrict__ partial) {
    double acc[kAdjMaxObs];
#pragma unroll
    for (int k = 0; k < kAdjMaxObs; k++) {
        acc[k] = 0;
    }
    v2<T> *lam = states[0];
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; g < nwork; g += stride) {
        const uint64_t base = expand(g, op.geo);
        v2<T> x[M], y[M], mu[M];
        if (TRAIN || update_lambda) {
#pragma unroll
            for (int j = 0; j < M; j++) {
                x[j] = lam[base | op.off[j]];
            }
        }
        if (TRAIN) {
            matvec<T, M>(op.g, 0, x, mu);
        }
        if (update_lambda) {
            matvec<T, M>(op.ud, op.ud_diag, x, y);
#pragma unroll
            for (int j = 0; j < M; j++) {
                lam[base | op.off[j]] = y[j];
            }
        }
        // observables in chunks: a chunk's loads are all issued before its
        // stores (the H_k may alias as far as the compiler knows)
#pragma unroll
        for (int k0 = 0; k0 < kAdjMaxObs; k0 += kAdjChunk) {
            v2<T> hx[kAdjChunk][M];
#pragma unroll
            for (int c = 0; c < kAdjChunk; c++) {
                if (k0 + c < nobs) {
                    const v2<T> *h = states[1 + k0 + c];
#pragma unroll
                    for (int j = 0; j < M; j++) {
                        hx[c][j] = h[base | op.off[j]];
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < kAdjChunk; c++) {
                if (k0 + c < nobs) {
                    if (TRAIN) {
                        double s = 0;
#pragma unroll
                        for (int j = 0; j < M; j++) {
                            // Im(conj(h) * mu)
                            s += double(hx[c][j].x) * double(mu[j].y) - double(hx[c][j].y) * double(mu[j].x);
                        }
                        acc[k0 + c] += s;
                    }
                    matvec<T, M>(op.ud, op.ud_diag, hx[c], y);
                    v2<T> *h = states[1 + k0 + c];
#pragma unroll
                    for (int j = 0; j < M; j++) {
                        h[base | op.off[j]] = y[j];
                    }
                }
            }
        }
    }
    if (TRAIN) {
        adj_block_totals(acc, nobs, partial + static_cast<uint64_t>(blockIdx.x) * nobs);
    }
}

// Same step for parity-phase ops (MultiRZ): per amplitude, U^dag and G are
// phases selected by the parity of i & mask.
template <class T, bool TRAIN>
__global__ __launch_bounds__(kBlock) void k_adjoint_parity(v2<T> *const *__restrict__ states,
                                                           int nobs, uint64_t len, uint64_t mask,
                                                           dcx<T> u0, dcx<T> u1, dcx<T> g0,
                                                           dcx<T> g1, int update_lambda,
                                                           double *__restrict__ partial) {
    double acc[kAdjMaxObs];
#pragma unroll
    for (int k = 0; k < kAdjMaxObs; k++) {
        acc[k] = 0;
    }
    v2<T> *lam = states[0];
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < len; i += stride) {
        const bool odd = __popcll(i & mask) & 1;
        const dcx<T> u = odd ? u1 : u0;
        v2<T> x;
        if (TRAIN || update_lambda) {
            x = lam[i];
        }
        v2<T> mu;
        if (TRAIN) {
            mu = cmul(odd ? g1 : g0, x);
        }
        if (update_lambda) {
            lam[i] = cmul(u, x);
        }
#pragma unroll
        for (int k0 = 0; k0 < kAdjMaxObs; k0 += kAdjChunk) {
            v2<T> hv[kAdjChunk];
#pragma unroll
            for (int c = 0; c < kAdjChunk; c++) {
                if (k0 + c < nobs) {
                    hv[c] = states[1 + k0 + c][i];
                }
            }
#pragma unroll
            for (int c = 0; c < kAdjChunk; c++) {
                if (k0 + c < nobs) {
                    if (TRAIN) {
                        acc[k0 + c] += double(hv[c].x) * double(mu.y) - double(hv[c].y) * double(mu.x);
                    }
                    states[1 + k0 + c][i] = cmul(u, hv[c]);
                }
            }
        }
    }
    if (TRAIN) {
        adj_block_totals(acc, nobs, partial + static_cast<uint64_t>(blockIdx.x) * nobs);
    }
}

// Trainable steps whose partial sums wait in their own slots of the partial
// buffer, reduced together by one launch (k_adjoint_final_many).
constexpr int kAdjFinalSlots = 64;
struct AdjFinal {
    double factor; // the generator's scale times -2 (sign of the inverse)
    int nblocks;   // blocks of the step's launch, each wrote nobs partials
    int nobs;      // observables of the launch
    int dst;       // jacobian offset of its first observable
};
struct AdjFinals {
    AdjFinal t[kAdjFinalSlots];
};

// For slot s = blockIdx.y and observable k = blockIdx.x:
// jac[dst + k] = factor * sum_b partial_s[b * nobs + k], 256 threads each
// summing a strided share of the blocks, then a fixed-order block sum.
__global__ __launch_bounds__(kBlock) void k_adjoint_final_many(const double *__restrict__ partial,
                                                               size_t slot_stride, const AdjFinals tab,
                                                               double *__restrict__ jac) {
    const AdjFinal t = tab.t[blockIdx.y];
    const int k = blockIdx.x;
    if (k >= t.nobs) {
        return; // the whole block
    }
    const double *p = partial + static_cast<size_t>(blockIdx.y) * slot_stride;
    double s = 0;
    for (int b = threadIdx.x; b < t.nblocks; b += kBlock) {
        s += p[static_cast<size_t>(b) * t.nobs + k];
    }
    const double2 r = block_sum2(double2{s, 0});
    if (threadIdx.x == 0) {
        jac[t.dst + k] = t.factor * r.x;
    }
}

// Fused step of the vector-Jacobian product's backward sweep
// (VectorJacobianProduct.hpp:77-155): per amplitude group of the op, load
// lambda's and mu's groups; if the op is trainable, accumulate the complex
// <mu|G lambda> (= <G mu|lambda>, G Hermitian and zero outside the op's
// geometry); store U^dag lambda and U^dag mu.  One launch per op, each state
// read and written once; partial[blockIdx.x] = the block's sum.
template <class T, int M, bool TRAIN>
__global__ __launch_bounds__(kBlock) void k_vjp_step(v2<T> *__restrict__ lam, v2<T> *__restrict__ mu,
                                                     uint64_t nwork, const AdjOp<T, M> op,
                                                     double2 *__restrict__ partial) {
    double2 acc{0, 0};
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; g < nwork; g += stride) {
        const uint64_t base = expand(g, op.geo);
        v2<T> x[M], h[M], y[M];
#pragma unroll
        for (int j = 0; j < M; j++) {
            x[j] = lam[base | op.off[j]];
            h[j] = mu[base | op.off[j]];
        }
        if (TRAIN) {
            v2<T> gx[M];
            matvec<T, M>(op.g, 0, x, gx);
#pragma unroll
            for (int j = 0; j < M; j++) {
                // conj(h) * gx
                acc.x += double(h[j].x) * double(gx[j].x) + double(h[j].y) * double(gx[j].y);
                acc.y += double(h[j].x) * double(gx[j].y) - double(h[j].y) * double(gx[j].x);
            }
        }
        matvec<T, M>(op.ud, op.ud_diag, x, y);
#pragma unroll
        for (int j = 0; j < M; j++) {
            lam[base | op.off[j]] = y[j];
        }
        matvec<T, M>(op.ud, op.ud_diag, h, y);
#pragma unroll
        for (int j = 0; j < M; j++) {
            mu[base | op.off[j]] = y[j];
        }
    }
    if (TRAIN) {
        const double2 r = block_sum2(acc);
        if (threadIdx.x == 0) {
            partial[blockIdx.x] = r;
        }
    }
}

// The same for parity-phase ops (MultiRZ): U^dag and G are phases selected
// by the parity of i & mask.
template <class T, bool TRAIN>
__global__ __launch_bounds__(kBlock) void k_vjp_parity(v2<T> *__restrict__ lam, v2<T> *__restrict__ mu,
                                                       uint64_t len, uint64_t mask, dcx<T> u0, dcx<T> u1,
                                                       dcx<T> g0, dcx<T> g1, double2 *__restrict__ partial) {
    double2 acc{0, 0};
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < len; i += stride) {
        const bool odd = __popcll(i & mask) & 1;
        const v2<T> x = lam[i], h = mu[i];
        if (TRAIN) {
            const v2<T> gx = cmul(odd ? g1 : g0, x);
            acc.x += double(h.x) * double(gx.x) + double(h.y) * double(gx.y);
            acc.y += double(h.x) * double(gx.y) - double(h.y) * double(gx.x);
        }
        const dcx<T> u = odd ? u1 : u0;
        lam[i] = cmul(u, x);
        mu[i] = cmul(u, h);
    }
    if (TRAIN) {
        const double2 r = block_sum2(acc);
        if (threadIdx.x == 0) {
            partial[blockIdx.x] = r;
        }
    }
}

// out[rows[b]] = sum over i < nblocks of partial[rows[b] * nblocks + i], in
// a fixed order: one block per row
__global__ __launch_bounds__(kBlock) void k_vjp_final(const double2 *__restrict__ partial, int nblocks,
                                                      const int *__restrict__ rows, double2 *__restrict__ out) {
    const int t = rows[blockIdx.x];
    double2 s{0, 0};
    for (int b = threadIdx.x; b < nblocks; b += kBlock) {
        const double2 p = partial[static_cast<size_t>(t) * nblocks + b];
        s.x += p.x;
        s.y += p.y;
    }
    const double2 r = block_sum2(s);
    if (threadIdx.x == 0) {
        out[t] = r;
    }
}

} // namespace plamd
