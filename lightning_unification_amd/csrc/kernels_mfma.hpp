// Fused 5-qubit blocks on a complex64 state with FP32 matrix cores.
//
// A fused block is a dense 32x32 complex matrix U on five bit positions.  For
// every group of 32 amplitudes (the five bits varying, all others fixed) the
// kernel computes psi' = U psi as the real product
//     [Re psi'; Im psi'] = [[Ur, -Ui], [Ui, Ur]] [Re psi; Im psi]
// with v_mfma_f32_32x32x2f32 (32x32 output, K = 2 per instruction).
//
// A workgroup owns a tile of 32 x 2^CB amplitudes: the 5 target bits and the
// CB lowest other bits ("column" bits), i.e. 2^CB groups, one wave per 32.  The tile moves
// HBM -> LDS -> HBM once, with whole 128-byte lines whenever the column bits
// include bits 0..3.  LDS holds the tile as [target index j][column c].  Wave
// w multiplies U into columns 32w .. 32w+31: for each K pair (j, j+1) a lane
// supplies A = U[row = lane % 32][j + lane / 32] and B = x[j + lane / 32][c],
// and the real and imaginary outputs accumulate in two chains whose MFMAs
// alternate (so consecutive MFMAs do not wait on one another).
//
// FP64 is kept where the precision form needs it: each workgroup accumulates
// the sum of |psi'|^2 over its tiles in double and writes one partial; a
// one-block kernel reduces the partials in a fixed order to the
// renormalisation factor that the next block (or the final scale pass)
// multiplies in.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels_reduce.hpp"

namespace plamd {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kF5ColBits = 7; // column bits per tile: 32 x 128 amplitudes, 4 waves

struct Fused5Geom {
    int pos[14];          // tile bit positions (5 targets + column bits), ascending
    uint32_t lds_of[14];  // LDS index contribution of each tile bit
};

// Per-lane matrix operands, host-packed: lane l, step s (K pair 2s, 2s+1):
// a[l][s] = {Ur[l % 32][2s + l / 32], Ui[l % 32][2s + l / 32]}.
struct Fused5Mat {
    float2 a[64][16];
};

template <int TB> __device__ __forceinline__ uint64_t f5_base(uint64_t q, const Fused5Geom &g) {
#pragma unroll
    for (int k = 0; k < TB; k++) {
        const int p = g.pos[k];
        q = ((q >> p) << (p + 1)) | (q & ((uint64_t{1} << p) - 1));
    }
    return q;
}

// CB column bits: a workgroup of 2^(CB+1) threads (2^(CB-5) waves) works on
// 32 x 2^CB tiles; thread t moves tile elements e = t + 2^(CB+1) r, r < 16.
// The grid is persistent: workgroup b takes the contiguous tile range
// [b T, (b+1) T), and the loads of tile i+1 are issued before tile i is
// multiplied and stored, so HBM stays busy while the matrix cores work.
template <int CB>
__global__ __launch_bounds__(2 << CB, 3) void k_fused5_mfma(float2 *__restrict__ sv, uint64_t ntiles,
                                                         uint64_t per_block, const Fused5Geom geo,
                                                         const Fused5Mat *__restrict__ mat,
                                                         const double *__restrict__ scale,
                                                         double *__restrict__ partial) {
    constexpr int NCOL = 1 << CB, LB = CB + 1, TB = CB + 5, NW = (2 << CB) / 64;
    __shared__ float2 tile[32 * NCOL];
    __shared__ double wsum[NW];
    const uint64_t q0 = static_cast<uint64_t>(blockIdx.x) * per_block;
    const uint64_t q1 = q0 + per_block < ntiles ? q0 + per_block : ntiles;
    const int tid = threadIdx.x;
    if (q0 >= q1) { // uniform per workgroup
        if (partial && tid == 0) {
            partial[blockIdx.x] = 0;
        }
        return;
    }
    // element e sets tile bit k when bit k of e is set (the low LB bits come
    // from tid, the next 4 from r): consecutive threads, consecutive addresses
    uint64_t a_lo = 0;
    uint32_t l_lo = 0;
#pragma unroll
    for (int k = 0; k < LB; k++) {
        if ((tid >> k) & 1) {
            a_lo |= uint64_t{1} << geo.pos[k];
            l_lo |= geo.lds_of[k];
        }
    }
    // tile element r of this thread: address / LDS offsets from the 4 r bits
    auto a_of = [&](int r) {
        uint64_t a = a_lo;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            a |= static_cast<uint64_t>((r >> k) & 1) << geo.pos[LB + k];
        }
        return a;
    };
    auto l_of = [&](int r) {
        uint32_t l = l_lo;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            l |= ((r >> k) & 1) ? geo.lds_of[LB + k] : 0U;
        }
        return l;
    };
    const int lane = tid & 63, wave = tid >> 6;
    const int col = wave * 32 + (lane & 31), half = lane >> 5;
    const float s = scale ? static_cast<float>(*scale) : 1.0f;

    double block_norm = 0;
    float2 v[16];
    uint64_t base = f5_base<TB>(q0, geo);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        v[r] = ld_nt<float>(sv + (base | a_of(r)));
    }
    for (uint64_t q = q0; q < q1; q++) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            tile[l_of(r)] = v[r];
        }
        __syncthreads();
        const uint64_t cur = base;
        if (q + 1 < q1) { // next tile's loads in flight during this tile's work
            base = f5_base<TB>(q + 1, geo);
#pragma unroll
            for (int r = 0; r < 16; r++) {
                v[r] = ld_nt<float>(sv + (base | a_of(r)));
            }
        }
        floatx16 re = {}, im = {}; // two chains, alternating
#pragma unroll
        for (int st = 0; st < 16; st++) {
            const float2 a = mat->a[lane][st];
            const float2 x = tile[(2 * st + half) * NCOL + col];
            const float ur = a.x * s, ui = a.y * s;
            re = __builtin_amdgcn_mfma_f32_32x32x2f32(ur, x.x, re, 0, 0, 0);
            im = __builtin_amdgcn_mfma_f32_32x32x2f32(ui, x.x, im, 0, 0, 0);
            re = __builtin_amdgcn_mfma_f32_32x32x2f32(-ui, x.y, re, 0, 0, 0);
            im = __builtin_amdgcn_mfma_f32_32x32x2f32(ur, x.y, im, 0, 0, 0);
        }
        // every lane of this wave has read its operands from columns it alone
        // writes below; the wave's LDS accesses complete in program order
        double nrm = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int row = 8 * (i >> 2) + 4 * half + (i & 3);
            float2 y;
            y.x = re[i];
            y.y = im[i];
            tile[row * NCOL + col] = y;
            nrm += static_cast<double>(y.x) * y.x + static_cast<double>(y.y) * y.y;
        }
        if (partial) {
            nrm = wave_sum(nrm);
            if (lane == 0) {
                wsum[wave] = nrm;
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; r++) {
            st_nt<float>(sv + (cur | a_of(r)), tile[l_of(r)]);
        }
        if (partial && tid == 0) {
            for (int w = 0; w < NW; w++) {
                block_norm += wsum[w];
            }
        }
        __syncthreads(); // the tile and wsum are rewritten by the next iteration
    }
    if (partial && tid == 0) {
        partial[blockIdx.x] = block_norm; // its tiles in order: a fixed summation order
    }
}

// scale_out = 1 / sqrt(sum of partial[0 .. count)), fixed summation order
__global__ __launch_bounds__(1024) void k_norm_scale(const double *__restrict__ partial,
                                                     uint64_t count, double *__restrict__ scale_out) {
    __shared__ double ws[16];
    double acc = 0;
    for (uint64_t i = threadIdx.x; i < count; i += 1024) {
        acc += partial[i];
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) {
        ws[threadIdx.x >> 6] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < 16; w++) {
            t += ws[w];
        }
        scale_out[0] = t > 0 ? 1.0 / sqrt(t) : 1.0;
    }
}

// psi *= scale (the last renormalisation of a fused run)
template <class T>
__global__ __launch_bounds__(kBlock) void k_scale(v2<T> *__restrict__ sv, uint64_t len,
                                                  const double *__restrict__ scale) {
    const T s = static_cast<T>(*scale);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < len; i += stride) {
        v2<T> a = ld_nt<T>(sv + i);
        a.x *= s;
        a.y *= s;
        st_nt<T>(sv + i, a);
    }
}

// partial[b] = sum over the block's i of |a_i - b_i|^2 and |a_i|^2, FP64,
// for states of any two precisions
template <class A, class B>
__global__ __launch_bounds__(kBlock) void k_diff_norm(const v2<A> *__restrict__ a,
                                                      const v2<B> *__restrict__ b, uint64_t len,
                                                      double2 *__restrict__ partial) {
    double2 acc{0, 0};
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < len;
         i += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const v2<A> x = a[i];
        const v2<B> y = b[i];
        const double dr = double(x.x) - double(y.x), di = double(x.y) - double(y.y);
        acc.x += dr * dr + di * di;
        acc.y += double(x.x) * double(x.x) + double(x.y) * double(x.y);
    }
    const double2 r = block_sum2(acc);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = r;
    }
}

} // namespace plamd
