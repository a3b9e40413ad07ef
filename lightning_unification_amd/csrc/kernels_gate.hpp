// Gate-application kernels for gfx950.
//
// Every gate reaches the device as a group op (lower.hpp): for each free index
// g the kernel expands g to a base index by inserting zero bits at the pinned
// positions, ORs in the fixed control bits and touches M amplitudes at
// base | off[j].  Consecutive lanes take consecutive g, so whenever every
// pinned position is >= 3 (complex128) / >= 4 (complex64) each wavefront
// load covers whole 128-byte lines; lower positions still read every byte
// once, with the line shared by neighbouring loads in L2.
//
// M <= 4: matrix and offsets travel in the kernel arguments (uniform,
// scalar-loaded).  8 <= M <= 32: matrix in device memory, staged into LDS per
// workgroup.  M > 32: one workgroup per group, amplitudes staged in LDS.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plamd {

template <class T> struct vec2;
template <> struct vec2<double> { using type = double2; };
template <> struct vec2<float> { using type = float2; };
template <class T> using v2 = typename vec2<T>::type;

template <class T> struct dcx {
    T r, i;
};

template <class T> __device__ __forceinline__ v2<T> cmul(dcx<T> a, v2<T> b) {
    v2<T> o;
    o.x = a.r * b.x - a.i * b.y;
    o.y = a.r * b.y + a.i * b.x;
    return o;
}
template <class T> __device__ __forceinline__ void cfma(v2<T> &acc, dcx<T> a, v2<T> b) {
    acc.x += a.r * b.x - a.i * b.y;
    acc.y += a.r * b.y + a.i * b.x;
}

struct Geom {
    uint64_t fixed;
    int nb;
    int pos[12];
};

// Insert a zero bit at each pinned position (ascending) and set the fixed bits.
__device__ __forceinline__ uint64_t expand(uint64_t g, const Geom &geo) {
#pragma unroll
    for (int j = 0; j < 12; j++) {
        if (j < geo.nb) {
            const int p = geo.pos[j];
            const uint64_t lo = g & ((uint64_t{1} << p) - 1);
            g = ((g >> p) << (p + 1)) | lo;
        }
    }
    return g | geo.fixed;
}

template <class T, int M> struct SmallOp {
    Geom geo;
    uint64_t off[M];
    dcx<T> mat[M * M];
    int diag;
};

constexpr int kBlock = 256;

// Small group op, UNROLL independent groups per thread per iteration (their
// loads are issued back to back to keep enough bytes in flight).
template <class T, int M, int UNROLL>
__global__ __launch_bounds__(kBlock) void k_group_small(v2<T> *__restrict__ sv, uint64_t nwork,
                                                        const SmallOp<T, M> op) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock * UNROLL;
    for (uint64_t g0 = static_cast<uint64_t>(blockIdx.x) * kBlock * UNROLL + threadIdx.x; g0 < nwork;
         g0 += stride) {
        uint64_t base[UNROLL];
        v2<T> v[UNROLL][M];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const uint64_t g = g0 + static_cast<uint64_t>(u) * kBlock;
            base[u] = g < nwork ? expand(g, op.geo) : ~uint64_t{0};
            if (base[u] != ~uint64_t{0}) {
#pragma unroll
                for (int j = 0; j < M; j++) {
                    v[u][j] = sv[base[u] | op.off[j]];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            if (base[u] == ~uint64_t{0}) {
                continue;
            }
            if (op.diag) {
#pragma unroll
                for (int j = 0; j < M; j++) {
                    sv[base[u] | op.off[j]] = cmul(op.mat[j * M + j], v[u][j]);
                }
            } else {
#pragma unroll
                for (int r = 0; r < M; r++) {
                    v2<T> acc = cmul(op.mat[r * M], v[u][0]);
#pragma unroll
                    for (int c = 1; c < M; c++) {
                        cfma(acc, op.mat[r * M + c], v[u][c]);
                    }
                    sv[base[u] | op.off[r]] = acc;
                }
            }
        }
    }
}

template <class T, int M> struct BigOp {
    Geom geo;
    uint64_t off[M];
    const dcx<T> *mat; // M*M in device memory
};

template <class T, int M>
__global__ __launch_bounds__(kBlock) void k_group_big(v2<T> *__restrict__ sv, uint64_t nwork,
                                                      const BigOp<T, M> op) {
    __shared__ dcx<T> m[M * M];
    for (int i = threadIdx.x; i < M * M; i += kBlock) {
        m[i] = op.mat[i];
    }
    __syncthreads();
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; g < nwork; g += stride) {
        const uint64_t base = expand(g, op.geo);
        v2<T> v[M];
#pragma unroll
        for (int j = 0; j < M; j++) {
            v[j] = sv[base | op.off[j]];
        }
#pragma unroll
        for (int r = 0; r < M; r++) {
            v2<T> acc = cmul(m[r * M], v[0]);
#pragma unroll
            for (int c = 1; c < M; c++) {
                cfma(acc, m[r * M + c], v[c]);
            }
            sv[base | op.off[r]] = acc;
        }
    }
}

// Large dense op (M up to 4096): one workgroup per group, amplitudes in LDS.
template <class T>
__global__ __launch_bounds__(kBlock) void k_group_huge(v2<T> *__restrict__ sv, uint64_t nwork,
                                                       Geom geo, int M,
                                                       const uint64_t *__restrict__ off,
                                                       const dcx<T> *__restrict__ mat) {
    extern __shared__ unsigned char smem_raw[];
    v2<T> *buf = reinterpret_cast<v2<T> *>(smem_raw);
    for (uint64_t g = blockIdx.x; g < nwork; g += gridDim.x) {
        const uint64_t base = expand(g, geo);
        for (int j = threadIdx.x; j < M; j += kBlock) {
            buf[j] = sv[base | off[j]];
        }
        __syncthreads();
        for (int r = threadIdx.x; r < M; r += kBlock) {
            v2<T> acc;
            acc.x = 0;
            acc.y = 0;
            for (int c = 0; c < M; c++) {
                cfma(acc, mat[static_cast<size_t>(r) * M + c], buf[c]);
            }
            sv[base | off[r]] = acc;
        }
        __syncthreads();
    }
}

// amp[i] *= phase[popcount(i & mask) & 1]   (MultiRZ and its generator)
template <class T>
__global__ __launch_bounds__(kBlock) void k_parity(v2<T> *__restrict__ sv, uint64_t len,
                                                   uint64_t mask, dcx<T> p0, dcx<T> p1) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < len; i += stride) {
        const dcx<T> p = (__popcll(i & mask) & 1) ? p1 : p0;
        sv[i] = cmul(p, sv[i]);
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_fill_basis(v2<T> *__restrict__ sv, uint64_t len,
                                                       uint64_t index) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < len; i += stride) {
        v2<T> z;
        z.x = (i == index) ? T(1) : T(0);
        z.y = 0;
        sv[i] = z;
    }
}

} // namespace plamd
