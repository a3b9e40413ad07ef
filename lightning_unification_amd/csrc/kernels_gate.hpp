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
// The loop leaves after the nb pinned positions (nb is uniform), so a
// one-qubit gate pays for one insertion, not for twelve guarded ones.
__device__ __forceinline__ uint64_t expand(uint64_t g, const Geom &geo) {
#pragma unroll
    for (int j = 0; j < 12; j++) {
        if (j >= geo.nb) {
            break;
        }
        const int p = geo.pos[j];
        const uint64_t lo = g & ((uint64_t{1} << p) - 1);
        g = ((g >> p) << (p + 1)) | lo;
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

// Streaming loads/stores: each amplitude is touched once per gate, so the
// accesses are marked non-temporal (no L2/MALL retention).
template <class T> __device__ __forceinline__ v2<T> ld_nt(const v2<T> *p) {
    v2<T> v;
    v.x = __builtin_nontemporal_load(&p->x);
    v.y = __builtin_nontemporal_load(&p->y);
    return v;
}
template <class T> __device__ __forceinline__ void st_nt(v2<T> *p, v2<T> v) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that XCD x
// walks one contiguous eighth of the index space (when the grid divides).
__device__ __forceinline__ uint64_t xcd_block() {
    const uint64_t b = blockIdx.x, nb = gridDim.x;
    if (nb & 7) {
        return b;
    }
    return (b & 7) * (nb >> 3) + (b >> 3);
}

// One-shot group op: thread per group, grid covers the work exactly once.
// DIAG (a diagonal matrix: each member scales itself) is a template
// parameter, so that each instantiation has straight-line loads and stores
// and every store keeps its non-temporal hint.
template <class T, int M, bool DIAG>
__global__ __launch_bounds__(kBlock) void k_group_flat(v2<T> *__restrict__ sv, uint64_t nwork,
                                                       const SmallOp<T, M> op) {
    const uint64_t g = xcd_block() * kBlock + threadIdx.x;
    if (g >= nwork) {
        return;
    }
    const uint64_t base = expand(g, op.geo);
    v2<T> v[M];
#pragma unroll
    for (int j = 0; j < M; j++) {
        v[j] = ld_nt<T>(sv + (base | op.off[j]));
    }
    if constexpr (DIAG) {
#pragma unroll
        for (int j = 0; j < M; j++) {
            st_nt<T>(sv + (base | op.off[j]), cmul(op.mat[j * M + j], v[j]));
        }
    } else {
#pragma unroll
        for (int r = 0; r < M; r++) {
            v2<T> acc = cmul(op.mat[r * M], v[0]);
#pragma unroll
            for (int c = 1; c < M; c++) {
                cfma(acc, op.mat[r * M + c], v[c]);
            }
            st_nt<T>(sv + (base | op.off[r]), acc);
        }
    }
}

// 2x2 op on target bit p < 6 with every other pinned bit >= 6: lane l owns
// one amplitude (fully coalesced wavefront loads), its partner lives in lane
// l ^ (1 << p) and is fetched with a cross-lane shuffle.  `geo` pins the
// control bits only; the target bit is a free bit of the thread index.
template <class T, bool DIAG>
__global__ __launch_bounds__(kBlock) void k_pair_lane(v2<T> *__restrict__ sv, uint64_t nelem,
                                                      Geom geo, int p, dcx<T> m00, dcx<T> m01,
                                                      dcx<T> m10, dcx<T> m11) {
    const uint64_t t = xcd_block() * kBlock + threadIdx.x;
    // nelem is a multiple of 64 here, so whole wavefronts are in or out
    if (t >= nelem) {
        return;
    }
    const uint64_t i = expand(t, geo);
    const v2<T> v = ld_nt<T>(sv + i);
    v2<T> o;
    o.x = __shfl_xor(v.x, 1 << p, 64);
    o.y = __shfl_xor(v.y, 1 << p, 64);
    const bool hi = (t >> p) & 1;
    v2<T> r;
    if constexpr (DIAG) {
        r = cmul(hi ? m11 : m00, v);
    } else if (hi) {
        r = cmul(m10, o);
        cfma(r, m11, v);
    } else {
        r = cmul(m00, v);
        cfma(r, m01, o);
    }
    st_nt<T>(sv + i, r);
}

// Tiled group op for gates with a pinned bit inside a 128-byte line (bit
// positions 0..2 for complex128), where per-group access would issue
// partial-line loads and stores.  A workgroup stages a tile of 256
// consecutive amplitudes times every combination of the pinned bits >= 8
// ("high" pins, at most kTileMaxHi of them) in LDS with full-line coalesced
// loads, applies every group that lies in the tile, and writes the whole tile
// back, so HBM sees whole lines in both directions.
constexpr int kTileLow = 8;   // tile spans amplitude bits 0..7
// A lowered op with at most four members pins at most four bits (the
// double excitation), and this path needs one of them below bit 3, so at
// most three pins are high.  LDS: 256 * 8 * 16 B = 32 KiB for complex128.
constexpr int kTileMaxHi = 3;

template <class T, int M> struct TileOp {
    Geom hi;                 // pins >= kTileLow (ascending), no fixed bits
    Geom lo;                 // pins < kTileLow (ascending), no fixed bits
    uint32_t fixed_lo;       // fixed bits below kTileLow
    uint32_t fixed_h;        // fixed high bits, as bits of the tile's h index
    uint32_t off_e[M];       // member offset: bits below kTileLow
    uint32_t off_h[M];       // member offset: high pins as h-index bits
    dcx<T> mat[M * M];
    int diag;
};

// NHI = number of high pins (op.hi.nb); the LDS tile holds 2^NHI * 256 amplitudes.
template <class T, int M, int NHI, bool DIAG>
__global__ __launch_bounds__(kBlock) void k_group_tile(v2<T> *__restrict__ sv, uint64_t ntiles,
                                                       const TileOp<T, M> op) {
    __shared__ v2<T> tile[(1 << NHI) << kTileLow];
    const uint64_t t = xcd_block();
    if (t >= ntiles) {
        return;
    }
    constexpr int nh = 1 << NHI;
    const uint64_t e = threadIdx.x; // kBlock == 1 << kTileLow
    const uint64_t base = expand((t << kTileLow) | e, op.hi);
    for (int h = 0; h < nh; h++) {
        uint64_t hb = 0;
        for (int k = 0; k < NHI; k++) {
            if ((h >> k) & 1) {
                hb |= uint64_t{1} << op.hi.pos[k];
            }
        }
        tile[(h << kTileLow) | e] = ld_nt<T>(sv + (base | hb));
    }
    __syncthreads();
    const int ngroups = kBlock >> op.lo.nb;
    for (int q = threadIdx.x; q < ngroups; q += kBlock) {
        const uint32_t be = static_cast<uint32_t>(expand(q, op.lo)) | op.fixed_lo;
        v2<T> v[M];
#pragma unroll
        for (int j = 0; j < M; j++) {
            v[j] = tile[((op.fixed_h | op.off_h[j]) << kTileLow) | (be | op.off_e[j])];
        }
        if constexpr (DIAG) {
#pragma unroll
            for (int j = 0; j < M; j++) {
                tile[((op.fixed_h | op.off_h[j]) << kTileLow) | (be | op.off_e[j])] =
                    cmul(op.mat[j * M + j], v[j]);
            }
        } else {
#pragma unroll
            for (int r = 0; r < M; r++) {
                v2<T> acc = cmul(op.mat[r * M], v[0]);
#pragma unroll
                for (int c = 1; c < M; c++) {
                    cfma(acc, op.mat[r * M + c], v[c]);
                }
                tile[((op.fixed_h | op.off_h[r]) << kTileLow) | (be | op.off_e[r])] = acc;
            }
        }
    }
    __syncthreads();
    for (int h = 0; h < nh; h++) {
        uint64_t hb = 0;
        for (int k = 0; k < NHI; k++) {
            if ((h >> k) & 1) {
                hb |= uint64_t{1} << op.hi.pos[k];
            }
        }
        st_nt<T>(sv + (base | hb), tile[(h << kTileLow) | e]);
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
