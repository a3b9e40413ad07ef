// Inverse-CDF sampling on the device.
//
// Rule (lightning_kokkos MeasurementsKokkos.hpp:700-741 + MeasuresFunctors.hpp:55-105):
// with C the exclusive prefix sum of p_k = |psi_k|^2 and C_N := 1, a uniform
// U maps to 0 if U <= C_1, else to the i with C_i < U <= C_{i+1}; i.e. to
// (#{k : C_k < U}) - 1, clamped to [0, N-1].
//
// Device plan (the state is read once; one sample reads 1 KiB of it):
//   (1) k_cdf_groups: FP64 sums of p_k over groups of kCdfGroup = 64
//       consecutive amplitudes (one wavefront loads four groups per step,
//       coalesced, and reduces them together);
//   (2) k_cdf_tile_sums, k_scan_exclusive, k_cdf_tile_scan: the exclusive
//       prefix of the group sums, gpre[g] = C_{64 g}, scanned in tiles of
//       kCdfTile groups (tile totals, one scan of the totals, then each
//       tile's own scan offset by its total's prefix, tpre[t] = gpre[t * kCdfTile]);
//   (3) k_cdf_search: one thread per sample bisects tpre, then gpre inside the
//       tile, for the last group with C_{64 g} < U; its wavefront then walks
//       its 64 samples one after another, each time loading the chosen group
//       (one amplitude per lane) and counting with a wavefront scan the C_k
//       below U.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels_reduce.hpp"

namespace plamd {

constexpr int kCdfGroup = 64;    // amplitudes per group
constexpr int kCdfTile = 4096;   // groups per scan tile
constexpr int kCdfSteps = 4;     // four-group steps per wavefront in k_cdf_groups

__device__ __forceinline__ double shfl_xor_d(double x, int o) { return __shfl_xor(x, o, 64); }

// gsum[g] = sum of |psi_k|^2, k in [64 g, 64 g + 64) (FP64)
template <class T>
__global__ __launch_bounds__(kBlock) void k_cdf_groups(const v2<T> *__restrict__ sv, uint64_t len,
                                                       uint64_t ngroups, double *__restrict__ gsum) {
    const uint64_t wave = static_cast<uint64_t>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const uint64_t g0 = wave * (4 * kCdfSteps);
    double p[4 * kCdfSteps];
#pragma unroll
    for (int k = 0; k < 4 * kCdfSteps; k++) {
        const uint64_t e = (g0 + k) * kCdfGroup + lane;
        p[k] = 0;
        if (e < len) {
            const v2<T> a = ld_nt<T>(&sv[e]);
            p[k] = double(a.x) * double(a.x) + double(a.y) * double(a.y);
        }
    }
#pragma unroll
    for (int st = 0; st < kCdfSteps; st++) {
        double *q = p + 4 * st;
        // four groups reduced together: after the xor-32 and xor-16 steps
        // lane row r (lanes 16 r .. 16 r + 15) holds partials of group r only
        const bool hi32 = lane >= 32, hi16 = (lane & 16) != 0;
        const double a0 = (hi32 ? q[2] : q[0]) + shfl_xor_d(hi32 ? q[0] : q[2], 32);
        const double a1 = (hi32 ? q[3] : q[1]) + shfl_xor_d(hi32 ? q[1] : q[3], 32);
        double v = (hi16 ? a1 : a0) + shfl_xor_d(hi16 ? a0 : a1, 16);
        v += shfl_xor_d(v, 8);
        v += shfl_xor_d(v, 4);
        v += shfl_xor_d(v, 2);
        v += shfl_xor_d(v, 1);
        const uint64_t g = g0 + 4 * st + (lane >> 4);
        if ((lane & 15) == 0 && g < ngroups) {
            gsum[g] = v;
        }
    }
}

// tsum[t] = sum of gsum over tile t (kCdfTile groups; 256 threads x 16)
__global__ __launch_bounds__(kBlock) void k_cdf_tile_sums(const double *__restrict__ gsum, uint64_t ngroups,
                                                          double *__restrict__ tsum) {
    const uint64_t lo = static_cast<uint64_t>(blockIdx.x) * kCdfTile + threadIdx.x * (kCdfTile / kBlock);
    double s = 0;
#pragma unroll
    for (int j = 0; j < kCdfTile / kBlock; j++) {
        if (lo + j < ngroups) {
            s += gsum[lo + j];
        }
    }
    const double2 r = block_sum2(double2{s, 0});
    if (threadIdx.x == 0) {
        tsum[blockIdx.x] = r.x;
    }
}

// gpre[g] = tpre[t] + exclusive prefix of gsum inside tile t
__global__ __launch_bounds__(kBlock) void k_cdf_tile_scan(const double *__restrict__ gsum, uint64_t ngroups,
                                                          const double *__restrict__ tpre,
                                                          double *__restrict__ gpre) {
    __shared__ double part[kBlock];
    constexpr int per = kCdfTile / kBlock;
    const uint64_t lo = static_cast<uint64_t>(blockIdx.x) * kCdfTile + threadIdx.x * per;
    double v[per];
    double s = 0;
#pragma unroll
    for (int j = 0; j < per; j++) {
        v[j] = lo + j < ngroups ? gsum[lo + j] : 0.0;
        s += v[j];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < 64) { // exclusive scan of the 256 thread totals by one wavefront
        const int lane = threadIdx.x;
        double t[4];
        double run = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            t[j] = part[4 * lane + j];
            run += t[j];
        }
        double incl = run;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const double y = __shfl_up(incl, o, 64);
            if (lane >= o) {
                incl += y;
            }
        }
        double ex = incl - run;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            part[4 * lane + j] = ex;
            ex += t[j];
        }
    }
    __syncthreads();
    double run = tpre[blockIdx.x] + part[threadIdx.x];
    if (threadIdx.x == 0) {
        run = tpre[blockIdx.x]; // the tile's first group starts exactly at its tile prefix
    }
#pragma unroll
    for (int j = 0; j < per; j++) {
        if (lo + j < ngroups) {
            gpre[lo + j] = run;
        }
        run += v[j];
    }
}

// Exclusive prefix of `count` block sums, one workgroup of 1024 threads.
// prefix[count] receives the total.
__global__ __launch_bounds__(1024) void k_scan_exclusive(const double *__restrict__ in,
                                                         uint64_t count,
                                                         double *__restrict__ prefix) {
    __shared__ double tot[1024];
    const uint64_t per = (count + 1023) / 1024;
    const uint64_t lo = threadIdx.x * per;
    const uint64_t hi = lo + per < count ? lo + per : count;
    double s = 0;
    for (uint64_t i = lo; i < hi; i++) {
        s += in[i];
    }
    tot[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = 0;
        for (int t = 0; t < 1024; t++) {
            const double v = tot[t];
            tot[t] = run;
            run += v;
        }
        prefix[count] = run;
    }
    __syncthreads();
    double run = tot[threadIdx.x];
    for (uint64_t i = lo; i < hi; i++) {
        prefix[i] = run;
        run += in[i];
    }
}

__device__ __forceinline__ double wave_incl_scan(double x, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double y = __shfl_up(x, o, 64);
        if (lane >= o) {
            x += y;
        }
    }
    return x;
}

// value of lane `lane` (uniform) of a 64-bit register
__device__ __forceinline__ uint64_t readlane_u64(uint64_t x, int lane) {
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(x)), lane));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(x >> 32)), lane));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ double readlane_d(double x, int lane) {
    return __longlong_as_double(static_cast<long long>(readlane_u64(static_cast<uint64_t>(__double_as_longlong(x)), lane)));
}

// idx[s] = (#{k : C_k < u[s]}) - 1 clamped to [0, len - 1]
template <class T>
__global__ __launch_bounds__(kBlock) void k_cdf_search(const v2<T> *__restrict__ sv, uint64_t len,
                                                       uint64_t ngroups, uint64_t ntiles,
                                                       const double *__restrict__ tpre,
                                                       const double *__restrict__ gpre,
                                                       const double *__restrict__ u, int64_t m,
                                                       int64_t *__restrict__ idx) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = s < m;
    const double U = valid ? u[s] : 0.0;
    // tiles whose first C is below U (tpre[0] = 0)
    uint64_t lo = 0, hi = ntiles;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (tpre[mid] < U) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    int64_t g = -1; // last group with C_{64 g} < U, -1: none (U <= 0)
    double base = 0;
    if (lo > 0) {
        const uint64_t t = lo - 1;
        uint64_t a = t * kCdfTile, b = a + kCdfTile < ngroups ? a + kCdfTile : ngroups;
        a += 1; // gpre[t * kCdfTile] = tpre[t] < U already
        while (a < b) {
            const uint64_t mid = (a + b) >> 1;
            if (gpre[mid] < U) {
                a = mid + 1;
            } else {
                b = mid;
            }
        }
        g = static_cast<int64_t>(a) - 1;
        base = gpre[g];
    }
    // the wavefront scans its samples' groups one after another
    const int64_t wave0 = s - lane;
    const int nv = m - wave0 < 64 ? static_cast<int>(m - wave0) : 64;
    int64_t my = 0;
    for (int i = 0; i < nv; i++) {
        const int64_t gi = static_cast<int64_t>(readlane_u64(static_cast<uint64_t>(g), i));
        if (gi < 0) {
            if (lane == i) {
                my = 0;
            }
            continue;
        }
        const double Ui = readlane_d(U, i);
        const double bi = readlane_d(base, i);
        const uint64_t e = static_cast<uint64_t>(gi) * kCdfGroup + lane;
        double p = 0;
        if (e < len) {
            const v2<T> a = sv[e];
            p = double(a.x) * double(a.x) + double(a.y) * double(a.y);
        }
        const double c = bi + (wave_incl_scan(p, lane) - p);
        const uint64_t below = __ballot(e < len && c < Ui);
        int64_t r = static_cast<int64_t>(gi) * kCdfGroup + __popcll(below) - 1;
        r = r < 0 ? 0 : (r > static_cast<int64_t>(len) - 1 ? static_cast<int64_t>(len) - 1 : r);
        if (lane == i) {
            my = r;
        }
    }
    if (valid) {
        idx[s] = my;
    }
}

// SplitMix64 uniforms: u_s = (mix(seed + (s+1)*golden) >> 11) * 2^-53.
__device__ __forceinline__ uint64_t splitmix_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void k_uniforms(uint64_t seed, int64_t m,
                                                     double *__restrict__ u) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (s < m) {
        const uint64_t w = splitmix_mix(seed + uint64_t(s + 1) * 0x9E3779B97F4A7C15ull);
        u[s] = double(w >> 11) * (1.0 / 9007199254740992.0);
    }
}

// samples[s*n + (n-1-j)] = (idx_s >> j) & 1  (MeasurementsLQubit.hpp:546-552)
__global__ __launch_bounds__(kBlock) void k_sample_bits(const int64_t *__restrict__ idx,
                                                        int64_t m, int n,
                                                        uint64_t *__restrict__ bits) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (t < m * n) {
        const int64_t s = t / n;
        const int col = int(t - s * n);
        bits[t] = (uint64_t(idx[s]) >> (n - 1 - col)) & 1U;
    }
}

} // namespace plamd
