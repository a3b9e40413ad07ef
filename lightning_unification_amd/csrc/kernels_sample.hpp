// Inverse-CDF sampling on the device.
//
// Rule (lightning_kokkos MeasurementsKokkos.hpp:700-741 + MeasuresFunctors.hpp:55-105):
// with C the exclusive prefix sum of p_k = |psi_k|^2 and C_N := 1, a uniform
// U maps to 0 if U <= C_1, else to the i with C_i < U <= C_{i+1}; i.e. to
// (#{k : C_k < U}) - 1, clamped to [0, N-1].
//
// Device plan: (1) one wavefront per sample block of kSampleBlock
// amplitudes sums its probabilities (FP64); (2) one workgroup turns the block
// sums into an exclusive prefix; (3) one wavefront per sample bisects the
// block prefix, then re-scans the chosen block in order with wavefront
// inclusive scans (coalesced loads) and counts the C_k below U.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels_reduce.hpp"

namespace plamd {

constexpr int kSampleBlock = 1024;

template <class T>
__global__ __launch_bounds__(kBlock) void k_sample_blocksum(const v2<T> *__restrict__ sv,
                                                            uint64_t len, int sb,
                                                            uint64_t nsb,
                                                            double *__restrict__ bsum) {
    const uint64_t wave = static_cast<uint64_t>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wave >= nsb) {
        return;
    }
    const uint64_t start = wave * sb;
    double acc = 0;
    for (int j = lane; j < sb; j += 64) {
        const uint64_t e = start + j;
        if (e < len) {
            const v2<T> a = sv[e];
            acc += double(a.x) * double(a.x) + double(a.y) * double(a.y);
        }
    }
    acc = wave_sum(acc);
    if (lane == 0) {
        bsum[wave] = acc;
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

template <class T>
__global__ __launch_bounds__(kBlock) void k_sample_search(const v2<T> *__restrict__ sv,
                                                          uint64_t len, int sb, uint64_t nsb,
                                                          const double *__restrict__ prefix,
                                                          const double *__restrict__ u,
                                                          int64_t m, int64_t *__restrict__ idx) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (s >= m) {
        return;
    }
    const double U = u[s];
    // number of sample blocks whose starting C is below U
    uint64_t lo = 0, hi = nsb; // invariant: prefix[<lo] < U, prefix[>=hi] >= U
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (prefix[mid] < U) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    if (lo == 0) {
        if (lane == 0) {
            idx[s] = 0;
        }
        return;
    }
    const uint64_t b = lo - 1;
    double run = prefix[b];
    int cnt = 0;
    const uint64_t start = b * sb;
    for (int j0 = 0; j0 < sb; j0 += 64) {
        const uint64_t e = start + j0 + lane;
        double p = 0;
        if (j0 + lane < sb && e < len) {
            const v2<T> a = sv[e];
            p = double(a.x) * double(a.x) + double(a.y) * double(a.y);
        }
        const double incl = wave_incl_scan(p, lane);
        const double ce = run + (incl - p);
        if (j0 + lane < sb && e < len && ce < U) {
            cnt++;
        }
        run += __shfl(incl, 63, 64);
    }
    double c = wave_sum(double(cnt));
    if (lane == 0) {
        int64_t i = int64_t(start) + int64_t(c) - 1;
        if (i < 0) {
            i = 0;
        }
        if (i > int64_t(len) - 1) {
            i = int64_t(len) - 1;
        }
        idx[s] = i;
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
