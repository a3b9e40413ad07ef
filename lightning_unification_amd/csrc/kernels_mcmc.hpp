// Metropolis-Hastings sampling of |psi|^2 (generate_samples_metropolis,
// MeasurementsLQubit.hpp:388-434, metropolis_step :613-640, proposals of
// TransitionKernels.hpp).
//
// The reference runs one sequential chain from index 0.  Here many
// independent chains run at once, one per thread: chain c starts at
// start[c], an inverse-CDF draw from |psi|^2 (so every chain is in the
// stationary distribution from its first step, however short it is),
// discards `burnin` steps, then records its next `per_chain` states as output
// rows c * per_chain ... (the last chain is cut at the shot count).  A step
// proposes an index, reads |psi|^2 there and accepts with probability
// min(1, p_new / p_old) (always when p_old = 0, as exp(log p_new - log 0)
// does in the reference).  Proposal kinds:
//   0: uniform over the indices with |psi_i| > epsilon of the precision
//      (the list is compacted in index order, so runs are reproducible);
//   1: pick a bit uniformly and a value uniformly; flip the bit when it
//      differs from the value, otherwise stay.
// Random words: SplitMix64, chain c keyed by mix(seed + (c + 1) * golden).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels_sample.hpp"

namespace plamd {

constexpr int kNzChunk = 4 * kBlock; // amplitudes per compaction block

template <class T> __device__ __forceinline__ bool nz_flag(const v2<T> &a, T eps) {
    return sqrt(double(a.x) * a.x + double(a.y) * a.y) > double(eps);
}

// counts[b] = number of flagged amplitudes in [b kNzChunk, (b+1) kNzChunk)
template <class T>
__global__ __launch_bounds__(kBlock) void k_nz_count(const v2<T> *__restrict__ sv, uint64_t len, T eps,
                                                     double *__restrict__ counts) {
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kNzChunk + 4 * threadIdx.x;
    double c = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (base + k < len && nz_flag(sv[base + k], eps)) {
            c += 1;
        }
    }
    const double2 s = block_sum2(double2{c, 0});
    if (threadIdx.x == 0) {
        counts[blockIdx.x] = s.x;
    }
}

// list[offset[b] + rank] = index, in index order
template <class T>
__global__ __launch_bounds__(kBlock) void k_nz_write(const v2<T> *__restrict__ sv, uint64_t len, T eps,
                                                     const double *__restrict__ offset,
                                                     uint64_t *__restrict__ list) {
    __shared__ int pre[kBlock];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kNzChunk + 4 * threadIdx.x;
    int f[4], c = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        f[k] = (base + k < len && nz_flag(sv[base + k], eps)) ? 1 : 0;
        c += f[k];
    }
    pre[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) { // exclusive scan over the block's 256 counts
        int run = 0;
        for (int t = 0; t < kBlock; t++) {
            const int v = pre[t];
            pre[t] = run;
            run += v;
        }
    }
    __syncthreads();
    uint64_t pos = static_cast<uint64_t>(offset[blockIdx.x]) + pre[threadIdx.x];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (f[k]) {
            list[pos++] = base + k;
        }
    }
}

struct McmcRng {
    uint64_t key, ctr = 0;
    __device__ uint64_t word() { return splitmix_mix(key + (++ctr) * 0x9E3779B97F4A7C15ull); }
    __device__ double uniform() { return double(word() >> 11) * (1.0 / 9007199254740992.0); }
    __device__ uint64_t below(uint64_t bound) { return __umul64hi(word(), bound); } // Lemire
};

template <class T>
__global__ __launch_bounds__(kBlock) void k_mcmc_chains(const v2<T> *__restrict__ sv, uint64_t len,
                                                        int n, int kind,
                                                        const uint64_t *__restrict__ nz, uint64_t nnz,
                                                        const int64_t *__restrict__ start, uint64_t burnin, uint64_t per_chain,
                                                        uint64_t nshots, uint64_t seed,
                                                        int64_t *__restrict__ idx_out) {
    const uint64_t c = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const uint64_t first = c * per_chain;
    if (first >= nshots) {
        return;
    }
    const uint64_t rec = per_chain < nshots - first ? per_chain : nshots - first;
    McmcRng rng{splitmix_mix(seed + (c + 1) * 0x9E3779B97F4A7C15ull)};
    uint64_t idx = static_cast<uint64_t>(start[c]);
    const v2<T> a0 = sv[idx];
    double p = double(a0.x) * a0.x + double(a0.y) * a0.y;
    for (uint64_t st = 0; st < burnin + rec; st++) {
        uint64_t t;
        if (kind == 0) {
            t = nz ? nz[rng.below(nnz)] : rng.below(len);
        } else {
            const int site = static_cast<int>(rng.below(static_cast<uint64_t>(n)));
            const uint64_t val = rng.below(2);
            t = (((idx >> site) & 1U) == val) ? idx : (idx ^ (uint64_t{1} << site));
        }
        const v2<T> at = sv[t];
        const double pt = double(at.x) * at.x + double(at.y) * at.y;
        const double u = rng.uniform();
        if (p == 0.0 || u * p < pt) {
            idx = t;
            p = pt;
        }
        if (st >= burnin) {
            idx_out[first + (st - burnin)] = static_cast<int64_t>(idx);
        }
    }
}

} // namespace plamd
