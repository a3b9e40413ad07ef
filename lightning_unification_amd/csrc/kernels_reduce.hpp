// Reduction kernels: inner products, Pauli-string expectation values, norms,
// probabilities and marginals.  All sums accumulate in FP64 in a fixed order
// (per-thread strided sums, wavefront shuffles, one LDS step, per-block
// partials, one final block), so results are run-to-run reproducible.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels_gate.hpp"

namespace plamd {

constexpr int kReduceBlocks = 1024; // max partials per reduction

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        x += __shfl_xor(x, o, 64);
    }
    return x;
}

// Sum of two doubles over a 256-thread block; result valid in thread 0.
template <int NT = kBlock> __device__ __forceinline__ double2 block_sum2(double2 v) {
    __shared__ double sx[NT / 64], sy[NT / 64];
    v.x = wave_sum(v.x);
    v.y = wave_sum(v.y);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        sx[wid] = v.x;
        sy[wid] = v.y;
    }
    __syncthreads();
    double2 r{0, 0};
    if (threadIdx.x == 0) {
        for (int w = 0; w < NT / 64; w++) {
            r.x += sx[w];
            r.y += sy[w];
        }
    }
    __syncthreads();
    return r;
}

// partial[b] = sum over the block's elements of conj(a_i) * b_i
template <class T>
__global__ __launch_bounds__(kBlock) void k_dot(const v2<T> *__restrict__ a,
                                                const v2<T> *__restrict__ b, uint64_t len,
                                                double2 *__restrict__ partial) {
    double2 acc{0, 0};
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < len; i += stride) {
        const v2<T> x = a[i], y = b[i];
        acc.x += double(x.x) * double(y.x) + double(x.y) * double(y.y);
        acc.y += double(x.x) * double(y.y) - double(x.y) * double(y.x);
    }
    const double2 r = block_sum2(acc);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = r;
    }
}

// Sum `count` partials (fixed order) into out[0]; one block.
static __global__ __launch_bounds__(kBlock) void k_reduce_final(const double2 *__restrict__ partial,
                                                         int count, double2 *__restrict__ out) {
    double2 acc{0, 0};
    for (int i = threadIdx.x; i < count; i += kBlock) {
        acc.x += partial[i].x;
        acc.y += partial[i].y;
    }
    const double2 r = block_sum2(acc);
    if (threadIdx.x == 0) {
        out[0] = r;
    }
}

// One Pauli word: x-mask (X and Y letters), zy-mask (Z and Y letters) and
// the number of Y letters mod 4.  Its matrix element is
//   <i|P|i^x> = (-i)^ny * (-1)^popcount(i & zy).
struct PauliTerm {
    uint64_t zy;
    double coeff;
    int ny;
    int pad;
};

// partial[b] = sum_t coeff_t * Re<psi|P_t|psi> for terms sharing one x-mask.
// x == 0: one pass over amplitudes.  x != 0: one pass over pairs (i, i^x)
// with the lowest x bit of i clear, each amplitude read once.
template <class T>
__global__ __launch_bounds__(kBlock) void k_pauli_expval(const v2<T> *__restrict__ sv,
                                                         uint64_t nwork, uint64_t x, int lowbit,
                                                         int nterms,
                                                         const PauliTerm *__restrict__ terms,
                                                         double2 *__restrict__ partial) {
    double acc = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; g < nwork; g += stride) {
        if (x == 0) {
            const v2<T> a = sv[g];
            const double p = double(a.x) * double(a.x) + double(a.y) * double(a.y);
            double s = 0;
            for (int t = 0; t < nterms; t++) {
                const double sg = (__popcll(g & terms[t].zy) & 1) ? -1.0 : 1.0;
                s += terms[t].coeff * sg;
            }
            acc += p * s;
        } else {
            const uint64_t lo = g & ((uint64_t{1} << lowbit) - 1);
            const uint64_t i = ((g >> lowbit) << (lowbit + 1)) | lo;
            const uint64_t j = i ^ x;
            const v2<T> a = sv[i], c = sv[j];
            // z = conj(a) * c
            const double zr = double(a.x) * double(c.x) + double(a.y) * double(c.y);
            const double zi = double(a.x) * double(c.y) - double(a.y) * double(c.x);
            double s = 0;
            for (int t = 0; t < nterms; t++) {
                const double si = (__popcll(i & terms[t].zy) & 1) ? -1.0 : 1.0;
                const double sj = (__popcll(j & terms[t].zy) & 1) ? -1.0 : 1.0;
                // Re[(-i)^ny (si z + sj conj z)]
                double v;
                switch (terms[t].ny & 3) {
                case 0: v = (si + sj) * zr; break;
                case 1: v = (si - sj) * zi; break;
                case 2: v = -(si + sj) * zr; break;
                default: v = -(si - sj) * zi; break;
                }
                s += terms[t].coeff * v;
            }
            acc += s;
        }
    }
    const double2 r = block_sum2(double2{acc, 0});
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = r;
    }
}

// dst[i] (+)= c * (-i)^ny * (-1)^popcount(i & zy) * src[i ^ x]
template <class T>
__global__ __launch_bounds__(kBlock) void k_apply_pauli(const v2<T> *__restrict__ src,
                                                        v2<T> *__restrict__ dst, uint64_t len,
                                                        uint64_t x, uint64_t zy, dcx<T> c,
                                                        int accumulate) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < len; i += stride) {
        const v2<T> s = src[i ^ x];
        dcx<T> f = c;
        if (__popcll(i & zy) & 1) {
            f.r = -f.r;
            f.i = -f.i;
        }
        const v2<T> t = cmul(f, s);
        if (accumulate) {
            v2<T> d = dst[i];
            d.x += t.x;
            d.y += t.y;
            dst[i] = d;
        } else {
            dst[i] = t;
        }
    }
}

// y[i] += c * x[i]   (scaleAndAdd, LinearAlgebra.hpp:776-839)
template <class T>
__global__ __launch_bounds__(kBlock) void k_axpy(const v2<T> *__restrict__ x, v2<T> *__restrict__ y,
                                                 uint64_t len, dcx<T> c) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < len; i += stride) {
        v2<T> d = y[i];
        const v2<T> t = cmul(c, x[i]);
        d.x += t.x;
        d.y += t.y;
        y[i] = d;
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_probs(const v2<T> *__restrict__ sv, uint64_t len,
                                                  double *__restrict__ out) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < len; i += stride) {
        const v2<T> a = sv[i];
        out[i] = double(a.x * a.x + a.y * a.y);
    }
}

struct MarginalGeom {
    int nw;
    int bitpos[32];   // bit position of caller wire j (MSB of the bin index first)
    Geom rest;        // pins the marginal wires (ascending), no fixed bits; nw <= 12
    uint64_t free_mask; // bits not among the marginal wires
};

// Scatter the bits of r into the set bits of mask, lowest first (pdep).
__device__ __forceinline__ uint64_t deposit(uint64_t r, uint64_t mask) {
    uint64_t out = 0;
    while (mask) {
        const uint64_t low = mask & (~mask + 1);
        if (r & 1U) {
            out |= low;
        }
        r >>= 1;
        mask ^= low;
    }
    return out;
}

__device__ __forceinline__ uint64_t bin_bits(uint64_t bin, const MarginalGeom &m) {
    uint64_t r = 0;
    for (int j = 0; j < m.nw; j++) {
        if ((bin >> (m.nw - 1 - j)) & 1U) {
            r |= uint64_t{1} << m.bitpos[j];
        }
    }
    return r;
}

// Marginal probabilities, bin bit j taken from bit position bitpos[j].  Grid: blockIdx.y
// = bin, blockIdx.x = chunk of the remaining index space; partial[bin][chunk].
template <class T>
__global__ __launch_bounds__(kBlock) void k_marginal(const v2<T> *__restrict__ sv,
                                                     uint64_t nrest, MarginalGeom m,
                                                     double *__restrict__ partial) {
    const uint64_t bin = blockIdx.y;
    const uint64_t bb = bin_bits(bin, m);
    double acc = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t r = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; r < nrest; r += stride) {
        const v2<T> a = sv[expand(r, m.rest) | bb];
        acc += double(a.x * a.x + a.y * a.y);
    }
    const double2 s = block_sum2(double2{acc, 0});
    if (threadIdx.x == 0) {
        partial[bin * gridDim.x + blockIdx.x] = s.x;
    }
}

// Marginal probabilities when L of the marginal bits lie below bit 3 (inside
// one 128-byte line of complex128 amplitudes).  One bin per block would read
// only part of every line; here blockIdx.y fixes the pattern of the other
// marginal bits (deposited into himask) and each thread reads all 2^L
// amplitudes of a line that differ in the low marginal bits (lomask),
// keeping one sum per pattern.  The bin of a bit pattern follows bitpos, as
// bin_bits does in reverse.  Partials are laid out as k_marginal's.
__device__ __forceinline__ uint64_t pattern_bin(uint64_t bits, const MarginalGeom &m) {
    uint64_t bin = 0;
    for (int j = 0; j < m.nw; j++) {
        bin |= ((bits >> m.bitpos[j]) & 1U) << (m.nw - 1 - j);
    }
    return bin;
}

template <class T, int L>
__global__ __launch_bounds__(kBlock) void k_marginal_lo(const v2<T> *__restrict__ sv, uint64_t nrest,
                                                        MarginalGeom m, uint64_t lomask, uint64_t himask,
                                                        double *__restrict__ partial) {
    constexpr int C = 1 << L;
    static_assert(L >= 1 && L <= 3, "low marginal bits lie in bits 0..2");
    uint64_t off[C];
    double acc[C];
#pragma unroll
    for (int c = 0; c < C; c++) {
        off[c] = deposit(static_cast<uint64_t>(c), lomask);
        acc[c] = 0;
    }
    const uint64_t hb = deposit(blockIdx.y, himask);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t r = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; r < nrest; r += stride) {
        const uint64_t base = expand(r, m.rest) | hb;
#pragma unroll
        for (int c = 0; c < C; c++) {
            const v2<T> a = sv[base | off[c]];
            acc[c] += double(a.x * a.x + a.y * a.y);
        }
    }
#pragma unroll
    for (int c = 0; c < C; c += 2) {
        const double2 s = block_sum2(double2{acc[c], acc[c + 1]});
        if (threadIdx.x == 0) {
            partial[pattern_bin(hb | off[c], m) * gridDim.x + blockIdx.x] = s.x;
            partial[pattern_bin(hb | off[c + 1], m) * gridDim.x + blockIdx.x] = s.y;
        }
    }
}

// Final per-bin sum over `chunks` partials (thread per bin).
static __global__ __launch_bounds__(kBlock) void k_marginal_final(const double *__restrict__ partial,
                                                           uint64_t nbins, int chunks,
                                                           double *__restrict__ out) {
    const uint64_t b = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (b < nbins) {
        double s = 0;
        for (int c = 0; c < chunks; c++) {
            s += partial[b * chunks + c];
        }
        out[b] = s;
    }
}

// Marginal when few amplitudes fall in each bin: thread per bin, sequential.
template <class T>
__global__ __launch_bounds__(kBlock) void k_marginal_small(const v2<T> *__restrict__ sv,
                                                           uint64_t nbins, uint64_t nrest,
                                                           MarginalGeom m,
                                                           double *__restrict__ out) {
    const uint64_t bin = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (bin >= nbins) {
        return;
    }
    const uint64_t bb = bin_bits(bin, m);
    double s = 0;
    for (uint64_t r = 0; r < nrest; r++) {
        const v2<T> a = sv[deposit(r, m.free_mask) | bb];
        s += double(a.x * a.x + a.y * a.y);
    }
    out[bin] = s;
}

// Sparse (CSR) Hamiltonian moments in one pass, y = H psi row by row
// (SparseLinAlg.hpp apply_Sparse_Matrix):
//   partial.x = sum_i Re(conj(psi_i) y_i)   (the expectation value)
//   partial.y = sum_i |y_i|^2               (<psi|H^dag H|psi>)
template <class T>
__global__ __launch_bounds__(kBlock) void k_sparse_moments(const v2<T> *__restrict__ sv, uint64_t len,
                                                           const int64_t *__restrict__ row_map,
                                                           const int64_t *__restrict__ cols,
                                                           const dcx<T> *__restrict__ vals, int64_t nnz,
                                                           unsigned *__restrict__ bad,
                                                           double2 *__restrict__ partial) {
    // The arrays are checked here, not on the host: a row whose range is not
    // 0 <= row_map[i] <= row_map[i + 1] <= nnz sets bit 0 of *bad, a column
    // index outside the state bit 1; neither is read through.
    double2 acc{0, 0};
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < len;
         i += static_cast<uint64_t>(gridDim.x) * kBlock) {
        const int64_t lo = row_map[i], hi = row_map[i + 1];
        if (lo < 0 || hi < lo || hi > nnz) {
            atomicOr(bad, 1U);
            continue;
        }
        double yr = 0, yi = 0;
        for (int64_t j = lo; j < hi; j++) {
            const int64_t c = cols[j];
            if (static_cast<uint64_t>(c) >= len) {
                atomicOr(bad, 2U);
                continue;
            }
            const dcx<T> a = vals[j];
            const v2<T> x = sv[c];
            yr += double(a.r) * double(x.x) - double(a.i) * double(x.y);
            yi += double(a.r) * double(x.y) + double(a.i) * double(x.x);
        }
        const v2<T> p = sv[i];
        acc.x += double(p.x) * yr + double(p.y) * yi;
        acc.y += yr * yr + yi * yi;
    }
    const double2 r = block_sum2(acc);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = r;
    }
}

// Several x-mask groups of Pauli terms from one read of the state.  A pass
// owns a bit set S = {0, 1, 2} plus up to kPauliTileHigh high bits; every
// x-mask of the pass lies inside S, so both amplitudes of every pair (i, i^x)
// are in the same tile of 2^|S| amplitudes (all combinations of S, the other
// bits fixed).  Tiles are staged in LDS once; each group then walks its pairs
// there.  A term's Z/Y parity splits into the tile-local bits (zyl, in the
// compressed tile index) and the bits outside S (zyo), constant per tile.
// High bits per tile: 10 (tiles of 2^13 amplitudes, 1024-thread workgroups)
// where a workgroup's LDS holds such a tile, else 8 (2^11, 256 threads).
constexpr int kPauliTileHigh = 8;
constexpr int kPauliTileHighMax = 10;
struct PauliTileGeom {
    int nhigh;                       // high bits in S
    int high[kPauliTileHighMax];     // their positions, ascending (all >= 3)
    int nlow;                        // low bits in S: min(3, n)
};
// The terms of a group are sorted into buckets by the part of zyl that the
// thread's own elements differ in (pauli_bucket); bucket k is the term range
// [tb[k], tb[k + 1]) (4 buckets for a pair group, 8 for the diagonal one).
struct PauliTileGroup {
    uint32_t xl;   // x-mask in the compressed tile index (0: diagonal terms)
    uint32_t used; // bit k: bucket k holds terms
    int tb[9];
};
// A term of a pair group contributes s_e (cr Re z + ci Im z) for the pair
// (e, f = e ^ x), z = conj(psi_e) psi_f and s_e = (-1)^popcount(e & zy): with
// sigma = (-1)^popcount(x & zy) (the sign of f relative to e),
// Re[(-i)^ny (s_e z + s_f conj z)] = s_e (cr Re z + ci Im z), where
// cr = coeff (1 + sigma) {1, 0, -1, 0}[ny mod 4] and
// ci = coeff (1 - sigma) {0, 1, 0, -1}[ny mod 4].  Diagonal terms: cr = coeff.
struct PauliTileTerm {
    uint64_t zyo; // Z/Y bits outside S
    double cr, ci;
    uint32_t zyl; // Z/Y bits inside S, compressed
    int pad;
};

constexpr int kPauliPerThread = 8; // amplitudes a thread loads per tile (2^(3 + HIGH) / 2^HIGH)
constexpr int kPauliPairs = kPauliPerThread / 2;

// Tile-local element e with a zero inserted at bit lb (the pair's low end).
__host__ __device__ inline uint32_t pauli_insert0(uint32_t e, int lb) {
    const uint32_t m = (1U << lb) - 1;
    return ((e & ~m) << 1) | (e & m);
}
// Thread t of a workgroup of NT = 2^HIGH threads owns the elements t + NT j
// (diagonal group, j < 8) or the pairs whose low ends are
// pauli_insert0(t + NT j, lb) (j < 4).  They differ from the first one in bits
// fixed by j alone, so the sign of a term at element j is its sign at the
// first element times (-1)^popcount(j & k), k made of the term's zyl bits at
// those positions: the bucket of the term.
__host__ __device__ inline int pauli_bucket(uint32_t xl, uint32_t zyl, int high) {
    const uint32_t nt = 1U << high;
    if (xl == 0) {
        return static_cast<int>((zyl >> high) & 7U);
    }
    int lb = 0;
    while (!((xl >> lb) & 1U)) {
        lb++;
    }
    const uint32_t d1 = pauli_insert0(nt, lb), d2 = pauli_insert0(2 * nt, lb);
    return ((zyl & d1) ? 1 : 0) | ((zyl & d2) ? 2 : 0);
}

// v with its sign flipped where bit 0 of par is set
__device__ __forceinline__ double pauli_flip(double v, uint32_t par) {
    return __longlong_as_double(__double_as_longlong(v) ^ (static_cast<unsigned long long>(par) << 63));
}

// One pass of the Pauli-sum expectation over the whole state.  Persistent
// workgroups walk the tiles q = blockIdx.x, blockIdx.x + gridDim.x, ...; each
// thread holds its share of the next tile in registers while the current one
// is evaluated from LDS, so the state streams in behind the arithmetic.
//
// Per group a thread forms its values (|psi_e|^2 of 8 elements, or z of 4
// pairs) and their Walsh-Hadamard transform W[k] = sum_j (-1)^popcount(j & k)
// v_j.  A term of bucket k then contributes (-1)^p (its coefficient times
// W[k]) with p its parity at the thread's first element (the tile's bits
// outside S included): one AND, one popcount, a sign flip and two FMAs per
// term and thread, not per pair.  Terms and groups are read by uniform index
// (scalar loads).
// HIGH: high bits of a full tile, 2^HIGH threads per workgroup.  FULL: every
// tile holds 2^(3 + HIGH) amplitudes (no element or pair is out of range).
// SIGNS: the pass's terms are terms[t0 .. t0 + nt) with nt <= 64; lane l
// keeps term t0 + l's bits outside S, and one ballot per tile gives every
// term's sign there as a bit of a scalar mask (else each term computes its
// own with a scalar popcount).
template <class T, int HIGH, bool FULL, bool SIGNS>
__global__ __launch_bounds__(1 << HIGH) void k_pauli_tile(const v2<T> *__restrict__ sv, uint64_t ntiles,
                                                       const PauliTileGeom geo, int ngroups,
                                                       const PauliTileGroup *__restrict__ groups,
                                                       const PauliTileTerm *__restrict__ terms, int t0, int nt,
                                                       double2 *__restrict__ partial) {
    extern __shared__ unsigned char smem_raw[];
    v2<T> *tile = reinterpret_cast<v2<T> *>(smem_raw);
    const int nS = geo.nlow + geo.nhigh;
    constexpr int NT = 1 << HIGH;
    const uint32_t tsz = FULL ? (1U << (3 + HIGH)) : (1U << nS);
    const uint32_t lowmask = (1U << geo.nlow) - 1;
    const uint32_t tid = threadIdx.x;
    // the tile-local part of each loaded element's state index
    uint64_t off[kPauliPerThread];
#pragma unroll
    for (int k = 0; k < kPauliPerThread; k++) {
        const uint32_t e = tid + k * NT;
        uint64_t g = e & lowmask;
        for (int b = 0; b < geo.nhigh; b++) {
            g |= static_cast<uint64_t>((e >> (geo.nlow + b)) & 1U) << geo.high[b];
        }
        off[k] = g;
    }
    // q spread over the bits outside S (zeros inserted at S, ascending)
    auto tile_base = [&](uint64_t q) {
        uint64_t base = q << geo.nlow;
        for (int k = 0; k < geo.nhigh; k++) {
            const int p = geo.high[k];
            base = ((base >> p) << (p + 1)) | (base & ((uint64_t{1} << p) - 1));
        }
        return base;
    };
    uint64_t zyo_lane = 0; // SIGNS: this lane's term's bits outside S
    if constexpr (SIGNS) {
        const int l = static_cast<int>(tid & 63U);
        if (l < nt) {
            zyo_lane = terms[t0 + l].zyo;
        }
    }
    v2<T> nxt[kPauliPerThread];
    uint64_t q = blockIdx.x;
    uint64_t base = tile_base(q);
    if (q < ntiles) {
#pragma unroll
        for (int k = 0; k < kPauliPerThread; k++) {
            if (FULL || tid + k * NT < tsz) {
                nxt[k] = ld_nt<T>(&sv[base | off[k]]);
            }
        }
    }
    double acc = 0;
    for (; q < ntiles; q += gridDim.x) {
        const uint64_t cur = base; // this tile's bits outside S
        __syncthreads();           // the previous tile is no longer read
#pragma unroll
        for (int k = 0; k < kPauliPerThread; k++) {
            if (FULL || tid + k * NT < tsz) {
                tile[tid + k * NT] = nxt[k];
            }
        }
        __syncthreads();
        const uint64_t qn = q + gridDim.x;
        if (qn < ntiles) { // next tile's loads in flight during this tile's arithmetic
            base = tile_base(qn);
#pragma unroll
            for (int k = 0; k < kPauliPerThread; k++) {
                if (FULL || tid + k * NT < tsz) {
                    nxt[k] = ld_nt<T>(&sv[base | off[k]]);
                }
            }
        }
        // SIGNS: bit l = parity of term t0 + l outside S on this tile
        uint64_t smask = 0;
        if constexpr (SIGNS) {
            smask = __ballot(__popcll(cur & zyo_lane) & 1);
        }
        // one bucket's terms against the transform value (wr, wi)
        auto bucket = [&](int tb0, int tb1, uint32_t e0, double wr, double wi) {
            for (int t = tb0; t < tb1; t++) {
                const PauliTileTerm tm = terms[t];
                const uint32_t po = SIGNS ? static_cast<uint32_t>((smask >> (t - t0)) & 1U)
                                          : static_cast<uint32_t>(__popcll(cur & tm.zyo));
                const uint32_t par = __popc(e0 & tm.zyl) + po;
                acc += pauli_flip(fma(tm.ci, wi, tm.cr * wr), par);
            }
        };
        for (int gi = 0; gi < ngroups; gi++) {
            const uint32_t xl = groups[gi].xl, used = groups[gi].used;
            if (xl == 0) {
                double w[8];
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const uint32_t e = tid + j * NT;
                    w[j] = 0;
                    if (FULL || e < tsz) {
                        const v2<T> a = tile[e];
                        w[j] = double(a.x) * double(a.x) + double(a.y) * double(a.y);
                    }
                }
                if (__popc(used) <= 3) {
                    // each used transform value on its own: seven FMAs with uniform signs
                    for (uint32_t u = used; u; u &= u - 1) {
                        const int k = __ffs(u) - 1;
                        double wk = w[0];
#pragma unroll
                        for (int j = 1; j < 8; j++) {
                            wk = fma((__popc(j & k) & 1) ? -1.0 : 1.0, w[j], wk);
                        }
                        bucket(groups[gi].tb[k], groups[gi].tb[k + 1], tid, wk, 0.0);
                    }
                    continue;
                }
#pragma unroll
                for (int h = 1; h < 8; h <<= 1) {
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        if (!(j & h)) {
                            const double u = w[j], v = w[j | h];
                            w[j] = u + v;
                            w[j | h] = u - v;
                        }
                    }
                }
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    bucket(groups[gi].tb[k], groups[gi].tb[k + 1], tid, w[k], 0.0);
                }
                continue;
            }
            const int lb = __ffs(xl) - 1;
            const uint32_t npairs = tsz >> 1;
            const uint32_t e0 = pauli_insert0(tid, lb);
            // the thread's pairs: low ends e0 ^ d_j, partners e0 ^ d_j ^ xl,
            // d_j = pauli_insert0(NT j, lb) uniform (its bits are not e0's)
            const uint32_t d1 = pauli_insert0(NT, lb), d2 = pauli_insert0(2 * NT, lb);
            const char *tb8 = reinterpret_cast<const char *>(tile);
            const uint32_t e0b = e0 * static_cast<uint32_t>(sizeof(v2<T>));
            double zr[kPauliPairs], zi[kPauliPairs];
#pragma unroll
            for (int j = 0; j < kPauliPairs; j++) {
                const uint32_t dj = ((j & 1) ? d1 : 0U) | ((j & 2) ? d2 : 0U);
                zr[j] = zi[j] = 0;
                if (FULL || tid + j * NT < npairs) {
                    const v2<T> a = *reinterpret_cast<const v2<T> *>(tb8 + (e0b ^ (dj * sizeof(v2<T>))));
                    const v2<T> c = *reinterpret_cast<const v2<T> *>(tb8 + (e0b ^ ((dj ^ xl) * sizeof(v2<T>))));
                    // z = conj(a) * c
                    zr[j] = double(a.x) * double(c.x) + double(a.y) * double(c.y);
                    zi[j] = double(a.x) * double(c.y) - double(a.y) * double(c.x);
                }
            }
            if (__popc(used) <= 2) {
                for (uint32_t u = used; u; u &= u - 1) {
                    const int k = __ffs(u) - 1;
                    const double s1 = (k & 1) ? -1.0 : 1.0, s2 = (k & 2) ? -1.0 : 1.0, s3 = s1 * s2;
                    const double wr = fma(s3, zr[3], fma(s2, zr[2], fma(s1, zr[1], zr[0])));
                    const double wi = fma(s3, zi[3], fma(s2, zi[2], fma(s1, zi[1], zi[0])));
                    bucket(groups[gi].tb[k], groups[gi].tb[k + 1], e0, wr, wi);
                }
                continue;
            }
#pragma unroll
            for (int h = 1; h < kPauliPairs; h <<= 1) {
#pragma unroll
                for (int j = 0; j < kPauliPairs; j++) {
                    if (!(j & h)) {
                        const double ur = zr[j], vr = zr[j | h], ui = zi[j], vi = zi[j | h];
                        zr[j] = ur + vr;
                        zr[j | h] = ur - vr;
                        zi[j] = ui + vi;
                        zi[j | h] = ui - vi;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < kPauliPairs; k++) {
                bucket(groups[gi].tb[k], groups[gi].tb[k + 1], e0, zr[k], zi[k]);
            }
        }
    }
    const double2 r = block_sum2<NT>(double2{acc, 0});
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = r;
    }
}

} // namespace plamd
