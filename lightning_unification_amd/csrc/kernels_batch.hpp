// Batched forward + adjoint for layered ansatz circuits of uneven size.
//
// Circuit family (the hot path's batched and surrogate workloads, and the
// plain 20-qubit workload): `depth` layers, each = for every wire w, the
// rotations of `pattern` (e.g. RY then RZ) with their own angles, followed by
// a CNOT entangler (ladder i->i+1 or ring i->(i+1) mod n).  Observable: a
// Pauli sum per circuit.  Output: E = <psi|H|psi> and dE/dtheta for every
// angle, the quantities AdjointJacobianLQubit.hpp:212-336 returns for K = 1.
//
// Every launch covers all circuits that still have work at that step; block
// b finds its circuit by bisection in a per-launch prefix table.  Rotations
// run on LDS tiles: a tile holds 2^T amplitudes spanning T bit positions (the
// 3 lowest bits -- one 128-byte line -- plus up to W-3 target bits), so one
// HBM pass applies the rotations of up to W wires.  The entangler is one
// GF(2)-linear gather pass (out[y] = in[src(y)], src bits are parities of y).
// The adjoint keeps lambda and H_lambda tiles in LDS and accumulates
// Im<H|G lambda> for every angle of the tile before undoing its rotation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels_reduce.hpp"

namespace plamd {

constexpr int kBatchMaxQubits = 30;
constexpr int kBatchMaxPattern = 4;
constexpr int kBatchMaxTile = 13; // widest LDS tile (bits)

struct BCirc {
    int n;           // qubits
    int depth;       // layers
    int nstage;      // rotation stages per layer
    int nterms;      // Pauli terms of the observable
    uint64_t off;    // amplitude offset of buffer 0 (buffers k at off + k * 2^n)
    uint32_t par;    // offset of this circuit's angles / gradients
    uint32_t term;   // offset of its Pauli terms
    uint32_t fwd_rows[kBatchMaxQubits]; // entangler gather:  src bit b = parity(y & fwd_rows[b])
    uint32_t adj_rows[kBatchMaxQubits]; // inverse entangler gather
};

struct BLaunch {
    const int *circ;        // active circuits of this launch
    const uint64_t *prefix; // exclusive block prefix (count + 1 entries)
    int count;
};

__device__ __forceinline__ int find_circ(const BLaunch &L, uint64_t b, uint64_t &local) {
    int lo = 0, hi = L.count - 1; // largest c with prefix[c] <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (L.prefix[mid] <= b) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    local = b - L.prefix[lo];
    return L.circ[lo];
}

// Tile geometry of rotation stage s with tile width W: stage 0 spans bits
// [0, min(n, W)) and rotates all of them; stage s > 0 (only when n > W)
// rotates the target bits [alo, ahi) = [W + (s-1)(W-3), +W-3) clipped to n,
// inside a full W-bit tile: bits {0,1,2} (one 128-byte line, for coalescing)
// plus W-3 consecutive high bits ending no later than n, so the last stage's
// tile also holds some bits of the stage before it, left untouched.  Returns
// T, fills bits[] (ascending) and [alo, ahi).
__device__ __forceinline__ int stage_bits(int n, int W, int s, int *bits, int &alo, int &ahi) {
    if (s == 0) {
        const int T = n < W ? n : W;
        for (int k = 0; k < T; k++) {
            bits[k] = k;
        }
        alo = 0;
        ahi = T;
        return T;
    }
    alo = W + (s - 1) * (W - 3);
    ahi = alo + (W - 3) < n ? alo + (W - 3) : n;
    const int hlo = alo < n - (W - 3) ? alo : n - (W - 3);
    int T = 0;
    for (int k = 0; k < 3; k++) {
        bits[T++] = k;
    }
    for (int b = hlo; b < hlo + (W - 3); b++) {
        bits[T++] = b;
    }
    return T;
}

template <class T> struct Rot2 {
    dcx<T> m[4];
};

// exp(-i theta/2 sigma_axis), axis 0 = X, 1 = Y, 2 = Z  (LM.hpp:541-615
// matrices), from cs = (cos(theta/2), sin(theta/2)); dagger negates sin.
template <class T> __device__ __forceinline__ Rot2<T> rot_matrix(int axis, double2 cs, bool dagger) {
    const double c = cs.x, s = dagger ? -cs.y : cs.y;
    Rot2<T> r;
    if (axis == 0) {
        r.m[0] = {T(c), T(0)};
        r.m[1] = {T(0), T(-s)};
        r.m[2] = {T(0), T(-s)};
        r.m[3] = {T(c), T(0)};
    } else if (axis == 1) {
        r.m[0] = {T(c), T(0)};
        r.m[1] = {T(-s), T(0)};
        r.m[2] = {T(s), T(0)};
        r.m[3] = {T(c), T(0)};
    } else {
        r.m[0] = {T(c), T(-s)};
        r.m[1] = {T(0), T(0)};
        r.m[2] = {T(0), T(0)};
        r.m[3] = {T(c), T(s)};
    }
    return r;
}

// cs[i] = (cos(theta_i / 2), sin(theta_i / 2)) for every angle of the batch,
// once per replay, so the rotation kernels do no transcendental math
__global__ __launch_bounds__(kBlock) void k_batch_trig(const double *__restrict__ params, uint64_t n,
                                                      double2 *__restrict__ cs) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) {
        double s, c;
        sincos(0.5 * params[i], &s, &c);
        cs[i] = double2{c, s};
    }
}

template <class T> __device__ __forceinline__ Rot2<T> mat2_mul(const Rot2<T> &a, const Rot2<T> &b) {
    Rot2<T> r;
    for (int i = 0; i < 2; i++) {
        for (int j = 0; j < 2; j++) {
            const dcx<T> x = a.m[2 * i], y = b.m[j], u = a.m[2 * i + 1], v = b.m[2 + j];
            r.m[2 * i + j] = {x.r * y.r - x.i * y.i + u.r * v.r - u.i * v.i,
                              x.r * y.i + x.i * y.r + u.r * v.i + u.i * v.r};
        }
    }
    return r;
}
template <class T>
__device__ __forceinline__ void apply2(const Rot2<T> &r, v2<T> &a, v2<T> &b) {
    v2<T> x = cmul(r.m[0], a);
    cfma(x, r.m[1], b);
    v2<T> y = cmul(r.m[2], a);
    cfma(y, r.m[3], b);
    a = x;
    b = y;
}

// (a, b) <- R_axis^(dagger?)(theta) (a, b) using the structure of each Pauli
// rotation (RX symmetric with imaginary off-diagonal, RY real, RZ diagonal):
// about half the arithmetic of a general 2x2 complex product.
template <class T>
__device__ __forceinline__ void rot_apply(int axis, double2 cs, bool dagger, v2<T> &a, v2<T> &b) {
    const T c = T(cs.x), s = dagger ? T(-cs.y) : T(cs.y);
    v2<T> x, y;
    if (axis == 0) { // [[c, -i s], [-i s, c]]
        x.x = c * a.x + s * b.y;
        x.y = c * a.y - s * b.x;
        y.x = s * a.y + c * b.x;
        y.y = c * b.y - s * a.x;
    } else if (axis == 1) { // [[c, -s], [s, c]]
        x.x = c * a.x - s * b.x;
        x.y = c * a.y - s * b.y;
        y.x = s * a.x + c * b.x;
        y.y = s * a.y + c * b.y;
    } else { // diag(c - i s, c + i s)
        x.x = c * a.x + s * a.y;
        x.y = c * a.y - s * a.x;
        y.x = c * b.x - s * b.y;
        y.y = c * b.y + s * b.x;
    }
    a = x;
    b = y;
}

// local tile index e -> global amplitude index.  Every stage's tile is bits
// {0, 1, 2} plus one contiguous run starting at bits[3] (stage 0: the run
// 3..T-1), so bits 0..2 of e stay in place and the rest shift to the run.
__device__ __forceinline__ uint64_t tile_index(uint64_t base, uint32_t e, const int *bits, int T) {
    const int hlo = T > 3 ? bits[3] : 3;
    return base | (e & 7U) | (static_cast<uint64_t>(e >> 3) << hlo);
}

// the tile's base: the block's local index spread over the non-tile bits
__device__ __forceinline__ uint64_t tile_base(uint64_t q, const int *bits, int T) {
    uint64_t g = q;
    for (int k = 0; k < T; k++) { // bits[] ascending: insert zeros in order
        const int p = bits[k];
        const uint64_t lo = g & ((uint64_t{1} << p) - 1);
        g = ((g >> p) << (p + 1)) | lo;
    }
    return g;
}

// source index of y under an entangler map: bit b = parity(y & rows[b])
__device__ __forceinline__ uint64_t gather_src(uint64_t y, const uint32_t *rows, int n) {
    uint64_t x = 0;
    for (int b = 0; b < n; b++) {
        x |= static_cast<uint64_t>(__popc(static_cast<uint32_t>(y) & rows[b]) & 1) << b;
    }
    return x;
}

// Forward rotation stage: for every target wire of the tile, the product of
// its pattern rotations (one 2x2 per wire) is applied in LDS.  gsrc >= 0: the
// previous layer's entangler is fused in -- the tile is gathered from buffer
// gsrc through the entangler's map and written to buffer buf (so the
// entangler costs no pass of its own).
template <class T>
__global__ __launch_bounds__(kBlock) void k_batch_rot(v2<T> *__restrict__ arena,
                                                      const BCirc *__restrict__ circs,
                                                      const BLaunch L, int W, int stage, int layer,
                                                      int buf, int gsrc, int npat, int4 pat,
                                                      const double2 *__restrict__ cs) {
    extern __shared__ unsigned char smem_raw[];
    v2<T> *tile = reinterpret_cast<v2<T> *>(smem_raw);
    uint64_t q;
    const BCirc &c = circs[find_circ(L, blockIdx.x, q)];
    int bits[16];
    int alo, ahi;
    const int Tn = stage_bits(c.n, W, stage, bits, alo, ahi);
    const uint64_t len = uint64_t{1} << c.n;
    v2<T> *sv = arena + c.off + buf * len;
    const uint64_t base = tile_base(q, bits, Tn);
    const uint32_t tsz = 1U << Tn;
    if (gsrc >= 0) {
        const v2<T> *src = arena + c.off + gsrc * len;
        for (uint32_t e = threadIdx.x; e < tsz; e += kBlock) {
            tile[e] = src[gather_src(tile_index(base, e, bits, Tn), c.fwd_rows, c.n)];
        }
    } else {
        for (uint32_t e = threadIdx.x; e < tsz; e += kBlock) {
            tile[e] = sv[tile_index(base, e, bits, Tn)];
        }
    }
    __syncthreads();
    const int axes[4] = {pat.x, pat.y, pat.z, pat.w};
    for (int k = 0; k < Tn; k++) {
        const int b = bits[k];
        if (b < alo || b >= ahi) {
            continue;
        }
        const int w = c.n - 1 - b;
        const double2 *th = cs + c.par + (static_cast<uint64_t>(layer) * c.n + w) * npat;
        Rot2<T> m = rot_matrix<T>(axes[0], th[0], false);
        for (int a = 1; a < npat; a++) {
            m = mat2_mul(rot_matrix<T>(axes[a], th[a], false), m);
        }
        for (uint32_t p = threadIdx.x; p < (tsz >> 1); p += kBlock) {
            const uint32_t lo = p & ((1U << k) - 1);
            const uint32_t i0 = ((p >> k) << (k + 1)) | lo, i1 = i0 | (1U << k);
            v2<T> x = tile[i0], y = tile[i1];
            apply2(m, x, y);
            tile[i0] = x;
            tile[i1] = y;
        }
        __syncthreads();
    }
    for (uint32_t e = threadIdx.x; e < tsz; e += kBlock) {
        sv[tile_index(base, e, bits, Tn)] = tile[e];
    }
}

// Entangler as a gather: dst[y] = src[src_index(y)] for the buffers given.
// adj = 0 forward map, 1 its inverse.  Up to two states per launch.
template <class T>
__global__ __launch_bounds__(kBlock) void k_batch_ring(v2<T> *__restrict__ arena,
                                                       const BCirc *__restrict__ circs,
                                                       const BLaunch L, int adj, int from0,
                                                       int to0, int from1, int to1) {
    uint64_t q;
    const BCirc &c = circs[find_circ(L, blockIdx.x, q)];
    const uint64_t len = uint64_t{1} << c.n;
    const uint64_t y = q * kBlock + threadIdx.x;
    if (y >= len) {
        return;
    }
    const uint64_t x = gather_src(y, adj ? c.adj_rows : c.fwd_rows, c.n);
    v2<T> *base = arena + c.off;
    base[to0 * len + y] = base[from0 * len + x];
    if (from1 >= 0) {
        base[to1 * len + y] = base[from1 * len + x];
    }
}

// |0...0> in buffer 0 of every circuit
template <class T>
__global__ __launch_bounds__(kBlock) void k_batch_init(v2<T> *__restrict__ arena,
                                                       const BCirc *__restrict__ circs,
                                                       const BLaunch L) {
    uint64_t q;
    const BCirc &c = circs[find_circ(L, blockIdx.x, q)];
    const uint64_t len = uint64_t{1} << c.n;
    const uint64_t y = q * kBlock + threadIdx.x;
    if (y < len) {
        v2<T> z;
        z.x = y == 0 ? T(1) : T(0);
        z.y = 0;
        arena[c.off + y] = z;
    }
}

// H psi into buffer 2 + f, E += Re<psi|H psi> and N += <psi|psi> (FP64
// atomics per block), where psi sits in buffer f = depth mod 2 when layers
// end with an entangler (each one flips the buffer) and f = 0 otherwise.
template <class T>
__global__ __launch_bounds__(kBlock) void k_batch_obs(v2<T> *__restrict__ arena,
                                                      const BCirc *__restrict__ circs,
                                                      const BLaunch L, int ent,
                                                      const PauliTerm *__restrict__ terms,
                                                      const uint64_t *__restrict__ xmasks,
                                                      double *__restrict__ energy,
                                                      double *__restrict__ norm2) {
    uint64_t q;
    const int ci = find_circ(L, blockIdx.x, q);
    const BCirc &c = circs[ci];
    const uint64_t len = uint64_t{1} << c.n;
    const int pbuf = ent ? (c.depth & 1) : 0;
    const int hbuf = 2 + pbuf;
    const uint64_t y = q * kBlock + threadIdx.x;
    double e = 0, pn = 0;
    if (y < len) {
        const v2<T> *psi = arena + c.off + pbuf * len;
        double hr = 0, hi = 0;
        uint64_t xm_prev = ~uint64_t{0};
        v2<T> a{};
        for (int t = 0; t < c.nterms; t++) { // terms arrive grouped by x-mask
            const PauliTerm &pt = terms[c.term + t];
            const uint64_t xm = xmasks[c.term + t];
            if (xm != xm_prev) {
                a = psi[y ^ xm];
                xm_prev = xm;
            }
            // coeff * (-i)^ny * (-1)^popcount(y & zy) * a
            double cr = pt.coeff, ci2 = 0;
            switch (pt.ny & 3) {
            case 1: ci2 = -cr; cr = 0; break;
            case 2: cr = -cr; break;
            case 3: ci2 = cr; cr = 0; break;
            default: break;
            }
            if (__popcll(y & pt.zy) & 1) {
                cr = -cr;
                ci2 = -ci2;
            }
            hr += cr * double(a.x) - ci2 * double(a.y);
            hi += cr * double(a.y) + ci2 * double(a.x);
        }
        v2<T> hv;
        hv.x = T(hr);
        hv.y = T(hi);
        arena[c.off + hbuf * len + y] = hv;
        const v2<T> p = psi[y];
        e = double(p.x) * hr + double(p.y) * hi;
        pn = double(p.x) * double(p.x) + double(p.y) * double(p.y);
    }
    const double2 s = block_sum2(double2{e, pn});
    if (threadIdx.x == 0) {
        atomicAdd(energy + ci, s.x);
        atomicAdd(norm2 + ci, s.y);
    }
}

// Adjoint rotation stage: lambda (buffer lb) and H (buffer hb) tiles in LDS.
// For each target wire, pattern rotations are undone last-to-first; before
// undoing rotation a, Im<H|G_a lambda> is accumulated for angle a
// (scale -1/2 and the -2 of the reference's update cancel: dE/dtheta =
// Im<H|G lambda>, AdjointJacobianLQubit.hpp:70-75).  Each wave reduces its
// share with shuffles into a per-wave LDS slot; after the last wire one pass
// adds the four waves' slots and issues one FP64 atomic per angle.
// glb >= 0: the layer's entangler is undone in the same pass -- for a
// circuit whose first adjoint stage of the layer this is (stage ==
// nstage - 1), lambda and H are gathered from buffers glb and ghb through the
// inverse map and written to lb and hb.
template <class T>
__global__ __launch_bounds__(kBlock) void k_batch_rot_adj(v2<T> *__restrict__ arena,
                                                          const BCirc *__restrict__ circs,
                                                          const BLaunch L, int W, int stage,
                                                          int layer, int lb, int hb, int glb, int ghb,
                                                          int npat, int4 pat, const double2 *__restrict__ cs,
                                                          double *__restrict__ grads) {
    extern __shared__ unsigned char smem_raw[];
    __shared__ double wpart[kBlock / 64][kBatchMaxTile][kBatchMaxPattern];
    uint64_t q;
    const BCirc &c = circs[find_circ(L, blockIdx.x, q)];
    int bits[16];
    int alo, ahi;
    const int Tn = stage_bits(c.n, W, stage, bits, alo, ahi);
    const uint32_t tsz = 1U << Tn;
    v2<T> *tl = reinterpret_cast<v2<T> *>(smem_raw);
    v2<T> *th_ = tl + tsz;
    const uint64_t len = uint64_t{1} << c.n;
    v2<T> *lam = arena + c.off + lb * len;
    v2<T> *hh = arena + c.off + hb * len;
    const uint64_t base = tile_base(q, bits, Tn);
    if (glb >= 0 && stage == c.nstage - 1) { // uniform per workgroup
        const v2<T> *ls = arena + c.off + glb * len, *hs = arena + c.off + ghb * len;
        for (uint32_t e = threadIdx.x; e < tsz; e += kBlock) {
            const uint64_t x = gather_src(tile_index(base, e, bits, Tn), c.adj_rows, c.n);
            tl[e] = ls[x];
            th_[e] = hs[x];
        }
    } else {
        for (uint32_t e = threadIdx.x; e < tsz; e += kBlock) {
            const uint64_t g = tile_index(base, e, bits, Tn);
            tl[e] = lam[g];
            th_[e] = hh[g];
        }
    }
    __syncthreads();
    const int axes[4] = {pat.x, pat.y, pat.z, pat.w};
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int slot = 0; // target wires done so far: slot j is bit alo + j
    for (int k = 0; k < Tn; k++) {
        const int b = bits[k];
        if (b < alo || b >= ahi) {
            continue;
        }
        const int w = c.n - 1 - b;
        const uint64_t pidx = c.par + (static_cast<uint64_t>(layer) * c.n + w) * npat;
        double acc[kBatchMaxPattern] = {0, 0, 0, 0};
        // letter loops run to the compile-time maximum with a guard, so the
        // per-letter arrays keep static indices and stay in registers
        double2 rcs[kBatchMaxPattern];
#pragma unroll
        for (int a = 0; a < kBatchMaxPattern; a++) {
            rcs[a] = a < npat ? cs[pidx + a] : double2{1, 0};
        }
        for (uint32_t p = threadIdx.x; p < (tsz >> 1); p += kBlock) {
            const uint32_t lo = p & ((1U << k) - 1);
            const uint32_t i0 = ((p >> k) << (k + 1)) | lo, i1 = i0 | (1U << k);
            v2<T> l0 = tl[i0], l1 = tl[i1], h0 = th_[i0], h1 = th_[i1];
#pragma unroll
            for (int a = kBatchMaxPattern - 1; a >= 0; a--) {
                if (a >= npat) {
                    continue;
                }
                // G lambda for G = X, Y or Z on this wire
                v2<T> g0, g1;
                if (axes[a] == 0) {
                    g0 = l1;
                    g1 = l0;
                } else if (axes[a] == 1) {
                    g0.x = l1.y; g0.y = -l1.x;   // -i * l1
                    g1.x = -l0.y; g1.y = l0.x;   //  i * l0
                } else {
                    g0 = l0;
                    g1.x = -l1.x; g1.y = -l1.y;
                }
                // Im(conj(h) g)
                acc[a] += double(h0.x) * double(g0.y) - double(h0.y) * double(g0.x) +
                          double(h1.x) * double(g1.y) - double(h1.y) * double(g1.x);
                rot_apply<T>(axes[a], rcs[a], true, l0, l1);
                rot_apply<T>(axes[a], rcs[a], true, h0, h1);
            }
            tl[i0] = l0;
            tl[i1] = l1;
            th_[i0] = h0;
            th_[i1] = h1;
        }
#pragma unroll
        for (int a = 0; a < kBatchMaxPattern; a++) {
            if (a < npat) {
                const double s = wave_sum(acc[a]);
                if (lane == 0) {
                    wpart[wave][slot][a] = s;
                }
            }
        }
        slot++;
        __syncthreads();
    }
    for (int t = threadIdx.x; t < slot * npat; t += kBlock) {
        const int sl = t / npat, a = t - sl * npat;
        double s = 0;
        for (int v = 0; v < kBlock / 64; v++) {
            s += wpart[v][sl][a];
        }
        const int w = c.n - 1 - (alo + sl);
        atomicAdd(grads + c.par + (static_cast<uint64_t>(layer) * c.n + w) * npat + a, s);
    }
    for (uint32_t e = threadIdx.x; e < tsz; e += kBlock) {
        const uint64_t g = tile_index(base, e, bits, Tn);
        lam[g] = tl[e];
        hh[g] = th_[e];
    }
}

// zero n doubles (graph-safe replacement for a memset)
__global__ __launch_bounds__(kBlock) void k_zero(double *__restrict__ p, uint64_t n) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) {
        p[i] = 0;
    }
}

} // namespace plamd
