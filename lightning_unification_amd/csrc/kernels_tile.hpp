// Multi-gate tile passes: cache blocking of a gate sequence on chip.
//
// A pass fixes a set of k "tile bits": the `low` lowest bit positions
// (contiguous runs of 2^low amplitudes in HBM) plus k - low higher positions
// hb[0], hb[1], ... (in any order: the pass's plan puts first the ones that
// become lane bits).  Every assignment of the other n - k bits selects one
// tile of 2^k amplitudes (local index e: bits 0..low-1 of e are the
// amplitude's bits 0..low-1, bit low + j of e is its bit hb[j]).  One
// wavefront loads a tile, applies every operation of the pass -- each lowered
// to a group op on the k-qubit local system, so the arithmetic is the
// per-gate kernels' -- and writes the tile back.  HBM sees one read and one
// write of the state per pass, however many gates the pass holds.
//
// The tile lives in registers: lane l holds v[r], r < R = 2^(k-6), the
// amplitude of local index l | r << 6.  Lane bits are local bits 0..5 (so a
// wavefront load covers runs of 2^low amplitudes), register bits are local
// bits 6..k-1.  Operations of a pass, compiled on the host:
//   diag     multiply every amplitude by the phase its index bits select;
//   parity   the same for a parity phase (MultiRZ);
//   p1/p2/p4 a group op whose member bits are register bits: each lane
//            updates its own registers, indices fixed at compile time;
//   lane     a 2x2 op whose member bit is a lane bit: one cross-lane
//            exchange per register;
//   mix1     a group op over one lane bit and one register bit (as a 4x4):
//            two exchanges per register pair;
//   mix2     a group op over two lane bits (as a 4x4): three exchanges per
//            register;
//   lds      any other op, applied by tp_group / tp_parity through LDS
//            (the launch then carries a tile of LDS).
// Control bits may be lane bits (a per-lane predicate) or register bits (a
// uniform predicate on the compile-time register index).  No workgroup
// barrier is needed outside the lds case: the wavefront is the workgroup.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels_gate.hpp"
#include "kernels_tile_ip.hpp"

#include <type_traits>

namespace plamd {

constexpr int kTpLanes = 64;   // one wavefront per tile
constexpr int kTpMinBits = 9;  // tile = 2^k amplitudes, k in [9, 12]
constexpr int kTpMaxBits = 12;
constexpr int kTpMaxLow = 6;   // the lane index spans local bits 0..5
constexpr int kTpMaxM = 8;     // amplitudes per group of an lds op
constexpr int kTpMaxPins = 6;  // pinned positions of a tiled op

template <class T> struct TpOp {
    int kind;                   // 0 group, 1 diagonal group, 2 parity phase
    int M;                      // 1, 2, 4 or 8 amplitudes per group
    int nb;                     // pinned local positions
    int pos[kTpMaxPins];        // ascending
    uint32_t fixed;             // local bits forced to 1 in every member
    uint32_t mask;              // parity: local bits whose parity picks the phase
    uint32_t off[kTpMaxM];      // member offsets (local)
    dcx<T> mat[kTpMaxM * kTpMaxM]; // M x M row-major; parity: [0] even, [1] odd
};

struct TpGeom {
    int k;                 // tile bits
    int low;               // local bits 0..low-1 = amplitude bits 0..low-1
    int hb[kTpMaxBits];    // amplitude bit of local bit low + j
    Geom outer;            // pins hb[] (ascending): tile index << low -> base index
};

template <class T, int M>
__device__ __forceinline__ void tp_group(v2<T> *tile, const TpOp<T> &op, int k, int tid) {
    const int ng = (1 << k) >> op.nb;
    uint32_t off[M];
#pragma unroll
    for (int j = 0; j < M; j++) {
        off[j] = op.off[j];
    }
    const bool diag = op.kind == 1;
    // ng is a power of two: with fewer groups than lanes, lane t takes group
    // t mod ng and lanes sharing a group write identical values, so the loop
    // and every access are the same for all lanes (no divergent branch)
#pragma unroll 1
    for (int q0 = 0; q0 < ng; q0 += kTpLanes) {
        uint32_t be = static_cast<uint32_t>((q0 + tid) & (ng - 1));
        for (int j = 0; j < op.nb; j++) {
            const int p = op.pos[j];
            be = ((be >> p) << (p + 1)) | (be & ((1U << p) - 1));
        }
        be |= op.fixed;
        v2<T> v[M];
#pragma unroll
        for (int j = 0; j < M; j++) {
            v[j] = tile[be | off[j]];
        }
        if (diag) {
#pragma unroll
            for (int j = 0; j < M; j++) {
                tile[be | off[j]] = cmul(op.mat[j * M + j], v[j]);
            }
        } else {
            // rows one at a time (not unrolled): each row's matrix entries are
            // loaded where they are used instead of all M^2 being held
#pragma unroll 1
            for (int r = 0; r < M; r++) {
                v2<T> acc = cmul(op.mat[r * M], v[0]);
#pragma unroll
                for (int c = 1; c < M; c++) {
                    cfma(acc, op.mat[r * M + c], v[c]);
                }
                uint32_t o = off[0];
#pragma unroll
                for (int j = 1; j < M; j++) {
                    o = r == j ? off[j] : o;
                }
                tile[be | o] = acc;
            }
        }
    }
}

template <class T>
__device__ __forceinline__ void tp_parity(v2<T> *tile, const TpOp<T> &op, int k, int tid) {
    const dcx<T> p0 = op.mat[0], p1 = op.mat[1];
#pragma unroll 1
    for (int e0 = 0; e0 < (1 << k); e0 += kTpLanes) { // 2^k >= 2^9: a whole number of wavefronts
        const int e = e0 + tid;
        tile[e] = cmul((__popc(static_cast<uint32_t>(e) & op.mask) & 1) ? p1 : p0, tile[e]);
    }
}

// Any op the register forms do not cover, applied to the tile staged in LDS
// (the tile is in LDS, not in registers, while it runs).
template <class T>
__device__ __forceinline__ void tp_lds_op(v2<T> *tile, const TpOp<T> *opp, int k, int tid) {
    const TpOp<T> &op = *opp;
    if (op.kind == 2) {
        tp_parity<T>(tile, op, k, tid);
    } else if (op.M == 2) {
        tp_group<T, 2>(tile, op, k, tid);
    } else if (op.M == 4) {
        tp_group<T, 4>(tile, op, k, tid);
    } else if (op.M == 1) {
        tp_group<T, 1>(tile, op, k, tid);
    } else {
        tp_group<T, 8>(tile, op, k, tid);
    }
}

// v from lane `src` (ds_bpermute on 32-bit words)
__device__ __forceinline__ float tp_bperm(float x, int src) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(x)));
}
__device__ __forceinline__ double tp_bperm(double x, int src) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_ds_bpermute(src << 2, static_cast<int>(b));
    const int hi = __builtin_amdgcn_ds_bpermute(src << 2, static_cast<int>(b >> 32));
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
template <class T> __device__ __forceinline__ v2<T> tp_lane_xor(const v2<T> &v, int src) {
    v2<T> o;
    o.x = tp_bperm(v.x, src);
    o.y = tp_bperm(v.y, src);
    return o;
}

// A register-resident amplitude: two scalars rather than the native 2-vector
// type, so each component is a register of its own that an update rewrites
// in place (a 2-vector component update rebuilds the whole register tuple,
// which costs copies at every merge of the interpreter's control flow).
template <class T> struct TpAmp {
    T x, y;
    TpAmp() = default;
    __device__ __forceinline__ TpAmp(const v2<T> &a) : x(a.x), y(a.y) {}
    __device__ __forceinline__ operator v2<T>() const {
        v2<T> o;
        o.x = x;
        o.y = y;
        return o;
    }
};

// A uniform matrix element read into SGPRs before a lane-dependent select
// picks it.  Without the empty asm the compiler sinks the load into a branch
// on the select's divergent condition, and that one divergent branch makes
// the whole interpreter loop a structurized region (copies of the register
// tile at every merge).
template <class T> __device__ __forceinline__ dcx<T> tp_ld(const dcx<T> &a) {
    dcx<T> r = a;
    asm volatile("" : "+s"(r.r), "+s"(r.i));
    return r;
}

enum TrCode : int {
    kTrDiag = 1, kTrParity = 2, kTrP1 = 3, kTrP2 = 4, kTrP4 = 5, kTrLds = 6, kTrLane = 7, kTrMix1 = 8, kTrMix2 = 9
};

template <class T> struct TrInst {
    int code;
    int ja, jb;              // p1: ja; p2/p4/diag: register bits ja < jb
    uint32_t rmask, rval;    // register-index condition; parity: register-index parity mask
    uint32_t tmask, tval;    // lane-index condition; parity: lane-index parity mask
    uint32_t lx, lx2;        // lane / mix1: the lane member bit; mix2: both, as lane-index masks
    int nr;                  // diag: member bits held in registers (0..2)
    int rcb[2];              // diag: phase-index bit fed by register ja / jb
    int tcb[2];              // diag: lane bit feeding phase-index bit i, or -1
    int lds_op;              // lds: index into the TpOp table
    dcx<T> mat[16];          // p1/p2/lane: 2x2; p4: 4x4; diag: phase by index c
                             // (bit i = member bit i); parity: even, odd;
                             // mix1: [8 o + 4 j + c] row j of lanes whose lane
                             // member bit is o, over (own j=0, own j=1,
                             // partner j=0, partner j=1); mix2: [4 o + d] the
                             // row of lanes with member pattern o, column o ^ d
};

// Complex products with every real operation fused: a * x in 4 operations,
// a * x + b * y in 8.
template <class T> __device__ __forceinline__ v2<T> tp_mul(dcx<T> a, v2<T> x) {
    v2<T> o;
    o.x = fma(a.r, x.x, -a.i * x.y);
    o.y = fma(a.r, x.y, a.i * x.x);
    return o;
}
template <class T> __device__ __forceinline__ void tp_mac(v2<T> &acc, dcx<T> a, v2<T> x) {
    acc.x = fma(a.r, x.x, fma(-a.i, x.y, acc.x));
    acc.y = fma(a.r, x.y, fma(a.i, x.x, acc.y));
}
template <class T> __device__ __forceinline__ v2<T> tp_mul2(dcx<T> a, v2<T> x, dcx<T> b, v2<T> y) {
    v2<T> o = tp_mul(b, y);
    tp_mac(o, a, x);
    return o;
}

// In-place 2x2 update of the register pair (x, y) by m[0..3]: LC selects
// the per-lane (VGPR) form of the matrix, otherwise it is uniform (SGPR).
template <class T, bool LC>
__device__ __forceinline__ void tr_pair2(TpAmp<T> &x, TpAmp<T> &y, const dcx<T> (&m)[16]) {
    if constexpr (std::is_same_v<T, double>) {
        if constexpr (LC) {
            ip_pair_v(x.x, x.y, y.x, y.y, m[0].r, m[0].i, m[1].r, m[1].i, m[2].r, m[2].i, m[3].r, m[3].i);
        } else {
            ip_pair_s(x.x, x.y, y.x, y.y, m[0].r, m[0].i, m[1].r, m[1].i, m[2].r, m[2].i, m[3].r, m[3].i);
        }
    } else {
        const v2<T> a = x, c = y;
        x = tp_mul2(m[0], a, m[1], c);
        y = tp_mul2(m[2], a, m[3], c);
    }
}

// Group ops whose member bits are register bits.  m: the op's matrix, the
// identity on lanes whose own condition fails.  CHK: the op has register-bit
// controls (a uniform test per register group).
template <class T, int KB, int J, bool CHK, bool LC>
__device__ __forceinline__ void tr_p1(TpAmp<T> (&v)[1 << KB], const TrInst<T> &in, const dcx<T> (&m)[16]) {
#pragma unroll
    for (int r = 0; r < (1 << KB); r++) {
        if ((r >> J) & 1) {
            continue;
        }
        if (CHK && (static_cast<uint32_t>(r) & in.rmask) != in.rval) {
            continue;
        }
        tr_pair2<T, LC>(v[r], v[r | (1 << J)], m);
    }
}

// members r | 1 << JA and r | 1 << JB (bits JA, JB of r clear)
template <class T, int KB, int JA, int JB, bool CHK, bool LC>
__device__ __forceinline__ void tr_p2(TpAmp<T> (&v)[1 << KB], const TrInst<T> &in, const dcx<T> (&m)[16]) {
#pragma unroll
    for (int r = 0; r < (1 << KB); r++) {
        if (((r >> JA) & 1) || ((r >> JB) & 1)) {
            continue;
        }
        if (CHK && (static_cast<uint32_t>(r) & in.rmask) != in.rval) {
            continue;
        }
        tr_pair2<T, LC>(v[r | (1 << JA)], v[r | (1 << JB)], m);
    }
}

// members r, r | 1 << JA, r | 1 << JB, r | both
template <class T, int KB, int JA, int JB, bool CHK, bool LC>
__device__ __forceinline__ void tr_p4(TpAmp<T> (&v)[1 << KB], const TrInst<T> &in, const dcx<T> (&m)[16]) {
#pragma unroll
    for (int r = 0; r < (1 << KB); r++) {
        if (((r >> JA) & 1) || ((r >> JB) & 1)) {
            continue;
        }
        if (CHK && (static_cast<uint32_t>(r) & in.rmask) != in.rval) {
            continue;
        }
        constexpr int a_ = 1 << JA, b_ = 1 << JB;
        const v2<T> x[4] = {v[r], v[r | a_], v[r | b_], v[r | a_ | b_]};
        v2<T> y[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            y[i] = tp_mul(m[4 * i], x[0]);
#pragma unroll
            for (int c = 1; c < 4; c++) {
                tp_mac(y[i], m[4 * i + c], x[c]);
            }
        }
        v[r] = y[0];
        v[r | a_] = y[1];
        v[r | b_] = y[2];
        v[r | a_ | b_] = y[3];
    }
}

// One independent test per register bit (no else chain): each case keeps
// its compile-time register indices.
// diagonal op: v[r] *= pt[c_r], c_r = the phase-index bits register r feeds
template <class T, int KB, bool CHK, int NR, int J0, int J1>
__device__ __forceinline__ void tr_diag(TpAmp<T> (&v)[1 << KB], const TrInst<T> &in, const dcx<T> (&pt)[4]) {
#pragma unroll
    for (int r = 0; r < (1 << KB); r++) {
        if (CHK && (static_cast<uint32_t>(r) & in.rmask) != in.rval) {
            continue;
        }
        int c = 0;
        if constexpr (NR >= 1) {
            c |= (r >> J0) & 1;
        }
        if constexpr (NR >= 2) {
            c |= ((r >> J1) & 1) << 1;
        }
        if constexpr (std::is_same_v<T, double>) {
            ip_scale(v[r].x, v[r].y, pt[c].r, pt[c].i);
        } else {
            v[r] = tp_mul(pt[c], v[r]);
        }
    }
}

template <class T, int KB, bool CHK, int J = 0>
__device__ __forceinline__ void tr_diag1_at(TpAmp<T> (&v)[1 << KB], const TrInst<T> &in, const dcx<T> (&pt)[4]) {
    if constexpr (J < KB) {
        if (in.ja == J) {
            tr_diag<T, KB, CHK, 1, J, 0>(v, in, pt);
        }
        tr_diag1_at<T, KB, CHK, J + 1>(v, in, pt);
    }
}

template <class T, int KB, bool CHK, int JA = 0, int JB = 1>
__device__ __forceinline__ void tr_diag2_at(TpAmp<T> (&v)[1 << KB], const TrInst<T> &in, const dcx<T> (&pt)[4]) {
    if constexpr (JB < KB) {
        if (in.ja == JA && in.jb == JB) {
            tr_diag<T, KB, CHK, 2, JA, JB>(v, in, pt);
        }
        if constexpr (JA + 1 < JB) {
            tr_diag2_at<T, KB, CHK, JA + 1, JB>(v, in, pt);
        } else {
            tr_diag2_at<T, KB, CHK, 0, JB + 1>(v, in, pt);
        }
    }
}

template <class T, int KB, bool CHK, bool LC, int J = 0>
__device__ __forceinline__ void tr_p1_at(TpAmp<T> (&v)[1 << KB], const TrInst<T> &in, const dcx<T> (&m)[16]) {
    if constexpr (J < KB) {
        if (in.ja == J) {
            tr_p1<T, KB, J, CHK, LC>(v, in, m);
        }
        tr_p1_at<T, KB, CHK, LC, J + 1>(v, in, m);
    }
}

template <class T, int KB, bool P4, bool CHK, bool LC, int JA = 0, int JB = 1>
__device__ __forceinline__ void tr_pair_at(TpAmp<T> (&v)[1 << KB], const TrInst<T> &in, const dcx<T> (&m)[16]) {
    if constexpr (JB < KB) {
        if (in.ja == JA && in.jb == JB) {
            if constexpr (P4) {
                tr_p4<T, KB, JA, JB, CHK, LC>(v, in, m);
            } else {
                tr_p2<T, KB, JA, JB, CHK, LC>(v, in, m);
            }
        }
        if constexpr (JA + 1 < JB) {
            tr_pair_at<T, KB, P4, CHK, LC, JA + 1, JB>(v, in, m);
        } else {
            tr_pair_at<T, KB, P4, CHK, LC, 0, JB + 1>(v, in, m);
        }
    }
}

// register-member group op with uniform (LC false) or per-lane matrix
template <class T, int KB, bool LC>
__device__ __forceinline__ void tr_pgroup(TpAmp<T> (&v)[1 << KB], const TrInst<T> &in, const dcx<T> (&m)[16]) {
    const bool chk = in.rmask != 0;
    if (in.code == kTrP1) {
        if (chk) {
            tr_p1_at<T, KB, true, LC>(v, in, m);
        } else {
            tr_p1_at<T, KB, false, LC>(v, in, m);
        }
    } else if (in.code == kTrP2) {
        if (chk) {
            tr_pair_at<T, KB, false, true, LC>(v, in, m);
        } else {
            tr_pair_at<T, KB, false, false, LC>(v, in, m);
        }
    } else if constexpr (!LC) { // kTrP4 carries no lane condition (the host sends those through LDS)
        if (chk) {
            tr_pair_at<T, KB, true, true, false>(v, in, m);
        } else {
            tr_pair_at<T, KB, true, false, false>(v, in, m);
        }
    }
}

// one lane member bit (lx) and one register member bit J: per register pair
// a lane exchanges both amplitudes with lane ^ lx and forms its own two.
// CHK: the op has register-bit controls (a uniform test per pair); lanes
// whose own condition fails carry identity weights
template <class T, int KB, int J, bool CHK>
__device__ __forceinline__ void tr_mix1(TpAmp<T> (&v)[1 << KB], const TrInst<T> &in, const dcx<T> (&w)[8],
                                        int src) {
#pragma unroll
    for (int r = 0; r < (1 << KB); r++) {
        if ((r >> J) & 1) {
            continue;
        }
        if (CHK && (static_cast<uint32_t>(r) & in.rmask) != in.rval) {
            continue;
        }
        constexpr int b = 1 << J;
        const v2<T> p0 = tp_lane_xor<T>(v[r], src), p1 = tp_lane_xor<T>(v[r | b], src);
        if constexpr (std::is_same_v<T, double>) {
            const double ww[16] = {w[0].r, w[0].i, w[1].r, w[1].i, w[2].r, w[2].i, w[3].r, w[3].i,
                                   w[4].r, w[4].i, w[5].r, w[5].i, w[6].r, w[6].i, w[7].r, w[7].i};
            ip_mix1(v[r].x, v[r].y, v[r | b].x, v[r | b].y, p0.x, p0.y, p1.x, p1.y, ww);
        } else {
            const v2<T> a0 = v[r], a1 = v[r | b];
            v2<T> y0 = tp_mul(w[0], a0), y1 = tp_mul(w[4], a0);
            tp_mac(y0, w[1], a1);
            tp_mac(y1, w[5], a1);
            tp_mac(y0, w[2], p0);
            tp_mac(y1, w[6], p0);
            tp_mac(y0, w[3], p1);
            tp_mac(y1, w[7], p1);
            v[r] = y0;
            v[r | b] = y1;
        }
    }
}

template <class T, int KB, bool CHK, int J = 0>
__device__ __forceinline__ void tr_mix1_at(TpAmp<T> (&v)[1 << KB], const TrInst<T> &in, const dcx<T> (&w)[8],
                                           int src) {
    if constexpr (J < KB) {
        if (in.ja == J) {
            tr_mix1<T, KB, J, CHK>(v, in, w, src);
        }
        tr_mix1_at<T, KB, CHK, J + 1>(v, in, w, src);
    }
}

// 2x2 op on a lane bit: each register takes its partner from lane ^ lx.
// CHK: register-bit controls; lanes whose own condition fails carry
// identity weights (a = 1, b = 0), so every lane runs the same code
template <class T, int KB, bool CHK>
__device__ __forceinline__ void tr_lane(TpAmp<T> (&v)[1 << KB], const TrInst<T> &in, dcx<T> a, dcx<T> b, int src) {
#pragma unroll
    for (int r = 0; r < (1 << KB); r++) {
        if (CHK && (static_cast<uint32_t>(r) & in.rmask) != in.rval) {
            continue;
        }
        const v2<T> o = tp_lane_xor<T>(v[r], src);
        if constexpr (std::is_same_v<T, double>) {
            ip_lane(v[r].x, v[r].y, o.x, o.y, a.r, a.i, b.r, b.i);
        } else {
            v[r] = tp_mul2(a, v[r], b, o);
        }
    }
}

// two lane member bits: three exchanges per register (same conventions)
template <class T, int KB, bool CHK>
__device__ __forceinline__ void tr_mix2(TpAmp<T> (&v)[1 << KB], const TrInst<T> &in, const dcx<T> (&w)[4], int lane,
                                        int l1, int l2) {
#pragma unroll
    for (int r = 0; r < (1 << KB); r++) {
        if (CHK && (static_cast<uint32_t>(r) & in.rmask) != in.rval) {
            continue;
        }
        const v2<T> s1 = tp_lane_xor<T>(v[r], lane ^ l1);
        const v2<T> s2 = tp_lane_xor<T>(v[r], lane ^ l2);
        const v2<T> s3 = tp_lane_xor<T>(v[r], lane ^ l1 ^ l2);
        if constexpr (std::is_same_v<T, double>) {
            const double ww[8] = {w[0].r, w[0].i, w[1].r, w[1].i, w[2].r, w[2].i, w[3].r, w[3].i};
            ip_mix2(v[r].x, v[r].y, s1.x, s1.y, s2.x, s2.y, s3.x, s3.y, ww);
        } else {
            v2<T> y = tp_mul(w[0], v[r]);
            tp_mac(y, w[1], s1);
            tp_mac(y, w[2], s2);
            tp_mac(y, w[3], s3);
            v[r] = y;
        }
    }
}

// The program of a pass applied to one tile held in registers (v[r]: the
// amplitude of local index lane | r << 6).  lds: the workgroup's LDS tile,
// used only by kTrLds operations.
template <class T, int KB>
__device__ __forceinline__ void tp_run(TpAmp<T> (&v)[1 << KB], const TrInst<T> *__restrict__ prog, int ninst,
                                       const TpOp<T> *__restrict__ lops, v2<T> *lds, int lane) {
    constexpr int R = 1 << KB;
    const int k = 6 + KB;
    for (int pc = 0; pc < ninst; pc++) {
        const TrInst<T> &in = prog[pc];
        const int code = in.code;
        if (code == kTrDiag) {
            // per-lane phases for each combination of the register-fed index bits
            const bool tc = (static_cast<uint32_t>(lane) & in.tmask) == in.tval;
            int ct = 0;
            if (in.tcb[0] >= 0) {
                ct |= (lane >> in.tcb[0]) & 1;
            }
            if (in.tcb[1] >= 0) {
                ct |= ((lane >> in.tcb[1]) & 1) << 1;
            }
            const dcx<T> d0 = tp_ld(in.mat[0]), d1 = tp_ld(in.mat[1]), d2 = tp_ld(in.mat[2]), d3 = tp_ld(in.mat[3]);
            const int nr = in.nr;
            dcx<T> pt[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                pt[q] = dcx<T>{T(1), T(0)};
                if (q < (1 << nr)) {
                    int c = ct;
                    if (q & 1) {
                        c |= 1 << in.rcb[0];
                    }
                    if (q & 2) {
                        c |= 1 << in.rcb[1];
                    }
                    // two-level select on the index bits (components kept apart)
                    const bool b0 = c & 1, b1 = c & 2;
                    const T lr = b0 ? d1.r : d0.r, li = b0 ? d1.i : d0.i;
                    const T hr = b0 ? d3.r : d2.r, hi = b0 ? d3.i : d2.i;
                    pt[q].r = tc ? (b1 ? hr : lr) : T(1);
                    pt[q].i = tc ? (b1 ? hi : li) : T(0);
                }
            }
            if (in.rmask != 0) { // register-bit condition
                if (nr == 0) {
                    tr_diag<T, KB, true, 0, 0, 0>(v, in, pt);
                } else if (nr == 1) {
                    tr_diag1_at<T, KB, true>(v, in, pt);
                } else {
                    tr_diag2_at<T, KB, true>(v, in, pt);
                }
            } else if (nr == 0) {
                tr_diag<T, KB, false, 0, 0, 0>(v, in, pt);
            } else if (nr == 1) {
                tr_diag1_at<T, KB, false>(v, in, pt);
            } else {
                tr_diag2_at<T, KB, false>(v, in, pt);
            }
        } else if (code == kTrParity) {
            const bool odd = __popc(static_cast<uint32_t>(lane) & in.tmask) & 1;
            const dcx<T> e0 = tp_ld(in.mat[0]), e1 = tp_ld(in.mat[1]);
            const dcx<T> p0{odd ? e1.r : e0.r, odd ? e1.i : e0.i};
            const dcx<T> p1{odd ? e0.r : e1.r, odd ? e0.i : e1.i};
#pragma unroll
            for (int r = 0; r < R; r++) {
                const bool rodd = __popc(static_cast<uint32_t>(r) & in.rmask) & 1;
                const dcx<T> ph{rodd ? p1.r : p0.r, rodd ? p1.i : p0.i};
                if constexpr (std::is_same_v<T, double>) {
                    ip_scale(v[r].x, v[r].y, ph.r, ph.i);
                } else {
                    v[r] = tp_mul(ph, v[r]);
                }
            }
        } else if (code == kTrLane) {
            // partner amplitude from lane ^ lx; every lane takes part in the exchange
            const bool hi = (static_cast<uint32_t>(lane) & in.lx) != 0;
            const bool on = (static_cast<uint32_t>(lane) & in.tmask) == in.tval;
            const dcx<T> m0 = tp_ld(in.mat[0]), m1 = tp_ld(in.mat[1]), m2 = tp_ld(in.mat[2]), m3 = tp_ld(in.mat[3]);
            // own and partner weights; identity where the lane condition fails
            const dcx<T> a{on ? (hi ? m3.r : m0.r) : T(1), on ? (hi ? m3.i : m0.i) : T(0)};
            const dcx<T> b{on ? (hi ? m2.r : m1.r) : T(0), on ? (hi ? m2.i : m1.i) : T(0)};
            const int src = lane ^ static_cast<int>(in.lx);
            if (in.rmask == 0) {
                tr_lane<T, KB, false>(v, in, a, b, src);
            } else {
                tr_lane<T, KB, true>(v, in, a, b, src);
            }
        } else if (code == kTrMix1) {
            const bool hi = (static_cast<uint32_t>(lane) & in.lx) != 0;
            const bool on = (static_cast<uint32_t>(lane) & in.tmask) == in.tval;
            dcx<T> w[8];
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const dcx<T> lo_ = tp_ld(in.mat[c]), hi_ = tp_ld(in.mat[8 + c]);
                // identity rows (own amplitude, weight 1) where the lane condition fails
                const T idr = (c == 0 || c == 5) ? T(1) : T(0);
                w[c] = dcx<T>{on ? (hi ? hi_.r : lo_.r) : idr, on ? (hi ? hi_.i : lo_.i) : T(0)};
            }
            const int src = lane ^ static_cast<int>(in.lx);
            if (in.rmask != 0) {
                tr_mix1_at<T, KB, true>(v, in, w, src);
            } else {
                tr_mix1_at<T, KB, false>(v, in, w, src);
            }
        } else if (code == kTrMix2) {
            // two lane member bits: three exchanges per register
            const int o = ((static_cast<uint32_t>(lane) & in.lx) ? 1 : 0) | ((static_cast<uint32_t>(lane) & in.lx2) ? 2 : 0);
            const bool on = (static_cast<uint32_t>(lane) & in.tmask) == in.tval;
            dcx<T> w[4];
#pragma unroll
            for (int d = 0; d < 4; d++) {
                const dcx<T> e0 = tp_ld(in.mat[d]), e1 = tp_ld(in.mat[4 + d]), e2 = tp_ld(in.mat[8 + d]),
                             e3 = tp_ld(in.mat[12 + d]);
                const bool b0 = o & 1, b1 = o & 2;
                const T lr = b0 ? e1.r : e0.r, li = b0 ? e1.i : e0.i;
                const T hr = b0 ? e3.r : e2.r, hi = b0 ? e3.i : e2.i;
                w[d] = dcx<T>{on ? (b1 ? hr : lr) : (d == 0 ? T(1) : T(0)), on ? (b1 ? hi : li) : T(0)};
            }
            const int l1 = static_cast<int>(in.lx), l2 = static_cast<int>(in.lx2);
            if (in.rmask == 0) {
                tr_mix2<T, KB, false>(v, in, w, lane, l1, l2);
            } else {
                tr_mix2<T, KB, true>(v, in, w, lane, l1, l2);
            }
        } else if (code == kTrLds) {
            v2<T> *const lp = lds + lane; // register r at lp[64 r]: immediate offsets
#pragma unroll
            for (int r = 0; r < R; r++) {
                lp[r * kTpLanes] = v[r];
            }
            __syncthreads();
            tp_lds_op<T>(lds, lops + in.lds_op, k, lane);
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; r++) {
                v[r] = lp[r * kTpLanes];
            }
            __syncthreads();
        } else { // kTrP1, kTrP2, kTrP4
            dcx<T> m[16];
            if (in.tmask == 0) {
#pragma unroll
                for (int b = 0; b < 16; b++) {
                    m[b] = in.mat[b];
                }
                tr_pgroup<T, KB, false>(v, in, m);
            } else {
                // kTrP1 / kTrP2: lanes whose own condition fails apply the identity
                const bool on = (static_cast<uint32_t>(lane) & in.tmask) == in.tval;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const T id = (b == 0 || b == 3) ? T(1) : T(0);
                    const dcx<T> mb = tp_ld(in.mat[b]);
                    m[b] = dcx<T>{on ? mb.r : id, on ? mb.i : T(0)};
                }
                tr_pgroup<T, KB, true>(v, in, m);
            }
        }
    }
}

// KB = k - 6: the number of local bits held in registers.  One wavefront
// per tile, one tile per wavefront: the hardware launches a new wavefront as
// soon as one retires.  (Measured at 30 qubits, complex128: 7.1 ms per pass;
// persistent wavefronts that prefetch their next tile into registers, at
// two per SIMD, took 7.9 ms and lower DRAM efficiency; forcing four per
// SIMD instead of three gained nothing.)  `lds_bytes` of dynamic LDS only
// when the pass has lds ops.
template <class T, int KB>
__global__ __launch_bounds__(kTpLanes) void k_tile_pass(v2<T> *__restrict__ sv, uint64_t ntiles,
                                                        const TpGeom g, const TrInst<T> *__restrict__ prog,
                                                        int ninst, const TpOp<T> *__restrict__ lops) {
    extern __shared__ unsigned char smem_raw[];
    v2<T> *lds = reinterpret_cast<v2<T> *>(smem_raw);
    constexpr int R = 1 << KB;
    const uint64_t t = xcd_block();
    if (t >= ntiles) {
        return;
    }
    const int lane = threadIdx.x;
    // HBM offsets of the lane's local bits 0..5 and of each register's bits
    uint64_t loff = static_cast<uint64_t>(lane) & ((uint64_t{1} << g.low) - 1);
    const int nlane_hi = 6 - g.low;
    for (int j = 0; j < nlane_hi; j++) {
        if ((lane >> (g.low + j)) & 1) {
            loff |= uint64_t{1} << g.hb[j];
        }
    }
    // a register's offset: uniform, recomputed where used (the empty asm keeps
    // the compiler from holding all of them through the program)
    auto roff = [&](int r) {
        uint64_t o = 0;
#pragma unroll
        for (int j = 0; j < KB; j++) {
            int hbj = g.hb[nlane_hi + j];
            asm volatile("" : "+s"(hbj));
            if ((r >> j) & 1) {
                o |= uint64_t{1} << hbj;
            }
        }
        return o;
    };
    v2<T> *const p = sv + (expand(t << g.low, g.outer) | loff);
    TpAmp<T> v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        v[r] = ld_nt<T>(p + roff(r));
    }
    tp_run<T, KB>(v, prog, ninst, lops, lds, lane);
#pragma unroll
    for (int r = 0; r < R; r++) {
        st_nt<T>(p + roff(r), v[r]);
    }
}
} // namespace plamd
