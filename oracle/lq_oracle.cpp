/*
 * CPU oracle: a restatement of Lightning-Qubit's hot-path semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, as the checker; the product
 * package never links or calls it (see oracle_api.h for the interface).
 *
 * What it restates, and what pins each part:
 *   gate kernels      GateImplementationsLM.hpp (default kernel for all gates
 *                     but Toffoli/CSWAP/DoubleExcitation*, which follow
 *                     GateImplementationsPI.hpp; AssignKernelMap_Default.cpp:41-231)
 *                     -> pinned against oracle/_ref (the reference compiled
 *                     in place) on every gate x wire tuple, against committed
 *                     golden vectors, and against closed-form matrices.
 *   generators        GateImplementationsLM.hpp:1355-1640, PauliGenerator.hpp:31-56,
 *                     GateImplementationsPI.cpp:129-205
 *                     -> pinned against oracle/_ref and the identity
 *                     dU/dtheta = i s G U by finite difference.
 *   measurements      MeasurementsLQubit.hpp:79-317,571-605 (copy-apply-dot)
 *                     -> pinned against oracle/_ref and brute-force dense algebra.
 *   CDF sampling      lightning_kokkos/measurements/MeasurementsKokkos.hpp:700-741,
 *                     MeasuresFunctors.hpp:55-105 (exclusive scan + bisection)
 *                     -> parity pinned by brute-force counting on small
 *                     states; the reference's LQ sampler (alias method,
 *                     unseedable) cannot be compared draw by draw.
 *   adjoint Jacobian  AdjointJacobianLQubit.hpp:212-336
 *                     -> pinned against oracle/_ref and parameter shift.
 *   VJP               VectorJacobianProduct.hpp:77-155
 *                     -> pinned against oracle/_ref, central differences of
 *                     the forward pass, and 2 Re VJP(O psi) = adjoint row.
 *   sparse H          MeasurementsLQubit.hpp:183-214,437-471 + SparseLinAlg.hpp
 *                     -> pinned against oracle/_ref and dense numpy algebra.
 * Wire w is bit (n-1-w) of the basis index (GateImplementationsLM.hpp:209).
 */
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_api.h"

namespace {

thread_local std::string g_err;

template <class P> using C = std::complex<P>;

enum Gate : int {
    Identity = 0, PauliX, PauliY, PauliZ, Hadamard, S, T, PhaseShift, RX, RY,
    RZ, Rot, CNOT, CY, CZ, SWAP, IsingXX, IsingXY, IsingYY, IsingZZ,
    ControlledPhaseShift, CRX, CRY, CRZ, CRot, SingleExcitation,
    SingleExcitationMinus, SingleExcitationPlus, Toffoli, CSWAP,
    DoubleExcitation, DoubleExcitationMinus, DoubleExcitationPlus, MultiRZ,
    GateEnd
};

/* gate arity and parameter count: core/src/gates/Constant.hpp:163-269 */
constexpr int kWires[GateEnd] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2,
                                 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 3, 3, 4, 4, 4, -1};
constexpr int kParams[GateEnd] = {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 3, 0, 0, 0, 0, 1,
                                  1, 1, 1, 1, 1, 1, 1, 3, 1, 1, 1, 0, 0, 1, 1, 1, 1};

void fail(const std::string &m) { throw std::runtime_error(m); }

/*
 * Enumerate groups of amplitudes touched by a k-wire gate: for every k-bit
 * free index `g`, insert zeros at the (ascending) bit positions of the gate's
 * wires.  Restates revWireParity (GateImplementationsLM.hpp:51-125), which
 * does the same with precomputed parity masks.
 */
struct Groups {
    size_t masks[8];
    int m;
    Groups(std::vector<size_t> rev) : masks{}, m(static_cast<int>(rev.size())) {
        std::sort(rev.begin(), rev.end());
        auto trailing = [](size_t b) { return b == 0 ? size_t{0} : (~size_t{0} >> (64 - b)); };
        auto leading = [&](size_t b) { return ~trailing(b); };
        masks[0] = trailing(rev[0]);
        for (int i = 1; i < m; i++) {
            masks[i] = leading(rev[i - 1] + 1) & trailing(rev[i]);
        }
        masks[m] = leading(rev[m - 1] + 1);
    }
    size_t base(size_t k) const {
        size_t r = 0;
        for (int i = 0; i <= m; i++) {
            r |= (k << i) & masks[i];
        }
        return r;
    }
};

template <class P> struct SV {
    C<P> *a;
    size_t n;
    size_t len() const { return size_t{1} << n; }
    size_t rev(size_t w) const {
        if (w >= n) {
            fail("wire index out of range");
        }
        return n - 1 - w;
    }
};

/* 1-qubit loop (GateImplementationsLM.hpp:204-249 index map). */
template <class P, class F> void for1(const SV<P> &s, const std::vector<size_t> &w, F f) {
    const size_t r = s.rev(w[0]);
    const Groups g({r});
    const size_t sh = size_t{1} << r;
    for (size_t k = 0; k < (s.len() >> 1); k++) {
        const size_t i0 = g.base(k);
        f(i0, i0 | sh);
    }
}
/* 2-qubit loop; wires[0] is the high bit of the 4-index (LM.hpp:251-325). */
template <class P, class F> void for2(const SV<P> &s, const std::vector<size_t> &w, F f) {
    const size_t r0 = s.rev(w[1]), r1 = s.rev(w[0]);
    if (r0 == r1) {
        fail("wires must be distinct");
    }
    const Groups g({r0, r1});
    const size_t s0 = size_t{1} << r0, s1 = size_t{1} << r1;
    for (size_t k = 0; k < (s.len() >> 2); k++) {
        const size_t i00 = g.base(k);
        f(i00, i00 | s0, i00 | s1, i00 | s0 | s1);
    }
}
/* k-qubit loop; member j of a group has wires[0] as its most significant
 * bit (GateIndices / generateBitPatterns, GateIndices.cpp:31-45). */
template <class P, class F> void forK(const SV<P> &s, const std::vector<size_t> &w, F f) {
    const size_t k = w.size();
    std::vector<size_t> revs(k), off(size_t{1} << k, 0);
    for (size_t i = 0; i < k; i++) {
        revs[i] = s.rev(w[i]);
    }
    {
        auto sorted = revs;
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) {
            fail("wires must be distinct");
        }
    }
    for (size_t j = 0; j < off.size(); j++) {
        for (size_t b = 0; b < k; b++) {
            if ((j >> (k - 1 - b)) & 1U) {
                off[j] |= size_t{1} << revs[b];
            }
        }
    }
    const Groups g(revs);
    std::vector<size_t> idx(off.size());
    for (size_t q = 0; q < (s.len() >> k); q++) {
        const size_t b = g.base(q);
        for (size_t j = 0; j < off.size(); j++) {
            idx[j] = b | off[j];
        }
        f(idx);
    }
}

template <class P> C<P> cexp(P x) { return {std::cos(x), std::sin(x)}; }

/* Rot(phi, theta, omega) = RZ(omega) RY(theta) RZ(phi): core/src/gates/Gates.hpp:354-363 */
template <class P> std::vector<C<P>> rot(P phi, P theta, P omega) {
    const P c = std::cos(theta / 2), s = std::sin(theta / 2);
    const P p = phi + omega, m = phi - omega;
    return {C<P>{std::cos(p / 2) * c, -std::sin(p / 2) * c},
            C<P>{-std::cos(m / 2) * s, -std::sin(m / 2) * s},
            C<P>{std::cos(m / 2) * s, -std::sin(m / 2) * s},
            C<P>{std::cos(p / 2) * c, std::sin(p / 2) * c}};
}

/* Dense matrices: LM applySingleQubitOp / applyTwoQubitOp (LM.hpp:204-325),
 * PI applyMultiQubitOp for >= 3 wires (GateImplementationsPI.hpp:250-290).
 * `inverse` applies the conjugate transpose. */
template <class P>
void apply_matrix(SV<P> &s, const C<P> *m, const std::vector<size_t> &w, bool inv) {
    if (w.empty()) {
        fail("Number of wires must be larger than 0");
    }
    if (w.size() > s.n) {
        fail("Assertion failed: num_qubits >= wires.size()");
    }
    C<P> *a = s.a;
    if (w.size() == 1) {
        for1(s, w, [&](size_t i0, size_t i1) {
            const C<P> v0 = a[i0], v1 = a[i1];
            if (inv) {
                a[i0] = std::conj(m[0]) * v0 + std::conj(m[2]) * v1;
                a[i1] = std::conj(m[1]) * v0 + std::conj(m[3]) * v1;
            } else {
                a[i0] = m[0] * v0 + m[1] * v1;
                a[i1] = m[2] * v0 + m[3] * v1;
            }
        });
        return;
    }
    if (w.size() == 2) {
        for2(s, w, [&](size_t i00, size_t i01, size_t i10, size_t i11) {
            const C<P> v[4] = {a[i00], a[i01], a[i10], a[i11]};
            const size_t o[4] = {i00, i01, i10, i11};
            for (int r = 0; r < 4; r++) {
                C<P> acc = inv ? std::conj(m[r]) * v[0] : m[4 * r] * v[0];
                for (int c = 1; c < 4; c++) {
                    acc += inv ? std::conj(m[4 * c + r]) * v[c] : m[4 * r + c] * v[c];
                }
                a[o[r]] = acc;
            }
        });
        return;
    }
    const size_t dim = size_t{1} << w.size();
    std::vector<C<P>> v(dim);
    forK(s, w, [&](const std::vector<size_t> &idx) {
        for (size_t j = 0; j < dim; j++) {
            v[j] = a[idx[j]];
        }
        for (size_t i = 0; i < dim; i++) {
            C<P> acc{0, 0};
            for (size_t j = 0; j < dim; j++) {
                acc += inv ? std::conj(m[j * dim + i]) * v[j] : m[i * dim + j] * v[j];
            }
            a[idx[i]] = acc;
        }
    });
}

/* Named gates: GateImplementationsLM.hpp:389-1352, LM.cpp:27-118 and
 * GateImplementationsPI.cpp:17-126 / PI.hpp:786-815. */
template <class P>
void apply_gate(SV<P> &s, int op, const std::vector<size_t> &w, bool inv,
                const std::vector<P> &prm) {
    if (op < 0 || op >= GateEnd) {
        fail("unknown gate id");
    }
    if (kWires[op] > 0 && static_cast<int>(w.size()) != kWires[op]) {
        fail("Assertion failed: wires.size() == " + std::to_string(kWires[op]));
    }
    if (static_cast<int>(prm.size()) != kParams[op]) {
        fail("Assertion failed: params.size() == lookup(gate_num_params, gate_op)");
    }
    C<P> *a = s.a;
    const P th = prm.empty() ? P{0} : prm[0];
    const P sgn = inv ? P{-1} : P{1};
    const P isq2 = static_cast<P>(M_SQRT1_2);
    switch (op) {
    case Identity:
        return;
    case PauliX:
        return for1(s, w, [&](size_t i0, size_t i1) { std::swap(a[i0], a[i1]); });
    case PauliY:
        return for1(s, w, [&](size_t i0, size_t i1) {
            const C<P> v0 = a[i0], v1 = a[i1];
            a[i0] = {v1.imag(), -v1.real()};
            a[i1] = {-v0.imag(), v0.real()};
        });
    case PauliZ:
        return for1(s, w, [&](size_t, size_t i1) { a[i1] *= P{-1}; });
    case Hadamard:
        return for1(s, w, [&](size_t i0, size_t i1) {
            const C<P> v0 = a[i0], v1 = a[i1];
            a[i0] = isq2 * v0 + isq2 * v1;
            a[i1] = isq2 * v0 - isq2 * v1;
        });
    case S: {
        const C<P> f{0, sgn};
        return for1(s, w, [&](size_t, size_t i1) { a[i1] *= f; });
    }
    case T: {
        const C<P> f{isq2, sgn * isq2};
        return for1(s, w, [&](size_t, size_t i1) { a[i1] *= f; });
    }
    case PhaseShift: {
        const C<P> f = inv ? std::exp(-C<P>(0, th)) : std::exp(C<P>(0, th));
        return for1(s, w, [&](size_t, size_t i1) { a[i1] *= f; });
    }
    case RX: {
        const P c = std::cos(th / 2);
        const P js = inv ? -std::sin(-th / 2) : std::sin(-th / 2);
        return for1(s, w, [&](size_t i0, size_t i1) {
            const C<P> v0 = a[i0], v1 = a[i1];
            a[i0] = c * v0 + C<P>{-v1.imag() * js, v1.real() * js};
            a[i1] = C<P>{-v0.imag() * js, v0.real() * js} + c * v1;
        });
    }
    case RY: {
        const P c = std::cos(th / 2);
        const P sn = inv ? -std::sin(th / 2) : std::sin(th / 2);
        return for1(s, w, [&](size_t i0, size_t i1) {
            const C<P> v0 = a[i0], v1 = a[i1];
            a[i0] = {c * v0.real() - sn * v1.real(), c * v0.imag() - sn * v1.imag()};
            a[i1] = {sn * v0.real() + c * v1.real(), sn * v0.imag() + c * v1.imag()};
        });
    }
    case RZ: {
        const C<P> f0{std::cos(th / 2), -std::sin(th / 2)};
        const C<P> f1{std::cos(th / 2), std::sin(th / 2)};
        const C<P> s0 = inv ? std::conj(f0) : f0, s1 = inv ? std::conj(f1) : f1;
        return for1(s, w, [&](size_t i0, size_t i1) {
            a[i0] *= s0;
            a[i1] *= s1;
        });
    }
    case Rot: {
        const auto m = inv ? rot<P>(-prm[2], -prm[1], -prm[0]) : rot<P>(prm[0], prm[1], prm[2]);
        return apply_matrix(s, m.data(), w, false);
    }
    case CNOT:
        return for2(s, w, [&](size_t, size_t, size_t i10, size_t i11) { std::swap(a[i10], a[i11]); });
    case CY:
        return for2(s, w, [&](size_t, size_t, size_t i10, size_t i11) {
            const C<P> v10 = a[i10];
            a[i10] = {a[i11].imag(), -a[i11].real()};
            a[i11] = {-v10.imag(), v10.real()};
        });
    case CZ:
        return for2(s, w, [&](size_t, size_t, size_t, size_t i11) { a[i11] *= P{-1}; });
    case SWAP:
        return for2(s, w, [&](size_t, size_t i01, size_t i10, size_t) { std::swap(a[i10], a[i01]); });
    case IsingXX: {
        const P cr = std::cos(th / 2), sj = sgn * std::sin(th / 2);
        return for2(s, w, [&](size_t i00, size_t i01, size_t i10, size_t i11) {
            const C<P> v00 = a[i00], v01 = a[i01], v10 = a[i10], v11 = a[i11];
            a[i00] = {cr * v00.real() + sj * v11.imag(), cr * v00.imag() - sj * v11.real()};
            a[i01] = {cr * v01.real() + sj * v10.imag(), cr * v01.imag() - sj * v10.real()};
            a[i10] = {cr * v10.real() + sj * v01.imag(), cr * v10.imag() - sj * v01.real()};
            a[i11] = {cr * v11.real() + sj * v00.imag(), cr * v11.imag() - sj * v00.real()};
        });
    }
    case IsingXY: {
        const P cr = std::cos(th / 2), sj = sgn * std::sin(th / 2);
        return for2(s, w, [&](size_t, size_t i01, size_t i10, size_t) {
            const C<P> v01 = a[i01], v10 = a[i10];
            a[i01] = {cr * v01.real() - sj * v10.imag(), cr * v01.imag() + sj * v10.real()};
            a[i10] = {cr * v10.real() - sj * v01.imag(), cr * v10.imag() + sj * v01.real()};
        });
    }
    case IsingYY: {
        const P cr = std::cos(th / 2), sj = sgn * std::sin(th / 2);
        return for2(s, w, [&](size_t i00, size_t i01, size_t i10, size_t i11) {
            const C<P> v00 = a[i00], v01 = a[i01], v10 = a[i10], v11 = a[i11];
            a[i00] = {cr * v00.real() - sj * v11.imag(), cr * v00.imag() + sj * v11.real()};
            a[i01] = {cr * v01.real() + sj * v10.imag(), cr * v01.imag() - sj * v10.real()};
            a[i10] = {cr * v10.real() + sj * v01.imag(), cr * v10.imag() - sj * v01.real()};
            a[i11] = {cr * v11.real() - sj * v00.imag(), cr * v11.imag() + sj * v00.real()};
        });
    }
    case IsingZZ: {
        const C<P> f0{std::cos(th / 2), -std::sin(th / 2)};
        const C<P> f1{std::cos(th / 2), std::sin(th / 2)};
        const C<P> s0 = inv ? std::conj(f0) : f0, s1 = inv ? std::conj(f1) : f1;
        return for2(s, w, [&](size_t i00, size_t i01, size_t i10, size_t i11) {
            a[i00] *= s0;
            a[i01] *= s1;
            a[i10] *= s1;
            a[i11] *= s0;
        });
    }
    case ControlledPhaseShift: {
        const C<P> f = inv ? std::exp(-C<P>(0, th)) : std::exp(C<P>(0, th));
        return for2(s, w, [&](size_t, size_t, size_t, size_t i11) { a[i11] *= f; });
    }
    case CRX: {
        const P c = std::cos(th / 2), js = sgn * std::sin(th / 2);
        return for2(s, w, [&](size_t, size_t, size_t i10, size_t i11) {
            const C<P> v10 = a[i10], v11 = a[i11];
            a[i10] = {c * v10.real() + js * v11.imag(), c * v10.imag() - js * v11.real()};
            a[i11] = {c * v11.real() + js * v10.imag(), c * v11.imag() - js * v10.real()};
        });
    }
    case CRY: {
        const P c = std::cos(th / 2), sn = sgn * std::sin(th / 2);
        return for2(s, w, [&](size_t, size_t, size_t i10, size_t i11) {
            const C<P> v10 = a[i10], v11 = a[i11];
            a[i10] = {c * v10.real() - sn * v11.real(), c * v10.imag() - sn * v11.imag()};
            a[i11] = {sn * v10.real() + c * v11.real(), sn * v10.imag() + c * v11.imag()};
        });
    }
    case CRZ: {
        const C<P> f0{std::cos(th / 2), -std::sin(th / 2)};
        const C<P> f1{std::cos(th / 2), std::sin(th / 2)};
        const C<P> s0 = inv ? std::conj(f0) : f0, s1 = inv ? std::conj(f1) : f1;
        return for2(s, w, [&](size_t, size_t, size_t i10, size_t i11) {
            a[i10] *= s0;
            a[i11] *= s1;
        });
    }
    case CRot: {
        const auto m = inv ? rot<P>(-prm[2], -prm[1], -prm[0]) : rot<P>(prm[0], prm[1], prm[2]);
        return for2(s, w, [&](size_t, size_t, size_t i10, size_t i11) {
            const C<P> v0 = a[i10], v1 = a[i11];
            a[i10] = m[0] * v0 + m[1] * v1;
            a[i11] = m[2] * v0 + m[3] * v1;
        });
    }
    case SingleExcitation:
    case SingleExcitationMinus:
    case SingleExcitationPlus: {
        const P c = std::cos(th / 2), sn = sgn * std::sin(th / 2);
        C<P> e{1, 0};
        if (op == SingleExcitationMinus) {
            e = inv ? std::exp(C<P>(0, th / 2)) : std::exp(-C<P>(0, th / 2));
        } else if (op == SingleExcitationPlus) {
            e = inv ? std::exp(-C<P>(0, th / 2)) : std::exp(C<P>(0, th / 2));
        }
        const bool phase = op != SingleExcitation;
        return for2(s, w, [&](size_t i00, size_t i01, size_t i10, size_t i11) {
            const C<P> v01 = a[i01], v10 = a[i10];
            if (phase) {
                a[i00] *= e;
            }
            a[i01] = c * v01 - sn * v10;
            a[i10] = sn * v01 + c * v10;
            if (phase) {
                a[i11] *= e;
            }
        });
    }
    case Toffoli:
        return forK(s, w, [&](const std::vector<size_t> &i) { std::swap(a[i[6]], a[i[7]]); });
    case CSWAP:
        return forK(s, w, [&](const std::vector<size_t> &i) { std::swap(a[i[5]], a[i[6]]); });
    case DoubleExcitation:
    case DoubleExcitationMinus:
    case DoubleExcitationPlus: {
        const P c = std::cos(th / 2), sn = sgn * std::sin(th / 2);
        C<P> e{1, 0};
        if (op == DoubleExcitationMinus) {
            e = inv ? std::exp(C<P>(0, th / 2)) : std::exp(-C<P>(0, th / 2));
        } else if (op == DoubleExcitationPlus) {
            e = inv ? std::exp(-C<P>(0, th / 2)) : std::exp(C<P>(0, th / 2));
        }
        const bool phase = op != DoubleExcitation;
        return forK(s, w, [&](const std::vector<size_t> &i) {
            const C<P> v3 = a[i[3]], v12 = a[i[12]];
            if (phase) {
                for (int j = 0; j < 16; j++) {
                    if (j != 3 && j != 12) {
                        a[i[j]] *= e;
                    }
                }
            }
            a[i[3]] = c * v3 - sn * v12;
            a[i[12]] = sn * v3 + c * v12;
        });
    }
    case MultiRZ: {
        const C<P> f0{std::cos(th / 2), -std::sin(th / 2)};
        const C<P> f1{std::cos(th / 2), std::sin(th / 2)};
        const C<P> sh[2] = {inv ? std::conj(f0) : f0, inv ? std::conj(f1) : f1};
        size_t mask = 0;
        for (size_t wi : w) {
            mask |= size_t{1} << s.rev(wi);
        }
        for (size_t k = 0; k < s.len(); k++) {
            a[k] *= sh[__builtin_popcountll(k & mask) & 1];
        }
        return;
    }
    default:
        fail("unsupported gate");
    }
}

/* Generators: overwrite psi <- G psi and return s with U(t) = exp(i s t G).
 * LM.hpp:1355-1640, PauliGenerator.hpp:31-56, LM.cpp:120-205,
 * PI.cpp:128-205 (DoubleExcitation*, PI is their default kernel). */
template <class P> P apply_generator(SV<P> &s, int op, const std::vector<size_t> &w) {
    C<P> *a = s.a;
    const C<P> I{0, 1};
    const C<P> Z{0, 0};
    if (op >= 0 && op < GateEnd && kWires[op] > 0 && static_cast<int>(w.size()) != kWires[op]) {
        fail("Assertion failed: wires.size() == " + std::to_string(kWires[op]));
    }
    switch (op) {
    case RX:
        apply_gate<P>(s, PauliX, w, false, {});
        return P(-0.5);
    case RY:
        apply_gate<P>(s, PauliY, w, false, {});
        return P(-0.5);
    case RZ:
        apply_gate<P>(s, PauliZ, w, false, {});
        return P(-0.5);
    case PhaseShift:
        for1(s, w, [&](size_t i0, size_t) { a[i0] = Z; });
        return P(1);
    case IsingXX:
        for2(s, w, [&](size_t i00, size_t i01, size_t i10, size_t i11) {
            std::swap(a[i00], a[i11]);
            std::swap(a[i10], a[i01]);
        });
        return P(-0.5);
    case IsingXY:
        for2(s, w, [&](size_t i00, size_t i01, size_t i10, size_t i11) {
            std::swap(a[i10], a[i01]);
            a[i00] = Z;
            a[i11] = Z;
        });
        return P(0.5);
    case IsingYY:
        for2(s, w, [&](size_t i00, size_t i01, size_t i10, size_t i11) {
            const C<P> v00 = a[i00];
            a[i00] = -a[i11];
            a[i11] = -v00;
            std::swap(a[i10], a[i01]);
        });
        return P(-0.5);
    case IsingZZ:
        for2(s, w, [&](size_t, size_t i01, size_t i10, size_t) {
            a[i10] *= P{-1};
            a[i01] *= P{-1};
        });
        return P(-0.5);
    case CRX:
        for2(s, w, [&](size_t i00, size_t i01, size_t i10, size_t i11) {
            a[i00] = Z;
            a[i01] = Z;
            std::swap(a[i10], a[i11]);
        });
        return P(-0.5);
    case CRY:
        for2(s, w, [&](size_t i00, size_t i01, size_t i10, size_t i11) {
            a[i00] = Z;
            a[i01] = Z;
            const C<P> v0 = a[i10];
            a[i10] = {a[i11].imag(), -a[i11].real()};
            a[i11] = {-v0.imag(), v0.real()};
        });
        return P(-0.5);
    case CRZ:
        for2(s, w, [&](size_t i00, size_t i01, size_t, size_t i11) {
            a[i00] = Z;
            a[i01] = Z;
            a[i11] *= P{-1};
        });
        return P(-0.5);
    case ControlledPhaseShift:
        for2(s, w, [&](size_t i00, size_t i01, size_t i10, size_t) {
            a[i00] = Z;
            a[i01] = Z;
            a[i10] = Z;
        });
        return P(1);
    case SingleExcitation:
        for2(s, w, [&](size_t i00, size_t i01, size_t i10, size_t i11) {
            a[i00] = Z;
            a[i01] *= I;
            a[i10] *= -I;
            a[i11] = Z;
            std::swap(a[i10], a[i01]);
        });
        return P(-0.5);
    case SingleExcitationMinus:
        for2(s, w, [&](size_t, size_t i01, size_t i10, size_t) {
            a[i01] *= I;
            a[i10] *= -I;
            std::swap(a[i10], a[i01]);
        });
        return P(-0.5);
    case SingleExcitationPlus:
        for2(s, w, [&](size_t i00, size_t i01, size_t i10, size_t i11) {
            a[i00] *= P{-1};
            a[i01] *= I;
            a[i10] *= -I;
            a[i11] *= P{-1};
            std::swap(a[i10], a[i01]);
        });
        return P(-0.5);
    case DoubleExcitation:
        forK(s, w, [&](const std::vector<size_t> &i) {
            const C<P> v3 = a[i[3]], v12 = a[i[12]];
            for (size_t j : i) {
                a[j] = Z;
            }
            a[i[3]] = -v12 * I;
            a[i[12]] = v3 * I;
        });
        return P(-0.5);
    case DoubleExcitationMinus:
        forK(s, w, [&](const std::vector<size_t> &i) {
            a[i[3]] *= I;
            a[i[12]] *= -I;
            std::swap(a[i[3]], a[i[12]]);
        });
        return P(-0.5);
    case DoubleExcitationPlus:
        forK(s, w, [&](const std::vector<size_t> &i) {
            for (size_t j : i) {
                a[j] *= P{-1};
            }
            a[i[3]] *= -I;
            a[i[12]] *= I;
            std::swap(a[i[3]], a[i[12]]);
        });
        return P(-0.5);
    case MultiRZ: {
        size_t mask = 0;
        for (size_t wi : w) {
            mask |= size_t{1} << s.rev(wi);
        }
        for (size_t k = 0; k < s.len(); k++) {
            a[k] *= static_cast<P>(1 - 2 * int(__builtin_popcountll(k & mask) % 2));
        }
        return P(-0.5);
    }
    default:
        fail("Cannot find a registered kernel for a given generator and kernel pair.");
    }
    return P(0);
}

/* innerProdC = sum_i conj(v1_i) v2_i (LinearAlgebra.hpp:152-213): a
 * sequential left fold below 2^20 elements, an OpenMP reduction at or above. */
template <class P> C<P> inner(const C<P> *v1, const C<P> *v2, size_t n) {
    auto mc = [](C<P> x, C<P> y) {
        return C<P>{x.real() * y.real() + x.imag() * y.imag(),
                    -x.imag() * y.real() + x.real() * y.imag()};
    };
    if (n < (size_t{1} << 20)) {
        C<P> r{0, 0};
        for (size_t i = 0; i < n; i++) {
            r = r + mc(v1[i], v2[i]);
        }
        return r;
    }
    P re = 0, im = 0;
#pragma omp parallel for reduction(+ : re, im) schedule(static)
    for (size_t i = 0; i < n; i++) {
        const C<P> t = mc(v1[i], v2[i]);
        re += t.real();
        im += t.imag();
    }
    return {re, im};
}

/* Observable tree: NamedObs / HermitianObs / TensorProdObs / Hamiltonian
 * (core/src/observables/Observables.hpp:98-417, LQ ObservablesLQubit.hpp:156-355). */
template <class P> struct Obs {
    int kind = 0; // 0 named, 1 hermitian, 2 tensor, 3 hamiltonian
    int op = 0;
    std::vector<size_t> wires;
    std::vector<C<P>> mat;
    std::vector<P> coeffs;
    std::vector<Obs> kids;

    void apply(SV<P> &s) const {
        switch (kind) {
        case 0:
            apply_gate<P>(s, op, wires, false, {});
            return;
        case 1:
            apply_matrix<P>(s, mat.data(), wires, false);
            return;
        case 2:
            for (const auto &k : kids) {
                k.apply(s);
            }
            return;
        default: {
            /* Hamiltonian: res = sum_t c_t (O_t psi) by copy-apply-axpy,
             * then psi <- res (ObservablesLQubit.hpp:156-190). */
            std::vector<C<P>> res(s.len(), C<P>{0, 0});
            std::vector<C<P>> tmp(s.len());
            for (size_t t = 0; t < kids.size(); t++) {
                std::copy(s.a, s.a + s.len(), tmp.begin());
                SV<P> ts{tmp.data(), s.n};
                kids[t].apply(ts);
                const C<P> c{coeffs[t], 0};
                for (size_t i = 0; i < s.len(); i++) {
                    res[i] += c * tmp[i];
                }
            }
            std::copy(res.begin(), res.end(), s.a);
        }
        }
    }
};

template <class P> struct ObsReader {
    const int *prog;
    const double *coeffs;
    const P *mats;
    Obs<P> factor() {
        Obs<P> o;
        o.kind = *prog++;
        o.op = *prog++;
        const int nw = *prog++;
        o.wires.assign(prog, prog + nw);
        prog += nw;
        if (o.kind == 1) {
            const size_t dim = size_t{1} << (2 * nw);
            for (size_t i = 0; i < dim; i++) {
                o.mat.emplace_back(mats[2 * i], mats[2 * i + 1]);
            }
            mats += 2 * dim;
        } else if (o.kind != 0) {
            fail("bad observable factor kind");
        }
        return o;
    }
    Obs<P> term() {
        const int nf = *prog++;
        if (nf == 1) {
            return factor();
        }
        Obs<P> t;
        t.kind = 2;
        for (int f = 0; f < nf; f++) {
            t.kids.push_back(factor());
        }
        std::vector<size_t> all;
        for (auto &k : t.kids) {
            for (auto w : k.wires) {
                if (std::find(all.begin(), all.end(), w) != all.end()) {
                    fail("All wires in observables must be disjoint.");
                }
                all.push_back(w);
            }
        }
        return t;
    }
    Obs<P> one() {
        const int is_ham = *prog++;
        const int nt = *prog++;
        if (!is_ham) {
            if (nt != 1) {
                fail("non-Hamiltonian observable must have one term");
            }
            return term();
        }
        Obs<P> h;
        h.kind = 3;
        for (int t = 0; t < nt; t++) {
            h.coeffs.push_back(static_cast<P>(*coeffs++));
            h.kids.push_back(term());
        }
        return h;
    }
};

template <class F> int guard(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
    }
    return -1;
}

template <class P> C<P> *cp(P *x) { return reinterpret_cast<C<P> *>(x); }
template <class P> const C<P> *cp(const P *x) { return reinterpret_cast<const C<P> *>(x); }

template <class P> struct Tape {
    std::vector<int> ids;
    std::vector<std::vector<size_t>> wires;
    std::vector<bool> inv;
    std::vector<std::vector<P>> params;
    Tape(int nops, const int *id, const int *nw, const int *w, const uint8_t *iv,
         const int *np, const double *pr) {
        for (int i = 0; i < nops; i++) {
            ids.push_back(id[i]);
            wires.emplace_back(w, w + nw[i]);
            w += nw[i];
            inv.push_back(iv[i] != 0);
            std::vector<P> p;
            for (int j = 0; j < np[i]; j++) {
                p.push_back(static_cast<P>(pr[j]));
            }
            pr += np[i];
            params.push_back(p);
        }
    }
};

template <class P>
int apply_ops(P *sv, int n, int nops, const int *ids, const int *nw, const int *w,
              const uint8_t *iv, const int *np, const double *pr) {
    return guard([&] {
        SV<P> s{cp(sv), static_cast<size_t>(n)};
        Tape<P> t(nops, ids, nw, w, iv, np, pr);
        for (int i = 0; i < nops; i++) {
            apply_gate(s, t.ids[i], t.wires[i], t.inv[i], t.params[i]);
        }
    });
}

/* Measurements::expval/var(Observable): MeasurementsLQubit.hpp:248-295,571-605 */
template <class P>
int expval_obs(const P *sv, int n, const int *prog, const double *coeffs, const P *mats,
               double *out, bool var) {
    return guard([&] {
        ObsReader<P> rd{prog + 1, coeffs, mats};
        const Obs<P> ob = rd.one();
        const size_t len = size_t{1} << n;
        std::vector<C<P>> bra(cp(sv), cp(sv) + len);
        SV<P> b{bra.data(), static_cast<size_t>(n)};
        ob.apply(b);
        if (!var) {
            *out = std::real(inner(bra.data(), cp(sv), len));
        } else {
            const P ms = std::real(inner(bra.data(), bra.data(), len));
            const P sm = static_cast<P>(std::pow(std::real(inner(bra.data(), cp(sv), len)), 2));
            *out = ms - sm;
        }
    });
}

/* Measurements::expval/var(name, wires): MeasurementsLQubit.hpp:167-181,298-317 */
template <class P>
int named(const P *sv, int n, int id, int nw, const int *w, double *out, bool var) {
    return guard([&] {
        const size_t len = size_t{1} << n;
        std::vector<C<P>> o(cp(sv), cp(sv) + len);
        SV<P> s{o.data(), static_cast<size_t>(n)};
        apply_gate<P>(s, id, std::vector<size_t>(w, w + nw), false, {});
        if (!var) {
            *out = std::real(inner(cp(sv), o.data(), len));
        } else {
            const P ms = std::real(inner(o.data(), o.data(), len));
            const P sm = static_cast<P>(std::pow(std::real(inner(cp(sv), o.data(), len)), 2));
            *out = ms - sm;
        }
    });
}

/* Measurements::expval(matrix, wires): MeasurementsLQubit.hpp:145-158 */
template <class P>
int expval_matrix(const P *sv, int n, const P *mat, int nw, const int *w, double *out) {
    return guard([&] {
        const size_t len = size_t{1} << n;
        std::vector<C<P>> o(cp(sv), cp(sv) + len);
        SV<P> s{o.data(), static_cast<size_t>(n)};
        apply_matrix<P>(s, cp(mat), std::vector<size_t>(w, w + nw), false);
        *out = std::real(inner(cp(sv), o.data(), len));
    });
}

/* Measurements::probs() and probs(wires): MeasurementsLQubit.hpp:79-135.
 * Marginal over sorted wires (generateBitPatterns order, GateIndices.cpp:31-45),
 * then transposed to the caller's order (Util.hpp:488-520). */
template <class P> int probs(const P *sv, int n, int nw, const int *w, P *out) {
    return guard([&] {
        const size_t len = size_t{1} << n;
        const C<P> *a = cp(sv);
        auto nrm = [](C<P> z) { return z.real() * z.real() + z.imag() * z.imag(); };
        if (nw < 0) {
            for (size_t i = 0; i < len; i++) {
                out[i] = nrm(a[i]);
            }
            return;
        }
        std::vector<size_t> wires(w, w + nw);
        for (auto x : wires) {
            if (x >= static_cast<size_t>(n)) {
                fail("wire index out of range");
            }
        }
        std::vector<size_t> order(nw);
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](size_t i, size_t j) { return wires[i] < wires[j]; });
        std::vector<size_t> sorted(nw);
        for (int i = 0; i < nw; i++) {
            sorted[i] = wires[order[i]];
        }
        auto patterns = [&](const std::vector<size_t> &qs) {
            std::vector<size_t> idx{0};
            for (auto it = qs.rbegin(); it != qs.rend(); ++it) {
                const size_t v = size_t{1} << (n - 1 - *it);
                const size_t cur = idx.size();
                for (size_t j = 0; j < cur; j++) {
                    idx.push_back(idx[j] + v);
                }
            }
            return idx;
        };
        std::vector<size_t> rest;
        for (size_t q = 0; q < static_cast<size_t>(n); q++) {
            if (std::find(sorted.begin(), sorted.end(), q) == sorted.end()) {
                rest.push_back(q);
            }
        }
        const auto ind = patterns(sorted);
        const auto offs = patterns(rest);
        std::vector<P> pr(ind.size(), P{0});
        for (size_t i = 0; i < ind.size(); i++) {
            for (auto o : offs) {
                pr[i] += nrm(a[ind[i] + o]);
            }
        }
        if (wires != sorted) {
            std::vector<P> tr(pr.size());
            for (size_t i = 0; i < pr.size(); i++) {
                size_t src = 0, x = i;
                for (int ax = nw - 1; ax >= 0; ax--) {
                    src += (x % 2) << (nw - 1 - order[ax]);
                    x /= 2;
                }
                tr[i] = pr[src];
            }
            pr = tr;
        }
        std::copy(pr.begin(), pr.end(), out);
    });
}

/* CDF sampling (Kokkos rule): p_k = re^2 + im^2, exclusive prefix sum
 * C_0 = 0, C_k = sum_{j<k} p_j (MeasurementsKokkos.hpp:710-722), and for a
 * uniform U the index 0 if U <= C_1, else the i with C_i < U <= C_{i+1}, where
 * C_N := 1 (MeasuresFunctors.hpp:81-98).  Prefix sums are accumulated in
 * double for both precisions. */
template <class P>
int sample_cdf(const P *sv, int n, int64_t m, const double *u, int64_t *idx) {
    return guard([&] {
        const size_t len = size_t{1} << n;
        const C<P> *a = cp(sv);
        std::vector<double> cdf(len);
        double acc = 0;
        for (size_t k = 0; k < len; k++) {
            cdf[k] = acc;
            const double re = a[k].real(), im = a[k].imag();
            acc += re * re + im * im;
        }
        for (int64_t s = 0; s < m; s++) {
            const double U = u[s];
            size_t index;
            if (len == 1 || U <= cdf[1]) {
                index = 0;
            } else {
                size_t lo = 1, hi = len;
                while (hi - lo > 1) {
                    const size_t mid = hi - ((hi - lo) >> 1U);
                    const double c = (mid == len) ? 1.0 : cdf[mid];
                    if (c < U) {
                        lo = mid;
                    } else {
                        hi = mid;
                    }
                }
                index = hi - 1;
            }
            idx[s] = static_cast<int64_t>(index);
        }
    });
}

/* AdjointJacobian::adjointJacobian, AdjointJacobianLQubit.hpp:212-336.
 * The OpenMP loops over observables are kept (they do not change results). */
template <class P>
int adjoint(const P *sv, int n, int nops, const int *ids, const int *nw, const int *w,
            const uint8_t *iv, const int *np, const double *pr, const int *prog,
            const double *coeffs, const P *mats, int ntp, const int *tp, P *jac) {
    return guard([&] {
        Tape<P> ops(nops, ids, nw, w, iv, np, pr);
        ObsReader<P> rd{prog, coeffs, mats};
        const int nobs = *rd.prog++;
        std::vector<Obs<P>> obs;
        for (int i = 0; i < nobs; i++) {
            obs.push_back(rd.one());
        }
        const size_t K = obs.size();
        if (ntp == 0) {
            return;
        }
        const size_t len = size_t{1} << n;
        std::vector<P> J(K * ntp, P{0});
        size_t num_par_ops = 0;
        for (auto &p : ops.params) {
            num_par_ops += p.empty() ? 0 : 1;
        }
        std::vector<C<P>> lambda(cp(sv), cp(sv) + len), mu(len);
        std::vector<std::vector<C<P>>> H(K, std::vector<C<P>>(len));
        SV<P> L{lambda.data(), static_cast<size_t>(n)};
        std::string err;
#pragma omp parallel for if (K > 1)
        for (size_t k = 0; k < K; k++) {
            try {
                std::copy(lambda.begin(), lambda.end(), H[k].begin());
                SV<P> hs{H[k].data(), static_cast<size_t>(n)};
                obs[k].apply(hs);
            } catch (const std::exception &e) {
#pragma omp critical
                err = e.what();
            }
        }
        if (!err.empty()) {
            fail(err);
        }
        int t_num = ntp - 1;
        size_t cur_param = num_par_ops - 1;
        int tp_pos = ntp - 1;
        for (int op = nops - 1; op >= 0; op--) {
            if (ops.params[op].size() > 1) {
                fail("The operation is not supported using the adjoint differentiation method");
            }
            if (ops.ids[op] == -1) {
                continue;
            }
            if (tp_pos < 0) {
                break;
            }
            std::copy(lambda.begin(), lambda.end(), mu.begin());
            apply_gate(L, ops.ids[op], ops.wires[op], !ops.inv[op], ops.params[op]);
            if (!ops.params[op].empty()) {
                if (cur_param == static_cast<size_t>(tp[tp_pos])) {
                    SV<P> M{mu.data(), static_cast<size_t>(n)};
                    const P scale = apply_generator(M, ops.ids[op], ops.wires[op]) *
                                    (ops.inv[op] ? P{-1} : P{1});
                    const size_t row = static_cast<size_t>(t_num) * K;
#pragma omp parallel for if (K > 1)
                    for (size_t k = 0; k < K; k++) {
                        J[row + k] = -2 * scale * std::imag(inner(H[k].data(), mu.data(), len));
                    }
                    t_num--;
                    tp_pos--;
                }
                cur_param--;
            }
#pragma omp parallel for if (K > 1)
            for (size_t k = 0; k < K; k++) {
                SV<P> hs{H[k].data(), static_cast<size_t>(n)};
                apply_gate(hs, ops.ids[op], ops.wires[op], !ops.inv[op], ops.params[op]);
            }
        }
        /* param-major -> observable-major (LinearAlgebra.hpp:473 Transpose) */
        for (size_t k = 0; k < K; k++) {
            for (int t = 0; t < ntp; t++) {
                jac[k * ntp + t] = J[t * K + k];
            }
        }
    });
}

} // namespace

/* VectorJacobianProduct::operator() (VectorJacobianProduct.hpp:77-155):
 * lambda = psi, mu = dy; reverse sweep; w_t = i s <G mu|lambda>. */
template <class P>
int vjp(const P *sv, int n, int nops, const int *ids, const int *nw, const int *w,
        const uint8_t *iv, const int *np, const double *pr, const P *dy, int ntp, const int *tp,
        P *out) {
    return guard([&] {
        Tape<P> ops(nops, ids, nw, w, iv, np, pr);
        if (ntp == 0) {
            return;
        }
        const size_t len = size_t{1} << n;
        size_t num_par_ops = 0;
        for (auto &p : ops.params) {
            num_par_ops += p.empty() ? 0 : 1;
        }
        std::vector<C<P>> lambda(cp(sv), cp(sv) + len), mu(cp(dy), cp(dy) + len), mud(len);
        SV<P> L{lambda.data(), static_cast<size_t>(n)}, M{mu.data(), static_cast<size_t>(n)};
        size_t cur_param = num_par_ops - 1;
        int tp_pos = ntp - 1;
        for (int op = nops - 1; op >= 0; op--) {
            if (ops.params[op].size() > 1) {
                fail("The operation is not supported using the adjoint differentiation method");
            }
            if (ops.ids[op] == -1) {
                continue;
            }
            if (tp_pos < 0) {
                break;
            }
            if (!ops.params[op].empty()) {
                if (cur_param == static_cast<size_t>(tp[tp_pos])) {
                    std::copy(mu.begin(), mu.end(), mud.begin());
                    SV<P> D{mud.data(), static_cast<size_t>(n)};
                    const P scale = apply_generator(D, ops.ids[op], ops.wires[op]) *
                                    (ops.inv[op] ? P{-1} : P{1});
                    const C<P> z = C<P>{0, scale} * inner(mud.data(), lambda.data(), len);
                    out[2 * tp_pos] = z.real();
                    out[2 * tp_pos + 1] = z.imag();
                    tp_pos--;
                }
                cur_param--;
            }
            apply_gate(L, ops.ids[op], ops.wires[op], !ops.inv[op], ops.params[op]);
            apply_gate(M, ops.ids[op], ops.wires[op], !ops.inv[op], ops.params[op]);
        }
    });
}

/* Measurements::expval / var of a CSR matrix (MeasurementsLQubit.hpp:183-214,
 * 437-471): y = H psi row by row (SparseLinAlg.hpp apply_Sparse_Matrix), then
 * Re <psi|y> and <y|y> - Re<y|psi>^2. */
template <class P>
int sparse(const P *sv, int n, const int64_t *rm, int64_t rms, const int64_t *en, const P *va,
           int64_t nnz, double *out, bool var) {
    return guard([&] {
        const size_t len = size_t{1} << n;
        if (static_cast<size_t>(rms) != len + 1) {
            fail("Statevector and Hamiltonian have incompatible sizes.");
        }
        (void)nnz;
        const C<P> *psi = cp(sv), *v = cp(va);
        std::vector<C<P>> y(len);
        for (size_t i = 0; i < len; i++) {
            C<P> acc = 0;
            for (int64_t j = rm[i]; j < rm[i + 1]; j++) {
                acc += v[j] * psi[en[j]];
            }
            y[i] = acc;
        }
        if (!var) {
            *out = std::real(inner(psi, y.data(), len));
        } else {
            const P ms = std::real(inner(y.data(), y.data(), len));
            const P sm = std::real(inner(y.data(), psi, len));
            *out = ms - sm * sm;
        }
    });
}

extern "C" {
const char *orc_last_error(void) { return g_err.c_str(); }
int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

#define ORC_DEF(SUF, REAL)                                                              \
    int orc_apply_ops_##SUF(REAL *sv, int n, int nops, const int *ids,                  \
                            const int *nwires, const int *wires,                        \
                            const uint8_t *inverse, const int *nparams,                 \
                            const double *params) {                                     \
        return apply_ops<REAL>(sv, n, nops, ids, nwires, wires, inverse, nparams,       \
                               params);                                                 \
    }                                                                                   \
    int orc_apply_matrix_##SUF(REAL *sv, int n, const REAL *mat, int nw,                \
                               const int *wires, int inverse) {                         \
        return guard([&] {                                                              \
            SV<REAL> s{cp(sv), static_cast<size_t>(n)};                                 \
            apply_matrix<REAL>(s, cp(mat), std::vector<size_t>(wires, wires + nw),      \
                               inverse != 0);                                           \
        });                                                                             \
    }                                                                                   \
    int orc_apply_generator_##SUF(REAL *sv, int n, int id, int nw, const int *wires,    \
                                  int adj, double *scale) {                             \
        (void)adj;                                                                      \
        return guard([&] {                                                              \
            SV<REAL> s{cp(sv), static_cast<size_t>(n)};                                 \
            *scale = apply_generator<REAL>(s, id, std::vector<size_t>(wires, wires + nw)); \
        });                                                                             \
    }                                                                                   \
    int orc_expval_##SUF(const REAL *sv, int n, const int *prog, const double *coeffs,  \
                         const REAL *mats, double *out) {                               \
        return expval_obs<REAL>(sv, n, prog, coeffs, mats, out, false);                 \
    }                                                                                   \
    int orc_var_##SUF(const REAL *sv, int n, const int *prog, const double *coeffs,     \
                      const REAL *mats, double *out) {                                  \
        return expval_obs<REAL>(sv, n, prog, coeffs, mats, out, true);                  \
    }                                                                                   \
    int orc_expval_named_##SUF(const REAL *sv, int n, int id, int nw, const int *wires, \
                               double *out) {                                           \
        return named<REAL>(sv, n, id, nw, wires, out, false);                           \
    }                                                                                   \
    int orc_var_named_##SUF(const REAL *sv, int n, int id, int nw, const int *wires,    \
                            double *out) {                                              \
        return named<REAL>(sv, n, id, nw, wires, out, true);                            \
    }                                                                                   \
    int orc_expval_matrix_##SUF(const REAL *sv, int n, const REAL *mat, int nw,         \
                                const int *wires, double *out) {                        \
        return expval_matrix<REAL>(sv, n, mat, nw, wires, out);                         \
    }                                                                                   \
    int orc_probs_##SUF(const REAL *sv, int n, int nw, const int *wires, REAL *out) {   \
        return probs<REAL>(sv, n, nw, wires, out);                                      \
    }                                                                                   \
    int orc_adjoint_##SUF(const REAL *sv, int n, int nops, const int *ids,              \
                          const int *nwires, const int *wires, const uint8_t *inverse,  \
                          const int *nparams, const double *params, const int *prog,    \
                          const double *coeffs, const REAL *mats, int ntp,              \
                          const int *tp, REAL *jac) {                                   \
        return adjoint<REAL>(sv, n, nops, ids, nwires, wires, inverse, nparams, params, \
                             prog, coeffs, mats, ntp, tp, jac);                         \
    }                                                                                   \
    int orc_sample_cdf_##SUF(const REAL *sv, int n, int64_t m, const double *u,         \
                             int64_t *idx) {                                            \
        return sample_cdf<REAL>(sv, n, m, u, idx);                                      \
    }

ORC_DEF(c128, double)
ORC_DEF(c64, float)

int orc_vjp_c128(const double *sv, int n, int nops, const int *ids, const int *nw, const int *w,
                 const uint8_t *iv, const int *np, const double *pr, const double *dy, int ntp,
                 const int *tp, double *out) {
    return vjp<double>(sv, n, nops, ids, nw, w, iv, np, pr, dy, ntp, tp, out);
}
int orc_vjp_c64(const float *sv, int n, int nops, const int *ids, const int *nw, const int *w,
                const uint8_t *iv, const int *np, const double *pr, const float *dy, int ntp,
                const int *tp, float *out) {
    return vjp<float>(sv, n, nops, ids, nw, w, iv, np, pr, dy, ntp, tp, out);
}
int orc_expval_sparse_c128(const double *sv, int n, const int64_t *rm, int64_t rms,
                           const int64_t *en, const double *va, int64_t nnz, double *out) {
    return sparse<double>(sv, n, rm, rms, en, va, nnz, out, false);
}
int orc_expval_sparse_c64(const float *sv, int n, const int64_t *rm, int64_t rms,
                          const int64_t *en, const float *va, int64_t nnz, double *out) {
    return sparse<float>(sv, n, rm, rms, en, va, nnz, out, false);
}
int orc_var_sparse_c128(const double *sv, int n, const int64_t *rm, int64_t rms,
                        const int64_t *en, const double *va, int64_t nnz, double *out) {
    return sparse<double>(sv, n, rm, rms, en, va, nnz, out, true);
}
int orc_var_sparse_c64(const float *sv, int n, const int64_t *rm, int64_t rms,
                       const int64_t *en, const float *va, int64_t nnz, double *out) {
    return sparse<float>(sv, n, rm, rms, en, va, nnz, out, true);
}
}
