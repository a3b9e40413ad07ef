// Host-side lowering of named gates, generators and dense matrices to the one
// shape every gate kernel executes: a "group op".
//
// A group op touches, for every assignment of the n - nb bits that are not
// in `pos`, the M amplitudes  base | fixed | off[j]  (base = the free bits
// with zeros inserted at `pos`), and replaces them by  mat * v  (M x M, row
// major).  Gates that only change part of their 2^k subspace (controlled
// gates, phases, swaps) get a smaller M and a nonzero `fixed`, so each kernel
// reads and writes exactly the amplitudes the gate can change.  MultiRZ
// lowers to a parity phase instead.
//
// The semantics restated here are Lightning-Qubit's gate kernels
// (simulators/lightning_qubit/gates/cpu_kernels/GateImplementationsLM.hpp:389-1640,
// GateImplementationsLM.cpp:27-205, GateImplementationsPI.cpp:17-205 and
// PauliGenerator.hpp:31-56); each case cites its source lines.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace plamd {

struct InvalidArg : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};
struct OutOfRange : std::out_of_range {
    using std::out_of_range::out_of_range;
};

enum Gate : int {
    Identity = 0, PauliX, PauliY, PauliZ, Hadamard, S, T, PhaseShift, RX, RY,
    RZ, Rot, CNOT, CY, CZ, SWAP, IsingXX, IsingXY, IsingYY, IsingZZ,
    ControlledPhaseShift, CRX, CRY, CRZ, CRot, SingleExcitation,
    SingleExcitationMinus, SingleExcitationPlus, Toffoli, CSWAP,
    DoubleExcitation, DoubleExcitationMinus, DoubleExcitationPlus, MultiRZ,
    GateCount
};

// core/src/gates/Constant.hpp:46-96 (names), :163-197 (wires), :234-269 (params)
inline const char *gate_name(int id) {
    static const char *names[GateCount] = {
        "Identity", "PauliX", "PauliY", "PauliZ", "Hadamard", "S", "T", "PhaseShift",
        "RX", "RY", "RZ", "Rot", "CNOT", "CY", "CZ", "SWAP", "IsingXX", "IsingXY",
        "IsingYY", "IsingZZ", "ControlledPhaseShift", "CRX", "CRY", "CRZ", "CRot",
        "SingleExcitation", "SingleExcitationMinus", "SingleExcitationPlus",
        "Toffoli", "CSWAP", "DoubleExcitation", "DoubleExcitationMinus",
        "DoubleExcitationPlus", "MultiRZ"};
    if (id < 0 || id >= GateCount) {
        throw OutOfRange("unknown gate id " + std::to_string(id));
    }
    return names[id];
}
inline int gate_id(const std::string &name) {
    for (int i = 0; i < GateCount; i++) {
        if (name == gate_name(i)) {
            return i;
        }
    }
    if (name == "QubitStateVector" || name == "BasisState") {
        return -1;
    }
    throw OutOfRange("unordered_map::at: unknown operation " + name);
}
inline int gate_wires(int id) {
    static const int w[GateCount] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2,
                                     2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 3, 3, 4, 4, 4, -1};
    return w[id];
}
inline int gate_nparams(int id) {
    static const int p[GateCount] = {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 3, 0, 0, 0, 0, 1,
                                     1, 1, 1, 1, 1, 1, 1, 3, 1, 1, 1, 0, 0, 1, 1, 1, 1};
    return p[id];
}
inline bool has_generator(int id) {
    switch (id) {
    case PhaseShift: case RX: case RY: case RZ: case IsingXX: case IsingXY:
    case IsingYY: case IsingZZ: case CRX: case CRY: case CRZ:
    case ControlledPhaseShift: case SingleExcitation: case SingleExcitationMinus:
    case SingleExcitationPlus: case DoubleExcitation: case DoubleExcitationMinus:
    case DoubleExcitationPlus: case MultiRZ:
        return true;
    default:
        return false;
    }
}

constexpr int kMaxPos = 12;

template <class P> struct GroupOp {
    enum Kind { Noop, Group, Parity } kind = Group;
    int M = 1;               // amplitudes per group
    int nb = 0;              // number of pinned bit positions
    int pos[kMaxPos] = {};   // ascending
    uint64_t fixed = 0;      // bits forced to 1 in every member
    std::vector<uint64_t> off;
    std::vector<std::complex<P>> mat; // M*M row-major
    bool diag = false;
    uint64_t parity_mask = 0;         // Parity kind
    std::complex<P> parity_phase[2];  // phase for even / odd parity

    bool same_geometry(const GroupOp &o) const {
        if (kind != o.kind || M != o.M || nb != o.nb || fixed != o.fixed) {
            return false;
        }
        if (kind == Parity) {
            return parity_mask == o.parity_mask;
        }
        for (int i = 0; i < nb; i++) {
            if (pos[i] != o.pos[i]) {
                return false;
            }
        }
        return off == o.off;
    }
};

template <class P> using cx = std::complex<P>;

// Geometry builder: `bits` are the bit positions of all wires the group pins
// (targets and controls), `fixed` the ones forced to 1, `off` the member
// offsets.
template <class P>
GroupOp<P> make_group(std::vector<int> bits, uint64_t fixed, std::vector<uint64_t> off,
                      std::vector<cx<P>> mat, bool diag = false) {
    GroupOp<P> g;
    g.kind = GroupOp<P>::Group;
    std::sort(bits.begin(), bits.end());
    if (std::adjacent_find(bits.begin(), bits.end()) != bits.end()) {
        throw InvalidArg("wires must be distinct");
    }
    if (bits.size() > static_cast<size_t>(kMaxPos)) {
        throw InvalidArg("too many wires for one operation on the GPU path");
    }
    g.nb = static_cast<int>(bits.size());
    for (int i = 0; i < g.nb; i++) {
        g.pos[i] = bits[i];
    }
    g.fixed = fixed;
    g.M = static_cast<int>(off.size());
    g.off = std::move(off);
    g.mat = std::move(mat);
    g.diag = diag;
    return g;
}

struct Wires {
    std::vector<int> bit; // bit position of each wire, in caller order
    Wires(int n, const int *w, int nw, int expect) {
        if (expect >= 0 && nw != expect) {
            throw InvalidArg("Assertion failed: wires.size() == " + std::to_string(expect));
        }
        for (int i = 0; i < nw; i++) {
            if (w[i] < 0 || w[i] >= n) {
                throw InvalidArg("wire index out of range");
            }
            bit.push_back(n - 1 - w[i]);
        }
    }
    uint64_t m(int i) const { return uint64_t{1} << bit[i]; }
};

// Rot(phi, theta, omega) = RZ(omega) RY(theta) RZ(phi): core/src/gates/Gates.hpp:354-363
template <class P> std::vector<cx<P>> rot_matrix(P phi, P theta, P omega) {
    const P c = std::cos(theta / 2), s = std::sin(theta / 2);
    const P p = phi + omega, m = phi - omega;
    return {cx<P>{std::cos(p / 2) * c, -std::sin(p / 2) * c},
            cx<P>{-std::cos(m / 2) * s, -std::sin(m / 2) * s},
            cx<P>{std::cos(m / 2) * s, -std::sin(m / 2) * s},
            cx<P>{std::cos(p / 2) * c, std::sin(p / 2) * c}};
}

// All 2^k member offsets of a k-wire subspace, wires[0] most significant
// (GateIndices generateBitPatterns, GateIndices.cpp:31-45).
inline std::vector<uint64_t> subspace_offsets(const Wires &w) {
    const size_t k = w.bit.size();
    std::vector<uint64_t> off(size_t{1} << k, 0);
    for (size_t j = 0; j < off.size(); j++) {
        for (size_t b = 0; b < k; b++) {
            if ((j >> (k - 1 - b)) & 1U) {
                off[j] |= w.m(static_cast<int>(b));
            }
        }
    }
    return off;
}

// Dense matrix on arbitrary wires; `inverse` takes the conjugate transpose
// (LM applySingleQubitOp/applyTwoQubitOp, LM.hpp:204-325; PI applyMultiQubitOp,
// PI.hpp:250-290).
template <class P>
GroupOp<P> lower_matrix(int n, const cx<P> *m, const int *wires, int nw, bool inverse) {
    if (nw <= 0) {
        throw InvalidArg("Number of wires must be larger than 0");
    }
    if (nw > n) {
        throw InvalidArg("Assertion failed: num_qubits >= wires.size()");
    }
    Wires w(n, wires, nw, -1);
    const size_t dim = size_t{1} << nw;
    std::vector<cx<P>> mat(dim * dim);
    for (size_t i = 0; i < dim; i++) {
        for (size_t j = 0; j < dim; j++) {
            mat[i * dim + j] = inverse ? std::conj(m[j * dim + i]) : m[i * dim + j];
        }
    }
    return make_group<P>(w.bit, 0, subspace_offsets(w), std::move(mat));
}

template <class P> cx<P> expi(P x) { return std::exp(cx<P>(0, x)); }

// Named gate -> group op (GateImplementationsLM.hpp:389-1352, LM.cpp:27-118,
// PI.cpp:17-126, PI.hpp:786-815).  `params` must hold gate_nparams(id) values.
template <class P>
GroupOp<P> lower_gate(int n, int id, const int *wires, int nw, bool inv, const P *prm, int np) {
    if (id < 0 || id >= GateCount) {
        throw OutOfRange("unknown gate id " + std::to_string(id));
    }
    if (np != gate_nparams(id)) {
        throw InvalidArg("Assertion failed: params.size() == lookup(gate_num_params, gate_op)");
    }
    const Wires w(n, wires, nw, gate_wires(id));
    const P th = np ? prm[0] : P(0);
    const P sgn = inv ? P(-1) : P(1);
    const P isq2 = static_cast<P>(M_SQRT1_2);
    const cx<P> O{0, 0}, I1{1, 0}, Im{0, 1};
    using V = std::vector<cx<P>>;
    // most gates are built from one 1-bit target and optional controls
    auto one = [&](int t, uint64_t fixed, std::vector<int> extra, V mat, bool diag = false) {
        std::vector<int> bits = extra;
        bits.push_back(w.bit[t]);
        return make_group<P>(bits, fixed, {0, w.m(t)}, std::move(mat), diag);
    };
    auto phase_on = [&](std::vector<int> idx, cx<P> ph) {
        std::vector<int> bits;
        uint64_t fixed = 0;
        for (int i : idx) {
            bits.push_back(w.bit[i]);
            fixed |= w.m(i);
        }
        return make_group<P>(bits, fixed, {0}, V{ph});
    };
    switch (id) {
    case Identity: { // LM.hpp:389-399
        GroupOp<P> g;
        g.kind = GroupOp<P>::Noop;
        return g;
    }
    case PauliX: // LM.hpp:401-418
        return one(0, 0, {}, V{O, I1, I1, O});
    case PauliY: // LM.hpp:420-439
        return one(0, 0, {}, V{O, -Im, Im, O});
    case PauliZ: // LM.hpp:441-455
        return phase_on({0}, cx<P>{-1, 0});
    case Hadamard: // LM.hpp:458-476
        return one(0, 0, {}, V{{isq2, 0}, {isq2, 0}, {isq2, 0}, {-isq2, 0}});
    case S: // LM.hpp:479-496
        return phase_on({0}, cx<P>{0, sgn});
    case T: // LM.hpp:499-517
        return phase_on({0}, cx<P>{isq2, sgn * isq2});
    case PhaseShift: // LM.hpp:520-538
        return phase_on({0}, inv ? std::exp(-cx<P>(0, th)) : std::exp(cx<P>(0, th)));
    case RX: { // LM.hpp:541-563: m01 = m10 = i * js, js = sin(-theta/2) (negated if inverse)
        const P c = std::cos(th / 2);
        const P js = inv ? -std::sin(-th / 2) : std::sin(-th / 2);
        return one(0, 0, {}, V{{c, 0}, {0, js}, {0, js}, {c, 0}});
    }
    case RY: { // LM.hpp:566-588
        const P c = std::cos(th / 2), s = sgn * std::sin(th / 2);
        return one(0, 0, {}, V{{c, 0}, {-s, 0}, {s, 0}, {c, 0}});
    }
    case RZ: { // LM.hpp:591-615
        cx<P> f0{std::cos(th / 2), -std::sin(th / 2)}, f1{std::cos(th / 2), std::sin(th / 2)};
        if (inv) {
            f0 = std::conj(f0);
            f1 = std::conj(f1);
        }
        return one(0, 0, {}, V{f0, O, O, f1}, true);
    }
    case Rot: { // LM.hpp:618-628
        auto m = inv ? rot_matrix<P>(-prm[2], -prm[1], -prm[0]) : rot_matrix<P>(prm[0], prm[1], prm[2]);
        return one(0, 0, {}, m);
    }
    case CNOT: // LM.hpp:633-656: swap |10> <-> |11>, wires[0] the control
        return one(1, w.m(0), {w.bit[0]}, V{O, I1, I1, O});
    case CY: // LM.hpp:659-685
        return one(1, w.m(0), {w.bit[0]}, V{O, -Im, Im, O});
    case CZ: // LM.hpp:688-708
        return phase_on({0, 1}, cx<P>{-1, 0});
    case SWAP: // LM.hpp:743-764: swap |01> <-> |10>
        return make_group<P>({w.bit[0], w.bit[1]}, 0, {w.m(1), w.m(0)}, V{O, I1, I1, O});
    case IsingXX: { // LM.hpp:767-809: cr*v - i*sj*(bit-flipped v)
        const P cr = std::cos(th / 2), sj = sgn * std::sin(th / 2);
        const cx<P> c{cr, 0}, x{0, -sj};
        return make_group<P>({w.bit[0], w.bit[1]}, 0, {0, w.m(1), w.m(0), w.m(0) | w.m(1)},
                             V{c, O, O, x, O, c, x, O, O, x, c, O, x, O, O, c});
    }
    case IsingXY: { // LM.hpp:813-852: only |01>, |10> change
        const P cr = std::cos(th / 2), sj = sgn * std::sin(th / 2);
        const cx<P> c{cr, 0}, x{0, sj};
        return make_group<P>({w.bit[0], w.bit[1]}, 0, {w.m(1), w.m(0)}, V{c, x, x, c});
    }
    case IsingYY: { // LM.hpp:855-897
        const P cr = std::cos(th / 2), sj = sgn * std::sin(th / 2);
        const cx<P> c{cr, 0}, p{0, sj}, q{0, -sj};
        return make_group<P>({w.bit[0], w.bit[1]}, 0, {0, w.m(1), w.m(0), w.m(0) | w.m(1)},
                             V{c, O, O, p, O, c, q, O, O, q, c, O, p, O, O, c});
    }
    case IsingZZ: { // LM.hpp:901-935
        cx<P> f0{std::cos(th / 2), -std::sin(th / 2)}, f1{std::cos(th / 2), std::sin(th / 2)};
        if (inv) {
            f0 = std::conj(f0);
            f1 = std::conj(f1);
        }
        return make_group<P>({w.bit[0], w.bit[1]}, 0, {0, w.m(1), w.m(0), w.m(0) | w.m(1)},
                             V{f0, O, O, O, O, f1, O, O, O, O, f1, O, O, O, O, f0}, true);
    }
    case ControlledPhaseShift: // LM.hpp:938-966
        return phase_on({0, 1}, inv ? std::exp(-cx<P>(0, th)) : std::exp(cx<P>(0, th)));
    case CRX: { // LM.hpp:969-1003
        const P c = std::cos(th / 2), js = sgn * std::sin(th / 2);
        return one(1, w.m(0), {w.bit[0]}, V{{c, 0}, {0, -js}, {0, -js}, {c, 0}});
    }
    case CRY: { // LM.hpp:1006-1038
        const P c = std::cos(th / 2), s = sgn * std::sin(th / 2);
        return one(1, w.m(0), {w.bit[0]}, V{{c, 0}, {-s, 0}, {s, 0}, {c, 0}});
    }
    case CRZ: { // LM.hpp:1041-1069
        cx<P> f0{std::cos(th / 2), -std::sin(th / 2)}, f1{std::cos(th / 2), std::sin(th / 2)};
        if (inv) {
            f0 = std::conj(f0);
            f1 = std::conj(f1);
        }
        return one(1, w.m(0), {w.bit[0]}, V{f0, O, O, f1}, true);
    }
    case CRot: { // LM.hpp:711-740
        auto m = inv ? rot_matrix<P>(-prm[2], -prm[1], -prm[0]) : rot_matrix<P>(prm[0], prm[1], prm[2]);
        return one(1, w.m(0), {w.bit[0]}, m);
    }
    case SingleExcitation: { // LM.cpp:27-52: rotation of |01>, |10>
        const P c = std::cos(th / 2), s = sgn * std::sin(th / 2);
        return make_group<P>({w.bit[0], w.bit[1]}, 0, {w.m(1), w.m(0)},
                             V{{c, 0}, {-s, 0}, {s, 0}, {c, 0}});
    }
    case SingleExcitationMinus:
    case SingleExcitationPlus: { // LM.cpp:55-118
        const P c = std::cos(th / 2), s = sgn * std::sin(th / 2);
        cx<P> e;
        if (id == SingleExcitationMinus) {
            e = inv ? std::exp(cx<P>(0, th / 2)) : std::exp(-cx<P>(0, th / 2));
        } else {
            e = inv ? std::exp(-cx<P>(0, th / 2)) : std::exp(cx<P>(0, th / 2));
        }
        return make_group<P>({w.bit[0], w.bit[1]}, 0, {0, w.m(1), w.m(0), w.m(0) | w.m(1)},
                             V{e, O, O, O, O, {c, 0}, {-s, 0}, O, O, {s, 0}, {c, 0}, O, O, O, O, e});
    }
    case Toffoli: // PI.hpp:786-800: swap |110> <-> |111>
        return one(2, w.m(0) | w.m(1), {w.bit[0], w.bit[1]}, V{O, I1, I1, O});
    case CSWAP: // PI.hpp:802-815: swap |101> <-> |110>
        return make_group<P>({w.bit[0], w.bit[1], w.bit[2]}, w.m(0), {w.m(2), w.m(1)},
                             V{O, I1, I1, O});
    case DoubleExcitation: { // PI.cpp:17-40: rotation of |0011>, |1100>
        const P c = std::cos(th / 2), s = sgn * std::sin(th / 2);
        return make_group<P>(w.bit, 0, {w.m(2) | w.m(3), w.m(0) | w.m(1)},
                             V{{c, 0}, {-s, 0}, {s, 0}, {c, 0}});
    }
    case DoubleExcitationMinus:
    case DoubleExcitationPlus: { // PI.cpp:42-126
        const P c = std::cos(th / 2), s = sgn * std::sin(th / 2);
        cx<P> e;
        if (id == DoubleExcitationMinus) {
            e = inv ? std::exp(cx<P>(0, th / 2)) : std::exp(-cx<P>(0, th / 2));
        } else {
            e = inv ? std::exp(-cx<P>(0, th / 2)) : std::exp(cx<P>(0, th / 2));
        }
        V m(256, O);
        for (int j = 0; j < 16; j++) {
            m[j * 16 + j] = e;
        }
        m[3 * 16 + 3] = {c, 0};
        m[3 * 16 + 12] = {-s, 0};
        m[12 * 16 + 3] = {s, 0};
        m[12 * 16 + 12] = {c, 0};
        return make_group<P>(w.bit, 0, subspace_offsets(w), std::move(m));
    }
    case MultiRZ: { // LM.hpp:1332-1352: phase by parity of the masked bits
        GroupOp<P> g;
        g.kind = GroupOp<P>::Parity;
        cx<P> f0{std::cos(th / 2), -std::sin(th / 2)}, f1{std::cos(th / 2), std::sin(th / 2)};
        g.parity_phase[0] = inv ? std::conj(f0) : f0;
        g.parity_phase[1] = inv ? std::conj(f1) : f1;
        for (size_t i = 0; i < w.bit.size(); i++) {
            if (g.parity_mask & w.m(static_cast<int>(i))) {
                throw InvalidArg("wires must be distinct");
            }
            g.parity_mask |= w.m(static_cast<int>(i));
        }
        return g;
    }
    default:
        throw OutOfRange("unknown gate id");
    }
}

// Generator G of a parametrised gate, U(t) = exp(i s t G); returns s.
// `restricted` = false gives the in-place generator of applyGenerator
// (the whole group subspace, zeros included); `restricted` = true gives G on
// the geometry lower_gate uses for the same gate, valid wherever G vanishes
// outside that geometry (used by the fused adjoint step).
// Sources: PauliGenerator.hpp:31-56, LM.hpp:1355-1640, LM.cpp:120-205,
// PI.cpp:128-205.
template <class P>
GroupOp<P> lower_generator(int n, int id, const int *wires, int nw, bool restricted, P *scale) {
    if (id < 0 || id >= GateCount || !has_generator(id)) {
        throw InvalidArg("Cannot find a registered kernel for a given generator and kernel pair.");
    }
    const Wires w(n, wires, nw, gate_wires(id));
    const cx<P> O{0, 0}, I1{1, 0}, N1{-1, 0}, Im{0, 1}, mIm{0, -1};
    using V = std::vector<cx<P>>;
    auto quad = [&](V m, bool diag = false) {
        return make_group<P>({w.bit[0], w.bit[1]}, 0, {0, w.m(1), w.m(0), w.m(0) | w.m(1)},
                             std::move(m), diag);
    };
    auto ctrl = [&](V m, bool diag = false) {
        return make_group<P>({w.bit[0], w.bit[1]}, w.m(0), {0, w.m(1)}, std::move(m), diag);
    };
    *scale = P(-0.5);
    switch (id) {
    case RX:
        return make_group<P>({w.bit[0]}, 0, {0, w.m(0)}, V{O, I1, I1, O});
    case RY:
        return make_group<P>({w.bit[0]}, 0, {0, w.m(0)}, V{O, mIm, Im, O});
    case RZ:
        return make_group<P>({w.bit[0]}, 0, {0, w.m(0)}, V{I1, O, O, N1}, true);
    case PhaseShift: // |1><1|
        *scale = P(1);
        if (restricted) {
            return make_group<P>({w.bit[0]}, w.m(0), {0}, V{I1});
        }
        return make_group<P>({w.bit[0]}, 0, {0, w.m(0)}, V{O, O, O, I1}, true);
    case IsingXX:
        return quad(V{O, O, O, I1, O, O, I1, O, O, I1, O, O, I1, O, O, O});
    case IsingXY: // (XX + YY)/2 acting on |01>, |10>; scale +1/2
        *scale = P(0.5);
        if (restricted) {
            return make_group<P>({w.bit[0], w.bit[1]}, 0, {w.m(1), w.m(0)}, V{O, I1, I1, O});
        }
        return quad(V{O, O, O, O, O, O, I1, O, O, I1, O, O, O, O, O, O});
    case IsingYY:
        return quad(V{O, O, O, N1, O, O, I1, O, O, I1, O, O, N1, O, O, O});
    case IsingZZ:
        return quad(V{I1, O, O, O, O, N1, O, O, O, O, N1, O, O, O, O, I1}, true);
    case CRX:
        if (restricted) {
            return ctrl(V{O, I1, I1, O});
        }
        return quad(V{O, O, O, O, O, O, O, O, O, O, O, I1, O, O, I1, O});
    case CRY:
        if (restricted) {
            return ctrl(V{O, mIm, Im, O});
        }
        return quad(V{O, O, O, O, O, O, O, O, O, O, O, mIm, O, O, Im, O});
    case CRZ:
        if (restricted) {
            return ctrl(V{I1, O, O, N1}, true);
        }
        return quad(V{O, O, O, O, O, O, O, O, O, O, I1, O, O, O, O, N1}, true);
    case ControlledPhaseShift:
        *scale = P(1);
        if (restricted) {
            return make_group<P>({w.bit[0], w.bit[1]}, w.m(0) | w.m(1), {0}, V{I1});
        }
        return quad(V{O, O, O, O, O, O, O, O, O, O, O, O, O, O, O, I1}, true);
    case SingleExcitation:
        if (restricted) {
            return make_group<P>({w.bit[0], w.bit[1]}, 0, {w.m(1), w.m(0)}, V{O, mIm, Im, O});
        }
        return quad(V{O, O, O, O, O, O, mIm, O, O, Im, O, O, O, O, O, O});
    case SingleExcitationMinus:
        return quad(V{I1, O, O, O, O, O, mIm, O, O, Im, O, O, O, O, O, I1});
    case SingleExcitationPlus:
        return quad(V{N1, O, O, O, O, O, mIm, O, O, Im, O, O, O, O, O, N1});
    case DoubleExcitation:
    case DoubleExcitationMinus:
    case DoubleExcitationPlus: {
        if (id == DoubleExcitation && restricted) {
            return make_group<P>(w.bit, 0, {w.m(2) | w.m(3), w.m(0) | w.m(1)}, V{O, mIm, Im, O});
        }
        const cx<P> rest = id == DoubleExcitation ? O : (id == DoubleExcitationMinus ? I1 : N1);
        V m(256, O);
        for (int j = 0; j < 16; j++) {
            m[j * 16 + j] = rest;
        }
        m[3 * 16 + 3] = O;
        m[12 * 16 + 12] = O;
        m[3 * 16 + 12] = mIm;
        m[12 * 16 + 3] = Im;
        return make_group<P>(w.bit, 0, subspace_offsets(w), std::move(m));
    }
    case MultiRZ: {
        GroupOp<P> g;
        g.kind = GroupOp<P>::Parity;
        g.parity_phase[0] = I1;
        g.parity_phase[1] = N1;
        for (size_t i = 0; i < w.bit.size(); i++) {
            if (g.parity_mask & w.m(static_cast<int>(i))) {
                throw InvalidArg("wires must be distinct");
            }
            g.parity_mask |= w.m(static_cast<int>(i));
        }
        return g;
    }
    default:
        break;
    }
    throw InvalidArg("Cannot find a registered kernel for a given generator and kernel pair.");
}

} // namespace plamd
