"""Pins the oracle with what the mathematics fixes, independently of the
Lightning-Qubit kernels it restates: every gate against its textbook matrix
embedded with numpy, every generator through U(t) = expm(i s t G), measurements
against dense linear algebra, CDF sampling against brute-force counting, the
adjoint Jacobian against exact parameter-shift, and the known-answer values
the reference's own tests hold.  CPU only."""
import itertools
import math

import numpy as np
import pytest
import scipy.linalg

from conftest import random_state
from lightning_unification_amd.circuits import Obs, Tape, pauli_word
from lightning_unification_amd.gates import GATE_NAMES, GATE_NPARAMS, GATE_WIRES, GENERATOR_GATES

I2 = np.eye(2)
X = np.array([[0, 1], [1, 0]], complex)
Y = np.array([[0, -1j], [1j, 0]])
Z = np.diag([1.0, -1.0]).astype(complex)
H = np.array([[1, 1], [1, -1]]) / math.sqrt(2)


def ctrl(u):
    k = u.shape[0]
    m = np.eye(2 * k, dtype=complex)
    m[k:, k:] = u
    return m


def rx(t):
    return scipy.linalg.expm(-0.5j * t * X)


def ry(t):
    return scipy.linalg.expm(-0.5j * t * Y)


def rz(t):
    return scipy.linalg.expm(-0.5j * t * Z)


def textbook(name, p):
    """Gate matrices from their definitions (wires[0] = most significant)."""
    c = lambda t: math.cos(t / 2)  # noqa: E731
    s = lambda t: math.sin(t / 2)  # noqa: E731
    if name == "Identity":
        return I2
    if name in ("PauliX", "PauliY", "PauliZ", "Hadamard"):
        return {"PauliX": X, "PauliY": Y, "PauliZ": Z, "Hadamard": H}[name]
    if name == "S":
        return np.diag([1, 1j])
    if name == "T":
        return np.diag([1, np.exp(1j * math.pi / 4)])
    if name == "PhaseShift":
        return np.diag([1, np.exp(1j * p[0])])
    if name in ("RX", "RY", "RZ"):
        return {"RX": rx, "RY": ry, "RZ": rz}[name](p[0])
    if name == "Rot":
        return rz(p[2]) @ ry(p[1]) @ rz(p[0])
    if name == "CNOT":
        return ctrl(X)
    if name == "CY":
        return ctrl(Y)
    if name == "CZ":
        return ctrl(Z)
    if name == "SWAP":
        return np.eye(4)[[0, 2, 1, 3]]
    if name == "IsingXX":
        return scipy.linalg.expm(-0.5j * p[0] * np.kron(X, X))
    if name == "IsingYY":
        return scipy.linalg.expm(-0.5j * p[0] * np.kron(Y, Y))
    if name == "IsingZZ":
        return scipy.linalg.expm(-0.5j * p[0] * np.kron(Z, Z))
    if name == "IsingXY":
        return scipy.linalg.expm(0.25j * p[0] * (np.kron(X, X) + np.kron(Y, Y)))
    if name == "ControlledPhaseShift":
        return np.diag([1, 1, 1, np.exp(1j * p[0])])
    if name in ("CRX", "CRY", "CRZ"):
        return ctrl({"CRX": rx, "CRY": ry, "CRZ": rz}[name](p[0]))
    if name == "CRot":
        return ctrl(textbook("Rot", p))
    if name.startswith("SingleExcitation"):
        ph = {"SingleExcitation": 1, "SingleExcitationMinus": np.exp(-0.5j * p[0]),
              "SingleExcitationPlus": np.exp(0.5j * p[0])}[name]
        m = np.diag([ph, 0, 0, ph]).astype(complex)
        m[1:3, 1:3] = [[c(p[0]), -s(p[0])], [s(p[0]), c(p[0])]]
        return m
    if name == "Toffoli":
        return ctrl(ctrl(X))
    if name == "CSWAP":
        return ctrl(textbook("SWAP", []))
    if name.startswith("DoubleExcitation"):
        ph = {"DoubleExcitation": 1, "DoubleExcitationMinus": np.exp(-0.5j * p[0]),
              "DoubleExcitationPlus": np.exp(0.5j * p[0])}[name]
        m = np.eye(16, dtype=complex) * ph
        m[3, 3], m[3, 12], m[12, 3], m[12, 12] = c(p[0]), -s(p[0]), s(p[0]), c(p[0])
        return m
    raise KeyError(name)


def multirz(t, k):
    par = np.array([bin(i).count("1") & 1 for i in range(1 << k)])
    return np.diag(np.exp(-0.5j * t * (1 - 2 * par)))


def embed_apply(state, u, wires, n):
    """Apply u (wires[0] = most significant local bit) to a state whose wire w
    is bit n-1-w of the index: tensor contraction on the (2,)*n view."""
    k = len(wires)
    t = state.reshape((2,) * n)
    t = np.moveaxis(t, wires, list(range(k)))
    t = (u @ t.reshape(1 << k, -1)).reshape((2,) * n)
    t = np.moveaxis(t, list(range(k)), wires)
    return t.reshape(-1)


def test_gates_match_textbook(oracle):
    n = 4
    rs = np.random.default_rng(11)
    for gi, g in enumerate(GATE_NAMES):
        arity = GATE_WIRES[gi]
        for k in ([1, 2, 3, 4] if arity == -1 else [arity]):
            for wires in itertools.permutations(range(n), k):
                p = list(rs.uniform(-7, 7, GATE_NPARAMS[gi]))
                u = multirz(p[0], k) if g == "MultiRZ" else textbook(g, p)
                s = random_state(n, int(rs.integers(1 << 30)))
                for inv in (False, True):
                    got = oracle.apply_tape(s.copy(), Tape().add(g, wires, p, inv))
                    exp = embed_apply(s, u.conj().T if inv else u, list(wires), n)
                    assert np.max(np.abs(got - exp)) < 1e-13, (g, wires, inv)


def test_generators_exponentiate(oracle):
    """U(t) = exp(i * s * t * G) with G read off the generator kernel and s its
    returned scale (Test_GateImplementations_Generator.cpp:98-151 checks the
    same identity by finite differences)."""
    for g in GENERATOR_GATES:
        gi = GATE_NAMES.index(g)
        k = 3 if GATE_WIRES[gi] == -1 else GATE_WIRES[gi]
        wires = list(range(k))
        G = np.zeros((1 << k, 1 << k), complex)
        scale = None
        for j in range(1 << k):
            e = np.zeros(1 << k, complex)
            e[j] = 1
            scale = oracle.apply_generator(e, g, wires)
            G[:, j] = e
        for t in (0.37, -1.9, 2.6):
            u = np.zeros_like(G)
            for j in range(1 << k):
                e = np.zeros(1 << k, complex)
                e[j] = 1
                u[:, j] = oracle.apply_tape(e, Tape().add(g, wires, [t]))
            assert np.max(np.abs(u - scipy.linalg.expm(1j * scale * t * G))) < 1e-12, g


def dense_pauli(word, wires, n):
    mats = {"X": X, "Y": Y, "Z": Z}
    full = np.array([[1.0 + 0j]])
    for w in range(n):
        m = mats[word[wires.index(w)]] if w in wires else I2
        full = np.kron(full, m)
    return full


def test_expval_var_probs_dense(oracle):
    n = 5
    s = random_state(n, 4)
    for word, wires in (("X", [0]), ("Y", [3]), ("Z", [4]), ("XYZ", [1, 4, 0]), ("YY", [2, 3])):
        O = dense_pauli(word, wires, n)
        e = np.vdot(s, O @ s).real
        assert abs(oracle.expval(s, pauli_word(word, wires)) - e) < 1e-14
        assert abs(oracle.var(s, pauli_word(word, wires)) - (np.vdot(O @ s, O @ s).real - e * e)) < 1e-14
    rs = np.random.default_rng(5)
    h = rs.normal(size=(4, 4)) + 1j * rs.normal(size=(4, 4))
    h = h + h.conj().T
    Hs = embed_apply(s, h, [3, 1], n)
    assert abs(oracle.expval(s, Obs.hermitian(h, [3, 1])) - np.vdot(s, Hs).real) < 1e-13
    ham = Obs.hamiltonian([0.4, -1.3], [pauli_word("XZ", [0, 2]), Obs.hermitian(h, [3, 1])])
    Hpsi = 0.4 * (dense_pauli("XZ", [0, 2], n) @ s) - 1.3 * Hs
    assert abs(oracle.expval(s, ham) - np.vdot(s, Hpsi).real) < 1e-13
    assert abs(oracle.var(s, ham) - (np.vdot(Hpsi, Hpsi).real - np.vdot(s, Hpsi).real ** 2)) < 1e-12
    p = np.abs(s) ** 2
    t = p.reshape((2,) * n)
    for wires in ([0], [4], [1, 3], [3, 1], [4, 0, 2], [1, 2, 0], [2, 0, 1, 3]):
        rest = tuple(w for w in range(n) if w not in wires)
        marg = t.sum(axis=rest)  # axes in ascending wire order
        # The reference orders output axis j as sorted(wires)[argsort(wires)[j]]
        # (marginal over sorted wires, then transpose_state_tensor with
        # sorting_indices, MeasurementsLQubit.hpp:99-135); its own table
        # (Test_MeasurementsBase.cpp:66-93, wires {1,2,0}) fixes this choice.
        order = list(np.argsort(wires, kind="stable"))
        exp = np.transpose(marg, order).reshape(-1)
        assert np.max(np.abs(oracle.probs(s, wires) - exp)) < 1e-15, wires


def test_cdf_sampling_brute_force(oracle):
    for n in (1, 3, 7):
        s = random_state(n, n)
        if n == 7:
            s[10:20] = 0
            s /= np.linalg.norm(s)
        p = np.abs(s) ** 2
        C = np.concatenate([[0.0], np.cumsum(p)[:-1]])  # exclusive prefix
        u = np.random.default_rng(n).uniform(size=3000)
        u[:3] = [0.0, 1e-300, 1 - 2 ** -53]
        idx = oracle.sample_cdf(s, u)
        exp = np.clip(np.array([np.count_nonzero(C < x) for x in u]) - 1, 0, len(p) - 1)
        # ties at a boundary may round either way in a different summation order
        bad = np.nonzero(idx != exp)[0]
        for b in bad:
            assert np.min(np.abs(np.cumsum(p) - u[b])) < 1e-12
        assert len(bad) == 0 or len(bad) < 3
    # distribution: empirical frequencies (Test_MeasurementsBase.cpp:380-425 style)
    s = random_state(4, 99)
    u = np.random.default_rng(0).uniform(size=200000)
    idx = oracle.sample_cdf(s, u)
    freq = np.bincount(idx, minlength=16) / len(u)
    assert np.max(np.abs(freq - np.abs(s) ** 2)) < 0.01


def param_shift(oracle, tape, obs, n):
    """Exact two-term shift for Pauli-generated rotations."""
    jac = []
    k = 0
    for i, (name, wires, inv, params) in enumerate(tape.ops):
        if not params:
            continue
        row = []
        for sh in (math.pi / 2, -math.pi / 2):
            t = Tape()
            t.ops = [list(o) for o in tape.ops]
            t.ops[i][3] = [params[0] + sh]
            t.ops = [tuple(o) for o in t.ops]
            s = np.zeros(1 << n, complex)
            s[0] = 1
            oracle.apply_tape(s, t)
            row.append([oracle.expval(s, o) for o in obs])
        jac.append((np.array(row[0]) - np.array(row[1])) / 2)
        k += 1
    return np.array(jac).T


def test_adjoint_vs_parameter_shift(oracle):
    from lightning_unification_amd.circuits import plain_layers
    n = 5
    tape, obs, tp = plain_layers(n, 2, seed=3)
    obs = obs + [pauli_word("XX", [0, 1]), Obs.hamiltonian([0.5, -0.7], [pauli_word("ZY", [2, 4]),
                                                                          Obs.named("PauliX", [3])])]
    s = np.zeros(1 << n, complex)
    s[0] = 1
    oracle.apply_tape(s, tape)
    jac = oracle.adjoint(s, tape, obs, tp)
    assert np.max(np.abs(jac - param_shift(oracle, tape, obs, n))) < 1e-12


def test_reference_known_answers(oracle):
    """Values held by the reference's tests (CS/algorithms/tests/
    Test_AdjointJacobian.cpp:147,175,315-317,361-366 and
    CS/measurements/tests/Test_MeasurementsBase.cpp:66-93,141-144)."""
    for p in (-math.pi, -0.4, 0.0, 1.3, math.pi):
        s = np.array([1, 0], complex)
        t = Tape().add("RX", [0], [p])
        oracle.apply_tape(s, t)
        assert abs(oracle.adjoint(s, t, [Obs.named("PauliZ", [0])], [0])[0, 0] + math.sin(p)) < 1e-14
        s = np.array([1, 0], complex)
        t = Tape().add("RY", [0], [p])
        oracle.apply_tape(s, t)
        assert abs(oracle.adjoint(s, t, [Obs.named("PauliX", [0])], [0])[0, 0] - math.cos(p)) < 1e-14
    param = [-math.pi / 7, math.pi / 5, 2 * math.pi / 3]
    t = Tape()
    for w in range(3):
        t.add("RX", [w], [param[w]])
    s = np.zeros(8, complex)
    s[0] = 1
    oracle.apply_tape(s, t)
    jac = oracle.adjoint(s, t, [pauli_word("ZZZ", [0, 1, 2])], [0, 1, 2])[0]
    np.testing.assert_allclose(jac, [-0.1755096592645253, 0.26478810666384334, -0.6312451595102775],
                               atol=1e-7)
    t = Tape()
    for name, w, p in (("RZ", [0], param[0]), ("RY", [0], param[1]), ("RZ", [0], param[2]),
                       ("CNOT", [0, 1], None), ("CNOT", [1, 2], None), ("RZ", [1], param[0]),
                       ("RY", [1], param[1]), ("RZ", [1], param[2])):
        t.add(name, w, [] if p is None else [p])
    s = np.zeros(8, complex)
    s[0] = 1
    oracle.apply_tape(s, t)
    jac = oracle.adjoint(s, t, [pauli_word("XXX", [0, 1, 2])], list(range(6)))[0]
    np.testing.assert_allclose(jac, [0.0, -0.674214427, 0.275139672, 0.275139672, -0.0129093062,
                                     0.323846156], atol=1e-7)
    # createNonTrivialState (TestHelpers.hpp:381-414): RX(phi), RY(phi) per wire
    s = np.zeros(8, complex)
    s[0] = 1
    t = Tape()
    ph = 0.7
    for w in range(3):
        t.add("RX", [w], [ph]).add("RY", [w], [ph])
        ph -= 0.2
    oracle.apply_tape(s, t)
    np.testing.assert_allclose(oracle.probs(s, [0, 1, 2]),
                               [0.67078706, 0.03062806, 0.0870997, 0.00397696, 0.17564072,
                                0.00801973, 0.02280642, 0.00104134], atol=1e-6)
    np.testing.assert_allclose(oracle.probs(s, [2, 1]), [0.84642778, 0.10990612, 0.0386478,
                                                         0.0050183], atol=1e-6)
    np.testing.assert_allclose(oracle.probs(s, [1, 2, 0]),
                               [0.67078706, 0.0870997, 0.17564072, 0.02280642, 0.03062806,
                                0.00397696, 0.00801973, 0.00104134], atol=1e-6)
    np.testing.assert_allclose(oracle.probs(s, [2, 0, 1]),
                               [0.67078706, 0.17564072, 0.03062806, 0.00801973, 0.0870997,
                                0.02280642, 0.00397696, 0.00104134], atol=1e-6)
    ref = {"PauliX": [0.49272486, 0.42073549, 0.28232124],
           "PauliY": [-0.64421768, -0.47942553, -0.29552020],
           "PauliZ": [0.58498357, 0.77015115, 0.91266780]}
    for name, vals in ref.items():
        for w in range(3):
            assert abs(oracle.expval(s, Obs.named(name, [w])) - vals[w]) < 1e-6


def test_mutation_is_caught(oracle):
    """A plausible mistake (transposed operand of a dense 2-qubit matrix)
    must fail the textbook check used above."""
    rs = np.random.default_rng(1)
    m = rs.normal(size=(4, 4)) + 1j * rs.normal(size=(4, 4))
    s = random_state(3, 2)
    good = oracle.apply_matrix(s.copy(), m, [2, 0])
    bad = oracle.apply_matrix(s.copy(), m.T, [2, 0])
    exp = embed_apply(s, m, [2, 0], 3)
    assert np.max(np.abs(good - exp)) < 1e-13
    assert np.max(np.abs(bad - exp)) > 1e-3
