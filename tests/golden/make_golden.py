"""Generate golden vectors from the reference's own Lightning-Qubit code.

Runs the reference C++ compiled in place (oracle/_ref/libref_lq.so, built by
`make -C oracle ref` from /root/reference) on seeded inputs and writes the
inputs and outputs as .npz fixtures next to this script.  The fixtures are
data only; the tests compare the oracle restatement and the GPU engine
against them without needing the reference tree.

    python tests/golden/make_golden.py
"""
import itertools
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from lightning_unification_amd.circuits import (Obs, Tape, encode_observables, pauli_word,  # noqa: E402
                                                plain_layers, random_hamiltonian, random_layers)
from lightning_unification_amd.gates import GATE_NAMES, GATE_NPARAMS, GATE_WIRES, GENERATOR_GATES  # noqa: E402
from oracle.pyoracle import Oracle  # noqa: E402


def seeded_state(n, seed):
    rs = np.random.default_rng(seed)
    v = rs.normal(size=1 << n) + 1j * rs.normal(size=1 << n)
    return v / np.linalg.norm(v)


def gates(ref):
    """Every gate x (up to 6) wire tuples x inverse on a 4-qubit state."""
    n = 4
    out = {}
    rs = np.random.default_rng(2024)
    k = 0
    for gi, g in enumerate(GATE_NAMES):
        arity = GATE_WIRES[gi]
        for kk in ([1, 2, 3, 4] if arity == -1 else [arity]):
            for wires in list(itertools.permutations(range(n), kk))[:6]:
                for inv in (0, 1):
                    p = rs.uniform(-7, 7, GATE_NPARAMS[gi])
                    s = seeded_state(n, 1000 + k)
                    t = Tape().add(g, wires, p, bool(inv))
                    o = ref.apply_tape(s.copy(), t)
                    key = f"{k:04d}"
                    out[f"in_{key}"] = s
                    out[f"out_{key}"] = o
                    out[f"meta_{key}"] = np.array([gi, inv] + list(wires), np.int64)
                    out[f"par_{key}"] = p
                    k += 1
    for g in GENERATOR_GATES:
        gi = GATE_NAMES.index(g)
        kk = 3 if GATE_WIRES[gi] == -1 else GATE_WIRES[gi]
        for wires in list(itertools.permutations(range(n), kk))[:4]:
            s = seeded_state(n, 5000 + k)
            o = s.copy()
            scale = ref.apply_generator(o, g, wires)
            key = f"{k:04d}"
            out[f"gin_{key}"] = s
            out[f"gout_{key}"] = o
            out[f"gmeta_{key}"] = np.array([gi] + list(wires), np.int64)
            out[f"gscale_{key}"] = np.array([scale])
            k += 1
    np.savez_compressed(os.path.join(HERE, "gates_n4.npz"), **out)


def measurements(ref):
    n = 6
    s = seeded_state(n, 77)
    rs = np.random.default_rng(7)
    h2 = rs.normal(size=(4, 4)) + 1j * rs.normal(size=(4, 4))
    h2 = h2 + h2.conj().T
    obs = [Obs.named("PauliX", [0]), Obs.named("PauliY", [3]), Obs.named("PauliZ", [5]),
           Obs.named("Hadamard", [2]), pauli_word("XZY", [1, 4, 0]), Obs.hermitian(h2, [2, 4]),
           Obs.tensor([Obs.hermitian(h2, [0, 1]), Obs.named("PauliX", [5])]),
           random_hamiltonian(n, 9, seed=5)]
    out = {"state": s}
    for i, ob in enumerate(obs):
        prog, coeffs, mats = encode_observables([ob])
        out[f"prog_{i}"] = prog
        out[f"coeffs_{i}"] = coeffs
        out[f"mats_{i}"] = mats
        out[f"expval_{i}"] = np.array([ref.expval(s, ob)])
        out[f"var_{i}"] = np.array([ref.var(s, ob)])
    wl = [[0], [5], [1, 3], [3, 1], [4, 0, 2], list(range(n))]
    for j, w in enumerate(wl):
        out[f"probs_w_{j}"] = np.array(w, np.int64)
        out[f"probs_{j}"] = ref.probs(s, w)
    np.savez_compressed(os.path.join(HERE, "measurements_n6.npz"), **out)
    return obs


def adjoint(ref):
    """Mixed tape with inverses, trainable subset, four observables."""
    n = 5
    gs = ["RX", "RY", "RZ", "CNOT", "CRX", "CRZ", "CRY", "IsingXX", "IsingZZ", "IsingYY",
          "IsingXY", "Hadamard", "T", "PhaseShift", "ControlledPhaseShift", "SingleExcitation",
          "SingleExcitationMinus", "SingleExcitationPlus", "MultiRZ", "DoubleExcitation",
          "DoubleExcitationPlus", "DoubleExcitationMinus", "Toffoli", "SWAP"]
    rs = np.random.default_rng(3)
    tape = Tape()
    for _ in range(40):
        g = gs[rs.integers(len(gs))]
        k = GATE_WIRES[GATE_NAMES.index(g)]
        k = 3 if k == -1 else k
        tape.add(g, rs.choice(n, k, replace=False).tolist(),
                 rs.uniform(0, 6.3, GATE_NPARAMS[GATE_NAMES.index(g)]), bool(rs.integers(2)))
    s = seeded_state(n, 31)
    ref.apply_tape(s, tape)
    obs = [Obs.named("PauliZ", [0]), pauli_word("XY", [1, 3]), random_hamiltonian(n, 6, seed=2),
           Obs.tensor([Obs.named("Hadamard", [4]), Obs.named("PauliZ", [2])])]
    npar = tape.num_params()
    tp = sorted(rs.choice(npar, npar - 3, replace=False).tolist())
    jac = ref.adjoint(s, tape, obs, tp)
    ids, nw, wires, inv, npa, par = tape.encode()
    prog, coeffs, mats = encode_observables(obs)
    np.savez_compressed(os.path.join(HERE, "adjoint_n5.npz"), state=s, ids=ids, nw=nw, wires=wires,
                        inv=inv, npar=npa, par=par, prog=prog, coeffs=coeffs, mats=mats,
                        tp=np.array(tp), jac=jac)


def plain_full(ref):
    """The plain-method workload at its full size: 20 qubits, 10 layers,
    <Z_w> for all wires and the full 20 x 600 Jacobian."""
    tape, obs, tp = plain_layers(20, 10, seed=1001)
    s = np.zeros(1 << 20, np.complex128)
    s[0] = 1
    ref.apply_tape(s, tape)
    ev = np.array([ref.expval(s, o) for o in obs])
    jac = ref.adjoint(s, tape, obs, tp)
    np.savez_compressed(os.path.join(HERE, "plain_n20.npz"), expvals=ev, jac=jac,
                        state_head=s[:64], state_norm=np.array([np.linalg.norm(s)]))


def random_circuit_small(ref):
    """The large-state gate mix (Hadamard, T, RX, RY, RZ, Rot, CRX, CRZ,
    IsingXX, IsingZZ, SWAP, Toffoli) at 16 qubits, with the 100-term
    Hamiltonian."""
    n = 16
    tape = random_layers(n, 50, seed=1002)
    s = np.zeros(1 << n, np.complex128)
    s[0] = 1
    ref.apply_tape(s, tape)
    ham = random_hamiltonian(n, 100, seed=1002)
    e = ref.expval(s, ham)
    idx = np.sort(np.random.default_rng(16).choice(1 << n, 4096, replace=False))
    np.savez_compressed(os.path.join(HERE, "random_n16.npz"), idx=idx, amps=s[idx],
                        expval=np.array([e]), norm=np.array([np.linalg.norm(s)]))


def cdf_indices_exact(s, u, chunk=1 << 24):
    """Inverse-CDF indices of the uniforms `u` (any order) on state `s`, with
    the prefix sums of |psi|^2 carried in extended precision so that the
    result does not depend on a summation order: index i with
    C_i < u <= C_{i+1}, C exclusive, index 0 for u <= C_1 and the last index
    for u above the total (the rule of MeasuresFunctors.hpp:81-98, where
    C_N := 1).  Also returns each uniform's distance to the nearer of its two
    boundaries, which tells a comparison where two summation orders may
    legitimately disagree."""
    order = np.argsort(u, kind="stable")
    us = u[order].astype(np.longdouble)
    idx = np.full(u.size, s.size - 1, np.int64)
    gap = np.full(u.size, np.inf)
    base = np.longdouble(0)
    j = 0
    for lo in range(0, s.size, chunk):
        a = s[lo:lo + chunk]
        p = (a.real.astype(np.float64) ** 2 + a.imag.astype(np.float64) ** 2).astype(np.longdouble)
        incl = base + np.cumsum(p)
        hi_j = int(np.searchsorted(us, incl[-1], side="right"))
        if hi_j > j:
            k = np.searchsorted(incl, us[j:hi_j], side="left")
            idx[order[j:hi_j]] = lo + k
            upper = incl[k]
            lower = np.where(k > 0, incl[np.maximum(k - 1, 0)], base)
            gap[order[j:hi_j]] = np.minimum(upper - us[j:hi_j], us[j:hi_j] - lower).astype(np.float64)
            j = hi_j
        base = incl[-1]
    return idx, gap


def c2_full(ref, n=30):
    """The single-large-state workload at its full size: the seeded 50-layer
    gate mix on 30 qubits (complex128, 16 GiB) run by the reference's own
    kernels, then the 100-term Hamiltonian term by term through the
    reference's Measurements::expval (copy, apply, innerProdC: 32 GiB peak),
    amplitudes at 4096 seeded indices and the inverse-CDF indices of the
    first 10^4 of the benchmark's 10^6 uniforms.  Scalars and small arrays
    only are written; the state is never stored."""
    import time
    from lightning_unification_amd.circuits import sample_uniforms
    tape = random_layers(n, 50, seed=1002)
    ham = random_hamiltonian(n, 100, seed=1002)
    s = np.zeros(1 << n, np.complex128)
    s[0] = 1
    t0 = time.time()
    for k in range(0, len(tape), 10):
        part = Tape()
        part.ops = tape.ops[k:k + 10]
        ref.apply_tape(s, part)
        print(f"forward {k + len(part)}/{len(tape)} gates, {time.time() - t0:.0f} s", flush=True)
    idx = np.sort(np.random.default_rng(30).choice(1 << n, 4096, replace=False))
    amps = s[idx].copy()
    norm = float(np.sqrt(np.vdot(s, s).real))
    u = sample_uniforms(10 ** 6, seed=1002)[:10 ** 4]
    sidx, gap = cdf_indices_exact(s, u)
    terms = []
    for t, term in enumerate(ham.terms):
        terms.append(ref.expval(s, term))
        print(f"term {t + 1}/{len(ham.terms)}: {terms[-1]:.17g}, {time.time() - t0:.0f} s", flush=True)
    terms = np.array(terms)
    coeffs = np.array(ham.coeffs)
    np.savez_compressed(os.path.join(HERE, f"c2_n{n}.npz"), idx=idx, amps=amps,
                        norm=np.array([norm]), term_expvals=terms, coeffs=coeffs,
                        expval=np.array([float(np.dot(coeffs, terms))]),
                        uniforms=u, sample_idx=sidx, sample_gap=gap)


if __name__ == "__main__":
    ref = Oracle(ref=True)
    if len(sys.argv) > 1 and sys.argv[1] == "c2_n30":
        c2_full(ref, 30)
        print("c2_n30 written to", HERE)
        sys.exit(0)
    gates(ref)
    measurements(ref)
    adjoint(ref)
    random_circuit_small(ref)
    plain_full(ref)
    print("fixtures written to", HERE)
