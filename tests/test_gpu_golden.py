"""GPU engine against golden vectors produced by the reference's own code
(tests/golden), including the plain workload at its full size (20 qubits,
10 layers, <Z_w> on all 20 wires and the 20 x 600 adjoint Jacobian), and
size-independent properties at 30 qubits (16 GiB complex128)."""
import os

import numpy as np
import pytest

from conftest import rel_err
from lightning_unification_amd.circuits import (Obs, Tape, plain_layers, random_hamiltonian,
                                                random_layers, sample_uniforms)
from lightning_unification_amd.gates import GATE_NAMES

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_golden_gates():
    from lightning_unification_amd import capi
    z = load("gates_n4.npz")
    keys = sorted(k[3:] for k in z.files if k.startswith("in_"))
    for key in keys:
        meta = z[f"meta_{key}"]
        gi, inv, wires = int(meta[0]), bool(meta[1]), [int(w) for w in meta[2:]]
        with capi.DeviceStateVector.from_array(z[f"in_{key}"]) as sv:
            sv.apply(GATE_NAMES[gi], wires, inv, list(z[f"par_{key}"]))
            assert rel_err(sv.state(), z[f"out_{key}"]) < 1e-12, (GATE_NAMES[gi], wires, inv)


def test_golden_plain_full_size():
    """Forward pass, 20 expectation values and the 20 x 600 Jacobian at the
    full plain-workload size against the reference's values (rel 1e-12)."""
    from lightning_unification_amd import capi
    z = load("plain_n20.npz")
    tape, obs, tp = plain_layers(20, 10, seed=1001)
    with capi.DeviceStateVector(20) as sv:
        sv.apply_tape(tape)
        st = sv.state()
        assert rel_err(st[:64], z["state_head"]) < 1e-12
        ev = np.array([sv.expval(o) for o in obs])
        assert rel_err(ev, z["expvals"]) < 1e-12
        jac = sv.adjoint_jacobian(tape, obs, tp)
    assert jac.shape == (20, 600)
    assert rel_err(jac, z["jac"]) < 1e-12


def test_golden_random_circuit_n16():
    from lightning_unification_amd import capi
    z = load("random_n16.npz")
    n = 16
    with capi.DeviceStateVector(n) as sv:
        sv.apply_tape(random_layers(n, 50, seed=1002))
        st = sv.state()
        assert rel_err(st[z["idx"]], z["amps"]) < 1e-12
        e = sv.expval(random_hamiltonian(n, 100, seed=1002))
    assert rel_err(e, z["expval"][0]) < 1e-12


def test_golden_adjoint_mixed():
    from lightning_unification_amd import capi
    z = load("adjoint_n5.npz")
    tape = Tape()
    wo = po = 0
    for i in range(len(z["ids"])):
        nw, npar = int(z["nw"][i]), int(z["npar"][i])
        tape.add(GATE_NAMES[int(z["ids"][i])], [int(w) for w in z["wires"][wo:wo + nw]],
                 list(z["par"][po:po + npar]), bool(z["inv"][i]))
        wo += nw
        po += npar
    from lightning_unification_amd.circuits import encode_observables  # noqa: F401
    import ctypes
    prog, coeffs, mats = z["prog"].astype(np.int32), z["coeffs"], z["mats"]
    tp = z["tp"].astype(np.int32)
    ids, nw, wires, inv, npa, par = tape.encode()
    p = lambda a: np.ascontiguousarray(a).ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    K = int(prog[0])
    jac = np.zeros(K * len(tp))
    with capi.DeviceStateVector.from_array(z["state"]) as sv:
        rc = capi.load().pl_adjoint_jacobian(sv.h, len(ids), p(ids), p(nw), p(wires), p(inv), p(npa),
                                             p(par), p(prog), p(coeffs),
                                             p(np.ascontiguousarray(mats).view(np.float64)),
                                             len(tp), p(tp), p(jac))
        assert rc == 0, capi.load().pl_last_error()
    assert rel_err(jac.reshape(K, -1), z["jac"]) < 1e-12


def test_30_qubit_properties():
    """At 30 qubits complex128 (16 GiB): the random 50-layer circuit keeps the
    norm, its inverse sequence returns |0...0>, <H> is real-bounded by
    sum |c|, probabilities sum to 1 and 10^5 samples follow the one-wire
    marginals."""
    from lightning_unification_amd import capi
    n = 30
    tape = random_layers(n, 50, seed=1002)
    ham = random_hamiltonian(n, 100, seed=1002)
    with capi.DeviceStateVector(n) as sv:
        sv.apply_tape(tape)
        one = Obs.named("Identity", [0])
        norm = sv.expval(one)
        assert abs(norm - 1.0) < 1e-12
        e = sv.expval(ham)
        assert abs(e) <= sum(abs(c) for c in ham.coeffs)
        marg = sv.probs([0])
        assert abs(marg.sum() - 1.0) < 1e-12
        u = sample_uniforms(100000, seed=1002)
        idx = sv.sample(u)
        top = (idx >> (n - 1)) & 1
        assert abs(top.mean() - marg[1]) < 0.01
        inv = Tape()
        for name, wires, inverse, params in reversed(tape.ops):
            inv.add(name, wires, params, not inverse)
        sv.apply_tape(inv)
        p0 = sv.probs([0, 1, 2, 3])
        assert abs(p0[0] - 1.0) < 1e-10
