"""The drop-in module driven exactly as Lightning-Qubit's device drives its
own (tests/lq_replay.py restates the device's calls), with the device
mirrors' residency checked through the transfer counters, and every result
checked against the oracle at 1e-12 (complex128)."""
import gc

import numpy as np
import pytest

from conftest import random_state, rel_err
from lightning_unification_amd.circuits import Obs as OObs
from lightning_unification_amd.circuits import Tape as OTape

import lq_replay as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ops():
    from lightning_unification_amd import lightning_qubit_ops
    return lightning_qubit_ops


def counters(ops):
    return np.array(ops.transfer_counters(), dtype=np.int64)


def oracle_tape(operations):
    """The oracle's tape for the replay's operations: state preparation
    dropped (it is applied to the oracle's input state) and Rot expanded
    into RZ, RY, RZ as the serialiser does."""
    t = OTape()
    for op in operations:
        if isinstance(op, (R.QubitStateVector, R.BasisState)):
            continue
        for e in op.expand():
            t.add(e.name, e.wires, e.parameters)
    return t


def oracle_obs(ob):
    if ob.kind == "named":
        return OObs.named(ob.name, ob.wires)
    if ob.kind == "hermitian":
        return OObs.hermitian(ob.matrix, ob.wires)
    if ob.kind == "tensor":
        return OObs.tensor([oracle_obs(f) for f in ob.factors])
    return OObs.hamiltonian(ob.coeffs, [oracle_obs(t) for t in ob.terms])


def workload(n, seed):
    rs = np.random.default_rng(seed)
    psi0 = random_state(n, seed)
    ops_ = [R.QubitStateVector(psi0, list(range(n)))]
    names = ["RX", "RY", "RZ", "Rot", "CNOT", "CRX", "IsingXX", "Hadamard", "T", "SWAP", "Toffoli",
             "CRZ", "IsingZZ", "PhaseShift"]
    arity = {"CNOT": 2, "CRX": 2, "IsingXX": 2, "SWAP": 2, "CRZ": 2, "IsingZZ": 2, "Toffoli": 3}
    npar = {"RX": 1, "RY": 1, "RZ": 1, "Rot": 3, "CRX": 1, "IsingXX": 1, "CRZ": 1, "IsingZZ": 1,
            "PhaseShift": 1}
    for _ in range(40):
        g = names[rs.integers(len(names))]
        w = rs.choice(n, arity.get(g, 1), replace=False).tolist()
        ops_.append(R.Op(g, w, rs.uniform(0, 2 * np.pi, npar.get(g, 0)).tolist()))
    h = rs.normal(size=(4, 4)) + 1j * rs.normal(size=(4, 4))
    h = h + h.conj().T
    observables = [
        R.Obs("named", "PauliZ", [0]),
        R.Obs("named", "PauliX", [n - 1]),
        R.Obs("tensor", factors=[R.Obs("named", "PauliX", [1]), R.Obs("named", "PauliZ", [3])]),
        R.Obs("hamiltonian", coeffs=[0.4, -1.3, 0.25],
              terms=[R.Obs("named", "PauliY", [2]),
                     R.Obs("tensor", factors=[R.Obs("named", "PauliZ", [0]),
                                              R.Obs("named", "PauliZ", [4])]),
                     R.Obs("hermitian", matrix=h, wires=[1, 5])]),
        R.Obs("hermitian", matrix=h, wires=[3, 2]),
    ]
    return psi0, ops_, observables


@pytest.mark.parametrize("c64", [False, True])
def test_apply_with_rotations_then_measure(ops, oracle, c64):
    n = 11
    dtype = np.complex64 if c64 else np.complex128
    tol = 5e-5 if c64 else 1e-12
    psi0, operations, observables = workload(n, 21)
    rotations = [R.Op("Hadamard", [2]), R.Op("RY", [5], [-np.pi / 2]), R.Op("RX", [7], [np.pi / 2])]
    dev = R.LQReplay(n, dtype)
    dev.apply(operations, rotations)
    pre = psi0.astype(dtype)
    oracle.apply_tape(pre, oracle_tape(operations))
    post = pre.copy()
    oracle.apply_tape(post, oracle_tape(rotations))
    # results are in the numpy buffers once apply returns (no explicit release)
    assert rel_err(dev.state, pre) < tol
    assert rel_err(np.ravel(dev._state), post) < tol
    # every measurement wraps the buffer afresh; none re-uploads it
    before = counters(ops)
    for ob in observables:
        assert abs(dev.expval(ob) - oracle.expval(pre, oracle_obs(ob))) < tol * 10
        assert abs(dev.var(ob) - oracle.var(pre, oracle_obs(ob))) < tol * 10
    probs = dev.probability_lightning([0, 4, 6])
    np.testing.assert_allclose(probs, oracle.probs(post, [0, 4, 6]), atol=tol)
    bits = dev.generate_samples(4000)
    assert bits.shape == (4000, n)
    moved = counters(ops) - before
    assert moved[0] == 0, f"re-wrapping unchanged buffers uploaded {moved[0]} bytes"
    assert moved[1] == 0
    # the samples come from the rotated state: marginals within 5 sigma
    p1 = oracle.probs(post, list(range(n)))
    marg = np.array([p1.reshape([2] * n).sum(axis=tuple(j for j in range(n) if j != w))[1]
                     for w in range(n)])
    sig = np.sqrt(marg * (1 - marg) / 4000) + 1e-3
    assert np.all(np.abs(bits.mean(axis=0) - marg) < 5 * sig)


@pytest.mark.parametrize("c64", [False, True])
@pytest.mark.parametrize("train_prep", [True, False])
def test_adjoint_jacobian_state_prep_and_rot(ops, oracle, c64, train_prep):
    """_process_jacobian_tape with a QubitStateVector first: tp_shift,
    record_tp_rows and the scatter into all_params columns, Rot expanded."""
    n = 8
    dtype = np.complex64 if c64 else np.complex128
    psi0, operations, observables = workload(n, 33)
    npar = sum(op.num_params for op in operations)  # the state is parameter 0
    rs = np.random.default_rng(5)
    trainable = sorted(rs.choice(np.arange(1, npar), npar // 2, replace=False).tolist())
    if train_prep:
        trainable = [0] + trainable
    tape = R.Tape(operations, observables, trainable)
    dev = R.LQReplay(n, dtype)
    jac = dev.adjoint_jacobian(tape)
    assert jac.shape == (len(observables), len(trainable))
    # expected: the oracle's adjoint over the expanded tape, every parameter
    fwd = psi0.astype(np.complex128)
    ot = oracle_tape(operations)
    oracle.apply_tape(fwd, ot)
    full = oracle.adjoint(fwd, ot, [oracle_obs(o) for o in observables], list(range(npar - 1)))
    exp = np.zeros_like(jac)
    for c, p in enumerate(trainable):
        if p == 0:
            continue  # the state preparation's column stays zero
        exp[:, c] = full[:, p - 1]
    assert rel_err(jac, exp) < (5e-5 if c64 else 1e-12)
    if not c64:
        # the column bookkeeping against central differences of the forward pass
        ham = oracle_obs(observables[3])
        for c in (1, len(trainable) // 2, len(trainable) - 1):
            p = trainable[c] - 1
            vals = []
            for d in (+1e-5, -1e-5):
                t2 = OTape()
                k = 0
                for name, wires, inv, params in ot.ops:
                    pp = list(params)
                    for j in range(len(pp)):
                        if k == p:
                            pp[j] += d
                        k += 1
                    t2.add(name, wires, pp, inv)
                s = psi0.astype(np.complex128)
                oracle.apply_tape(s, t2)
                vals.append(oracle.expval(s, ham))
            assert abs((vals[0] - vals[1]) / 2e-5 - jac[3, c]) < 1e-6


def test_rewrap_moves_no_bytes_and_sync(ops, oracle):
    n = 14
    s = random_state(n, 2)
    buf = s.copy()
    nbytes = buf.nbytes
    c0 = counters(ops)
    sv = ops.StateVectorC128(buf)
    assert (counters(ops) - c0)[0] == nbytes
    sv.RX([3], False, [0.7])
    sv.sync()  # written back without releasing the wrapper
    exp = oracle.apply_tape(s.copy(), OTape().add("RX", [3], [0.7]))
    assert rel_err(buf, exp) < 1e-12
    del sv
    c1 = counters(ops)
    sv2 = ops.StateVectorC128(buf)
    M = ops.MeasurementsC128(sv2)
    e = M.expval("PauliZ", [3])
    assert abs(e - oracle.expval_named(exp, "PauliZ", [3])) < 1e-12
    assert (counters(ops) - c1)[0] == 0
    del M, sv2


@pytest.mark.parametrize("where", ["interior", "first", "last", "whole"])
def test_host_writes_invalidate_the_device_copy(ops, oracle, where):
    n = 13
    buf = random_state(n, 4)
    sv = ops.StateVectorC128(buf)
    del sv
    if where == "interior":
        buf[1234] = 0.25 + 0.5j
    elif where == "first":
        buf[0] *= -1
    elif where == "last":
        buf[-1] = 0
    else:
        np.copyto(buf, random_state(n, 5))
    c0 = counters(ops)
    sv = ops.StateVectorC128(buf)
    M = ops.MeasurementsC128(sv)
    assert abs(M.expval("PauliX", [2]) - oracle.expval_named(buf, "PauliX", [2])) < 1e-12
    assert (counters(ops) - c0)[0] == buf.nbytes
    # a write while a wrapper is alive is seen by that wrapper's next call
    buf[77] = 0.1
    assert abs(M.expval("PauliY", [6]) - oracle.expval_named(buf, "PauliY", [6])) < 1e-12


def test_conflicting_writes_raise(ops):
    buf = random_state(12, 6)
    sv = ops.StateVectorC128(buf)
    sv.Hadamard([1], False, [])
    buf[100] = 0.0  # the device copy holds the Hadamard result, the host this write
    with pytest.raises(RuntimeError):
        sv.RX([0], False, [0.1])
    del sv


def test_mirror_released_with_its_owner(ops):
    gc.collect()
    base = ops.mirror_count()
    buf = random_state(10, 7)
    sv = ops.StateVectorC128(buf)
    del sv
    assert ops.mirror_count() == base + 1
    del buf
    gc.collect()
    assert ops.mirror_count() == base


def test_copy_then_gate_on_source():
    """pl_sv_copy orders the copy against later work on the source."""
    from lightning_unification_amd import capi
    n = 24
    s = random_state(n, 9)
    with capi.DeviceStateVector.from_array(s) as a, capi.DeviceStateVector(n) as b:
        b.copy_from(a)
        a.apply("Hadamard", [0])
        a.apply("RX", [5], params=[0.3])
        assert np.array_equal(b.state(), s)
