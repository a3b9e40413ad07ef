"""GPU parity: every gate, generator and dense matrix through the C-ABI
against the oracle (Lightning-Qubit semantics) on the same seeded states.
Tolerance: 1e-12 relative for complex128; 2e-5 for complex64."""
import itertools

import numpy as np
import pytest

from conftest import random_state, rel_err
from lightning_unification_amd.circuits import Tape
from lightning_unification_amd.gates import GATE_NAMES, GATE_NPARAMS, GATE_WIRES, GENERATOR_GATES

pytestmark = pytest.mark.gpu

TOL = {np.complex128: 1e-12, np.complex64: 2e-5}


def _capi():
    from lightning_unification_amd import capi
    return capi


def wire_tuples(n, k, limit=None):
    """Ordered k-tuples of distinct wires: all of them, or `limit` distinct
    ones drawn at random (without listing all n!/(n-k)! of them first)."""
    count = 1
    for i in range(k):
        count *= n - i
    if not limit or count <= limit:
        return list(itertools.permutations(range(n), k))
    rs = np.random.default_rng(k * 100 + n)
    seen = set()
    while len(seen) < limit:
        seen.add(tuple(int(w) for w in rs.permutation(n)[:k]))
    return sorted(seen)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [1, 3, 7, 11, 13])
def test_all_gates_all_wires(oracle, dtype, n):
    capi = _capi()
    rs = np.random.default_rng(n)
    worst = 0.0
    for gi, g in enumerate(GATE_NAMES):
        k = GATE_WIRES[gi]
        ks = [kk for kk in ([1, 2, 3, n] if k == -1 else [k]) if kk <= n]
        for kk in ks:
            for wires in wire_tuples(n, kk, limit=40):
                for inv in (False, True):
                    p = list(rs.uniform(-7, 7, GATE_NPARAMS[gi]))
                    s0 = random_state(n, int(rs.integers(1 << 30)), dtype)
                    exp = oracle.apply_tape(s0.copy(), Tape().add(g, wires, p, inv))
                    sv = capi.DeviceStateVector.from_array(s0)
                    sv.apply(g, wires, inv, p)
                    got = sv.state()
                    sv.close()
                    e = rel_err(got, exp)
                    worst = max(worst, e)
                    assert e < TOL[dtype], (g, wires, inv, e)
    print("worst rel err", worst)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_four_wire_gates_on_the_tile_path(oracle, dtype):
    """Two-member group ops with one member bit inside a 128-byte line and
    three at bit 8 or above: the LDS tile kernel with four high pins (13
    qubits: wire w is bit 12 - w, so wires 10..12 are bits 0..2 and wires
    0..4 are bits 8..12)."""
    capi = _capi()
    n = 13
    rs = np.random.default_rng(413)
    tuples = [(0, 1, 2, 12), (12, 0, 3, 1), (4, 11, 2, 0), (1, 3, 10, 2), (2, 4, 0, 11)]
    for g in ("DoubleExcitation", "DoubleExcitationMinus", "DoubleExcitationPlus"):
        for wires in tuples:
            for inv in (False, True):
                p = list(rs.uniform(-7, 7, 1))
                s0 = random_state(n, int(rs.integers(1 << 30)), dtype)
                exp = oracle.apply_tape(s0.copy(), Tape().add(g, wires, p, inv))
                with capi.DeviceStateVector.from_array(s0) as sv:
                    sv.apply(g, wires, inv, p)
                    assert rel_err(sv.state(), exp) < TOL[dtype], (g, wires, inv)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_generators(oracle, dtype):
    capi = _capi()
    n = 6
    for g in GENERATOR_GATES:
        k = 3 if g == "MultiRZ" else GATE_WIRES[GATE_NAMES.index(g)]
        for wires in wire_tuples(n, k, limit=12):
            s0 = random_state(n, sum(map(ord, g)) * 131 + sum(wires) * 7 + len(wires), dtype)
            a = s0.copy()
            sa = oracle.apply_generator(a, g, wires)
            sv = capi.DeviceStateVector.from_array(s0)
            sb = sv.apply_generator(g, wires)
            assert sa == sb, (g, sa, sb)
            assert rel_err(sv.state(), a) < TOL[dtype], (g, wires)
            sv.close()


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6])
def test_matrices(oracle, dtype, k):
    capi = _capi()
    n = 7
    rs = np.random.default_rng(k)
    for wires in wire_tuples(n, k, limit=6):
        m = rs.normal(size=(1 << k, 1 << k)) + 1j * rs.normal(size=(1 << k, 1 << k))
        m = m.astype(dtype)
        for inv in (False, True):
            s0 = random_state(n, int(rs.integers(1 << 30)), dtype)
            exp = oracle.apply_matrix(s0.copy(), m, wires, inv)
            sv = capi.DeviceStateVector.from_array(s0)
            sv.apply_matrix(m, wires, inv)
            scale = max(1.0, float(np.abs(m).sum(axis=1).max()))
            assert rel_err(sv.state(), exp, floor=scale) < TOL[dtype] * 4, (k, wires, inv)
            sv.close()


def test_errors():
    capi = _capi()
    sv = capi.DeviceStateVector(3)
    with pytest.raises(ValueError):
        sv.apply("RX", [0], False, [])          # wrong parameter count
    with pytest.raises(ValueError):
        sv.apply("CNOT", [0, 0])                # repeated wire
    with pytest.raises(ValueError):
        sv.apply("RX", [5], False, [0.1])       # wire out of range
    with pytest.raises(ValueError):
        sv.apply("CNOT", [0])                   # wrong arity
    with pytest.raises(ValueError):
        sv.apply_matrix(np.eye(2), [], False)   # no wires
    sv.close()
