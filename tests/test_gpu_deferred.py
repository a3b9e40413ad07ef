"""Deferred gates (pl_sv_set_deferred): pl_sv_apply_gate / pl_sv_apply_ops
queue their gates, which run as tile passes before anything else uses the
state.  Every way of using the state must see what gate-by-gate application
gives: the oracle at 1e-12 (complex128), and the eager path's own results for
measurements.  Calls that overwrite the whole state drop the queue."""
import numpy as np
import pytest

from conftest import random_state, rel_err
from lightning_unification_amd.circuits import Obs, pauli_word, random_hamiltonian, random_layers
from test_tile_plan import without_prep, zoo_tape

pytestmark = pytest.mark.gpu


def _capi():
    from lightning_unification_amd import capi
    return capi


def _pair(n, dtype, tape, seed=5):
    """An eager and a deferred state from the same start, the tape applied
    to both (the deferred one only queued)."""
    capi = _capi()
    psi = random_state(n, seed, dtype)
    eager = capi.DeviceStateVector.from_array(psi)
    eager.apply_tape(tape)
    lazy = capi.DeviceStateVector.from_array(psi)
    lazy.set_deferred(True)
    lazy.apply_tape(tape)
    return psi, eager, lazy


@pytest.mark.parametrize("n,dtype", [(6, np.complex128), (12, np.complex128), (14, np.complex128),
                                     (13, np.complex64)])
def test_deferred_tape_equals_oracle(oracle, n, dtype):
    tape = zoo_tape(n, 40 + n)
    gates = without_prep(tape)
    print(f"deferred tape: n={n} {np.dtype(dtype).name} {len(gates.ops)} gates", flush=True)
    psi, eager, lazy = _pair(n, dtype, tape)
    with eager, lazy:
        assert lazy.deferred_count() == len(gates.ops)
        got = lazy.state()  # the download applies the queue
        assert lazy.deferred_count() == 0
        exp = oracle.apply_tape(psi.astype(np.complex128), gates)
        tol = 1e-12 if dtype == np.complex128 else 2e-5
        assert rel_err(got, exp) < tol
        assert rel_err(got, eager.state()) < tol


def _use_results(sv, n):
    """Named uses of a state; each returns something comparable."""
    ham = random_hamiltonian(n, 12, seed=9)
    u = np.random.default_rng(3).uniform(size=64)
    return {
        "norm2": lambda: sv.norm2(),
        "expval": lambda: sv.expval(ham),
        "var": lambda: sv.var(pauli_word("XZ", [1, 4])),
        "expval_named": lambda: sv.expval_named("PauliY", [2]),
        "probs": lambda: sv.probs([0, 3, 5]),
        "sample": lambda: sv.sample(u),
        "synchronize": lambda: (sv.synchronize(), sv.state())[1],
        "device_ptr": lambda: (sv.device_ptr(), sv.state())[1],
        "event_record": lambda: (sv.event_record(0), sv.state())[1],
        "apply_matrix": lambda: (sv.apply_matrix(np.eye(2) * 1j, [2]), sv.state())[1],
        "apply_generator": lambda: (sv.apply_generator("RX", [1]), sv.state())[1],
    }


@pytest.mark.parametrize("use", ["norm2", "expval", "var", "expval_named", "probs", "sample",
                                 "synchronize", "device_ptr", "event_record", "apply_matrix",
                                 "apply_generator"])
def test_each_use_applies_the_queue_first(use):
    n = 11
    tape = random_layers(n, 4, seed=21)
    print(f"deferred use: {use}", flush=True)
    _, eager, lazy = _pair(n, np.complex128, tape)
    with eager, lazy:
        a = _use_results(eager, n)[use]()
        b = _use_results(lazy, n)[use]()
        assert lazy.deferred_count() == 0
        if use == "sample":
            assert np.array_equal(a, b)
        else:
            assert rel_err(np.asarray(b), np.asarray(a)) < 1e-12


def test_adjoint_and_copy_and_distance_see_the_queue():
    capi = _capi()
    from test_gpu_adjoint import random_tape
    n = 10
    tape = random_tape(n, 40, np.random.default_rng(4))  # gates the adjoint method takes
    obs = [Obs.named("PauliZ", [0]), pauli_word("XY", [2, 7])]
    tp = list(range(min(6, len(tape.ops))))
    print("deferred: adjoint, copy, distance", flush=True)
    _, eager, lazy = _pair(n, np.complex128, tape)
    with eager, lazy, capi.DeviceStateVector(n, np.complex128) as dst:
        # the Jacobian of the queued state: its forward state is the queue applied
        ja = eager.adjoint_jacobian(tape, obs, tp)
        jb = lazy.adjoint_jacobian(tape, obs, tp)
        assert rel_err(jb, ja) < 1e-12
        lazy.apply_tape(tape)
        eager.apply_tape(tape)
        assert lazy.deferred_count() == len(tape.ops)
        dst.copy_from(lazy)
        assert lazy.deferred_count() == 0
        assert rel_err(dst.state(), eager.state()) < 1e-12
        lazy.apply_tape(tape)
        eager.apply_tape(tape)
        assert eager.distance(lazy) < 1e-12
        assert lazy.deferred_count() == 0


@pytest.mark.parametrize("how", ["upload", "set_basis_state", "reset", "copy_into"])
def test_overwriting_the_state_drops_the_queue(how):
    capi = _capi()
    n = 10
    tape = random_layers(n, 3, seed=8)
    other = random_state(n, 12)
    with capi.DeviceStateVector(n, np.complex128) as sv, \
            capi.DeviceStateVector.from_array(other) as src:
        sv.set_deferred(True)
        sv.apply_tape(tape)
        assert sv.deferred_count() == len(tape.ops)
        if how == "upload":
            sv.upload(other)
            exp = other
        elif how == "set_basis_state":
            sv.set_basis_state(37)
            exp = np.zeros(1 << n, complex)
            exp[37] = 1
        elif how == "reset":
            sv.reset()
            exp = np.zeros(1 << n, complex)
            exp[0] = 1
        else:
            sv.copy_from(src)
            exp = other
        assert sv.deferred_count() == 0
        assert np.array_equal(sv.state(), exp)


def test_errors_at_the_call_and_queue_kept(oracle):
    capi = _capi()
    n = 9
    psi = random_state(n, 2)
    with capi.DeviceStateVector.from_array(psi) as sv:
        sv.set_deferred(True)
        sv.apply("RX", [1], params=[0.3])
        with pytest.raises((ValueError, IndexError)):
            sv.apply("CNOT", [0, 9])  # wire out of range
        with pytest.raises((ValueError, IndexError)):
            sv.apply("RY", [2], params=[0.1, 0.2])  # parameter count
        with pytest.raises((ValueError, IndexError)):
            sv.apply("CZ", [3, 3])  # repeated wire
        assert sv.deferred_count() == 1
        t = random_layers(n, 1, seed=1)
        t.ops = [("RX", [1], False, [0.3])]
        assert rel_err(sv.state(), oracle.apply_tape(psi.copy(), t)) < 1e-12


def test_queue_applies_itself_at_4096_gates(oracle):
    capi = _capi()
    n = 10
    rs = np.random.default_rng(17)
    t = random_layers(n, 1, seed=1)
    t.ops = []
    for _ in range(5000):
        t.add("RX", [int(rs.integers(n))], [float(rs.uniform(-3, 3))])
    psi = random_state(n, 4)
    with capi.DeviceStateVector.from_array(psi) as sv:
        sv.set_deferred(True)
        sv.apply_tape(t)
        assert sv.deferred_count() == 5000 - 4096
        assert rel_err(sv.state(), oracle.apply_tape(psi.copy(), t)) < 1e-12


def test_turning_deferral_off_applies_the_queue(oracle):
    capi = _capi()
    n = 12
    tape = random_layers(n, 2, seed=6)
    psi = random_state(n, 7)
    with capi.DeviceStateVector.from_array(psi) as sv:
        sv.set_deferred(True)
        sv.apply_tape(tape)
        sv.set_deferred(False)
        assert sv.deferred_count() == 0
        sv.apply("Hadamard", [3])
        assert sv.deferred_count() == 0
        t = random_layers(n, 2, seed=6)
        t.add("Hadamard", [3])
        assert rel_err(sv.state(), oracle.apply_tape(psi.copy(), t)) < 1e-12


def test_allocation_free_expval_refuses_a_queue():
    """pl_expval_noalloc must not allocate, and applying queued gates may:
    with gates queued it fails (ValueError) and leaves them queued; after
    pl_sv_synchronize it equals pl_expval."""
    n = 11
    ham = random_hamiltonian(n, 12, seed=9)
    tape = random_layers(n, 3, seed=2)
    _, eager, lazy = _pair(n, np.complex128, tape)
    with eager, lazy:
        with pytest.raises(ValueError):
            lazy.expval_noalloc(ham)
        assert lazy.deferred_count() == len(tape.ops)
        lazy.synchronize()
        assert rel_err(lazy.expval_noalloc(ham), eager.expval(ham)) < 1e-12
