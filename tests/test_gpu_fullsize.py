"""The large configurations at their full size.

30 qubits, complex128 (16 GiB): the seeded 50-layer gate mix through the
per-gate path, checked against scalars the reference's own Lightning-Qubit
kernels produced at the same size (tests/golden/c2_n30.npz, written by
tests/golden/make_golden.py from the reference compiled in place):
amplitudes at 4096 seeded indices, each of the 100 Pauli terms' expectation
and the Hamiltonian's, and the inverse-CDF indices of the first 10^4 of the
benchmark's uniforms (bit-exact outside the tie band).

33 qubits, complex64 (64 GiB, indices past 2^32): the same gate mix through
the fused matrix-core path and the tile-pass path, with properties no
reference run at this size is needed for: the norm, and the inverse tape
taking the state back to |0...0>."""
import os

import numpy as np
import pytest

from lightning_unification_amd.circuits import Obs, Tape, random_hamiltonian, random_layers

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c2_n30.npz")


def _capi():
    from lightning_unification_amd import capi
    return capi


def _gather(sv, idx):
    """Amplitudes of a device state at the given indices, one small
    device-to-host copy each (the whole state would be 16 GiB)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    item = np.dtype(sv.dtype).itemsize
    base = sv.device_ptr()
    out = np.empty(len(idx), sv.dtype)
    sv.synchronize()
    for k, i in enumerate(idx):
        rc = hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data + k * item), ctypes.c_void_p(base + int(i) * item),
                           item, 2)  # hipMemcpyDeviceToHost
        assert rc == 0
    return out


@pytest.mark.skipif(not os.path.exists(GOLDEN), reason="c2_n30.npz not generated")
def test_c2_30_qubits_against_the_reference():
    capi = _capi()
    g = np.load(GOLDEN)
    n = 30
    tape = random_layers(n, 50, seed=1002)
    ham = random_hamiltonian(n, 100, seed=1002)
    print("c2 full size: forward, 990 gates on 30 qubits", flush=True)
    with capi.DeviceStateVector(n, np.complex128) as sv:
        sv.apply_tape(tape)
        amps = _gather(sv, g["idx"])
        err = np.linalg.norm(amps - g["amps"]) / np.linalg.norm(g["amps"])
        assert err < 1e-12, err
        assert abs(np.sqrt(sv.norm2()) - g["norm"][0]) < 1e-12
        print("c2 full size: 100 terms", flush=True)
        terms = np.array([sv.expval(t) for t in ham.terms])
        assert np.max(np.abs(terms - g["term_expvals"])) < 1e-12
        e = sv.expval(ham)
        assert abs(e - g["expval"][0]) < 1e-12 * max(1.0, float(np.sum(np.abs(g["coeffs"]))))
        print("c2 full size: samples", flush=True)
        idx = sv.sample(g["uniforms"])
        bad = np.nonzero(idx != g["sample_idx"])[0]
        # the prefix sums differ from the extended-precision ones only by
        # rounding: indices may differ only for uniforms within 1e-12 of a
        # boundary
        assert np.all(g["sample_gap"][bad] <= 1e-12), (bad, g["sample_gap"][bad])


def _inverse(tape):
    inv = Tape()
    for name, wires, inverse, params in reversed(tape.ops):
        inv.add(name, wires, params, not inverse)
    return inv


def _z_all(n):
    return Obs.hamiltonian([1.0 / n] * n, [Obs.named("PauliZ", [w]) for w in range(n)])


@pytest.mark.parametrize("path", ["fused", "tiled"])
def test_c3_33_qubits_norm_and_inverse(path):
    capi = _capi()
    n = 33
    tape = random_layers(n, 50, seed=1002)
    print(f"c3 full size ({path}): forward, {len(tape)} gates on 33 qubits, complex64", flush=True)
    with capi.DeviceStateVector(n, np.complex64) as sv:
        if path == "fused":
            sv.apply_tape_fused(tape, 5, renorm=False)
        else:
            sv.apply_tape_tiled(tape)
        drift = abs(sv.norm2() - 1.0)
        assert drift < 1e-4, drift
        # a state this deep in the circuit is far from a basis state
        assert abs(sv.expval(_z_all(n))) < 0.05
        print(f"c3 full size ({path}): inverse tape", flush=True)
        if path == "fused":
            sv.apply_tape_fused(_inverse(tape), 5, renorm=True)
        else:
            sv.apply_tape_tiled(_inverse(tape), renorm=True)
        assert abs(sv.norm2() - 1.0) < 1e-6  # renormalised in FP64
        # back at |0...0>: every <Z_w> = 1 and amplitude 0 carries the norm
        assert sv.expval(_z_all(n)) > 1 - 1e-4
        p0 = sv.probs([0, 16, 32])
        assert p0[0] > 1 - 1e-4
