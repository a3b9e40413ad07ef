"""Uploads and downloads of whole states (pl_sv_upload / pl_sv_download).
From 8 MiB on they go in 64 MiB chunks through two pinned slots that every
handle of the process shares; the results must be the caller's bytes,
exactly, with device work before a download included."""
import numpy as np
import pytest

from conftest import random_state, rel_err
from lightning_unification_amd.circuits import random_layers

pytestmark = pytest.mark.gpu


def _capi():
    from lightning_unification_amd import capi
    return capi


# 128 MiB (two full chunks), 32 MiB (one short chunk), 8 MiB (the threshold),
# 4 MiB (one plain copy)
@pytest.mark.parametrize("n,dtype", [(23, np.complex128), (22, np.complex64), (19, np.complex128),
                                     (18, np.complex128)])
def test_round_trip_is_exact(n, dtype):
    capi = _capi()
    psi = random_state(n, 70 + n, dtype)
    print(f"transfer: n={n} {np.dtype(dtype).name} {psi.nbytes >> 20} MiB", flush=True)
    with capi.DeviceStateVector.from_array(psi) as sv:
        assert np.array_equal(sv.state(), psi)
        other = psi[::-1].copy()
        sv.upload(other)
        assert np.array_equal(sv.state(), other)


def test_download_follows_the_gates(oracle):
    capi = _capi()
    n = 23
    psi = random_state(n, 5)
    t = random_layers(n, 1, seed=3)
    t.ops = [("Hadamard", [n - 1], False, []), ("RX", [0], False, [0.4]), ("CNOT", [0, n - 1], False, [])]
    print("transfer: gates then download, n=23", flush=True)
    with capi.DeviceStateVector.from_array(psi) as sv:
        sv.apply_tape(t)
        got = sv.state()
    exp = oracle.apply_tape(psi.copy(), t)
    assert rel_err(got, exp) < 1e-12


def test_two_handles_share_the_slots():
    capi = _capi()
    n = 22
    a0, b0 = random_state(n, 1), random_state(n, 2)
    print("transfer: two handles, n=22", flush=True)
    with capi.DeviceStateVector.from_array(a0) as a, capi.DeviceStateVector.from_array(b0) as b:
        for k in range(3):
            a.upload(b0 if k % 2 else a0)
            b.upload(a0 if k % 2 else b0)
            assert np.array_equal(a.state(), b0 if k % 2 else a0)
            assert np.array_equal(b.state(), a0 if k % 2 else b0)


def test_large_outputs_and_inputs(oracle):
    """The measurements' large host arrays take the same chunked path:
    probabilities of every basis state (16 MiB at 21 qubits, 32 MiB at 22),
    marginal probabilities of 21 of 22 wires (16 MiB), 10^6 uniforms in and
    indices out (8 MB each), and 2 x 10^5 bit rows of 22 qubits (35 MB)."""
    capi = _capi()
    for n in (21, 22):
        psi = random_state(n, 40 + n)
        print(f"transfer: probs n={n}", flush=True)
        with capi.DeviceStateVector.from_array(psi) as sv:
            assert rel_err(sv.probs(), oracle.probs(psi)) < 1e-12
            if n == 22:
                w = list(range(1, n))
                assert rel_err(sv.probs(w), oracle.probs(psi, w)) < 1e-12
                u = np.sort(np.random.default_rng(9).uniform(size=1_000_000))
                print("transfer: 10^6 samples", flush=True)
                got = sv.sample(u)
                assert np.array_equal(got, oracle.sample_cdf(psi, u))
                bits = sv.generate_samples(200_000, seed=11)
                assert bits.shape == (200_000, n) and bits.nbytes > 8 << 20
                # sample s depends on (seed, s) alone: the first rows of the
                # chunked copy equal a call small enough for a plain copy
                head = sv.generate_samples(1000, seed=11)
                assert head.nbytes < 8 << 20
                assert np.array_equal(bits[:1000], head)
                assert np.all((bits == 0) | (bits == 1))


def test_module_mirror_of_a_large_buffer(oracle):
    """The pybind module's device mirror of a 20-qubit (16 MiB) numpy buffer,
    driven as the device does: the first wrap uploads through the pinned
    slots, the gates run on the device, and the write-back at the wrapper's
    release fills the (write-watched) buffer through them.  A host write
    afterwards is seen, and the next wrap uploads it again."""
    from lightning_unification_amd import lightning_qubit_ops as ops
    from lightning_unification_amd.circuits import Tape, plain_layers
    from test_gpu_pybind import apply_lightning
    n = 20
    tape, _, _ = plain_layers(n, 1, seed=8)
    state = random_state(n, 9)
    buf = state.copy()
    print("transfer: module mirror, n=20", flush=True)
    up0, down0 = ops.transfer_counters()
    apply_lightning(ops, buf, tape)
    assert rel_err(buf, oracle.apply_tape(state.copy(), tape)) < 1e-12
    up1, down1 = ops.transfer_counters()
    assert (up1 - up0, down1 - down0) == (buf.nbytes, buf.nbytes)
    buf[:] = state  # a host write into the mirrored buffer
    apply_lightning(ops, buf, Tape().add("Hadamard", [0]))
    assert rel_err(buf, oracle.apply_tape(state.copy(), Tape().add("Hadamard", [0]))) < 1e-12
    up2, _ = ops.transfer_counters()
    assert up2 - up1 == buf.nbytes


@pytest.mark.parametrize("wires", [[0, 5, 2], [17, 3, 9], [16, 17, 1], [15, 17, 0, 16],
                                   [2, 17, 15, 7, 16, 0]])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_marginals_with_low_wires(oracle, wires, dtype):
    """Marginal probabilities whose wires include 0 to 3 of the three lowest
    bits (wires n-1, n-2, n-3: the amplitudes of one 128-byte line), wires in
    any order, at 18 qubits (2^12 or more amplitudes per bin)."""
    capi = _capi()
    n = 18
    psi = random_state(n, 60 + len(wires), dtype)
    print(f"marginal: wires {wires} {np.dtype(dtype).name}", flush=True)
    with capi.DeviceStateVector.from_array(psi) as sv:
        got = sv.probs(wires)
    exp = oracle.probs(psi.astype(np.complex128), wires)
    assert rel_err(got, exp) < (1e-12 if dtype == np.complex128 else 1e-6)
