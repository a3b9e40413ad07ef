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
