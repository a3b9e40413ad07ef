"""GPU parity of the precision path (pl_sv_apply_ops_fused): fused dense
blocks against the oracle's unfused run.  complex128 (dense-matrix kernels)
at 1e-12; complex64 on FP32 matrix cores (n >= 12) against the complex128
oracle, with the error stated in the tolerance; renormalisation; the FP64
norm and distance helpers."""
import numpy as np
import pytest

from conftest import random_state, rel_err
from lightning_unification_amd.circuits import Tape, random_layers
from test_fusion import mixed_tape, without_prep

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [3, 7, 13])
@pytest.mark.parametrize("max_wires", [1, 3, 5])
def test_fused_c128_matches_oracle(oracle, n, max_wires):
    from lightning_unification_amd import capi
    tape = mixed_tape(n, 40 + n) if n >= 6 else random_layers(n, 6, seed=n)
    psi = random_state(n, 3)
    exp = oracle.apply_tape(psi.copy(), without_prep(tape))
    print(f"fused c128 n={n} w={max_wires}", flush=True)
    sv = capi.DeviceStateVector.from_array(psi)
    nb = sv.apply_tape_fused(tape, max_wires)
    assert 0 < nb <= len(tape)
    assert np.max(np.abs(sv.state() - exp)) < 1e-12
    sv.close()


@pytest.mark.parametrize("n", [12, 13, 15])
def test_fused_c64_mfma(oracle, n):
    """Matrix-core blocks: against the complex128 oracle (FP32 arithmetic,
    error ~1e-6) and against the complex64 dense-matrix kernels."""
    from lightning_unification_amd import capi
    tape = random_layers(n, 8, seed=100 + n)
    tape.add("MultiRZ", list(range(7)), [0.77])
    psi = random_state(n, 4)
    exp = oracle.apply_tape(psi.copy(), tape)
    print(f"fused c64 mfma n={n}", flush=True)
    a = capi.DeviceStateVector.from_array(psi.astype(np.complex64))
    nb = a.apply_tape_fused(tape, 5)
    b = capi.DeviceStateVector.from_array(psi.astype(np.complex64))
    b.apply_tape_fused(tape, 5, mfma=False)
    sa, sb = a.state(), b.state()
    err = np.max(np.abs(sa - exp))
    assert err < 2e-5 * np.max(np.abs(exp)) * np.sqrt(nb), err
    assert np.max(np.abs(sa - sb)) < 2e-5 * np.max(np.abs(exp)) * np.sqrt(nb)
    # distance helper against numpy, mixed precision
    ref = capi.DeviceStateVector.from_array(exp)
    d = a.distance(ref)
    assert abs(d - np.linalg.norm(sa.astype(np.complex128) - exp)) < 1e-9
    s128 = sa.astype(np.complex128)
    assert abs(a.norm2() - np.vdot(s128, s128).real) < 1e-9
    for s in (a, b, ref):
        s.close()


def test_fused_renormalises(oracle):
    from lightning_unification_amd import capi
    n = 13
    tape = random_layers(n, 6, seed=7)
    psi = random_state(n, 8) * 1.25  # deliberately not unit norm
    exp = oracle.apply_tape(psi.copy(), tape) / 1.25
    for dt, mfma in ((np.complex64, True), (np.complex64, False), (np.complex128, True)):
        print(f"renorm {np.dtype(dt).name} mfma={mfma}", flush=True)
        sv = capi.DeviceStateVector.from_array(psi.astype(dt))
        sv.apply_tape_fused(tape, 5, renorm=True, mfma=mfma)
        tol = 1e-12 if dt == np.complex128 else 1e-6
        assert abs(sv.norm2() - 1.0) < tol
        assert rel_err(sv.state(), exp) < (1e-12 if dt == np.complex128 else 5e-5)
        sv.close()


def test_fused_rejects_bad_width():
    from lightning_unification_amd import capi
    sv = capi.DeviceStateVector(4)
    with pytest.raises(ValueError):
        sv.apply_tape_fused(Tape().add("RX", [0], [0.1]), 6)
    with pytest.raises(ValueError):
        sv.apply_tape_fused(Tape().add("RX", [5], [0.1]), 5)
    sv.close()


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_fused_long_tape_reuses_matrix_slots(oracle, dtype):
    """More blocks than the executor's 256 matrix slots: it waits for the
    queued blocks and starts the slots over part-way through (both the
    matrix-core blocks of complex64 and the dense-matrix blocks of
    complex128)."""
    from lightning_unification_amd import capi
    n = 12
    tape = random_layers(n, 330, seed=55)  # 2569 operations, 319 blocks
    psi = random_state(n, 9)
    exp = oracle.apply_tape(psi.copy(), tape)
    print(f"fused long tape: {len(tape.ops)} ops {np.dtype(dtype).name}", flush=True)
    with capi.DeviceStateVector.from_array(psi.astype(dtype)) as sv:
        nb = sv.apply_tape_fused(tape, 5, renorm=dtype == np.complex64)
        got = sv.state()
    assert nb > 300
    tol = 1e-12 if dtype == np.complex128 else 2e-5 * np.sqrt(nb)
    assert np.max(np.abs(got - exp)) < tol * max(1.0, np.max(np.abs(exp)))
