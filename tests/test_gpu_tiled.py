"""GPU parity of the tile-pass executor (pl_sv_apply_ops_tiled) against the
oracle's run of the same tape on the same seeded state, for every gate kind
(random wires, parameters and inverse flags), every legal low-bit count and
tile sizes 2^9 .. 2^12; and against the per-gate GPU path on the 20-qubit
version of the 50-layer random circuit.  Tolerance: 1e-12 relative for
complex128; complex64 accumulates rounding over ~200 gates, 1e-4."""
import numpy as np
import pytest

from conftest import random_state, rel_err
from lightning_unification_amd.circuits import random_layers
from test_tile_plan import without_prep, zoo_tape

pytestmark = pytest.mark.gpu

TOL = {np.complex128: 1e-12, np.complex64: 1e-4}


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,k,low", [(9, 9, 0), (10, 9, 3), (11, 10, 6), (12, 11, 5), (12, 12, 6),
                                     (13, 12, 1), (14, 12, 4), (14, 11, 2)])
def test_tiled_matches_oracle(oracle, dtype, n, k, low):
    from lightning_unification_amd import capi
    tape = zoo_tape(n, 200 + n + k + low)
    psi = random_state(n, 7 + n, dtype)
    exp = oracle.apply_tape(psi.astype(np.complex128), without_prep(tape))
    print(f"tiled n={n} k={k} low={low} {np.dtype(dtype).name}", flush=True)
    with capi.DeviceStateVector.from_array(psi) as sv:
        if dtype == np.complex128 and k > 11:
            # a 2^12-amplitude complex128 tile does not fit a wavefront's registers
            with pytest.raises(ValueError):
                sv.apply_tape_tiled(tape, k, low)
            assert rel_err(sv.state(), psi) == 0  # nothing applied
            return
        npass = sv.apply_tape_tiled(tape, k, low)
        got = sv.state()
    assert npass == len(capi.tile_plan(n, tape, k, low))
    assert rel_err(got, exp) < TOL[dtype]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("k,low", [(9, 1), (12, 6), (11, 4), (10, 3)])
def test_tiled_matches_per_gate_path_20q(dtype, k, low):
    from lightning_unification_amd import capi
    n = 20
    tape = random_layers(n, 50, seed=1002)
    print(f"tiled 20q k={k} low={low} {np.dtype(dtype).name}", flush=True)
    if dtype == np.complex128 and k > 11:
        pytest.skip("complex128 tiles hold at most 2^11 amplitudes (covered by the oracle test)")
    with capi.DeviceStateVector(n, dtype) as a, capi.DeviceStateVector(n, dtype) as b:
        a.apply_tape(tape)
        npass = b.apply_tape_tiled(tape, k, low)
        assert npass < len(tape.ops) // 5
        d = a.distance(b)
        na = a.norm2()
    assert abs(na - 1) < (1e-12 if dtype == np.complex128 else 1e-4)
    assert d < (1e-12 if dtype == np.complex128 else 1e-4), d


def test_tiled_errors():
    from lightning_unification_amd import capi
    tape = random_layers(10, 2, seed=3)
    with capi.DeviceStateVector(10) as sv:
        for k, low in ((11, 3), (8, 3), (10, 7), (13, 3)):
            with pytest.raises(Exception):
                sv.apply_tape_tiled(tape, k, low)
        # nothing was applied by the failed calls
        st = sv.state()
        assert abs(st[0] - 1) < 1e-15


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [5, 8, 9, 12])
def test_tiled_defaults(oracle, dtype, n):
    """tile_bits = 0 / low_bits = -1: the precision's defaults, clamped to the
    state; below 9 qubits the tape runs gate by gate."""
    from lightning_unification_amd import capi
    tape = zoo_tape(n, 300 + n)
    psi = random_state(n, 11 + n, dtype)
    exp = oracle.apply_tape(psi.astype(np.complex128), without_prep(tape))
    print(f"tiled defaults n={n} {np.dtype(dtype).name}", flush=True)
    with capi.DeviceStateVector.from_array(psi) as sv:
        sv.apply_tape_tiled(tape)
        got = sv.state()
    assert rel_err(got, exp) < TOL[dtype]


def test_tiled_long_tape_refills_program_buffers(oracle):
    """9000 gates in one call: the executor's program buffers hold 4096
    operations' worth, so it waits for the queued passes and starts them
    over (more than once) part-way through the tape."""
    from lightning_unification_amd import capi
    n = 12
    tape = zoo_tape(n, 77, count=8900)
    psi = random_state(n, 31)
    exp = oracle.apply_tape(psi.copy(), without_prep(tape))
    print(f"tiled long tape: {len(tape.ops)} ops", flush=True)
    with capi.DeviceStateVector.from_array(psi) as sv:
        npass = sv.apply_tape_tiled(tape, 11, 3)
        got = sv.state()
    assert npass == len(capi.tile_plan(n, tape, 11, 3))
    assert rel_err(got, exp) < 1e-12
