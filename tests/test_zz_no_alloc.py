"""pl_expval_noalloc against pl_expval through the C-ABI: the same Pauli-sum
kernels with the term tables and partial sums in the library's fixed device
array instead of allocated buffers."""
import numpy as np
import pytest

from conftest import random_state, rel_err
from lightning_unification_amd.circuits import Obs, pauli_word, random_hamiltonian

pytestmark = pytest.mark.gpu


def _capi():
    from lightning_unification_amd import capi
    return capi


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_noalloc_matches_expval_two_calls(dtype):
    capi = _capi()
    tol = 1e-12 if dtype == np.complex128 else 1e-6
    print(f"noalloc first call: 20 qubits, 100 terms, {np.dtype(dtype).name}", flush=True)
    with capi.DeviceStateVector.from_array(random_state(20, 1, dtype)) as big:
        ham = random_hamiltonian(20, 100, seed=3)
        a = big.expval_noalloc(ham)
        b = big.expval(ham)
        assert rel_err(a, b) < tol
    print("noalloc second call: 12 qubits, 9 terms", flush=True)
    with capi.DeviceStateVector.from_array(random_state(12, 2, dtype)) as small:
        ham = random_hamiltonian(12, 9, seed=4)
        a = small.expval_noalloc(ham)
        b = small.expval(ham)
        assert rel_err(a, b) < tol
        word = pauli_word("XZY", [0, 5, 11])
        assert rel_err(small.expval_noalloc(word), small.expval(word)) < tol
        assert rel_err(small.expval_noalloc(Obs.named("PauliZ", [3])),
                       small.expval(Obs.named("PauliZ", [3]))) < tol


def test_noalloc_rejects_what_needs_scratch():
    capi = _capi()
    with capi.DeviceStateVector.from_array(random_state(8, 5)) as sv:
        h = np.array([[1.0, 0.5], [0.5, -1.0]])
        with pytest.raises(ValueError):
            sv.expval_noalloc(Obs.hermitian(h, [2]))
        # more Pauli terms than the fixed scratch holds (32 bytes per term
        # against 2 MiB)
        many = random_hamiltonian(8, 100000, seed=6)
        with pytest.raises(ValueError):
            sv.expval_noalloc(many)
        # the array is still usable afterwards
        ham = random_hamiltonian(8, 20, seed=7)
        assert rel_err(sv.expval_noalloc(ham), sv.expval(ham)) < 1e-12
