"""GPU parity for measurements through the C-ABI: expval / var (named,
Hermitian, tensor products, Hamiltonians), probs, probs(wires) and
inverse-CDF sampling against the oracle on the same seeded states."""
import numpy as np
import pytest

from conftest import random_state, rel_err
from lightning_unification_amd.circuits import Obs, pauli_word, random_hamiltonian

pytestmark = pytest.mark.gpu

TOL = {np.complex128: 1e-12, np.complex64: 2e-5}


def _capi():
    from lightning_unification_amd import capi
    return capi


def observables(n, rs, dtype):
    h2 = rs.normal(size=(4, 4)) + 1j * rs.normal(size=(4, 4))
    h2 = (h2 + h2.conj().T).astype(dtype)
    h1 = rs.normal(size=(2, 2)) + 1j * rs.normal(size=(2, 2))
    h1 = (h1 + h1.conj().T).astype(dtype)
    return [
        Obs.named("PauliZ", [n - 1]),
        Obs.named("PauliX", [0]),
        Obs.named("PauliY", [1]),
        Obs.named("Hadamard", [2]),
        Obs.named("Identity", [1]),
        pauli_word("XYZ", [0, 3, n - 1]),
        pauli_word("YY", [2, 1]),
        Obs.hermitian(h1, [3]),
        Obs.hermitian(h2, [4, 1]),
        Obs.tensor([Obs.hermitian(h2, [0, 2]), Obs.named("PauliY", [n - 1])]),
        Obs.tensor([Obs.named("Hadamard", [1]), Obs.named("PauliZ", [0])]),
        random_hamiltonian(n, 25, seed=3),
        Obs.hamiltonian([0.3, -1.2, 0.7], [Obs.hermitian(h2, [1, 3]), pauli_word("ZZ", [0, n - 1]),
                                           Obs.named("Hadamard", [2])]),
    ]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [5, 12])
def test_expval_var(oracle, dtype, n):
    capi = _capi()
    rs = np.random.default_rng(n)
    s = random_state(n, 10 + n, dtype)
    sv = capi.DeviceStateVector.from_array(s)
    for ob in observables(n, rs, dtype):
        for var in (False, True):
            exp = oracle.expval(s, ob, var)
            got = sv.var(ob) if var else sv.expval(ob)
            assert rel_err(got, exp) < TOL[dtype] * 10, (ob.kind, ob.name, var, got, exp)
    for name in ("PauliX", "PauliY", "PauliZ", "Hadamard"):
        for var in (False, True):
            exp = oracle.expval_named(s, name, [2], var)
            got = sv.expval_named(name, [2], var)
            assert rel_err(got, exp) < TOL[dtype] * 10, (name, var)
    m = rs.normal(size=(4, 4)) + 1j * rs.normal(size=(4, 4))
    m = (m + m.conj().T).astype(dtype)
    assert rel_err(sv.expval_matrix(m, [3, 0]), oracle.expval_matrix(s, m, [3, 0]),
                   floor=4.0) < TOL[dtype] * 10
    sv.close()


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [4, 13])
def test_probs(oracle, dtype, n):
    capi = _capi()
    s = random_state(n, 77 + n, dtype)
    sv = capi.DeviceStateVector.from_array(s)
    for wires in [None, [0], [n - 1], [2, 0], [0, 1, 2], [n - 1, 1, 0], [1, 2, 0], [2, 0, 1],
                  [3, 0, 2, 1], list(range(n)), list(reversed(range(n))),
                  list(range(1, n)) + [0]]:
        exp = oracle.probs(s, wires)
        got = sv.probs(wires)
        assert got.shape == exp.shape
        assert np.max(np.abs(got - exp)) < TOL[dtype] * 10, wires
    sv.close()


def boundary_ties(cdf_incl, u, idx_a, idx_b, band=1e-12):
    """Indices may differ only where u lies within `band` of a CDF boundary."""
    bad = np.nonzero(idx_a != idx_b)[0]
    for s in bad:
        lo = min(idx_a[s], idx_b[s])
        hi = max(idx_a[s], idx_b[s])
        edges = cdf_incl[lo:hi]
        if not np.any(np.abs(edges - u[s]) <= band):
            return s
    return None


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [1, 6, 14, 20, 23])
def test_sampling_bit_exact(oracle, dtype, n):
    capi = _capi()
    s = random_state(n, 5 + n, dtype)
    if n == 6:
        s[5:9] = 0  # zero-probability runs
        s /= np.linalg.norm(s)
    rs = np.random.default_rng(n)
    u = rs.uniform(size=20000)
    u[:4] = [0.0, 1e-300, 1.0 - 2 ** -53, 0.5]
    sv = capi.DeviceStateVector.from_array(s.astype(dtype))
    got = sv.sample(u)
    exp = oracle.sample_cdf(s.astype(dtype), u)
    p = np.abs(s.astype(np.complex128)) ** 2
    cdf_incl = np.cumsum(p)
    assert boundary_ties(cdf_incl, u, got, exp) is None
    bits = sv.generate_samples(500, seed=123)
    from lightning_unification_amd.circuits import Rng
    idx = oracle.sample_cdf(s.astype(dtype), Rng(123).uniform(500))
    exp_bits = (idx[:, None] >> np.arange(n - 1, -1, -1)[None, :]) & 1
    assert np.array_equal(bits.astype(np.int64), exp_bits)
    sv.close()


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [11, 12, 14])
def test_pauli_sum_tiled_passes(oracle, dtype, n):
    """Many Pauli words: x-masks packed into tiled passes (partial tiles at
    11 and 12 qubits, full ones at 14), words too wide for a tile, Z-only
    words, and many words sharing one x-mask with different Z/Y bits (so the
    per-thread Walsh-Hadamard transform runs in full as well as for one or
    two buckets), against the oracle's term-by-term evaluation."""
    from lightning_unification_amd import capi
    from lightning_unification_amd.circuits import Obs, pauli_word
    rs = np.random.default_rng(77 + n)
    words, coeffs = [], []
    for k in range(60):
        wt = int(rs.integers(1, min(n, 13))) if k % 5 else int(rs.integers(1, 4))
        wires = rs.choice(n, wt, replace=False).tolist()
        letters = "".join("XYZ"[int(rs.integers(3))] for _ in range(wt))
        if k % 7 == 0:
            letters = "Z" * wt
        words.append(pauli_word(letters, wires))
        coeffs.append(float(rs.normal()))
    # one x-mask (X on wires 0 and n-1) with Z factors on every other wire
    # subset drawn at random, and Z-only words on random subsets
    for k in range(24):
        zs = [w for w in range(1, n - 1) if rs.uniform() < 0.5]
        words.append(pauli_word("X" + "Z" * len(zs) + "X", [0] + zs + [n - 1]))
        coeffs.append(float(rs.normal()))
        zw = [w for w in range(n) if rs.uniform() < 0.4] or [0]
        words.append(pauli_word("Z" * len(zw), zw))
        coeffs.append(float(rs.normal()))
    ham = Obs.hamiltonian(coeffs, words)
    s = random_state(n, 5, dtype)
    exp = oracle.expval(s.astype(np.complex128), ham)
    sv = capi.DeviceStateVector.from_array(s)
    got = sv.expval(ham)
    assert abs(got - exp) < (1e-12 if dtype == np.complex128 else 2e-5) * max(1.0, abs(exp))
    sv.close()
