"""Seeded synthetic workloads and the flat encodings that cross the C-ABI.

RNG: SplitMix64 words (fully specified, vectorisable), u = (w >> 11) * 2^-53 in
[0, 1), angles 2*pi*u, integers by Lemire's multiply-shift, N(0,1) by
Box-Muller.  Every consumer (GPU path, oracle, fixture scripts) draws the same
numbers from the same seed.

Encodings (see include/pl_amd.h):
  tape        -> ids[], nwires[], wires[], inverse[], nparams[], params[]
  observables -> int32 program, float64 coeffs, complex matrices pool
"""
import math

import numpy as np

from .gates import GATE_ID, GATE_NPARAMS, GATE_WIRES, gate_id

_GOLDEN = 0x9E3779B97F4A7C15
_M64 = (1 << 64) - 1


def _mix(z):
    """SplitMix64 finaliser on a uint64 numpy array."""
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


class Rng:
    """SplitMix64 stream.  ``Rng(seed)`` and ``Rng(seed, stream=j)`` (a
    sub-stream keyed by SplitMix64(seed, j)) are fully deterministic."""

    def __init__(self, seed, stream=None):
        s = seed & _M64
        if stream is not None:
            s = int(_mix(np.array([(s + (stream + 1) * _GOLDEN) & _M64], np.uint64))[0])
        self.state = s

    def words(self, k):
        k = int(k)
        with np.errstate(over="ignore"):
            ctr = (np.arange(1, k + 1, dtype=np.uint64) * np.uint64(_GOLDEN)
                   + np.uint64(self.state))
            out = _mix(ctr)
        self.state = (self.state + k * _GOLDEN) & _M64
        return out

    def uniform(self, k):
        return (self.words(k) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def angle(self):
        return 2.0 * math.pi * float(self.uniform(1)[0])

    def integer(self, bound):
        """Uniform integer in [0, bound) (Lemire multiply-shift)."""
        w = int(self.words(1)[0])
        return (w * bound) >> 64

    def normal(self):
        u1, u2 = self.uniform(2)
        return math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2.0 * math.pi * u2)

    def shuffle(self, seq):
        seq = list(seq)
        for i in range(len(seq) - 1, 0, -1):
            j = self.integer(i + 1)
            seq[i], seq[j] = seq[j], seq[i]
        return seq


class Tape:
    """An ordered list of named operations (the OpsData of
    core/src/algorithms/JacobianData.hpp:40-200)."""

    def __init__(self):
        self.ops = []  # (name, wires, inverse, params)

    def add(self, name, wires, params=(), inverse=False):
        self.ops.append((name, list(wires), bool(inverse), [float(p) for p in params]))
        return self

    def __len__(self):
        return len(self.ops)

    def num_params(self):
        return sum(1 for op in self.ops if op[3])

    def encode(self):
        ids = np.array([gate_id(o[0]) for o in self.ops], np.int32)
        nw = np.array([len(o[1]) for o in self.ops], np.int32)
        wires = np.array([w for o in self.ops for w in o[1]], np.int32)
        inv = np.array([1 if o[2] else 0 for o in self.ops], np.uint8)
        npar = np.array([len(o[3]) for o in self.ops], np.int32)
        par = np.array([p for o in self.ops for p in o[3]], np.float64)
        return ids, nw, wires, inv, npar, par


class Obs:
    """Observable tree mirroring NamedObs / HermitianObs / TensorProdObs /
    Hamiltonian (core/src/observables/Observables.hpp:98-417)."""

    def __init__(self, kind, name=None, wires=(), matrix=None, factors=(), coeffs=(), terms=()):
        self.kind = kind
        self.name = name
        self.wires = list(wires)
        self.matrix = None if matrix is None else np.asarray(matrix, np.complex128)
        self.factors = list(factors)
        self.coeffs = [float(c) for c in coeffs]
        self.terms = list(terms)

    @staticmethod
    def named(name, wires):
        return Obs("named", name=name, wires=wires)

    @staticmethod
    def hermitian(matrix, wires):
        return Obs("herm", matrix=matrix, wires=wires)

    @staticmethod
    def tensor(factors):
        return Obs("tensor", factors=factors)

    @staticmethod
    def hamiltonian(coeffs, terms):
        return Obs("ham", coeffs=coeffs, terms=terms)

    def _factors(self):
        if self.kind in ("named", "herm"):
            return [self]
        if self.kind == "tensor":
            return self.factors
        raise ValueError("Hamiltonian terms must be Named, Hermitian or TensorProd observables")

    @staticmethod
    def _emit_term(term, prog, mats):
        fs = term._factors()
        prog.append(len(fs))
        for f in fs:
            if f.kind == "named":
                prog += [0, GATE_ID[f.name], len(f.wires)] + f.wires
            else:
                prog += [1, 0, len(f.wires)] + f.wires
                mats.append(f.matrix.reshape(-1))

    def emit(self, prog, coeffs, mats):
        if self.kind == "ham":
            prog += [1, len(self.terms)]
            for c, t in zip(self.coeffs, self.terms):
                coeffs.append(c)
                Obs._emit_term(t, prog, mats)
        else:
            prog += [0, 1]
            Obs._emit_term(self, prog, mats)


def encode_observables(obs_list):
    """-> (int32 program, float64 coeffs, complex128 matrix pool)."""
    prog, coeffs, mats = [len(obs_list)], [], []
    for ob in obs_list:
        ob.emit(prog, coeffs, mats)
    pool = np.concatenate(mats) if mats else np.zeros(1, np.complex128)
    return (np.array(prog, np.int32), np.array(coeffs if coeffs else [0.0], np.float64),
            np.ascontiguousarray(pool))


def pauli_word(word, wires):
    """Tensor product of Pauli letters, e.g. pauli_word('XZ', [3, 7])."""
    names = {"X": "PauliX", "Y": "PauliY", "Z": "PauliZ"}
    fs = [Obs.named(names[c], [w]) for c, w in zip(word, wires)]
    return fs[0] if len(fs) == 1 else Obs.tensor(fs)


# ----------------------------------------------------------------------------
# Workloads
# ----------------------------------------------------------------------------

def plain_layers(n=20, layers=10, seed=1001):
    """Plain method: RX, RY, RZ on every wire (angles in that order), then a
    CNOT ladder i -> i+1; <Z_w> for every wire; every parameter trainable."""
    rng = Rng(seed)
    tape = Tape()
    for _ in range(layers):
        for w in range(n):
            for g in ("RX", "RY", "RZ"):
                tape.add(g, [w], [rng.angle()])
        for i in range(n - 1):
            tape.add("CNOT", [i, i + 1])
    obs = [Obs.named("PauliZ", [w]) for w in range(n)]
    return tape, obs, list(range(tape.num_params()))


LARGE_GATESET = ["Hadamard", "T", "RX", "RY", "RZ", "Rot", "CRX", "CRZ", "IsingXX",
                 "IsingZZ", "SWAP", "Toffoli"]


def random_layers(n=30, layers=50, seed=1002, gateset=LARGE_GATESET):
    """Single large state vector workload.  Each layer shuffles the wires and
    walks the list, drawing gates uniformly from `gateset` restricted to
    arity <= wires left, so each layer touches every wire at most once."""
    rng = Rng(seed)
    tape = Tape()
    for _ in range(layers):
        order = rng.shuffle(range(n))
        pos = 0
        while pos < n:
            left = n - pos
            allowed = [g for g in gateset if GATE_WIRES[GATE_ID[g]] <= left]
            g = allowed[rng.integer(len(allowed))]
            k = GATE_WIRES[GATE_ID[g]]
            wires = order[pos:pos + k]
            pos += k
            params = [rng.angle() for _ in range(GATE_NPARAMS[GATE_ID[g]])]
            tape.add(g, wires, params)
    return tape


def random_hamiltonian(n=30, terms=100, seed=1002, stream=7):
    """Sum of random Pauli strings of weight 1..4 with N(0,1) coefficients."""
    rng = Rng(seed, stream)
    coeffs, words = [], []
    for _ in range(terms):
        wt = 1 + rng.integer(4)
        wires = rng.shuffle(range(n))[:wt]
        letters = "".join("XYZ"[rng.integer(3)] for _ in range(wt))
        coeffs.append(rng.normal())
        words.append(pauli_word(letters, wires))
    return Obs.hamiltonian(coeffs, words)


def sample_uniforms(m, seed=1002, stream=9):
    return Rng(seed, stream).uniform(m)


ENTANGLERS = {"none": 0, "ladder": 1, "ring": 2}


def layered_tape(n, depth, pattern, entangler, params):
    """The layered-ansatz family in tape order: per layer, for each wire the
    rotations of `pattern` (e.g. "YZ" = RY then RZ) with consecutive angles
    from `params`, then CNOT(i, i+1) for a ladder / CNOT(i, (i+1) mod n) for
    a ring."""
    tape = Tape()
    k = 0
    for _ in range(depth):
        for w in range(n):
            for a in pattern:
                tape.add("R" + a, [w], [params[k]])
                k += 1
        if entangler != "none" and n > 1:
            last = n if entangler == "ring" else n - 1
            for i in range(last):
                tape.add("CNOT", [i, (i + 1) % n])
    return tape


def hea(n, depth, rng):
    """Hardware-efficient ansatz: RY, RZ on every wire, then a CNOT ring."""
    tape = Tape()
    for _ in range(depth):
        for w in range(n):
            tape.add("RY", [w], [rng.angle()])
            tape.add("RZ", [w], [rng.angle()])
        if n > 1:
            for i in range(n):
                tape.add("CNOT", [i, (i + 1) % n])
    return tape


def tfim(n, J=1.0, h=1.0):
    """Periodic transverse-field Ising model -J sum Z_i Z_{i+1} - h sum X_i."""
    coeffs, terms = [], []
    for i in range(n):
        coeffs.append(-J)
        terms.append(Obs.tensor([Obs.named("PauliZ", [i]), Obs.named("PauliZ", [(i + 1) % n])]))
    for i in range(n):
        coeffs.append(-h)
        terms.append(Obs.named("PauliX", [i]))
    return Obs.hamiltonian(coeffs, terms)


def batch_circuits(count=4096, seed=1004, nmin=8, nmax=20, dmin=4, dmax=12):
    """Batching workload: `count` HEA circuits with uneven qubit counts and
    depths, each with its own TFIM (J=1, h ~ U[0.5, 1.5])."""
    out = []
    for j in range(count):
        rng = Rng(seed, j)
        n = nmin + rng.integer(nmax - nmin + 1)
        d = dmin + rng.integer(dmax - dmin + 1)
        h = 0.5 + float(rng.uniform(1)[0])
        out.append((n, hea(n, d, rng), tfim(n, 1.0, h)))
    return out


def batch_shapes(count=4096, seed=1004, nmin=8, nmax=20, dmin=4, dmax=12):
    """The same workload as batch_circuits without building tapes:
    -> (qubits[], depths[], fields h[], flat angles in tape order).  Angles
    are the draws hea() makes (RY then RZ per wire per layer)."""
    ns, ds, hs, ps = [], [], [], []
    for j in range(count):
        rng = Rng(seed, j)
        n = nmin + rng.integer(nmax - nmin + 1)
        d = dmin + rng.integer(dmax - dmin + 1)
        hs.append(0.5 + float(rng.uniform(1)[0]))
        ns.append(n)
        ds.append(d)
        ps.append(2.0 * math.pi * rng.uniform(2 * n * d))
    return ns, ds, hs, np.concatenate(ps)
