"""Gate fusion plan (pl_fuse_tape, host only): applying the fused blocks in
order must reproduce the oracle's unfused run of the same tape, for every
fusion width, with and without padding to 5-bit blocks; blocks respect the
width and are unitary; every operation (state preparation aside) is applied
exactly once, and an operation applied before an earlier one of the tape
shares no wire with it (they commute)."""
import numpy as np
import pytest

from conftest import random_state
from lightning_unification_amd.circuits import Tape, random_layers


def apply_block(state, n, bits, U):
    """state <- U on bit positions `bits` (block-local index bit m = bits[m])."""
    k = len(bits)
    idx = np.arange(1 << n)
    rest = idx.copy()
    for b in bits:
        rest &= ~(1 << b)
    local = np.zeros_like(idx)
    for m, b in enumerate(bits):
        local |= ((idx >> b) & 1) << m
    out = np.zeros_like(state)
    for jp in range(1 << k):
        src = rest.copy()
        for m, b in enumerate(bits):
            src |= ((jp >> m) & 1) << b
        out += U[local, jp] * state[src]
    return out


def mixed_tape(n, seed):
    rs = np.random.default_rng(seed)
    t = random_layers(n, 4, seed=seed)
    t.add("MultiRZ", list(range(min(n, 6))), [0.3])       # wider than 5: stays unfused
    t.add("CSWAP", [0, n - 1, 1], inverse=True)
    t.add("DoubleExcitation", [1, 2, n - 2, n - 1], [rs.uniform(0, 6)])
    t.add("CRot", [2, 0], rs.uniform(0, 6, 3))
    t.ops.append(("BasisState", [0], False, []))            # state preparation: skipped
    return t


@pytest.mark.parametrize("max_wires", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("pad_to", [0, 5])
def test_fused_blocks_reproduce_tape(oracle, max_wires, pad_to):
    from lightning_unification_amd import capi
    from lightning_unification_amd.gates import GATE_ID
    n = 7
    tape = mixed_tape(n, 11 + max_wires)
    psi = random_state(n, 5)
    exp = oracle.apply_tape(psi.copy(), without_prep(tape))
    blocks = capi.fuse_tape(n, tape, max_wires, pad_to)
    got = psi.copy()
    applied = []
    for ops, bits, U in blocks:
        applied += ops
        if bits is None:
            assert len(ops) == 1
            t1 = Tape()
            t1.ops = [tape.ops[ops[0]]]
            got = oracle.apply_tape(got, t1)
            continue
        assert len(bits) == max(pad_to, len(bits)) and len(bits) <= max(max_wires, pad_to)
        assert bits == sorted(bits)
        assert np.allclose(U @ U.conj().T, np.eye(U.shape[0]), atol=1e-13)
        got = apply_block(got, n, bits, U)
    gates = [i for i, op in enumerate(tape.ops) if op[0] in GATE_ID]
    assert sorted(applied) == gates
    check_overtaking(tape, applied)
    assert np.max(np.abs(got - exp)) < 1e-12


def check_overtaking(tape, applied):
    """Operations applied out of tape order act on disjoint wires."""
    pos = {i: k for k, i in enumerate(applied)}
    for a in applied:
        for b in applied:
            if a < b and pos[a] > pos[b]:
                assert not set(tape.ops[a][1]) & set(tape.ops[b][1]), (a, b)


def without_prep(tape):
    """The tape minus its state-preparation ops (the oracle restates
    applyOperation, which has none; applyOperations skips them)."""
    from lightning_unification_amd.gates import GATE_ID
    t = Tape()
    t.ops = [op for op in tape.ops if op[0] in GATE_ID]
    return t


def test_fusion_counts_and_errors():
    from lightning_unification_amd import capi
    tape = random_layers(30, 50, seed=1002)
    blocks = capi.fuse_tape(30, tape, 5)
    applied = [i for b in blocks for i in b[0]]
    assert sorted(applied) == list(range(len(tape)))
    check_overtaking(tape, applied)
    # moving operations past commuting ones packs about six per 5-wire block
    # (consecutive operations alone give about three)
    assert len(blocks) < len(tape) / 5
    with pytest.raises(ValueError):
        capi.fuse_tape(30, tape, 6)
    with pytest.raises(ValueError):
        capi.fuse_tape(3, Tape().add("RX", [3], [0.1]), 5)
