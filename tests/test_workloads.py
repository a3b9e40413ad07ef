"""Seeded workload generators: the tape-free batch description must draw the
same circuits and angles as the tape builder, and layered_tape must
reproduce the HEA / plain-layer families."""
import numpy as np

from lightning_unification_amd.circuits import (Rng, batch_circuits, batch_shapes, hea,
                                                layered_tape, plain_layers)


def test_batch_shapes_match_tapes():
    circ = batch_circuits(40, nmax=14)
    ns, ds, hs, ps = batch_shapes(40, nmax=14)
    k = 0
    for (n, tape, ham), nn, d, h in zip(circ, ns, ds, hs):
        assert n == nn
        ang = [op[3][0] for op in tape.ops if op[3]]
        assert len(ang) == 2 * n * d
        assert np.array_equal(ang, ps[k:k + len(ang)])
        assert ham.coeffs[-1] == -h
        rebuilt = layered_tape(n, d, "YZ", "ring", ps[k:k + len(ang)])
        assert rebuilt.ops == tape.ops
        k += len(ang)
    assert k == ps.size


def test_layered_tape_plain_family():
    tape, _, _ = plain_layers(5, 3, seed=4)
    ang = [op[3][0] for op in tape.ops if op[3]]
    assert layered_tape(5, 3, "XYZ", "ladder", ang).ops == tape.ops


def test_hea_ring_small():
    t = hea(2, 1, Rng(3))
    assert [op[0] for op in t.ops] == ["RY", "RZ", "RY", "RZ", "CNOT", "CNOT"]
    assert [op[1] for op in t.ops[4:]] == [[0, 1], [1, 0]]
