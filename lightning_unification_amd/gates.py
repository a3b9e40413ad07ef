"""Gate vocabulary of the Lightning-Qubit hot path.

Ids follow ``Pennylane::Gates::GateOperation`` (core/src/gates/GateOperation.hpp:27-70);
arity and parameter counts follow core/src/gates/Constant.hpp:163-269.  The ids
are the integers that cross the C-ABI in ``include/pl_amd.h``.
"""

GATE_NAMES = [
    "Identity", "PauliX", "PauliY", "PauliZ", "Hadamard", "S", "T", "PhaseShift",
    "RX", "RY", "RZ", "Rot", "CNOT", "CY", "CZ", "SWAP", "IsingXX", "IsingXY",
    "IsingYY", "IsingZZ", "ControlledPhaseShift", "CRX", "CRY", "CRZ", "CRot",
    "SingleExcitation", "SingleExcitationMinus", "SingleExcitationPlus",
    "Toffoli", "CSWAP", "DoubleExcitation", "DoubleExcitationMinus",
    "DoubleExcitationPlus", "MultiRZ",
]
GATE_ID = {name: i for i, name in enumerate(GATE_NAMES)}

# -1 = any number of wires (MultiRZ)
GATE_WIRES = [1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2,
              2, 2, 2, 2, 2, 3, 3, 4, 4, 4, -1]
GATE_NPARAMS = [0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 3, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1,
                1, 1, 3, 1, 1, 1, 0, 0, 1, 1, 1, 1]

# Gates with a generator (GeneratorOperation, GateOperation.hpp:75-98).
GENERATOR_GATES = [
    "PhaseShift", "RX", "RY", "RZ", "IsingXX", "IsingXY", "IsingYY", "IsingZZ",
    "CRX", "CRY", "CRZ", "ControlledPhaseShift", "SingleExcitation",
    "SingleExcitationMinus", "SingleExcitationPlus", "DoubleExcitation",
    "DoubleExcitationMinus", "DoubleExcitationPlus", "MultiRZ",
]

# Op id the adjoint sweep skips (AdjointJacobianLQubit.hpp:287-290).
STATE_PREP_ID = -1
STATE_PREP_NAMES = ("QubitStateVector", "BasisState")


def gate_id(name):
    if name in STATE_PREP_NAMES:
        return STATE_PREP_ID
    return GATE_ID[name]


def touched_fraction(name):
    """Fraction f_g of amplitudes a gate can change (its algorithmic bytes
    are 2 * f_g * 2^n * sizeof(complex))."""
    quarter = {"CZ", "ControlledPhaseShift", "Toffoli", "CSWAP"}
    half = {"PauliZ", "S", "T", "PhaseShift", "CNOT", "CY", "SWAP", "CRX", "CRY",
            "CRZ", "CRot", "SingleExcitation", "IsingXY"}
    if name == "Identity":
        return 0.0
    if name in quarter:
        return 0.25
    if name in half:
        return 0.5
    if name == "DoubleExcitation":
        return 0.125
    return 1.0
