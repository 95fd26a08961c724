"""Case lists of the reference's tests (qibojit v0.1.17), shared by the CPU
oracle tests and the GPU parity tests."""

# tests/test_gates.py:18-31
APPLY_GATE = [
    (5, 4, []),
    (4, 2, []),
    (3, 0, []),
    (8, 5, []),
    (3, 0, [1, 2]),
    (4, 3, [0, 1, 2]),
    (5, 3, [1]),
    (5, 2, [1, 4]),
    (6, 3, [0, 2, 5]),
    (6, 3, [0, 2, 4, 5]),
]
# tests/test_gates.py:52
ONE_QUBIT_BASE = [(4, 1), (6, 5)]
# tests/test_gates.py:69-81
PAULI = [
    (3, 0, []),
    (4, 3, []),
    (5, 2, []),
    (3, 1, []),
    (3, 0, [1]),
    (4, 3, [0, 1]),
    (5, 2, [1, 3, 4]),
]
# tests/test_gates.py:96-99
ZPOW = [(3, 0, []), (3, 2, [1]), (3, 2, [0, 1]), (6, 1, [0, 2, 4])]
# tests/test_gates.py:112-125
TWO_QUBIT = [
    (5, [3, 4], []),
    (4, [2, 0], []),
    (2, [0, 1], []),
    (8, [6, 3], []),
    (3, [0, 1], [2]),
    (4, [1, 3], [0]),
    (5, [2, 3], [1, 4]),
    (5, [3, 1], [0, 2]),
    (6, [2, 5], [0, 1, 3, 4]),
]
# tests/test_gates.py:144
TWO_QUBIT_BASE = [(5, [3, 4]), (4, [2, 0])]
# tests/test_gates.py:163-176
CONTROL_SINGLE = [
    (2, [0, 1]),
    (3, [0, 2]),
    (4, [1, 3]),
    (3, [1, 2]),
    (4, [0, 2]),
    (4, [2, 3]),
    (5, [3, 4]),
    (6, [1, 4]),
]
CONTROL_SINGLE_GATES = ["CY", "CH", "CSX", "CSXDG"]
# tests/test_gates.py:188-194
CCZ = [(3, 2, [0, 1]), (4, 2, [1, 3])]
# tests/test_gates.py:206-216
DEUTSCH = [(3, [0, 1, 2]), (4, [1, 2, 3]), (4, [0, 2, 3]), (4, [2, 3, 1]), (5, [3, 4, 2]), (6, [1, 2, 4])]
# tests/test_gates.py:229-241
SWAP = [
    (2, [0, 1], []),
    (3, [0, 2], []),
    (4, [1, 3], []),
    (3, [1, 2], [0]),
    (4, [0, 2], [1]),
    (4, [2, 3], [0]),
    (5, [3, 4], [1, 2]),
    (6, [1, 4], [0, 2, 5]),
]
# tests/test_gates.py:253-267
FSIM = [
    (3, [0, 1], []),
    (4, [2, 0], []),
    (3, [1, 2], [0]),
    (4, [0, 1], [2]),
    (5, [0, 1], [2]),
    (5, [3, 4], [2]),
    (4, [0, 3], [1]),
    (4, [3, 2], [0]),
    (5, [1, 4], [2]),
    (6, [1, 3], [0, 4]),
    (6, [5, 0], [1, 2, 3]),
]
FSIM_PHI = 0.4321
# tests/test_gates.py:282
FANOUT = [2, 3, 4]
# tests/test_gates.py:295-311
MULTI_QUBIT = [
    (3, [0, 1, 2], []),
    (4, [2, 1, 3], []),
    (5, [0, 2, 3], []),
    (8, [2, 6, 3], []),
    (5, [0, 2, 3, 4], []),
    (8, [0, 4, 2, 5, 7], []),
    (7, [0, 2, 4, 3, 6, 5], []),
    (8, [0, 4, 2, 3, 5, 7, 1], []),
    (4, [2, 1, 3], [0]),
    (5, [0, 2, 3], [1]),
    (8, [2, 6, 3], [4, 7]),
    (5, [0, 2, 3, 4], [1]),
    (8, [0, 4, 2, 5, 7], [1, 3]),
]
# tests/test_gates.py:332-343
MULTI_QUBIT_LARGE = [
    (10, [0, 4, 2, 5, 9], [1, 3, 7, 8]),
    (20, [10, 8, 13], []),
    (20, [11, 19, 13, 4], []),
    (20, [12, 14, 2, 5, 17], []),
    (20, [12, 17, 3], [10]),
    (20, [19, 6, 13], [14, 7]),
    (20, [0, 19, 3, 14], [1]),
    (20, [0, 4, 18, 5, 17], [10, 3]),
]
# tests/test_gates.py:349
MULTI_QUBIT_BASE = [(5, [2, 3, 4]), (4, [2, 0, 1])]
# tests/test_gates.py:372-380
PARAMETRIZED_ON_CIRCUIT = [
    ("CRX", {"theta": 0.1}),
    ("CRY", {"theta": 0.1}),
    ("CRZ", {"theta": 0.1}),
    ("CU1", {"theta": 0.1}),
    ("CU2", {"phi": 0.1, "lam": 0.2}),
    ("CU3", {"theta": 0.1, "phi": 0.2, "lam": 0.3}),
    ("fSim", {"theta": 0.1, "phi": 0.2}),
]
# tests/test_ops.py:52-63
COLLAPSE = [
    (2, [0], [1]),
    (2, [1], [0]),
    (3, [1], [1]),
    (4, [1, 3], [1, 0]),
    (5, [1, 2, 4], [0, 1, 1]),
    (15, [4, 7], [0, 0]),
    (16, [8, 12, 15], [1, 0, 1]),
]


# ------------------------------------------------------------------------
# Cases beyond the reference's lists: the kernels the 30-qubit gate sweep
# runs, at their strides.


def sweep_placement(n):
    """The sweep's gates with its placement, as (label, gate name, targets,
    controls, parameters): H, X, Z, RZ, CNOT, SWAP, fSim, CU3 and a 3-qubit
    dense gate on target qubits 0, n // 2 and n - 1 (bit strides 2^(n-1),
    2^(n-1-n//2), 1); two-qubit gates on (t, t + 1), or (n - 2, n - 1) at
    t = n - 1; the control of CNOT / CU3 is the neighbouring qubit; the
    3-qubit gate on (t, t + 1, t + 2), or the three last qubits."""
    out = []
    theta = 0.731
    for t in (0, n // 2, n - 1):
        p2 = [t, t + 1] if t < n - 1 else [t - 1, t]
        ctrl = t + 1 if t < n - 1 else t - 1
        h3 = [t, t + 1, t + 2] if t + 2 < n else [n - 3, n - 2, n - 1]
        out += [
            (f"H({t})", "H", [t], [], ()),
            (f"X({t})", "X", [t], [], ()),
            (f"Z({t})", "Z", [t], [], ()),
            (f"RZ({t})", "RZ", [t], [], (theta,)),
            (f"CNOT({ctrl},{t})", "CNOT", [t], [ctrl], ()),
            (f"SWAP{tuple(p2)}", "SWAP", p2, [], ()),
            (f"fSim{tuple(p2)}", "fSim", p2, [], (0.3, 0.4)),
            (f"CU3({ctrl},{t})", "CU3", [t], [ctrl], (0.1, 0.2, 0.3)),
            (f"Haar3{tuple(h3)}", "Unitary", h3, [], ()),
        ]
    return out


# uncontrolled 3-target gates on the three lowest bits (qubits n-3..n-1) in
# several target orders: the LDS-staged chunk kernel takes them from 11 qubits
LOW3 = [
    (11, [8, 9, 10]),
    (11, [10, 8, 9]),
    (16, [15, 13, 14]),
    (16, [13, 14, 15]),
    (22, [21, 20, 19]),
    (22, [20, 19, 21]),
]

# one-target gates whose target bit (n - 1 - q) is >= 6, so the strided
# kernel runs instead of the in-wave exchange; complex64 moves two
# amplitudes per lane when bit 0 (qubit n - 1) is free, one otherwise
HIGH_BIT_ONE = [
    (12, 0, []),  # bit 11
    (12, 5, []),  # bit 6
    (12, 3, [11]),  # bit 8, control on bit 0
    (12, 2, [11, 7]),  # bit 9, controls on bits 0 and 4
    (12, 5, [0]),  # bit 6, control on bit 11
    (12, 4, [8]),  # bit 7, control inside the wave's 6 bits
    (16, 1, [15, 3]),  # bit 14
    (22, 0, []),  # bit 21
    (22, 10, [21]),  # bit 11, control on bit 0
    (22, 15, [2, 9]),  # bit 6
    (24, 0, []),  # bit 23 (complex64: one amplitude per lane)
    (24, 1, [0]),  # bit 22, control on bit 23
    (24, 13, [23]),  # bit 10 (complex64: one amplitude per lane), control on bit 0
]

# two-target gates at strided and in-wave target bits
HIGH_BIT_TWO = [
    (12, [0, 5], []),  # bits 11, 6
    (12, [5, 0], []),
    (12, [3, 9], []),  # bits 8, 2
    (12, [11, 2], []),  # bits 0, 9
    (12, [1, 4], [11]),  # bits 10, 7, control on bit 0
    (12, [6, 2], [0, 10]),  # bits 5, 9, controls on bits 11, 1
    (16, [14, 0], [7]),  # bits 1, 15
    (22, [0, 21], []),  # bits 21, 0
    (22, [10, 11], [21, 3]),  # bits 11, 10
    # both targets inside the wave's 6 bits (the in-wave kernels, including
    # the one-shuffle form for SWAP / fSim), controls inside and above it
    (12, [10, 11], [8]),  # bits 1, 0, control on bit 3
    (12, [11, 9], [0]),  # bits 0, 2, control on bit 11
]
