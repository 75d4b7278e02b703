"""Seeded synthetic circuits for the density-matrix engine.

A circuit is a ``GateTable``: parallel arrays with one row per API gate,
the same record the reference keeps per gate (op, qb0..qb4, theta, phi,
lambda; src/dmsim_amdgpu_omp.hpp:108-206) plus a 32-double coefficient row
used only by the two matrix extensions (op 38 = arbitrary 2x2, op 39 =
arbitrary 4x4, the reference's C1_GATE/C2_GATE, src/dmsim_cpu_omp.hpp:971-1088).

Workloads (sizes and gate mixes as the benchmark configurations define them):
  random_mixed   n qubits, uniform over {H,X,Y,Z,S,T,RX,RY,RZ,U3,CX,CZ,SWAP}
  brickwall_su4  20 brick-wall layers of Haar SU(4) + random U3/CX tail
  fusion_mix     70 % U3, 30 % {CX, RZZ, CU1}
  hea_batch      hardware-efficient ansatz circuits + Pauli-string Hamiltonians
  structured     basis state, QFT, adder block of CCX/C3X/C4X/CSWAP, inverse QFT
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# enum OP order of the reference (src/dmsim_cpu_omp.hpp:43-49) + extensions.
OP_NAMES = [
    "U3", "U2", "U1", "CX", "ID", "X", "Y", "Z", "H", "S",
    "SDG", "T", "TDG", "RX", "RY", "RZ", "CZ", "CY", "SWAP", "CH",
    "CCX", "CSWAP", "CRX", "CRY", "CRZ", "CU1", "CU3", "RXX", "RZZ", "RCCX",
    "RC3X", "C3X", "C3SQRTX", "C4X", "R", "SRN", "W", "RYY",
    "C1", "C2",
]
OP = {name: i for i, name in enumerate(OP_NAMES)}

# number of qubit operands per op (factories, src/dmsim_cpu_omp.hpp:546-733)
ARITY = {name: 1 for name in OP_NAMES}
for _n in ("CX", "CZ", "CY", "SWAP", "CH", "CRX", "CRY", "CRZ", "CU1", "CU3",
           "RXX", "RZZ", "RYY", "C2"):
    ARITY[_n] = 2
for _n in ("CCX", "CSWAP", "RCCX"):
    ARITY[_n] = 3
for _n in ("RC3X", "C3X", "C3SQRTX"):
    ARITY[_n] = 4
ARITY["C4X"] = 5

# which of (theta, phi, lambda) each parametrised factory fills
PARAM_SLOTS = {
    "U3": ("theta", "phi", "lambda"), "U2": ("phi", "lambda"), "U1": ("lambda",),
    "RX": ("theta",), "RY": ("theta",), "RZ": ("phi",),
    "CRX": ("lambda",), "CRY": ("lambda",), "CRZ": ("lambda",), "CU1": ("lambda",),
    "CU3": ("theta", "phi", "lambda"), "RXX": ("theta",), "RZZ": ("theta",),
    "RYY": ("theta",), "R": ("theta",),
}
_SLOT = {"theta": 0, "phi": 1, "lambda": 2}


@dataclass
class GateTable:
    n_qubits: int
    ops: list = field(default_factory=list)
    qbs: list = field(default_factory=list)
    prms: list = field(default_factory=list)
    coefs: list = field(default_factory=list)

    def add(self, name: str, qubits, params=(), coef=None):
        """Append one gate; ``params`` fill the factory's slots in order."""
        op = OP[name]
        qubits = list(qubits)
        if len(qubits) != ARITY[name]:
            raise ValueError(f"{name} takes {ARITY[name]} qubits")
        qb = qubits + [0] * (5 - len(qubits))
        prm = [0.0, 0.0, 0.0]
        for slot, v in zip(PARAM_SLOTS.get(name, ()), params):
            prm[_SLOT[slot]] = float(v)
        c = np.zeros(32)
        if coef is not None:
            c[: len(coef)] = coef
        self.ops.append(op)
        self.qbs.append(qb)
        self.prms.append(prm)
        self.coefs.append(c)
        return self

    def extend(self, other: "GateTable"):
        """Append every gate of another table (same qubit count)."""
        self.ops += list(other.ops)
        self.qbs += [list(q) for q in other.qbs]
        self.prms += [list(p) for p in other.prms]
        self.coefs += [np.array(c) for c in other.coefs]
        return self

    def add_matrix(self, qubits, mat):
        """Arbitrary 1- or 2-qubit matrix; qubits[0] is the high bit of the
        matrix index (C2_GATE convention: pos2 = base + 2^qubit1)."""
        mat = np.asarray(mat, dtype=complex)
        k = len(qubits)
        if mat.shape != (1 << k, 1 << k) or k not in (1, 2):
            raise ValueError("matrix must be 2x2 (1 qubit) or 4x4 (2 qubits)")
        flat = np.empty(2 * mat.size)
        flat[0::2] = mat.real.ravel()
        flat[1::2] = mat.imag.ravel()
        name = "C1" if k == 1 else "C2"
        self.ops.append(OP[name])
        self.qbs.append(list(qubits) + [0] * (5 - k))
        self.prms.append([0.0, 0.0, 0.0])
        c = np.zeros(32)
        c[: flat.size] = flat
        self.coefs.append(c)
        return self

    def __len__(self):
        return len(self.ops)

    def arrays(self):
        """(ops int32[ng], qb uint32[ng,5], prm float64[ng,3], coef float64[ng,32])"""
        ng = len(self.ops)
        ops = np.asarray(self.ops, dtype=np.int32).reshape(ng)
        qb = np.asarray(self.qbs, dtype=np.uint32).reshape(ng, 5)
        prm = np.asarray(self.prms, dtype=np.float64).reshape(ng, 3)
        coef = np.asarray(self.coefs, dtype=np.float64).reshape(ng, 32)
        return (np.ascontiguousarray(ops), np.ascontiguousarray(qb),
                np.ascontiguousarray(prm), np.ascontiguousarray(coef))

    def take(self, indices):
        """A table of the gates at the given positions, in that order."""
        t = GateTable(self.n_qubits)
        idx = list(indices)
        t.ops = [self.ops[i] for i in idx]
        t.qbs = [self.qbs[i] for i in idx]
        t.prms = [self.prms[i] for i in idx]
        t.coefs = [self.coefs[i] for i in idx]
        return t

    def slice(self, lo, hi):
        t = GateTable(self.n_qubits)
        t.ops, t.qbs = self.ops[lo:hi], self.qbs[lo:hi]
        t.prms, t.coefs = self.prms[lo:hi], self.coefs[lo:hi]
        return t


def _distinct(rng, n, k):
    return [int(x) for x in rng.choice(n, size=k, replace=False)]


MIXED_SET = ["H", "X", "Y", "Z", "S", "T", "RX", "RY", "RZ", "U3", "CX", "CZ", "SWAP"]


def random_mixed(n: int, n_gates: int, seed: int, gate_set=MIXED_SET) -> GateTable:
    """Gates uniform over ``gate_set``; operands uniform distinct qubits;
    angles uniform in [0, 2*pi)."""
    rng = np.random.default_rng(seed)
    t = GateTable(n)
    for _ in range(n_gates):
        name = gate_set[int(rng.integers(len(gate_set)))]
        k = ARITY[name]
        qs = _distinct(rng, n, k)
        params = [float(rng.uniform(0.0, 2 * math.pi)) for _ in PARAM_SLOTS.get(name, ())]
        if name in ("C1", "C2"):
            t.add_matrix(qs, haar_unitary(rng, 1 << k))
        else:
            t.add(name, qs, params)
    return t


def haar_unitary(rng, d: int) -> np.ndarray:
    """Haar-random SU(d): complex Gaussian, Householder QR (LAPACK), phases
    fixed by diag(R)/|diag(R)|, then det normalised to 1."""
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / math.sqrt(2.0)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    q = q * ph[None, :]
    det = np.linalg.det(q)
    q = q / det ** (1.0 / d)
    return q


def brickwall_su4(n: int = 14, layers: int = 20, tail: int = 2000, seed: int = 14,
                  cx_frac: float = 0.5) -> GateTable:
    """Brick-wall Haar SU(4) layers (even: (0,1),(2,3)..; odd: (1,2),(3,4)..)
    applied through the 4x4 matrix extension, then ``tail`` gates: U3 with
    probability 1-cx_frac, else CX, on uniform distinct qubits."""
    rng = np.random.default_rng(seed)
    t = GateTable(n)
    for layer in range(layers):
        start = layer % 2
        for a in range(start, n - 1, 2):
            t.add_matrix([a, a + 1], haar_unitary(rng, 4))
    for _ in range(tail):
        if rng.uniform() < cx_frac:
            t.add("CX", _distinct(rng, n, 2))
        else:
            t.add("U3", _distinct(rng, n, 1),
                  [float(x) for x in rng.uniform(0.0, 2 * math.pi, size=3)])
    return t


def fusion_mix(n: int = 16, n_gates: int = 4000, seed: int = 16) -> GateTable:
    """70 % U3 (three uniform angles), 30 % uniform over {CX, RZZ, CU1} on
    uniform distinct pairs."""
    rng = np.random.default_rng(seed)
    t = GateTable(n)
    for _ in range(n_gates):
        if rng.uniform() < 0.7:
            t.add("U3", _distinct(rng, n, 1),
                  [float(x) for x in rng.uniform(0.0, 2 * math.pi, size=3)])
        else:
            name = ["CX", "RZZ", "CU1"][int(rng.integers(3))]
            params = [float(rng.uniform(0.0, 2 * math.pi))] if name != "CX" else []
            t.add(name, _distinct(rng, n, 2), params)
    return t


def hea_circuit(n: int, depth: int, rng) -> GateTable:
    """Hardware-efficient ansatz: per layer RY(theta) on every qubit, RZ(theta)
    on every qubit, CX ladder (0,1),(1,2),..,(n-2,n-1)."""
    t = GateTable(n)
    for _ in range(depth):
        for q in range(n):
            t.add("RY", [q], [float(rng.uniform(0.0, 2 * math.pi))])
        for q in range(n):
            t.add("RZ", [q], [float(rng.uniform(0.0, 2 * math.pi))])
        for q in range(n - 1):
            t.add("CX", [q, q + 1])
    return t


def random_paulis(n: int, count: int, rng):
    """``count`` Pauli strings (chars IXYZ, qubit 0 first) with N(0,1) weights."""
    strings = ["".join("IXYZ"[int(c)] for c in rng.integers(4, size=n)) for _ in range(count)]
    weights = rng.standard_normal(count)
    return strings, weights


def hea_batch(n_circuits: int = 256, seed: int = 4, n_range=(6, 11), depth_range=(10, 40),
              n_paulis: int = 32):
    """Batch of HEA circuits with uneven qubit counts and depths, each with its
    own random Hamiltonian.  Returns a list of (GateTable, strings, weights)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_circuits):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
        circ = hea_circuit(n, depth, rng)
        strings, weights = random_paulis(n, n_paulis, rng)
        out.append((circ, strings, weights))
    return out


def qft(t: GateTable, qubits, inverse=False):
    """Textbook QFT with H and CU1(pi/2^k), without the final swaps."""
    qs = list(qubits)
    seq = []
    for j, a in enumerate(qs):
        seq.append(("H", [a], []))
        for k, b in enumerate(qs[j + 1:], start=1):
            seq.append(("CU1", [b, a], [math.pi / (1 << k)]))
    if inverse:
        for name, q, p in reversed(seq):
            t.add(name, q, [-x for x in p])
    else:
        for name, q, p in seq:
            t.add(name, q, p)
    return t


def structured(n: int = 15, seed: int = 0) -> GateTable:
    """Random basis state |x>, QFT, a ripple-carry style block of
    CCX/C3X/C4X/CSWAP on random operand qubits, inverse QFT."""
    rng = np.random.default_rng(seed)
    t = GateTable(n)
    x = int(rng.integers(1 << n))
    for q in range(n):
        if (x >> q) & 1:
            t.add("X", [q])
    qft(t, range(n))
    kinds = ["CCX", "C3X", "C4X", "CSWAP"]
    for _ in range(2 * n):
        name = kinds[int(rng.integers(len(kinds)))]
        t.add(name, _distinct(rng, n, ARITY[name]))
    qft(t, range(n), inverse=True)
    return t


def adder_n10() -> tuple[GateTable, GateTable]:
    """The reference's 10-qubit Cuccaro adder example
    (example/adder_n10_omp.py:38-80) as two segments run back to back:
    cin=0, A=0001, B=1111; every shot reads 1000000010 (= 514)."""
    first = GateTable(10)
    for q in (1, 5, 6, 7, 8):
        first.add("X", [q])
    second = GateTable(10)

    def majority(a, b, c):
        second.add("CX", [c, b]).add("CX", [c, a]).add("CCX", [a, b, c])

    def unmaj(a, b, c):
        second.add("CCX", [a, b, c]).add("CX", [c, a]).add("CX", [a, b])

    majority(0, 5, 1)
    majority(1, 6, 2)
    majority(2, 7, 3)
    majority(3, 8, 4)
    second.add("CX", [4, 9])
    unmaj(3, 8, 4)
    unmaj(2, 7, 3)
    unmaj(1, 6, 2)
    unmaj(0, 5, 1)
    return first, second
