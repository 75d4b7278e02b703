"""Pins the oracle (oracle/dmsim_oracle.c) before anything is checked against it.

1. Golden vectors written by tests/golden/gen_golden.py from the reference's
   own CPU path (src/dmsim_cpu_omp.hpp compiled in place): bit-exact.
2. The live reference build (oracle/_ref), where present: bit-exact on fresh
   random circuits, including multi-threaded reference runs.
3. What the mathematics fixes, independently of the reference: textbook gate
   matrices (up to global phase) for the gates that are textbook, unitarity,
   rho = M rho0 M^dagger by dense brute force, trace/Hermiticity, closed-form
   expectations (RX, the deuteron ansatz table of
   xacc/tests/DmSimCpuAcceleratorTester.cpp:36-40, Bell), measurement
   statistics, and the reference's documented deviations (C4X, C3SQRTX, R, SRN).
"""
import math
import os

import numpy as np
import pytest

import workloads as W
from conftest import GOLDEN

S2I = 0.70710678118654752440


def load(name):
    return np.load(os.path.join(GOLDEN, name))


def table_from(d, prefix=""):
    return (np.ascontiguousarray(d[prefix + "ops"]), np.ascontiguousarray(d[prefix + "qb"]),
            np.ascontiguousarray(d[prefix + "prm"]), np.ascontiguousarray(d[prefix + "coef"]))


def oracle_gate_matrix(oracle, k, op, qubits, params=(0.0, 0.0, 0.0), coef=None):
    """M with row r = M e_r as the forward pass applies it (identity rows in)."""
    dim = 1 << k
    re = np.eye(dim).ravel().copy()
    im = np.zeros(dim * dim)
    qb = np.zeros((1, 5), dtype=np.uint32)
    qb[0, : len(qubits)] = qubits
    c = np.zeros((1, 32)) if coef is None else np.asarray(coef, dtype=np.float64).reshape(1, 32)
    t = (np.array([op], dtype=np.int32), qb, np.array([params], dtype=np.float64), c)
    oracle.forward(k, dim, re, im, t)
    return (re + 1j * im).reshape(dim, dim).T


def equal_up_to_phase(a, b, tol=1e-12):
    idx = np.unravel_index(np.argmax(np.abs(b)), b.shape)
    ph = a[idx] / b[idx]
    return abs(abs(ph) - 1) < tol and np.allclose(a, ph * b, atol=tol)


# ------------------------------------------------------------------ golden ----

def test_golden_gate_matrices_bitwise(oracle):
    d = load("gate_matrices.npz")
    prm = tuple(d["params"])
    for name in W.OP_NAMES:
        k = W.ARITY[name]
        for tag in ("fwd", "rev"):
            coef = np.zeros(32)
            if name in ("C1", "C2"):
                u = d[f"{name}_{tag}_input"]
                coef[0: 2 * u.size: 2] = u.real.ravel()
                coef[1: 2 * u.size: 2] = u.imag.ravel()
            m = oracle_gate_matrix(oracle, k, W.OP[name], list(d[f"{name}_{tag}_qubits"]),
                                   prm, coef)
            assert np.array_equal(m, d[f"{name}_{tag}"]), (name, tag)


def test_golden_circuits_n6_bitwise(oracle):
    d = load("circuits_n6.npz")
    for s in range(4):
        re, im = oracle.sim(6, table_from(d, f"c{s}_"))
        assert np.array_equal(re, d[f"c{s}_re"]) and np.array_equal(im, d[f"c{s}_im"]), s


def test_golden_mixed_n8_bitwise_and_shots(oracle):
    import hashlib
    d = load("mixed_n8.npz")
    for s in range(1, 9):
        t = table_from(d, f"s{s}_")
        re, im = oracle.sim(8, t)
        assert hashlib.sha256(re.astype("<f8").tobytes()).hexdigest() == str(d[f"s{s}_sha_re"])
        assert hashlib.sha256(im.astype("<f8").tobytes()).hexdigest() == str(d[f"s{s}_sha_im"])
        shots, _ = oracle.measure(8, re, 64, int(d[f"s{s}_shot_seed"]))
        assert np.array_equal(shots, d[f"s{s}_shots"]), s
    re, _ = oracle.sim(8, table_from(d, "s1_"))
    assert np.array_equal(re, d["s1_re"])


def test_rowwise_forward_against_golden(oracle):
    """orc_sim_rowwise (the forward passes row by row) reproduces the golden
    vectors and the adder's shots bit for bit."""
    import hashlib
    d = load("circuits_n6.npz")
    for s in range(4):
        re, im = oracle.sim(6, table_from(d, f"c{s}_"), rowwise=True)
        assert np.array_equal(re, d[f"c{s}_re"]) and np.array_equal(im, d[f"c{s}_im"]), s
    d = load("mixed_n8.npz")
    for s in range(1, 9):
        re, im = oracle.sim(8, table_from(d, f"s{s}_"), rowwise=True)
        assert hashlib.sha256(re.astype("<f8").tobytes()).hexdigest() == str(d[f"s{s}_sha_re"])
        assert hashlib.sha256(im.astype("<f8").tobytes()).hexdigest() == str(d[f"s{s}_sha_im"])
    d = load("adder_n10.npz")
    re, _ = oracle.sim(10, table_from(d), rowwise=True)
    assert hashlib.sha256(re.astype("<f8").tobytes()).hexdigest() == str(d["sha_re"])


def test_rowwise_forward_equals_gatewise(oracle):
    """Loop order only: every op (with C1/C2 matrices), continued runs, and
    several thread counts give identical bits."""
    for n, seed, threads_set in ((5, 1, (1, 3, 8)), (9, 2, (8,)), (10, 3, (5,))):
        gate_set = [g for g in W.OP_NAMES if W.ARITY[g] <= n]
        t = W.random_mixed(n, 120, 700 + seed, gate_set=gate_set).arrays()
        t2 = W.brickwall_su4(n=n, layers=2, tail=30, seed=seed).arrays()
        for threads in threads_set:
            oracle.set_threads(threads)
            a = oracle.sim(n, t2, oracle.sim(n, t))
            b = oracle.sim(n, t2, oracle.sim(n, t, rowwise=True), rowwise=True)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (n, threads)
    oracle.set_threads(1)


def test_measure_diag_equals_measure(oracle):
    """orc_measure_diag on rho's diagonal draws exactly what orc_measure draws
    on rho (and thereby the reference's golden shots)."""
    d = load("mixed_n8.npz")
    for s in range(1, 9):
        re, _ = oracle.sim(8, table_from(d, f"s{s}_"))
        seed = int(d[f"s{s}_shot_seed"])
        diag = re.reshape(256, 256).diagonal().copy()
        a, ta = oracle.measure(8, re, 64, seed)
        b, tb = oracle.measure_diag(8, diag, 64, seed)
        assert np.array_equal(a, b) and np.array_equal(b, d[f"s{s}_shots"]) and ta == tb, s
    d = load("adder_n10.npz")
    re, _ = oracle.sim(10, table_from(d))
    shots, _ = oracle.measure_diag(10, re.reshape(1024, 1024).diagonal().copy(), 10,
                                   int(d["shot_seed"]))
    assert list(shots) == list(d["shots"])
    # a diagonal with its mass moved must change the draw (the scan reads diag)
    diag = np.zeros(256)
    diag[17] = 1.0
    assert set(oracle.measure_diag(8, diag, 32, 5)[0]) == {17}


def test_golden_tables_regenerate():
    """The workload generator still produces the circuits the fixtures hold."""
    d = load("mixed_n8.npz")
    for s in (1, 5):
        ops, qb, prm, coef = W.random_mixed(8, 1000, s).arrays()
        assert np.array_equal(ops, d[f"s{s}_ops"]) and np.array_equal(qb, d[f"s{s}_qb"])
        assert np.array_equal(prm, d[f"s{s}_prm"])


def test_golden_adder_known_answer(oracle):
    """README:228-247: cin=0, A=0001, B=1111 -> every shot reads 1000000010."""
    d = load("adder_n10.npz")
    re, im = oracle.sim(10, table_from(d))
    import hashlib
    assert hashlib.sha256(re.astype("<f8").tobytes()).hexdigest() == str(d["sha_re"])
    shots, total = oracle.measure(10, re, 10, int(d["shot_seed"]))
    assert list(shots) == [514] * 10 == list(d["shots"])
    assert abs(total - 1) < 1e-12


# ------------------------------------------------------- live reference ----

@pytest.mark.parametrize("n", [2, 3, 5, 7])
def test_live_reference_bitwise(oracle, reference, n):
    gate_set = [g for g in W.OP_NAMES if W.ARITY[g] <= n]
    for seed in range(3):
        t = W.random_mixed(n, 120, 9000 + 10 * n + seed, gate_set=gate_set).arrays()
        re, im = oracle.sim(n, t)
        rr, ri, _ = reference.sim(n, t)
        assert np.array_equal(re, rr) and np.array_equal(im, ri), seed


def test_live_reference_threads_and_shots(oracle, reference):
    t = W.random_mixed(6, 400, 77).arrays()
    re, im = oracle.sim(6, t)
    for threads in (1, 2, 8):
        rr, ri, shots = reference.sim(6, t, threads=threads, seed=4242, reps=100)
        assert np.array_equal(re, rr) and np.array_equal(im, ri)
        mine, _ = oracle.measure(6, re, 100, 4242)
        assert np.array_equal(mine, shots)


# --------------------------------------------------------- mathematics ----

def _textbook():
    I = np.eye(2)
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1, -1]).astype(complex)
    H = np.array([[1, 1], [1, -1]]) / math.sqrt(2)
    th, ph, la = 0.7, 1.3, -0.4

    def rx(t):
        return np.array([[math.cos(t / 2), -1j * math.sin(t / 2)],
                         [-1j * math.sin(t / 2), math.cos(t / 2)]])

    def ry(t):
        return np.array([[math.cos(t / 2), -math.sin(t / 2)], [math.sin(t / 2), math.cos(t / 2)]])

    def u3(t, p, l):
        return np.array([[math.cos(t / 2), -np.exp(1j * l) * math.sin(t / 2)],
                         [np.exp(1j * p) * math.sin(t / 2), np.exp(1j * (p + l)) * math.cos(t / 2)]])

    def ctrl(u):  # qubits [c, t]: c is the high bit of the matrix index
        m = np.eye(4, dtype=complex)
        m[2:, 2:] = u
        return m

    swap = np.eye(4)[[0, 2, 1, 3]]
    zz = np.diag([1, -1, -1, 1])
    tb = {
        "U3": u3(th, ph, la), "U2": u3(math.pi / 2, ph, la), "U1": np.diag([1, np.exp(1j * la)]),
        "ID": I, "X": X, "Y": Y, "Z": Z, "H": H, "S": np.diag([1, 1j]),
        "SDG": np.diag([1, -1j]), "T": np.diag([1, np.exp(1j * math.pi / 4)]),
        "TDG": np.diag([1, np.exp(-1j * math.pi / 4)]), "RX": rx(th), "RY": ry(th),
        "RZ": np.diag([np.exp(-1j * ph / 2), np.exp(1j * ph / 2)]),
        "W": np.array([[1, -1j], [-1j, 1]]) / math.sqrt(2),
        "CX": ctrl(X), "CZ": ctrl(Z), "CY": ctrl(Y), "CH": ctrl(H), "SWAP": swap,
        "CRX": ctrl(rx(la)), "CRY": ctrl(ry(la)),
        "CRZ": ctrl(np.diag([np.exp(-1j * la / 2), np.exp(1j * la / 2)])),
        "CU1": ctrl(np.diag([1, np.exp(1j * la)])), "CU3": ctrl(u3(th, ph, la)),
        "RZZ": np.diag(np.exp(-1j * th / 2 * np.diag(zz))),
        "RXX": _expm_pauli(np.kron(X, X), th), "RYY": _expm_pauli(np.kron(Y, Y), th),
    }
    ccx = np.eye(8, dtype=complex)
    ccx[6:, 6:] = X
    tb["CCX"] = ccx
    cswap = np.eye(8, dtype=complex)
    cswap[4:, 4:] = swap
    tb["CSWAP"] = cswap
    c3x = np.eye(16, dtype=complex)
    c3x[14:, 14:] = X
    tb["C3X"] = c3x
    return tb


def _expm_pauli(p, t):
    return math.cos(t / 2) * np.eye(4) - 1j * math.sin(t / 2) * p


def test_textbook_gates(oracle):
    tb = _textbook()
    for name, ref in tb.items():
        k = W.ARITY[name]
        m = oracle_gate_matrix(oracle, k, W.OP[name], list(range(k - 1, -1, -1)), (0.7, 1.3, -0.4))
        # qubits listed high bit first: operand 0 is the matrix's high bit
        assert equal_up_to_phase(m, ref), name


def test_documented_deviations(oracle):
    """The reference reproduces qelib1 literally: C4X is not a 4-controlled X,
    C3SQRTX is C3-(sqrt X)^dagger, R(theta) scales |1> by i*theta, SRN is not
    complex-linear."""
    c4x = oracle_gate_matrix(oracle, 5, W.OP["C4X"], [4, 3, 2, 1, 0])
    perm = (np.abs(np.abs(c4x) - 1) < 1e-12).sum(axis=0) == 1
    assert perm.sum() < 32 and np.allclose(c4x.conj().T @ c4x, np.eye(32), atol=1e-12)
    c3s = oracle_gate_matrix(oracle, 4, W.OP["C3SQRTX"], [3, 2, 1, 0])
    sx = np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]]) / 2
    want = np.eye(16, dtype=complex)
    want[14:, 14:] = sx.conj().T
    assert equal_up_to_phase(c3s, want)
    r = oracle_gate_matrix(oracle, 1, W.OP["R"], [0], (0.3, 0, 0))
    assert np.allclose(r, np.diag([1, 0.3j]))


def test_unitarity_of_named_gates(oracle):
    d = load("gate_matrices.npz")
    for name in W.OP_NAMES:
        if name in ("R", "SRN"):
            continue
        m = d[f"{name}_fwd"]
        assert np.allclose(m.conj().T @ m, np.eye(m.shape[0]), atol=1e-12), name


@pytest.mark.parametrize("n", [2, 3, 4])
def test_dense_brute_force(oracle, n):
    """One run equals conj(M rho0 M^dagger), M composed from textbook-free
    per-gate matrices in numpy (orientation and composition pinned)."""
    t = W.random_mixed(n, 60, 31 + n, gate_set=[g for g in W.OP_NAMES
                                               if W.ARITY[g] <= n and g != "SRN"])
    ops, qb, prm, coef = t.arrays()
    dim = 1 << n
    M = np.eye(dim, dtype=complex)
    for g in range(len(ops)):
        one = (ops[g:g + 1], qb[g:g + 1], prm[g:g + 1], coef[g:g + 1])
        re = np.eye(dim).ravel().copy()
        im = np.zeros(dim * dim)
        oracle.forward(n, dim, re, im, one)
        M = (re + 1j * im).reshape(dim, dim).T @ M
    rho0 = np.zeros((dim, dim))
    rho0[0, 0] = 1
    want = np.conj(M @ rho0 @ M.conj().T)
    re, im = oracle.sim(n, t.arrays())
    got = (re + 1j * im).reshape(dim, dim)
    assert np.allclose(got, want, atol=1e-12)


def test_trace_hermiticity_and_state_vector(oracle):
    n = 7
    t = W.random_mixed(n, 500, 3).arrays()
    re, im = oracle.sim(n, t)
    rho = (re + 1j * im).reshape(1 << n, 1 << n)
    assert abs(np.trace(rho) - 1) < 1e-12
    assert np.max(np.abs(rho - rho.conj().T)) < 1e-12
    psi = oracle.sv(n, t)
    assert np.max(np.abs(rho - np.outer(psi.conj(), psi))) < 1e-12


def test_rx_closed_form(oracle):
    """xacc/nvidia_omp/tests/DmSimAcceleratorTester.cpp:69-93: <Z> = cos(theta)."""
    for th in np.linspace(-math.pi, math.pi, 9):
        t = W.GateTable(1).add("RX", [0], [th]).arrays()
        re, _ = oracle.sim(1, t)
        assert abs((re[0] - re[3]) - math.cos(th)) < 1e-14


def test_deuteron_table(oracle):
    """xacc/tests/DmSimCpuAcceleratorTester.cpp:20-49: <Z0 Z1> of the ansatz
    equals sin(t); the reference's table holds it to 6 digits."""
    table = [0.0, -0.324699, -0.614213, -0.837166, -0.9694, -0.996584, -0.915773, -0.735724,
             -0.475947, -0.164595, 0.164595, 0.475947, 0.735724, 0.915773, 0.996584, 0.9694,
             0.837166, 0.614213, 0.324699, 0.0]
    for tval, want in zip(np.linspace(-math.pi, math.pi, 20), table):
        g = W.GateTable(2).add("X", [0]).add("RY", [1], [tval]).add("CX", [1, 0])
        g.add("H", [0]).add("H", [1])
        re, _ = oracle.sim(2, g.arrays())
        diag = re.reshape(4, 4).diagonal()
        zz = diag[0] - diag[1] - diag[2] + diag[3]
        assert abs(zz - want) < 1e-6 and abs(zz - math.sin(tval)) < 1e-13


def test_bell_and_shot_statistics(oracle):
    """xacc/tests/DmSimCpuAcceleratorTester.cpp:51-108: empty -> 00; X0 CX01 ->
    11; Bell -> only 00/11, about half each."""
    re, _ = oracle.sim(2, W.GateTable(2).arrays())
    assert set(oracle.measure(2, re, 100, 1)[0]) == {0}
    re, _ = oracle.sim(2, W.GateTable(2).add("X", [0]).add("CX", [0, 1]).arrays())
    assert set(oracle.measure(2, re, 100, 1)[0]) == {3}
    re, _ = oracle.sim(2, W.GateTable(2).add("H", [0]).add("CX", [0, 1]).arrays())
    shots, _ = oracle.measure(2, re, 4000, 12345)
    assert set(shots) <= {0, 3} and abs(np.mean(shots == 0) - 0.5) < 0.05


def test_planted_mistake_is_caught(oracle):
    """A wrong coefficient (H with 0.7071 instead of 1/sqrt 2, as a C1 matrix)
    must break bit-exactness against the golden H matrix."""
    d = load("gate_matrices.npz")
    coef = np.zeros(32)
    coef[[0, 2, 4]] = 0.7071
    coef[6] = -0.7071
    m = oracle_gate_matrix(oracle, 1, W.OP["C1"], [0], coef=coef)
    assert not np.array_equal(m, d["H_fwd"])
    assert np.allclose(m, d["H_fwd"], atol=1e-4)
