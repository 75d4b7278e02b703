"""GPU runs of OpenQASM circuits through the front-end: the plain method
bit-exact against the oracle on the same parsed gate table, the fused method
within 1e-10, the 15-qubit circuits against the pure-state oracle, and the
Cuccaro adder's closed-form result."""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
QASM = os.path.join(HERE, "golden", "qasm")


def load(name):
    with open(os.path.join(QASM, name)) as f:
        return f.read()


def arrays(dm, src):
    nq, ops, qb, prm = dm.qasm_parse(src)
    coef = np.zeros((len(ops), 32))
    return nq, (np.ascontiguousarray(ops), np.ascontiguousarray(qb), np.ascontiguousarray(prm),
                coef)


def run(dm, src, n, method):
    s = dm.Sim(n)
    s.set_method(method)
    s.append_qasm(src)
    s.upload()
    s.run()
    return s


@pytest.mark.parametrize("name", ["adder_n9.qasm", "pea3pi8_n5.qasm", "w_state_n3.qasm",
                                  "qec_n5.qasm", "grover_n3.qasm", "sat_n10.qasm",
                                  "vqe_uccsd_n8.qasm"])
def test_qasm_circuits_vs_oracle(oracle, name):
    import dmsim_amd as dm
    src = load(name)
    n, tab = arrays(dm, src)
    print(f"qasm {name}: n={n}, {len(tab[0])} gates", flush=True)
    ore, oim = oracle.sim(n, tab)
    s = run(dm, src, n, dm.METHOD_PLAIN_LAUNCH)
    re, im = s.get_dm()
    assert np.array_equal(re, ore) and np.array_equal(im, oim)
    if n >= 4:
        f = run(dm, src, n, dm.METHOD_FUSED)
        fre, fim = f.get_dm()
        assert max(np.abs(fre - ore).max(), np.abs(fim - oim).max()) <= 1e-10


def test_adder_result_and_shots(oracle):
    """cin = 0, a = 0001, b = 1111: b + a = 10000, so every shot reads a = 0001
    (qubit 1) and cout = 1 (qubit 9): 2 + 512 = 514."""
    import dmsim_amd as dm
    src = load("adder_n9.qasm")
    s = run(dm, src, 10, dm.METHOD_FUSED)
    shots, total = s.measure(100, 7)
    assert np.all(shots == 514) and abs(total - 1) < 1e-12


@pytest.mark.parametrize("name", ["qft_n15.qasm", "bv_n15.qasm"])
def test_qasm_n15_vs_state_vector(oracle, name):
    import dmsim_amd as dm
    src = load(name)
    n, tab = arrays(dm, src)
    s = run(dm, src, n, dm.METHOD_FUSED)
    psi = oracle.sv(n, tab)
    diag = s.get_diag()
    assert np.max(np.abs(diag - np.abs(psi) ** 2)) <= 1e-10
    rng = np.random.default_rng(3)
    dim = 1 << n
    i = rng.integers(0, dim, 2048, dtype=np.uint64)
    j = rng.integers(0, dim, 2048, dtype=np.uint64)
    got = s.get_entries(i * np.uint64(dim) + j)
    assert np.max(np.abs(got - np.conj(psi[i]) * psi[j])) <= 1e-10


def test_vqe_uccsd_n8_timing():
    """The circuit of the reference's published single-GPU number (10 808
    gates in 249.3 ms on a V100, README.md:47), timed here."""
    import dmsim_amd as dm
    src = load("vqe_uccsd_n8.qasm")
    for method in (dm.METHOD_FUSED, dm.METHOD_PLAIN):
        s = run(dm, src, 8, method)
        s.reset_dm()
        t0 = time.perf_counter()
        s.run()
        wall = time.perf_counter() - t0
        print(f"vqe_uccsd_n8 method {method}: {s.last_run_stats()['ms']:.2f} ms device, "
              f"{wall * 1e3:.2f} ms wall (reference, V100: 249.3 ms)", flush=True)
        assert wall < 0.2493
