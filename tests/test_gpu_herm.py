"""Hermitian block sweeps on the GPU (k_fused4_herm, which reads and transforms
only the tiles on and above the diagonal and writes each result with its
conjugate transpose, and k_relabel, which restores rho's bit order): against
the oracle (<= 1e-10, trace and Hermiticity <= 1e-12), against in-order fused
plans and the plain method, over continued runs, runs that start from the
plain method's conj-transposed orientation, plans mixing channels and
five-qubit blocks (relabelled before each), and graph replays."""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-10
FUSED = 2
UNITARY_SET = [g for g in W.OP_NAMES if g not in ("SRN", "R", "C4X")]


def dmx():
    import dmsim_amd
    return dmsim_amd


def workload(n, seed):
    """Brick-wall SU(4) layers, then a mix rich in X/CX/SWAP (trailing index
    maps) and every unitary op of up to four qubits."""
    rng = np.random.default_rng(seed)
    t = W.brickwall_su4(n=n, layers=3, tail=0, seed=seed)
    gate_set = [g for g in UNITARY_SET if W.ARITY[g] <= 4]
    t2 = W.random_mixed(n, 150, seed + 1, gate_set=gate_set)
    for _ in range(120):
        r = rng.uniform()
        if r < 0.35:
            t2.add("CX", [int(x) for x in rng.choice(n, 2, replace=False)])
        elif r < 0.5:
            t2.add("SWAP", [int(x) for x in rng.choice(n, 2, replace=False)])
        elif r < 0.6:
            t2.add("X", [int(rng.integers(n))])
        else:
            t2.add("U3", [int(rng.integers(n))], [float(v) for v in rng.uniform(0, 6.2, 3)])
    t.extend(t2)
    return t


def make(n, tables, herm=2, profiling=True):
    s = dmx().Sim(n)
    s.set_method(FUSED)
    s.set_fusion_hermitian(herm)
    s.set_profiling(profiling)
    for t in tables:
        s.append_table(t)
    s.upload()
    return s


def check_close(re, im, re_ref, im_ref, n):
    err = max(np.abs(re - re_ref).max(), np.abs(im - im_ref).max())
    assert err <= TOL, err
    rho = (re + 1j * im).reshape(1 << n, 1 << n)
    assert abs(np.trace(rho) - np.trace((re_ref + 1j * im_ref).reshape(1 << n, 1 << n))) < 1e-12
    assert np.max(np.abs(rho - rho.conj().T)) < 1e-12


@pytest.mark.parametrize("n", [8, 9, 10, 12])
def test_herm_vs_oracle_two_runs(oracle, n):
    t = workload(n, 9100 + n)
    s = make(n, [t])
    info = s.plan_info()
    assert info["herm_blocks"] > 0 and info["relabels"] == 1 and info["perm_blocks"] > 0
    s.run()
    prof = s.profile()
    assert prof["k_fused4_herm"]["calls"] == info["herm_blocks"]
    assert prof["k_relabel"]["calls"] == 1
    re, im = s.get_dm()
    ore, oim = oracle.sim(n, t.arrays())
    check_close(re, im, ore, oim, n)
    s.run()  # continues from the first run's result
    ore, oim = oracle.sim(n, t.arrays(), (ore, oim))
    re, im = s.get_dm()
    check_close(re, im, ore, oim, n)


@pytest.mark.parametrize("n", [9, 12])
def test_herm_matches_in_order_plans(n):
    t = workload(n, 9200 + n)
    a = make(n, [t], herm=2)
    b = make(n, [t], herm=0)
    for s in (a, b):
        s.run()
    ra, ia = a.get_dm()
    rb, ib = b.get_dm()
    assert max(np.abs(ra - rb).max(), np.abs(ia - ib).max()) <= TOL
    assert a.plan_info()["herm_blocks"] > 0 and b.plan_info()["herm_blocks"] == 0


def test_herm_after_plain_run(oracle):
    """The plain method leaves rho conj-transposed (the same Hermitian matrix);
    a Hermitian fused run continues from it."""
    n = 10
    t = workload(n, 9300)
    s = dmx().Sim(n)
    s.set_method(1)
    s.append_table(t)
    s.upload()
    s.run()
    s.set_method(FUSED)
    s.set_fusion_hermitian(2)
    s.clear_circuit()
    s.append_table(t)
    s.upload()
    assert s.plan_info()["herm_blocks"] > 0
    s.run()
    ore, oim = oracle.sim(n, t.arrays())
    ore, oim = oracle.sim(n, t.arrays(), (ore, oim))
    re, im = s.get_dm()
    check_close(re, im, ore, oim, n)


def test_herm_with_channels_and_five_qubit_blocks():
    """Channels and five-qubit blocks (C4X) after Hermitian sweeps run on
    whatever index bits hold their qubits: one relabelling, at the end.
    Against the plain method, over two runs."""
    dm = dmx()
    n = 10
    rng = np.random.default_rng(9400)
    chunks = []
    for c in range(4):
        t = workload(n, 9400 + c)
        if c % 2:
            t.add("C4X", [int(x) for x in rng.choice(n, 5, replace=False)])
        ops = dm.depolarizing(0.1) if c % 2 else dm.amplitude_damping(0.2)
        chunks.append((t, int(rng.integers(n)), ops))

    def build(method, herm):
        s = dm.Sim(n)
        s.set_method(method)
        s.set_fusion_hermitian(herm)
        for t, q, ops in chunks:
            s.append_table(t)
            s.append_kraus(q, ops)
        s.upload()
        s.run()
        s.run()
        return s

    ref = build(1, 0)
    got = build(FUSED, 2)
    info = got.plan_info()
    assert info["herm_blocks"] > 0 and info["relabels"] == 1
    ra, ia = ref.get_dm()
    rb, ib = got.get_dm()
    assert max(np.abs(ra - rb).max(), np.abs(ia - ib).max()) <= TOL
    rho = (rb + 1j * ib).reshape(1 << n, 1 << n)
    assert abs(np.trace(rho) - 1) < 1e-12 and np.max(np.abs(rho - rho.conj().T)) < 1e-12


def test_herm_graph_replays_match_eager():
    """Profiling off: the run is captured into a HIP graph and replayed; every
    replay equals the eager run bit for bit, also after a plain run in
    between."""
    n = 10
    t = workload(n, 9500)
    eager, graph = make(n, [t], profiling=True), make(n, [t], profiling=False)
    for step in range(4):
        if step == 2:
            for s in (eager, graph):
                s.set_method(1)
                s.run()
                s.set_method(FUSED)
        for s in (eager, graph):
            s.run()
        a, b = eager.get_dm(), graph.get_dm()
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), step


def test_herm_default_from_twelve_qubits():
    t = workload(12, 9600)
    s = make(12, [t], herm=1)
    assert s.plan_info()["herm_blocks"] > 0
    t = workload(10, 9601)
    s = make(10, [t], herm=1)
    assert s.plan_info()["herm_blocks"] == 0
