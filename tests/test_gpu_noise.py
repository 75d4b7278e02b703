"""GPU checks of single-qubit noise channels (rho -> sum_i K_i rho K_i^dagger):
dense NumPy evolution of the same sequence of matrix gates and channels is the
reference (there is no channel in the reference code, so the bar is the
mathematics), in the plain and fused methods, across continued runs (which
start from the conj-transposed orientation), plus closed forms."""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


def dmx():
    import dmsim_amd
    return dmsim_amd


def embed(n, qs, u):
    """Full 2^n x 2^n matrix of u acting on qubits qs (qs[0] = most
    significant bit of u's index)."""
    k = len(qs)
    full = np.zeros((1 << n, 1 << n), dtype=complex)
    for x in range(1 << n):
        sub_in = sum(((x >> q) & 1) << (k - 1 - j) for j, q in enumerate(qs))
        rest = x & ~sum(1 << q for q in qs)
        for so in range(1 << k):
            y = rest | sum(((so >> (k - 1 - j)) & 1) << q for j, q in enumerate(qs))
            full[y, x] += u[so, sub_in]
    return full


def random_program(n, length, seed):
    dm = dmx()
    rng = np.random.default_rng(seed)
    chans = [lambda: dm.depolarizing(rng.uniform(0, 0.3)),
             lambda: dm.amplitude_damping(rng.uniform(0, 0.5)),
             lambda: dm.phase_damping(rng.uniform(0, 0.5))]
    prog = []
    for _ in range(length):
        r = rng.uniform()
        if r < 0.3:
            prog.append(("kraus", int(rng.integers(n)), chans[int(rng.integers(3))]()))
        else:
            k = 1 if r < 0.6 else 2
            qs = [int(x) for x in rng.choice(n, size=k, replace=False)]
            prog.append(("gate", qs, W.haar_unitary(rng, 1 << k)))
    return prog


def apply_dense(n, prog, rho):
    for kind, q, op in prog:
        if kind == "gate":
            m = embed(n, q, op)
            rho = m @ rho @ m.conj().T
        else:
            rho = sum(embed(n, [q], k) @ rho @ embed(n, [q], k).conj().T for k in op)
    return rho


def load(sim, prog):
    for kind, q, op in prog:
        if kind == "gate":
            sim.append_matrix(q, op)
        else:
            sim.append_kraus(q, op)


def got_rho(sim, n):
    re, im = sim.get_dm()
    return np.conj((re + 1j * im).reshape(1 << n, 1 << n))  # get_dm holds conj(rho)


@pytest.mark.parametrize("method", [0, 1, 2])
@pytest.mark.parametrize("n", [4, 5])
def test_channels_vs_dense(n, method):
    dm = dmx()
    s = dm.Sim(n)
    s.set_method(method)
    rho = np.zeros((1 << n, 1 << n), dtype=complex)
    rho[0, 0] = 1
    for seg in range(3):  # continued runs: orientation alternates in the plain method
        prog = random_program(n, 25, 100 * n + 10 * method + seg)
        s.clear_circuit()
        load(s, prog)
        s.upload()
        s.run()
        rho = apply_dense(n, prog, rho)
        got = got_rho(s, n)
        tol = 1e-12 if method != 2 else 1e-10
        assert np.max(np.abs(got - rho)) <= tol, (seg, np.max(np.abs(got - rho)))
        assert abs(np.trace(got) - 1) <= 1e-12
        assert np.max(np.abs(got - got.conj().T)) <= 1e-12


def test_fused_plan_keeps_channels_as_barriers():
    dm = dmx()
    n = 6
    prog = random_program(n, 60, 7)
    s = dm.Sim(n)
    s.set_method(dm.METHOD_FUSED)
    load(s, prog)
    s.upload()
    s.run()
    st = s.last_run_stats()
    info = s.plan_info()
    n_chan = sum(1 for p in prog if p[0] == "kraus")
    assert st["launches"] == info["fused_blocks"] + info["fused5_blocks"] + n_chan
    rho = np.zeros((1 << n, 1 << n), dtype=complex)
    rho[0, 0] = 1
    assert np.max(np.abs(got_rho(s, n) - apply_dense(n, prog, rho))) <= 1e-10


@pytest.mark.parametrize("method", [0, 2])
def test_closed_forms(method):
    dm = dmx()
    n = 4
    g, p = 0.3, 0.12
    s = dm.Sim(n)
    s.set_method(method)
    s.append(W.OP["X"], [1])
    s.append_kraus(1, dm.amplitude_damping(g))   # |1> decays: P(1) = 1 - g
    s.append_kraus(2, dm.depolarizing(p))        # |0>: P(1) = 2p/3
    s.append(W.OP["H"], [3])
    s.append_kraus(3, dm.phase_damping(0.36))    # coherence scaled by sqrt(1 - 0.36)
    s.upload()
    s.run()
    rho = got_rho(s, n)
    d = np.real(np.diag(rho))
    p1 = lambda q: d[[i for i in range(1 << n) if (i >> q) & 1]].sum()
    assert abs(p1(1) - (1 - g)) <= 1e-14
    assert abs(p1(2) - 2 * p / 3) <= 1e-14
    # <X_3> = 2 Re rho[i, i ^ 8] summed over i with bit 3 clear
    ex = 2 * sum(np.real(rho[i, i | 8]) for i in range(1 << n) if not (i >> 3) & 1)
    assert abs(ex - 0.8) <= 1e-14


def test_channels_mixed_methods_continuation():
    """plain run (leaves rho conj-transposed), fused run (its channels use the
    transposed-orientation records), plain run again."""
    dm = dmx()
    n = 5
    s = dm.Sim(n)
    rho = np.zeros((1 << n, 1 << n), dtype=complex)
    rho[0, 0] = 1
    for seg, method in enumerate([0, 2, 1, 2]):
        prog = random_program(n, 25, 900 + seg)
        s.set_method(method)
        s.clear_circuit()
        load(s, prog)
        s.upload()
        s.run()
        rho = apply_dense(n, prog, rho)
    assert np.max(np.abs(got_rho(s, n) - rho)) <= 1e-10


def test_channels_between_permuting_sweeps():
    """12 qubits, CX/SWAP/X-rich segments separated by channels: the fused plan
    with out-of-place permuting sweeps (the planes swap after each) keeps the
    channels in place and matches the in-place fused plan and the plain
    method, over two consecutive runs (the second starts from the plain
    method's conj-transposed orientation for the plain simulation)."""
    dm = dmx()
    n = 12
    rng = np.random.default_rng(1212)
    chunks = []
    for c in range(6):
        t = W.GateTable(n)
        for _ in range(40):
            r = rng.uniform()
            if r < 0.5:
                t.add("CX", [int(x) for x in rng.choice(n, 2, replace=False)])
            elif r < 0.6:
                t.add("SWAP", [int(x) for x in rng.choice(n, 2, replace=False)])
            elif r < 0.7:
                t.add("X", [int(rng.integers(n))])
            else:
                t.add("U3", [int(rng.integers(n))], [float(v) for v in rng.uniform(0, 6.2, 3)])
        ops = dm.depolarizing(0.1) if c % 2 else dm.amplitude_damping(0.2)
        chunks.append((t, int(rng.integers(n)), ops))

    def make(method, permute):
        s = dm.Sim(n)
        s.set_method(method)
        s.set_fusion_permute(permute)
        for t, q, ops in chunks:
            s.append_table(t)
            s.append_kraus(q, ops)
        s.upload()
        s.run()
        s.run()
        return s

    plain = make(1, False)
    ref = got_rho(plain, n)
    launches = {}
    for permute in (False, True):
        s = make(2, permute)
        got = got_rho(s, n)
        assert np.max(np.abs(got - ref)) <= 1e-10, permute
        assert abs(np.trace(got) - 1) <= 1e-12
        launches[permute] = s.last_run_stats()["launches"]
    assert launches[True] < launches[False]
