"""GPU checks of the batched executor (many uneven circuits, one graph launch).

* each circuit's rho is bit-identical to the single-circuit fused run (same
  plan, same kernel body);
* energies sum_s w_s Re Tr(P_s rho) agree with dense numpy Pauli operators on
  the oracle's rho (n <= 8) and with the state-vector oracle on the full
  256-circuit workload (n = 6..11);
* graph replays are identical to one another and to an eager run.
"""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu

PAULI = {
    "I": np.eye(2), "X": np.array([[0, 1], [1, 0]], dtype=complex),
    "Y": np.array([[0, -1j], [1j, 0]]), "Z": np.diag([1.0, -1.0]).astype(complex),
}


def dense_pauli(s):
    """Operator with qubit 0 = least significant bit (s[0] acts on qubit 0)."""
    op = np.eye(1, dtype=complex)
    for ch in reversed(s):
        op = np.kron(op, PAULI[ch])
    return op


def sv_energy(psi, strings, weights):
    n = int(np.log2(psi.size))
    k = np.arange(psi.size)
    e = 0.0
    for s, w in zip(strings, weights):
        x = sum(1 << q for q, c in enumerate(s) if c in "XY")
        zy = sum(1 << q for q, c in enumerate(s) if c in "ZY")
        ny = s.count("Y")
        par = np.array([bin(v).count("1") & 1 for v in (k & zy)])
        phase = (1j ** ny) * np.where(par == 1, -1.0, 1.0)
        e += w * np.real(np.sum(np.conj(psi[k ^ x]) * phase * psi[k]))
        assert n == len(s)
    return e


def make_batch(items):
    import dmsim_amd
    b = dmsim_amd.Batch()
    for circ, strings, weights in items:
        b.add_circuit(circ, strings, weights)
    b.upload()
    return b


def small_items(seed=3, count=10):
    rng = np.random.default_rng(seed)
    items = []
    for _ in range(count):
        n = int(rng.integers(4, 9))
        circ = W.hea_circuit(n, int(rng.integers(2, 8)), rng)
        strings, weights = W.random_paulis(n, 8, rng)
        items.append((circ, strings, weights))
    return items


def test_batch_matches_oracle_dense(oracle):
    items = small_items()
    b = make_batch(items)
    b.run()
    en = b.energies()
    for cid, (circ, strings, weights) in enumerate(items):
        n = circ.n_qubits
        re, im = oracle.sim(n, circ.arrays())
        rho = np.conj((re + 1j * im).reshape(1 << n, 1 << n))
        want = sum(w * np.real(np.trace(dense_pauli(s) @ rho)) for s, w in zip(strings, weights))
        assert abs(en[cid] - want) < 1e-10, cid
        gre, gim = b.get_dm(cid)
        assert max(np.abs(gre - re).max(), np.abs(gim - im).max()) < 1e-10


@pytest.mark.parametrize("autotune,herm", [(True, True), (False, True), (True, False)])
def test_batch_bitwise_equals_single_fused_run(autotune, herm):
    """Tile widths only regroup the 16 x 16 tiles: the batch, autotuned or
    not, reproduces the single-circuit fused run bit for bit -- with
    Hermitian sweeps for its circuits of 10 and 11 qubits, the run planned
    with Hermitian sweeps (dmx_set_fusion_hermitian 2), without, the in-order
    run."""
    import dmsim_amd
    items = small_items(seed=9, count=6)
    rng = np.random.default_rng(19)
    for n in (8, 9, 10, 11):
        strings, weights = W.random_paulis(n, 4, rng)
        items.append((W.hea_circuit(n, 6, rng), strings, weights))
    b = dmsim_amd.Batch()
    b.set_autotune(autotune)
    b.set_hermitian(herm)
    for circ, strings, weights in items:
        b.add_circuit(circ, strings, weights)
    b.upload()
    for circ, _, _ in items:
        m = b.tile_bits(circ.n_qubits)
        n = circ.n_qubits
        assert (5 if n > 4 else 4) <= m <= min(8, n) and (autotune or m == min(8, n))
    b.run()
    for cid, (circ, _, _) in enumerate(items):
        s = dmsim_amd.Sim(circ.n_qubits)
        s.set_method(dmsim_amd.METHOD_FUSED)
        s.set_fusion_hermitian(2 if herm and circ.n_qubits >= 10 else 0)
        s.append_table(circ)
        s.upload()
        if herm and circ.n_qubits >= 10:
            assert s.plan_info()["herm_blocks"] > 0
        s.run()
        re, im = s.get_dm()
        bre, bim = b.get_dm(cid)
        assert np.array_equal(re, bre) and np.array_equal(im, bim), cid


def test_graph_replays_match_eager():
    import dmsim_amd
    items = small_items(seed=5, count=8)
    b = make_batch(items)
    b.set_graph(False)
    b.run()
    eager = b.energies().copy()
    b.set_graph(True)
    for _ in range(4):
        b.run()
        assert np.array_equal(b.energies(), eager)
        # eager work of the program between replays
        s = dmsim_amd.Sim(5)
        s.append(W.OP["H"], [0])
        s.upload()
        s.run()
        s.get_dm()
    st = b.stats()
    assert st["circuits"] == 8 and st["gates"] == sum(len(c) for c, _, _ in items)


def test_full_hea_batch_against_state_vector(oracle):
    """The 256-circuit workload (n uniform in 6..11, depth 10..40, 32 Pauli
    strings each): every energy against the pure-state oracle."""
    items = W.hea_batch(n_circuits=256, seed=4)
    b = make_batch(items)
    b.run()
    b.run()  # replay from |0><0| again
    en = b.energies()
    for cid, (circ, strings, weights) in enumerate(items):
        psi = oracle.sv(circ.n_qubits, circ.arrays())
        want = sv_energy(psi, strings, weights)
        assert abs(en[cid] - want) < 1e-9, (cid, en[cid], want)


def test_batch_rejects_bad_input():
    import dmsim_amd
    b = dmsim_amd.Batch()
    with pytest.raises(ValueError):
        b.add_circuit(W.GateTable(3).add("H", [0]))  # below 4 qubits
    t = W.GateTable(5).add("H", [0])
    cid = b.add_circuit(t)
    with pytest.raises(ValueError):
        import ctypes as C
        dmsim_amd._check(dmsim_amd.load().dmx_batch_add_pauli(b._h, cid, b"XYZ", 1.0))
    with pytest.raises(dmsim_amd.DmxError):
        b.run()  # before upload


def test_batch_waves_match_one_wave(oracle):
    """A pool limit that splits the batch into several waves (each reusing the
    pool inside one run) gives the same energies and diagonals, bit for bit,
    as one wave; rho stays available for the last wave's circuits only."""
    import dmsim_amd
    items = small_items(seed=41, count=12)

    def build(limit):
        b = dmsim_amd.Batch()
        b.set_pool_limit(limit)
        for circ, strings, weights in items:
            b.add_circuit(circ, strings, weights)
        b.upload()
        b.run()
        return b

    one = build(0)
    many = build(1 << 20)  # 1 MiB: a few 8-qubit states per wave
    assert one.waves() == 1 and many.waves() >= 3
    assert np.array_equal(one.energies(), many.energies())
    last = None
    for cid, (circ, _, _) in enumerate(items):
        assert np.array_equal(one.get_diag(cid), many.get_diag(cid))
        psi = oracle.sv(circ.n_qubits, circ.arrays())
        assert np.max(np.abs(one.get_diag(cid) - np.abs(psi) ** 2)) <= 1e-10
        last = cid
    re1, im1 = one.get_dm(last)
    re2, im2 = many.get_dm(last)
    assert np.array_equal(re1, re2) and np.array_equal(im1, im2)
    with pytest.raises(dmsim_amd.DmxError):
        many.get_dm(0)
    for _ in range(2):  # graph replays
        many.run()
        assert np.array_equal(one.energies(), many.energies())
    one.close()
    many.close()


def test_batch_runs_c4x_circuits(oracle):
    """Circuits with C4X (a five-qubit gate) run in a batch as the gates the
    reference composes it from: diagonals against the state-vector oracle."""
    import dmsim_amd
    b = dmsim_amd.Batch()
    tables = [W.structured(n, seed) for n, seed in ((9, 1), (10, 2), (11, 3))]
    assert all(W.OP["C4X"] in t.ops for t in tables)
    for t in tables:
        b.add_circuit(t)
    b.upload()
    b.run()
    for cid, t in enumerate(tables):
        psi = oracle.sv(t.n_qubits, t.arrays())
        assert np.max(np.abs(b.get_diag(cid) - np.abs(psi) ** 2)) <= 1e-10
    b.close()


@pytest.mark.parametrize("learned", [False, True])
def test_batch_large_circuits_run_on_their_own(oracle, learned):
    """Circuits of 12 or more qubits run on their own inside their wave,
    through the single-circuit engine's fused plan (Hermitian block sweeps, a
    C4X kept as one five-qubit block), greedy or with the cost model: rho
    bit-identical to the circuit's own fused run; energies and diagonals
    against the state-vector oracle; the small circuit of the same wave still
    runs batched."""
    import costdata
    import dmsim_amd
    model = costdata.default_model() if learned else None
    rng = np.random.default_rng(1212)
    tables = [W.structured(12, 6), W.structured(9, 5), W.random_mixed(12, 300, seed=12),
              W.structured(13, 8)]
    items = []
    for t in tables:
        strings, weights = W.random_paulis(t.n_qubits, 6, rng)
        items.append((t, strings, weights))
    print(f"batch: large circuits on their own, learned={learned}", flush=True)
    b = dmsim_amd.Batch()
    if model is not None:
        b.set_cost_model(model)
    for circ, strings, weights in items:
        b.add_circuit(circ, strings, weights)
    b.upload()
    b.run()
    assert b.waves() == 1
    en = b.energies()
    for cid, (t, strings, weights) in enumerate(items):
        n = t.n_qubits
        psi = oracle.sv(n, t.arrays())
        assert np.max(np.abs(b.get_diag(cid) - np.abs(psi) ** 2)) <= 1e-10, cid
        assert abs(en[cid] - sv_energy(psi, strings, weights)) < 1e-9, cid
        if n < 12:
            continue
        s = dmsim_amd.Sim(n)
        s.set_method(dmsim_amd.METHOD_FUSED)
        if model is not None:
            s.set_cost_model(model)
        s.append_table(t)
        s.upload()
        assert s.plan_info()["herm_blocks"] > 0
        s.run()
        re, im = s.get_dm()
        bre, bim = b.get_dm(cid)
        assert np.array_equal(re, bre) and np.array_equal(im, bim), cid
        s.close()
    b.run()  # graph replay
    assert np.array_equal(en, b.energies())
    b.close()
    if model is not None:
        model.close()


def test_batch_cache_groups_are_bit_identical(oracle):
    """Circuits packed into groups that run all their block steps before the
    next group (a 1 MiB budget: many groups) give energies and rho bit for bit
    equal to whole-wave steps; the timed schedule (autotuning) picks one of
    the budgets and agrees as well."""
    import dmsim_amd
    rng = np.random.default_rng(77)
    items = []
    for _ in range(14):
        n = int(rng.integers(6, 11))
        circ = W.hea_circuit(n, int(rng.integers(4, 12)), rng)
        strings, weights = W.random_paulis(n, 6, rng)
        items.append((circ, strings, weights))

    def build(budget):
        b = dmsim_amd.Batch()
        if budget is not None:
            b.set_group_bytes(budget)
        for circ, strings, weights in items:
            b.add_circuit(circ, strings, weights)
        b.upload()
        b.run()
        return b

    whole, grouped, timed = build(0), build(1 << 20), build(None)
    sw, sg, st = whole.schedule(), grouped.schedule(), timed.schedule()
    print(f"batch schedules: whole {sw}, 1 MiB {sg}, timed {st}", flush=True)
    assert sw["groups"] == 1 and sg["groups"] > 4 and sg["group_bytes"] == 1 << 20
    assert st["group_bytes"] in (0, 16 << 20, 32 << 20, 64 << 20, 128 << 20)
    for other in (grouped, timed):
        assert np.array_equal(whole.energies(), other.energies())
        for cid in range(len(items)):
            a, b = whole.get_dm(cid), other.get_dm(cid)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), cid
    for cid, (circ, strings, weights) in enumerate(items):
        psi = oracle.sv(circ.n_qubits, circ.arrays())
        assert abs(whole.energies()[cid] - sv_energy(psi, strings, weights)) < 1e-9
    for b in (whole, grouped, timed):
        b.close()


def test_batch_hermitian_circuits_in_waves(oracle):
    """10- and 11-qubit circuits (batched Hermitian sweeps) split over several
    waves by a pool limit give energies and diagonals bit for bit equal to one
    wave, and the state-vector oracle's to 1e-10."""
    import dmsim_amd
    rng = np.random.default_rng(1011)
    items = []
    for n in (10, 11, 10, 11, 10):
        strings, weights = W.random_paulis(n, 6, rng)
        items.append((W.hea_circuit(n, int(rng.integers(4, 10)), rng), strings, weights))

    def build(limit):
        b = dmsim_amd.Batch()
        b.set_pool_limit(limit)
        for circ, strings, weights in items:
            b.add_circuit(circ, strings, weights)
        b.upload()
        b.run()
        return b

    one = build(0)
    many = build(16 << 22)  # one 11-qubit state per pool: five waves
    assert one.waves() == 1 and many.waves() == 5
    assert np.array_equal(one.energies(), many.energies())
    for cid, (circ, strings, weights) in enumerate(items):
        assert np.array_equal(one.get_diag(cid), many.get_diag(cid)), cid
        psi = oracle.sv(circ.n_qubits, circ.arrays())
        assert np.max(np.abs(one.get_diag(cid) - np.abs(psi) ** 2)) <= 1e-10
        assert abs(one.energies()[cid] - sv_energy(psi, strings, weights)) < 1e-9
    one.close()
    many.close()
