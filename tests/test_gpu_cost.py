"""GPU checks of the learned fusion cost model: the HIP network kernels against
the NumPy restatement (oracle/mlp_oracle.py), training, the timing-data
generator, and the learned planner judged by unfused (plain) execution for
both numerics and launch-time predictions."""
import math
import os
import sys

import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dm():
    import dmsim_amd
    return dmsim_amd


@pytest.fixture(scope="module")
def costdata():
    import costdata
    return costdata


def _params(seed, scale=0.2):
    return np.random.default_rng(seed).normal(0.0, scale, mo.P).astype(np.float32)


def test_forward_matches_oracle(dm):
    print("cost: forward vs oracle", flush=True)
    p = _params(1)
    model = dm.CostModel(1)
    model.set_params(p, 0.0, 1.0)
    X = np.random.default_rng(2).uniform(0, 1, (777, mo.F)).astype(np.float32)
    got = model.predict(X)
    ref = mo.forward(p.astype(np.float64), X.astype(np.float64))
    # FP32 sums of at most 64 terms: relative error ~1e-6
    assert np.max(np.abs(got - ref)) <= 1e-4 * max(1.0, np.max(np.abs(ref)))
    # the de-standardisation is affine in the stored mean and deviation
    model.set_params(p, 2.0, 3.0)
    assert np.allclose(model.predict(X), 3.0 * got + 2.0, rtol=1e-6, atol=1e-5)


@pytest.mark.parametrize("count", [1, 63, 64, 1000])
def test_gradient_matches_oracle(dm, count):
    print(f"cost: gradient vs oracle, count={count}", flush=True)
    p = _params(3)
    model = dm.CostModel(3)
    model.set_params(p, 0.0, 1.0)
    rng = np.random.default_rng(count)
    X = rng.uniform(0, 1, (count, mo.F)).astype(np.float32)
    y = rng.normal(0, 1, count).astype(np.float32)
    g = model.gradient(X, y)
    ref = mo.loss_grad(p.astype(np.float64), X.astype(np.float64), y.astype(np.float64))
    assert np.linalg.norm(g - ref) <= 1e-4 * np.linalg.norm(ref)
    assert np.max(np.abs(g - ref)) <= 1e-4 * np.max(np.abs(ref))
    # a second call starts from a zeroed accumulator (atomic order may differ)
    assert np.allclose(model.gradient(X, y), g, rtol=1e-5, atol=1e-7)


def test_adam_step_matches_oracle(dm):
    print("cost: one Adam step vs oracle", flush=True)
    p = _params(5)
    model = dm.CostModel(5)
    model.set_params(p, 0.0, 1.0)
    rng = np.random.default_rng(6)
    X = rng.uniform(0, 1, (256, mo.F)).astype(np.float32)
    y = rng.normal(0, 1, 256).astype(np.float32)
    lr = 1e-3
    model.train(X, y, epochs=1, batch=256, lr=lr, seed=0)  # one full-batch step
    got, mu, dev = model.get_params()
    ts = (y.astype(np.float64) - mu) / dev
    g = mo.loss_grad(p.astype(np.float64), X.astype(np.float64), ts)
    ref, _, _ = mo.adam_step(p.astype(np.float64), g, np.zeros(mo.P), np.zeros(mo.P), lr, 1)
    clear = np.abs(g) > 1e-3 * np.max(np.abs(g))  # sign of g unambiguous in FP32
    assert np.max(np.abs(got - ref)[clear]) <= 1e-6
    assert np.max(np.abs(got - p)) <= lr * 1.001


def test_training_fits_a_smooth_target(dm):
    print("cost: training on a synthetic target", flush=True)
    rng = np.random.default_rng(8)
    X = rng.uniform(0, 1, (3000, mo.F)).astype(np.float32)
    y = (1.5 * X[:, 0] - X[:, 3] + np.sin(3 * X[:, 5]) + 2 * X[:, 7] * X[:, 8]).astype(np.float32)
    model = dm.CostModel(8)
    before = np.mean((model.predict(X[2500:]) - y[2500:]) ** 2)
    loss = model.train(X[:2500], y[:2500], epochs=150, batch=64, lr=2e-3, seed=1)
    after = np.mean((model.predict(X[2500:]) - y[2500:]) ** 2)
    assert loss < 0.01
    assert after < 0.01 * np.var(y) and after < 0.01 * before


@pytest.fixture(scope="module")
def fitted(dm, costdata):
    """Timings of launch configurations at 6..10 qubits and a model fitted to
    80 % of them."""
    print("cost: collecting launch timings n=6..10", flush=True)
    X, y, meta = costdata.collect([6, 7, 8, 9, 10], seed=16)
    perm = np.random.default_rng(0).permutation(len(y))
    cut = int(0.8 * len(y))
    model, loss = costdata.fit(X[perm[:cut]], y[perm[:cut]], seed=16, epochs=400)
    return model, X[perm[cut:]], y[perm[cut:]], meta


def test_timing_predictions_hold_out(fitted, costdata):
    model, Xt, yt, meta = fitted
    err = costdata.rel_error(model, Xt, yt)
    print(f"held-out |t_pred/t - 1|: median {np.median(err):.3f} p90 "
          f"{np.quantile(err, 0.9):.3f} over {len(yt)}", flush=True)
    assert np.median(err) < 0.15
    assert np.quantile(err, 0.9) < 0.5


def _time_runs(sim, runs=5):
    ms = []
    for _ in range(runs):
        sim.reset_dm()
        sim.run()
        ms.append(sim.last_run_stats()["ms"])
    return float(np.median(ms))


@pytest.mark.parametrize("n", [6, 8, 10])
def test_learned_plan_judged_by_unfused(dm, fitted, n):
    import workloads as W
    model = fitted[0]
    table = W.random_mixed(n, 300, seed=100 + n)
    print(f"cost: learned plan vs greedy vs plain, n={n}", flush=True)

    def make(method, mdl=None, tune=False):
        s = dm.Sim(n)
        s.set_method(method)
        # in-place sweeps only
        s.set_fusion_permute(False)
        s.set_fusion_autotune(tune)
        if mdl is not None:
            s.set_cost_model(mdl)
        s.append_table(table)
        s.upload()
        s.run()
        return s

    plain = make(dm.METHOD_PLAIN_LAUNCH)
    greedy = make(dm.METHOD_FUSED)               # every block fused at the default width
    tuned = make(dm.METHOD_FUSED, tune=True)     # the same, at the measured fastest width
    learned = make(dm.METHOD_FUSED, model)
    re0, im0 = plain.get_dm()
    for s in (greedy, tuned, learned):
        re, im = s.get_dm()
        assert max(np.abs(re - re0).max(), np.abs(im - im0).max()) <= 1e-10
    info = learned.plan_info()
    assert info["predicted_us"] > 0 and info["greedy_us"] > 0
    t_greedy = _time_runs(greedy)
    t_tuned = _time_runs(tuned)
    t_learned = _time_runs(learned)
    print(f"n={n}: greedy {t_greedy:.3f} ms (predicted {info['greedy_us'] / 1e3:.3f}), tuned "
          f"{t_tuned:.3f} ms (width {tuned.tuned_tile_bits()}), learned {t_learned:.3f} ms "
          f"(predicted {info['predicted_us'] / 1e3:.3f}), plan {info}", flush=True)
    # the model's estimates of its own plan and of the default-width greedy
    # plan (when the groupings agree) against their measured run times
    assert abs(info["predicted_us"] / 1e3 / t_learned - 1.0) < 0.35
    if info["grouping"] == greedy.plan_info()["grouping"]:
        assert abs(info["greedy_us"] / 1e3 / t_greedy - 1.0) < 0.35
    # the learned plan is not slower than the measured-width greedy plan
    # beyond timing noise
    assert t_learned <= 1.10 * t_tuned


def test_five_qubit_launch_timing(dm):
    """The five-qubit block sweep is timed like the other kinds, and bad
    configurations are rejected before any launch."""
    with dm.Sim(9) as s:
        us = s.time_launch(dm.COST_FUSED5, 7, 0b11100110, 0, -1, 0, 5)
        assert 0 < us < 1e4
        us5 = s.time_launch(dm.COST_FUSED5, 5, 0x1F, 0, -1, 0, 5)
        assert 0 < us5 < 1e4
        for m, qmask in ((5, 0b11100110), (6, 0xF), (9, 0x1F)):
            with pytest.raises(ValueError):
                s.time_launch(dm.COST_FUSED5, m, qmask, 0, -1, 0, 1)


@pytest.mark.parametrize("n", [8, 10])
def test_learned_plan_with_five_qubit_blocks(dm, costdata, n):
    """Blocks of up to five qubits: the learned plan with the shipped model
    (fitted to this GPU's launches at 4..16 qubits), which may split a
    five-qubit block into four-qubit ones or take the four-qubit plan, against
    the greedy plan and unfused execution."""
    import workloads as W
    model = costdata.default_model()
    assert model is not None
    table = W.random_mixed(n, 300, seed=200 + n)
    print(f"cost: five-qubit learned plan vs greedy vs plain, n={n}", flush=True)

    def make(method, mdl=None):
        s = dm.Sim(n)
        s.set_method(method)
        s.set_fusion_max_k(5)
        if mdl is not None:
            s.set_cost_model(mdl)
        s.append_table(table)
        s.upload()
        s.run()
        return s

    plain = make(dm.METHOD_PLAIN_LAUNCH)
    greedy = make(dm.METHOD_FUSED)  # measured tile width (the default without a model)
    learned = make(dm.METHOD_FUSED, model)
    re0, im0 = plain.get_dm()
    for s in (greedy, learned):
        re, im = s.get_dm()
        assert max(np.abs(re - re0).max(), np.abs(im - im0).max()) <= 1e-10
    gi, info = greedy.plan_info(), learned.plan_info()
    assert gi["fused5_blocks"] > 0
    t_greedy = _time_runs(greedy)
    t_learned = _time_runs(learned)
    print(f"n={n}: greedy {t_greedy:.3f} ms plan {gi}; learned {t_learned:.3f} ms "
          f"(predicted {info['predicted_us'] / 1e3:.3f}) plan {info}", flush=True)
    assert abs(info["predicted_us"] / 1e3 / t_learned - 1.0) < 0.35
    assert t_learned <= 1.10 * t_greedy


@pytest.mark.parametrize("herm", [0, 2])
def test_learned_plan_five_qubit_gate_at_max_k4(dm, costdata, herm):
    """A five-qubit gate (C4X) forces a five-qubit block even when blocks are
    capped at four qubits: the learned plan at max_k = 4 keeps it (there is
    no narrower plan to compare with) and matches the plain method."""
    import workloads as W
    model = costdata.default_model()
    n = 9
    table = W.random_mixed(n, 120, seed=77)
    table.add("C4X", [0, 3, 5, 7, 8])
    table.extend(W.random_mixed(n, 60, seed=78))
    print(f"cost: learned plan with C4X at max_k=4, n={n}, herm={herm}", flush=True)
    ref = dm.Sim(n)
    ref.set_method(dm.METHOD_PLAIN_LAUNCH)
    ref.append_table(table)
    ref.upload()
    ref.run()
    s = dm.Sim(n)
    s.set_method(dm.METHOD_FUSED)
    s.set_fusion_max_k(4)
    s.set_fusion_hermitian(herm)
    s.set_cost_model(model)
    s.append_table(table)
    s.upload()
    assert s.plan_info()["fused5_blocks"] == 1
    s.run()
    re0, im0 = ref.get_dm()
    re, im = s.get_dm()
    assert max(np.abs(re - re0).max(), np.abs(im - im0).max()) <= 1e-10


def test_learned_plan_is_deterministic(dm, costdata):
    """The learned planner builds its candidate groupings on host threads of
    their own: two uploads of one circuit give the same plan (grouping, sweep
    counts, predicted time) and bit-identical rho."""
    import workloads as W
    model = costdata.default_model()
    n = 12
    table = W.brickwall_su4(n=n, layers=6, tail=300, seed=5)
    print(f"cost: learned plan twice, n={n}, {len(table)} gates", flush=True)
    plans, states = [], []
    for _ in range(2):
        s = dm.Sim(n)
        s.set_method(dm.METHOD_FUSED)
        s.set_cost_model(model)
        s.append_table(table)
        s.upload()
        plans.append((s.plan_sweeps(), s.plan_info()))
        s.run()
        states.append(s.get_dm())
        s.close()
    assert plans[0] == plans[1]
    assert plans[0][0]["herm_blocks"] > 0
    assert np.array_equal(states[0][0], states[1][0])
    assert np.array_equal(states[0][1], states[1][1])


def _qasm(name):
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "qasm", name)) as f:
        return f.read()


@pytest.mark.parametrize("case", ["vqe_uccsd_n8", "random_n8", "brickwall_n14_k4"])
def test_learned_beats_measured_greedy(dm, costdata, case):
    """What the trained network gains over the method without it: the greedy
    plan with its tile width measured on this GPU at upload (the fallback
    without a model).  With blocks of up to five qubits the network keeps or
    splits each five-qubit block and compares whole plans, and it chooses
    among candidate groupings by predicted time; the greedy plan takes every
    five-qubit block and the grouping with fewer sweeps.  Both are judged by
    the unfused run's result (1e-10)."""
    import workloads as W
    model = costdata.default_model()
    assert model is not None
    if case == "vqe_uccsd_n8":
        n, k, src, table, bound = 8, 5, _qasm("vqe_uccsd_n8.qasm"), None, 0.8
    elif case == "random_n8":
        n, k, src, table, bound = 8, 5, None, W.random_mixed(8, 300, seed=108), 0.8
    else:
        # the greedy plan's frontier grouping already looks ahead deeply at
        # this size; the network's candidate groupings save a few sweeps
        n, k, src, table, bound = 14, 4, None, W.brickwall_su4(n=14, layers=20, tail=2000,
                                                              seed=14), 1.0
    print(f"cost: learned vs measured-width greedy, {case}", flush=True)

    def make(method, mdl=None):
        s = dm.Sim(n)
        s.set_method(method)
        s.set_fusion_max_k(k)
        if mdl is not None:
            s.set_cost_model(mdl)
        if src is not None:
            s.append_qasm(src)
        else:
            s.append_table(table)
        s.upload()
        s.run()
        return s

    greedy = make(dm.METHOD_FUSED)
    learned = make(dm.METHOD_FUSED, model)
    ref = make(dm.METHOD_PLAIN_LAUNCH) if n < 14 else None
    a, b = greedy.get_dm(), learned.get_dm()
    assert max(np.abs(a[0] - b[0]).max(), np.abs(a[1] - b[1]).max()) <= 1e-10
    if ref is not None:
        r = ref.get_dm()
        assert max(np.abs(r[0] - b[0]).max(), np.abs(r[1] - b[1]).max()) <= 1e-10
        ref.close()
    del a, b
    t_greedy = _time_runs(greedy, runs=7)
    t_learned = _time_runs(learned, runs=7)
    gi, li = greedy.plan_info(), learned.plan_info()
    print(f"{case}: greedy {t_greedy:.3f} ms (width {greedy.tuned_tile_bits()}, plan {gi}); "
          f"learned {t_learned:.3f} ms (plan {li}): {t_learned / t_greedy:.3f}", flush=True)
    assert t_learned <= bound * t_greedy
    if n == 14:
        assert li["herm_blocks"] < gi["herm_blocks"]
    greedy.close()
    learned.close()
    model.close()
