"""CPU checks of the cost model's NumPy restatement and of the host-side
feature encoding (no GPU needed)."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def _params(rng, scale=0.3):
    return rng.normal(0.0, scale, mo.P)


def test_forward_closed_form():
    # only hidden unit 0 of each layer is live: f(x) = 2 relu(x0) + 0.5
    W1 = np.zeros((mo.H, mo.F)); W1[0, 0] = 1.0
    W2 = np.eye(mo.H)
    w3 = np.zeros(mo.H); w3[0] = 2.0
    p = mo.pack(W1, np.zeros(mo.H), W2, np.zeros(mo.H), w3, 0.5)
    X = np.zeros((3, mo.F)); X[:, 0] = [3.0, -1.0, 0.25]
    assert np.array_equal(mo.forward(p, X), [6.5, 0.5, 1.0])


def test_forward_layer_order():
    # a transposed W2 changes the result: W2[1][0] routes unit 0 into unit 1
    W1 = np.zeros((mo.H, mo.F)); W1[0, 0] = 1.0
    W2 = np.zeros((mo.H, mo.H)); W2[1, 0] = 3.0
    w3 = np.zeros(mo.H); w3[1] = 1.0
    p = mo.pack(W1, np.zeros(mo.H), W2, np.zeros(mo.H), w3, 0.0)
    X = np.zeros((1, mo.F)); X[0, 0] = 2.0
    assert mo.forward(p, X)[0] == 6.0


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    p = _params(rng)
    X = rng.uniform(0, 1, (40, mo.F))
    y = rng.normal(0, 1, 40)
    g = mo.loss_grad(p, X, y)
    idx = np.concatenate([rng.choice(mo.P, 300, replace=False), np.arange(mo.P - 65, mo.P)])
    h = 1e-6
    for i in idx:
        e = np.zeros(mo.P); e[i] = h
        fd = (mo.loss(p + e, X, y) - mo.loss(p - e, X, y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_adam_first_step_is_lr_times_sign():
    g = np.array([3.0, -0.5, 1e-3, -7.0])
    p, m, v = mo.adam_step(np.zeros(4), g, np.zeros(4), np.zeros(4), lr=0.01, step=1)
    assert np.allclose(p, -0.01 * np.sign(g), rtol=1e-4)


def test_feature_encoding():
    dm = pytest.importorskip("dmsim_amd")
    # fused, n = 14, block on qubits 0..3, m = 8: column bits 0..7 contiguous
    f = dm.cost_features(14, dm.COST_FUSED, m=8, qmask=0xF)
    assert f[0] == np.float32(14 / 16) and f[1] == 1 and f[8] == 1.0
    assert f[9] == np.float32(np.log2(4096) / 16)    # grid capped at 4096 workgroups
    assert f[10] == 1.0 and f[11] == np.float32(3 / 16) and f[12] == 0.0 and f[13] == 1.0
    # block on qubits 10..13: four low free bits join it
    f = dm.cost_features(14, dm.COST_FUSED, m=8, qmask=0x3C00)
    assert f[10] == 0.0 and f[13] == 0.5 and f[11] == np.float32(13 / 16)
    # n = 6, m = 4: 2^(2 + 2) super-tiles; column bits = the block's qubits 0, 3, 4, 5
    f = dm.cost_features(6, dm.COST_FUSED, m=4, qmask=0b111001)
    assert f[9] == np.float32(4 / 16) and f[13] == np.float32(1 / 8)
    with pytest.raises(ValueError):  # m = 4 needs qubit 0 in the block
        dm.cost_features(6, dm.COST_FUSED, m=4, qmask=0b111100)
    # five-qubit block on qubits 1, 2, 5, 6, 7 of n = 9 with m = 7: column bits
    # 0, 1, 2, 5, 6, 7 plus 3; 2^(4 + 2) super-tiles
    f = dm.cost_features(9, dm.COST_FUSED5, m=7, qmask=0b11100110)
    assert f[6] == 1 and f[1:6].sum() == 0 and f[8] == np.float32(7 / 8)
    assert f[9] == np.float32(6 / 16) and f[10] == np.float32(2 / 4)
    assert f[11] == np.float32(7 / 16) and f[12] == np.float32(1 / 16) and f[13] == np.float32(4 / 8)
    with pytest.raises(ValueError):  # m = 5 needs qubit 0 in the block
        dm.cost_features(9, dm.COST_FUSED5, m=5, qmask=0b11100110)
    with pytest.raises(ValueError):  # four qubits are not a five-qubit block
        dm.cost_features(9, dm.COST_FUSED5, m=6, qmask=0xF)
    # slow-index cx: flattened bits shifted by n
    f = dm.cost_features(10, dm.COST_CX, target=2, control=5, slow=1)
    assert f[4] == 1 and f[11] == np.float32(12 / 32) and f[13] == np.float32(16 / 32)
    assert f[10] == 0.0
    # Hermitian sweep, n = 14: 16 * 64 * 65 / 2 items, grid capped at 4096
    f = dm.cost_features(14, dm.COST_HERM, m=8, qmask=0b10010100001)
    assert f[7] == 1 and f[1:7].sum() == 0 and f[8] == 1.0
    assert f[9] == np.float32(12 / 16) and f[10] == np.float32(1 / 4)
    assert f[11] == np.float32(10 / 16) and f[12] == 0.0
    with pytest.raises(ValueError):
        dm.cost_features(14, dm.COST_HERM, m=7, qmask=0xF)
    f = dm.cost_features(10, dm.COST_DIAG, target=3, slow=0)
    assert f[3] == 1 and f[10] == 1.0 and f[13] == 0.0
    with pytest.raises(ValueError):
        dm.cost_features(10, dm.COST_FUSED, m=8, qmask=0x7)
    with pytest.raises(ValueError):
        dm.cost_features(10, dm.COST_CX, target=2, control=2)


def test_shipped_model_on_recorded_timings():
    """The shipped parameters (fitted to the launch timings recorded on an
    MI355X, tests/golden/launch_timings_gfx950.npz) evaluated by the NumPy
    restatement: the feature encoding still reproduces the recorded features,
    and the model's launch-time estimates track the measurements."""
    import os
    dm = pytest.importorskip("dmsim_amd")
    import costdata
    here = os.path.dirname(os.path.abspath(__file__))
    z = np.load(os.path.join(here, "golden", "launch_timings_gfx950.npz"))
    X, y, meta = z["X"], z["y"], z["meta"]
    for row, x in zip(meta[::7], X[::7]):
        n, kind, m, qmask, t, c, slow = (int(v) for v in row)
        assert np.array_equal(dm.cost_features(n, kind, m, qmask, t, c, slow), x)
    w = np.load(costdata.DEFAULT_MODEL)
    pred = mo.forward(w["params"], X) * float(w["dev"]) + float(w["mean"])
    err = np.abs(np.exp(pred - y) - 1.0)
    assert np.median(err) < 0.10 and np.quantile(err, 0.9) < 0.35
