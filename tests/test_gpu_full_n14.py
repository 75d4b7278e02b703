"""Full-rho parity at the n = 14 workload: 20 Haar-SU(4) brick-wall layers plus
2000 U3/CX gates (seed 14), every one of the 2 x 4^14 entries of the GPU
result against the density-matrix oracle (orc_sim with row-wise forward
passes, bit-identical to orc_sim).

  plain, both methods: bit for bit;
  fused (greedy, permuting blocks) and learned-fused: max|d rho| <= 1e-10;
  every result: trace within 1e-12 of the oracle's, Hermiticity
  max|A - A^dagger| <= 1e-12 over all entries;
  shots under a shared seed equal the oracle's sampler run on the oracle's
  own rho (src/dmsim_cpu_omp.hpp:474-515).
"""
import os

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu

N = 14
DIM = 1 << N
TOL = 1e-10
SHOT_SEED = 99
SHOTS = 200


@pytest.fixture(scope="module")
def workload():
    return W.brickwall_su4(n=N, layers=20, tail=2000, seed=14)


@pytest.fixture(scope="module")
def oracle_rho(oracle, workload):
    print(f"oracle: n={N}, {len(workload)} gates, row-wise forward passes", flush=True)
    oracle.set_threads(min(os.cpu_count() or 1, 16))
    try:
        re, im = oracle.sim(N, workload.arrays(), rowwise=True)
    finally:
        oracle.set_threads(1)
    shots, total = oracle.measure(N, re, SHOTS, SHOT_SEED)
    return re, im, shots, total


def run_engine(table, method, model=None):
    import dmsim_amd
    s = dmsim_amd.Sim(N)
    s.set_method(method)
    if model is not None:
        s.set_cost_model(model)
    s.append_table(table)
    s.upload()
    s.run()
    return s


def max_abs_diff(a, b, chunk=1 << 24):
    m = 0.0
    for k in range(0, a.size, chunk):
        m = max(m, float(np.max(np.abs(a[k:k + chunk] - b[k:k + chunk]))))
    return m


def hermiticity(re, im, rows=512):
    """max |A_ij - conj(A_ji)| over all entries, in bands of rows."""
    R = re.reshape(DIM, DIM)
    I = im.reshape(DIM, DIM)
    m = 0.0
    for r in range(0, DIM, rows):
        m = max(m, float(np.max(np.abs(R[r:r + rows] - R[:, r:r + rows].T))),
                float(np.max(np.abs(I[r:r + rows] + I[:, r:r + rows].T))))
    return m


def trace(re):
    return float(re.reshape(DIM, DIM).trace())


def check_common(s, re, im, oracle_rho):
    ore, _, oshots, _ = oracle_rho
    assert abs(trace(re) - trace(ore)) < 1e-12
    assert hermiticity(re, im) <= 1e-12
    shots, _ = s.measure(SHOTS, SHOT_SEED)
    assert np.array_equal(shots, oshots)


@pytest.mark.parametrize("method", [1, 0])  # one launch per primitive, persistent loop
def test_plain_full_rho_bitwise(workload, oracle_rho, method):
    s = run_engine(workload, method)
    print(f"plain method {method}: {s.last_run_stats()['ms']:.0f} ms", flush=True)
    re, im = s.get_dm()
    ore, oim = oracle_rho[0], oracle_rho[1]
    assert np.array_equal(re, ore), max_abs_diff(re, ore)
    assert np.array_equal(im, oim), max_abs_diff(im, oim)
    check_common(s, re, im, oracle_rho)
    s.close()


@pytest.mark.parametrize("learned", [False, True])
def test_fused_full_rho(workload, oracle_rho, learned):
    import costdata
    import dmsim_amd
    model = None
    if learned:
        model = dmsim_amd.CostModel()
        model.load(costdata.DEFAULT_MODEL)
    s = run_engine(workload, 2, model)
    info = s.plan_info()
    print(f"fused (learned={learned}): {s.last_run_stats()['ms']:.0f} ms, plan {info}",
          flush=True)
    re, im = s.get_dm()
    err = max(max_abs_diff(re, oracle_rho[0]), max_abs_diff(im, oracle_rho[1]))
    assert err <= TOL, err
    check_common(s, re, im, oracle_rho)
    s.close()
    if model is not None:
        model.close()
