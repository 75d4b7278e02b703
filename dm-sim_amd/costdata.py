"""Timing-data generator and trainer for the fused planner's cost model.

For each qubit count n it times, on this GPU, the launch configurations the
planner chooses between (``dmx_time_launch``): the dense 4- and 5-qubit block
sweeps at every tile width for a spread of qubit sets, and the primitive sweeps of
every kind on every target bit, on both index sides.  Targets are
ln(microseconds).  It then fits the model (``dmx_cost_train``) and writes its
parameters as an .npz that ``CostModel.load`` reads.

    python dm-sim_amd/costdata.py --n 4-14 --out cost_model.npz [--data data.npz]
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import dmsim_amd as dm  # noqa: E402

DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cost_model_gfx950.npz")


def launch_configs(n, rng, fused_sets=10):
    """(kind, m, qmask, target, control, slow) tuples worth timing at n qubits."""
    out = []
    if n >= 4:
        sets = {0xF, ((1 << 4) - 1) << (n - 4), 0b11 | (0b11 << (n - 2))}
        while len(sets) < min(fused_sets, math.comb(n, 4)):
            q = rng.choice(n, 4, replace=False)
            sets.add(int(sum(1 << int(x) for x in q)))
        for qmask in sorted(sets):
            for m in range(4 if qmask & 1 else 5, min(8, n) + 1):  # m = 4 needs qubit 0
                out.append((dm.COST_FUSED, m, qmask, 0, -1, 0))
    if n >= 5:
        sets = {0x1F, ((1 << 5) - 1) << (n - 5), 0b111 | (0b11 << (n - 2))}
        while len(sets) < min(fused_sets, math.comb(n, 5)):
            q = rng.choice(n, 5, replace=False)
            sets.add(int(sum(1 << int(x) for x in q)))
        for qmask in sorted(sets):
            for m in range(5 if qmask & 1 else 6, min(8, n) + 1):  # m = 5 needs qubit 0
                out.append((dm.COST_FUSED5, m, qmask, 0, -1, 0))
    if n >= 8:  # Hermitian block sweeps (tile width fixed at 8)
        sets = {0xF, ((1 << 4) - 1) << (n - 4), 0b11 | (0b11 << (n - 2))}
        while len(sets) < min(fused_sets, math.comb(n, 4)):
            q = rng.choice(n, 4, replace=False)
            sets.add(int(sum(1 << int(x) for x in q)))
        for qmask in sorted(sets):
            out.append((dm.COST_HERM, 8, qmask, 0, -1, 0))
    for slow in (0, 1):
        for t in range(n):
            out.append((dm.COST_DENSE, 0, 0, t, -1, slow))
            out.append((dm.COST_DIAG, 0, 0, t, -1, slow))
            if n >= 2:
                c = int(rng.choice([q for q in range(n) if q != t]))
                out.append((dm.COST_CX, 0, 0, t, c, slow))
                lo = int(rng.choice([q for q in range(n) if q != t]))
                out.append((dm.COST_QUAD, 0, 0, t, lo, slow))
    return out


def est_us(n, kind):
    full = 32.0 * 4 ** n
    by = full / 2 if kind in (dm.COST_DIAG, dm.COST_CX) else full
    if kind == dm.COST_FUSED5:
        by *= 2  # MFMA-heavier than one streaming pass
    if kind == dm.COST_HERM:
        by *= 0.55  # the upper tiles read, half the entries written
    return 4.0 + by / 4.0e6


def collect(ns, seed=16, budget_us=3000.0, progress=None):
    """Times every configuration of launch_configs(n) for n in ns.
    Returns X float32[S, F], y float32[S] (ln microseconds), meta int32[S, 7]."""
    rng = np.random.default_rng(seed)
    X, y, meta = [], [], []
    for n in ns:
        t0 = time.time()
        with dm.Sim(n) as sim:
            for (kind, m, qmask, t, c, slow) in launch_configs(n, rng):
                reps = int(min(200, max(3, budget_us / est_us(n, kind))))
                us = sim.time_launch(kind, m, qmask, t, c, slow, reps)
                X.append(dm.cost_features(n, kind, m, qmask, t, c, slow))
                y.append(math.log(us))
                meta.append((n, kind, m, qmask, t, c, slow))
        if progress:
            progress(f"n={n}: {sum(1 for r in meta if r[0] == n)} launches timed "
                     f"in {time.time() - t0:.1f} s")
    return (np.array(X, dtype=np.float32), np.array(y, dtype=np.float32),
            np.array(meta, dtype=np.int64))


def fit(X, y, seed=16, epochs=600, batch=64, lr=2e-3):
    model = dm.CostModel(seed)
    loss = model.train(X, y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return model, loss


def rel_error(model, X, y):
    """|exp(pred) / exp(y) - 1| per sample."""
    return np.abs(np.exp(model.predict(X) - y) - 1.0)


def default_model():
    """The model fitted to this GPU's timings, shipped with the package, or
    None if the file is absent (the planner then stays greedy)."""
    if not os.path.exists(DEFAULT_MODEL):
        return None
    model = dm.CostModel()
    model.load(DEFAULT_MODEL)
    return model


def _range(s):
    if "-" in s:
        a, b = s.split("-")
        return list(range(int(a), int(b) + 1))
    return [int(x) for x in s.split(",")]


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--n", default="4-14", help="qubit counts, e.g. 4-14 or 6,8,10")
    ap.add_argument("--seed", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=600)
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--data", default=None, help="also save the raw samples here (.npz)")
    args = ap.parse_args()
    X, y, meta = collect(_range(args.n), args.seed, progress=lambda s: print(s, flush=True))
    if args.data:
        np.savez(args.data, X=X, y=y, meta=meta)
    rng = np.random.default_rng(args.seed)
    perm = rng.permutation(len(y))
    cut = int(0.8 * len(y))
    tr, te = perm[:cut], perm[cut:]
    model, loss = fit(X[tr], y[tr], args.seed, args.epochs)
    err = rel_error(model, X[te], y[te])
    print(f"samples {len(y)}  train loss {loss:.4f}  held-out |t_pred/t - 1|: "
          f"median {np.median(err):.3f}  p90 {np.quantile(err, 0.9):.3f}", flush=True)
    # refit on everything for the shipped parameters
    model, loss = fit(X, y, args.seed, args.epochs)
    model.save(args.out)
    print(f"saved {args.out} (loss {loss:.4f})", flush=True)


if __name__ == "__main__":
    main()
