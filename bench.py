"""Benchmark: rho gate-updates/s at n = 14, FP64, on one MI355X.

Workload (one "step" = one run of the whole circuit on rho resident in HBM):
20 brick-wall layers of Haar SU(4) on adjacent pairs (130 gates, through the
4x4 matrix extension) followed by 2000 gates, U3 or CX with equal
probability, on uniform distinct qubits; seed 14; rho = |0><0| at the start,
each step continues from the previous state (the reference's
clear_circuit/upload/run continuation semantics).

Prints one JSON line.  `value` is API gate-updates/s of the headline method
(whole-job, device-resident data); `roofline` describes the dominant kernel;
`cpu_baseline` times the oracle's CPU restatement on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dm-sim_amd"))

METRIC = "ρ gate-updates/s at n=14 FP64: fused+batched vs plain gate-by-gate; HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed_runs(sim, steps, warmup):
    for _ in range(warmup):
        sim.run()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.run()
        dev_ms.append(sim.last_run_stats()["ms"])
    wall = time.perf_counter() - t0
    return wall, dev_ms, sim.last_run_stats()


def cpu_baseline(table, n, seconds_hint=15.0):
    """The oracle (C restatement of the reference CPU path) on the first gates
    of the same workload, from |0><0|, on this host's cores."""
    import numpy as np
    import workloads as W
    from oracle.oracle import Oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    o = Oracle()
    o.set_threads(threads)
    # skip the SU(4) prefix: the sample is 12 gates of the U3/CX tail
    start = 130
    sample = table.slice(start, start + 12)
    t0 = time.perf_counter()
    o.sim(n, sample.arrays())
    dt = time.perf_counter() - t0
    return {
        "value": len(sample) / dt,
        "unit": "gate-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle sim() of gates {start}..{start + len(sample) - 1} (U3/CX) of the "
                  f"n={n} workload from |0><0|, {dt:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=14)
    ap.add_argument("--methods", default="fused,plain_launch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import dmsim_amd
    import workloads as W

    n = args.n
    table = W.brickwall_su4(n=n, layers=20, tail=2000, seed=14)
    gates = len(table)
    num = 4 ** n
    results = {}
    for mname in args.methods.split(","):
        method = {"plain": 0, "plain_launch": 1, "fused": 2}[mname]
        sim = dmsim_amd.Sim(n)
        sim.set_method(method)
        sim.append_table(table)
        sim.upload()
        log(f"[bench] {mname}: n={n} gates={gates} warmup={args.warmup} steps={args.steps}")
        wall, dev_ms, stats = timed_runs(sim, args.steps, args.warmup)
        ms_step = wall * 1e3 / args.steps
        dev_avg = sum(dev_ms) / len(dev_ms)
        results[mname] = {
            "gate_updates_per_s": gates * args.steps / wall,
            "ms_per_step": ms_step,
            "device_ms_per_step": dev_avg,
            "launches_per_step": stats["launches"],
            "sweeps_per_step": stats["sweeps"],
        }
        log(f"[bench] {mname}: {ms_step:.1f} ms/step, device {dev_avg:.1f} ms, "
            f"{results[mname]['gate_updates_per_s']:.1f} gate-updates/s")
        sim.close()

    best = max(results, key=lambda k: results[k]["gate_updates_per_s"])
    r = results[best]
    # plain method: two sweeps per gate-update, 32*4^n bytes each (64*4^n B/gate)
    alg_bytes = 64.0 * num * gates
    achieved = alg_bytes / (r["device_ms_per_step"] * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": r["gate_updates_per_s"],
        "unit": "gate-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"brickwall_su4 n={n}: 20 Haar-SU(4) brick-wall layers + 2000 U3/CX, "
                        f"seed 14",
            "n_qubits": n, "gates": gates, "method": best,
        },
        "methods": results,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "plain sweep family (64*4^n B per gate-update over the run's device time)",
        },
    }
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(table, n)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
