"""Benchmark: rho gate-updates/s at n = 14, FP64, on one MI355X.

Workload (one "step" = one run of the whole circuit on rho resident in HBM):
20 brick-wall layers of Haar SU(4) on adjacent pairs (130 gates, through the
4x4 matrix extension) followed by 2000 gates, U3 or CX with equal
probability, on uniform distinct qubits; seed 14; rho = |0><0| at the start,
each step continues from the previous state (the reference's
clear_circuit/upload/run continuation semantics).

Prints one JSON line.  `value` is API gate-updates/s of the headline method
(whole-job, device-resident data).  `roofline` describes the kernel class that
takes the most device time in the headline method: its algorithmic bytes per
launch over its mean launch time, both from HIP events recorded around every
launch on the run's stream during the timed steps.  `cpu_baseline` times the
oracle's CPU restatement of the reference path on a bounded sample.

  python bench.py                      # default: fused (greedy and learned plans), plain
                                       # at n = 14, then the batched, n = 16 and n = 15 lines
  python bench.py --calibrate-copy     # one 2 x 2 GiB device copy (counter scaling)
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
# HBM bytes per launch over algorithmic bytes per launch, per fused sweep kernel,
# measured by rocprofv3 FETCH_SIZE and WRITE_SIZE runs of this workload scaled by
# a copy of known size; a constant here, from the summaries named beside it.
TRAFFIC_RATIO = {
    "k_fused4_herm": (4.363e9 / 4.343e9, "profiles/k_fused4_herm_n14_hbm_counters.md"),
    "k_relabel": (8.595e9 / 8.590e9, "profiles/k_fused4_herm_n14_hbm_counters.md"),
    "k_fused4": (8.591e9 / 8.590e9, "profiles/k_fused4_n14_hbm_counters.md"),
    "k_fused4_perm": (8.591e9 / 8.590e9, "profiles/k_fused4_n14_hbm_counters.md"),
}
METHODS = {"plain": 0, "plain_launch": 1, "fused": 2, "fused_learned": 2, "fused_learned5": 2,
           "fused_inorder": 2}  # the last only on request: in-order sweeps (profiling runs)
W_OP_C2 = 39  # the 4x4 matrix extension op (workloads.OP["C2"])


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def roofline_of(prof, n, counters_n=14):
    """Roofline of the kernel class with the most device time in a profile
    (sum of per-launch HIP-event times and algorithmic bytes): its bytes per
    launch over its mean launch time against the HBM peak.  traffic: the
    counter-measured HBM/algorithmic ratio of that kernel times its bytes per
    launch, where the ratio was measured at this n (else null)."""
    kname, kv = max(prof.items(), key=lambda kv: kv[1]["ms"])
    per_launch = kv["bytes"] / kv["calls"]
    mean_s = kv["ms"] / kv["calls"] * 1e-3
    achieved = per_launch / mean_s / 1e9
    ratio = TRAFFIC_RATIO.get(kname) if n == counters_n else None
    return {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": ratio[0] * per_launch if ratio else None,
        "traffic_source": (f"measured ratio from {ratio[1]} times this run's algorithmic bytes"
                           if ratio else None),
        "kernel": kname,
        "bytes_per_launch": per_launch,
        "mean_launch_us": mean_s * 1e6,
        "launches": kv["calls"],
    }


def plan_summary(info):
    keys = ("fused_blocks", "fused5_blocks", "plain_blocks", "herm_blocks", "relabels",
            "perm_blocks", "grouping", "predicted_us")
    return {k: info[k] for k in keys if k in info}


def fusion_n16(model, n=16, n_gates=4000, seed=16):
    """The learned-fusion configuration: n = 16 (rho 64 GiB), 4000 gates, 70 %
    U3 and 30 % CX/RZZ/CU1 on uniform pairs.  Blocks of up to five qubits,
    planned by the shipped cost model and by the greedy heuristic (no model:
    grouping by sweep count, tile width measured at upload); one timed run of
    each from |0><0| (a run takes seconds)."""
    import dmsim_amd
    import workloads as W

    table = W.fusion_mix(n=n, n_gates=n_gates, seed=seed)
    out = {"workload": f"fusion_mix n={n}: {len(table)} gates (70% U3, 30% CX/RZZ/CU1), "
                       f"seed {seed}, blocks of up to 5 qubits",
           "gates": len(table)}
    for label in ("learned", "greedy"):
        s = dmsim_amd.Sim(n)
        s.set_method(METHODS["fused"])
        s.set_fusion_max_k(5)
        if label == "learned":
            s.set_cost_model(model)
        s.append_table(table)
        s.upload()
        wall, dev_ms, stats, prof = timed_runs(s, 1, 0)
        out[label] = {
            "gate_updates_per_s": len(table) / wall,
            "ms_per_run": wall * 1e3,
            "device_ms": dev_ms[0],
            "launches": stats["launches"],
            "plan": plan_summary(s.plan_info()),
            "roofline": roofline_of(prof, n),
        }
        log(f"[bench] fusion n={n} {label}: {wall * 1e3:.0f} ms, "
            f"{len(table) / wall:.0f} gate-updates/s, plan {out[label]['plan']}")
        s.close()
    out["learned_vs_greedy"] = (out["learned"]["gate_updates_per_s"]
                                / out["greedy"]["gate_updates_per_s"])
    out["roofline"] = out["learned"]["roofline"]
    return out


def structured_n15(model, n=15, seeds=16):
    """The combined structured configuration: n = 15 (rho 16 GiB), per seed a
    random basis state, QFT, a block of CCX/C3X/C4X/CSWAP, inverse QFT.  All
    16 seeds through the learned-fused method (blocks of up to five qubits,
    one engine, each circuit uploaded, then its reset and run timed), all 16
    as one batch with the same cost model (in waves that fit HBM: 16 states
    of 17.2 GB, doubled for out-of-place sweeps, do not fit at once; each
    15-qubit circuit runs on its own in its wave through its learned
    Hermitian plan), and seed 0 through the plain method (one launch per
    primitive)."""
    import dmsim_amd
    import workloads as W

    tables = [W.structured(n, seed) for seed in range(seeds)]
    gates = sum(len(t) for t in tables)
    out = {"workload": f"structured n={n} x{seeds} seeds (QFT, CCX/C3X/C4X/CSWAP block, "
                       f"inverse QFT)", "gates": gates}
    s = dmsim_amd.Sim(n)
    s.set_method(METHODS["fused"])
    s.set_fusion_max_k(5)
    s.set_cost_model(model)
    s.set_profiling(True)
    wall = dev = 0.0
    prof = {}
    for t in tables:
        s.clear_circuit()
        s.append_table(t)
        s.upload()
        t0 = time.perf_counter()
        s.reset_dm()
        s.run()
        wall += time.perf_counter() - t0
        dev += s.last_run_stats()["ms"]
        for k, v in s.profile().items():
            acc = prof.setdefault(k, {"calls": 0, "ms": 0.0, "bytes": 0.0})
            for f in acc:
                acc[f] += v[f]
    s.close()
    out["learned_fused"] = {"gate_updates_per_s": gates / wall, "ms_total": wall * 1e3,
                            "device_ms_total": dev, "roofline": roofline_of(prof, n)}
    log(f"[bench] structured n={n} x{seeds} learned-fused: {wall * 1e3:.0f} ms, "
        f"{gates / wall:.0f} gate-updates/s")
    b = dmsim_amd.Batch()
    b.set_cost_model(model)
    for t in tables:
        b.add_circuit(t)
    b.upload()
    t0 = time.perf_counter()
    b.run()
    bwall = time.perf_counter() - t0
    st = b.stats()
    achieved = st["sweep_bytes"] / (st["ms"] * 1e-3) / 1e9
    out["batched"] = {
        "gate_updates_per_s": gates / bwall, "ms_total": bwall * 1e3, "device_ms": st["ms"],
        "waves": b.waves(), "launches": st["launches"],
        # the bytes of the batch's sweeps over the run's device time, which
        # also holds the resets and the diagonals' copies
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_run": st["sweep_bytes"], "device_ms_per_run": st["ms"]},
    }
    log(f"[bench] structured n={n} x{seeds} batched: {b.waves()} waves, {bwall * 1e3:.0f} ms, "
        f"{gates / bwall:.0f} gate-updates/s")
    b.close()
    p = dmsim_amd.Sim(n)
    p.set_method(METHODS["plain_launch"])
    p.append_table(tables[0])
    p.upload()
    pwall, pdev, _, pprof = timed_runs(p, 1, 0)
    p.close()
    out["plain_seed0"] = {"gate_updates_per_s": len(tables[0]) / pwall, "ms": pwall * 1e3,
                          "device_ms": pdev[0], "roofline": roofline_of(pprof, n)}
    log(f"[bench] structured n={n} seed 0 plain: {pwall * 1e3:.0f} ms, "
        f"{len(tables[0]) / pwall:.1f} gate-updates/s")
    plain = out["plain_seed0"]["gate_updates_per_s"]
    out["learned_fused_vs_plain"] = out["learned_fused"]["gate_updates_per_s"] / plain
    out["batched_vs_plain"] = out["batched"]["gate_updates_per_s"] / plain
    return out


def timed_runs(sim, steps, warmup):
    for _ in range(warmup):
        sim.run()
    sim.set_profiling(True)
    dev_ms = []
    prof = {}
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.run()
        dev_ms.append(sim.last_run_stats()["ms"])
        for k, v in sim.profile().items():
            acc = prof.setdefault(k, {"calls": 0, "ms": 0.0, "bytes": 0.0})
            for f in acc:
                acc[f] += v[f]
    wall = time.perf_counter() - t0
    sim.set_profiling(False)
    return wall, dev_ms, sim.last_run_stats(), prof


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(table, n, stride=8):
    """The oracle (C restatement of the reference CPU path, gate after gate
    over all rows as src/dmsim_cpu_omp.hpp:367-447 runs it) on every
    stride-th gate of the same workload -- SU(4) matrices and U3/CX in the
    workload's proportions -- from |0><0|, on this host's cores."""
    from oracle.oracle import Oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    o = Oracle()
    o.set_threads(threads)
    sample = table.take(range(0, len(table), stride))
    n_c2 = sum(1 for op in sample.ops if op == W_OP_C2)
    t0 = time.perf_counter()
    o.sim(n, sample.arrays())
    dt = time.perf_counter() - t0
    out = {
        "value": len(sample) / dt,
        "unit": "gate-updates/s",
        "cores": threads,
        "cpu": cpu_model(),
        "kind": "port",
        "sample": f"oracle sim() of every {stride}th gate of the n={n} workload ({len(sample)} "
                  f"gates: {n_c2} SU(4), {len(sample) - n_c2} U3/CX) from |0><0| with {threads} "
                  f"OpenMP threads: {dt:.1f} s",
    }
    out["reference"] = reference_baseline(table, n)
    return out


def reference_baseline(table, n, stride=256):
    """The reference's own CPU path (src/dmsim_cpu_omp.hpp compiled from its
    sources into oracle/_ref, where that build exists): one run from |0><0| of
    every stride-th gate of the workload, forward pass, its block transpose and
    backward pass on one thread (its multi-thread path has no entry for the
    4x4 matrices of the SU(4) gates).  None where oracle/_ref was not built."""
    from oracle.oracle import Reference

    if not Reference.available():
        return None
    sample = table.take(range(0, len(table), stride))
    n_c2 = sum(1 for op in sample.ops if op == W_OP_C2)
    t0 = time.perf_counter()
    Reference().sim(n, sample.arrays())
    dt = time.perf_counter() - t0
    return {
        "value": len(sample) / dt,
        "unit": "gate-updates/s",
        "cores": 1,
        "cpu": cpu_model(),
        "kind": "reference",
        "sample": f"oracle/_ref (src/dmsim_cpu_omp.hpp compiled from its sources) on every "
                  f"{stride}th gate of the n={n} workload ({len(sample)} gates: {n_c2} SU(4), "
                  f"{len(sample) - n_c2} U3/CX) from |0><0|, one thread: {dt:.1f} s",
    }


def batched(steps, warmup, seed=4, n_circuits=256, trace=False):
    """The batched configuration: 256 hardware-efficient-ansatz circuits
    (n uniform in 6..11, depth 10..40, RY/RZ layers + CX ladder), 32 random
    Pauli strings each; one run = every circuit from |0><0| plus every energy,
    one graph launch.  Timed with the autotuned tile widths and with the
    default widths, and compared with the same circuits run one after another
    (each already uploaded; only its reset and run are timed) through the
    single-circuit plain and fused methods."""
    import dmsim_amd
    import workloads as W

    items = W.hea_batch(n_circuits=n_circuits, seed=seed)
    gates = sum(len(c) for c, _, _ in items)
    out = {
        "workload": f"{n_circuits} HEA circuits, n in 6..11, depth 10..40, 32 Pauli strings, "
                    f"seed {seed}",
        "gates": gates,
    }
    for tuned in ((True,) if trace else (True, False)):
        b = dmsim_amd.Batch()
        b.set_autotune(tuned)
        if trace:  # eager launches, which a kernel-trace profiler records one by one
            b.set_graph(False)
        for circ, strings, weights in items:
            b.add_circuit(circ, strings, weights)
        b.upload()
        for _ in range(warmup):
            b.run()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.run()
        wall = (time.perf_counter() - t0) / steps
        st = b.stats()
        key = "batched" if tuned else "batched_default_tiles"
        out[f"{key}_gate_updates_per_s"] = gates / wall
        out[f"{key}_ms_per_run"] = wall * 1e3
        out[f"{key}_device_ms"] = st["ms"]
        if tuned:
            out["launches_per_run"] = st["launches"]
            out["tile_bits"] = {n: b.tile_bits(n) for n in range(6, 12)}
            out["schedule"] = b.schedule()
            # the batched run's block sweeps: 32 * 4^n bytes each, over the
            # run's device time (which also holds the reset and the energies)
            out["roofline"] = {
                "bound": "hbm", "achieved": st["sweep_bytes"] / (st["ms"] * 1e-3) / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": st["sweep_bytes"] / (st["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "bytes_per_run": st["sweep_bytes"], "device_ms_per_run": st["ms"],
                "note": "one circuit's rho (<= 67 MB at 11 qubits) fits the 256 MB Infinity "
                        "Cache, but a launch sweeps all circuits of one width (up to 3.3 GB): "
                        "the 10- and 11-qubit steps stream from HBM",
            }
        b.close()
    if trace:
        return out
    for mname in ("plain", "plain_launch", "fused"):
        sims = {}
        dt = 0.0
        for circ, _, _ in items:
            n = circ.n_qubits
            if n not in sims:
                sims[n] = dmsim_amd.Sim(n)
                sims[n].set_method(METHODS[mname])
            s = sims[n]
            s.clear_circuit()
            s.append_table(circ)
            s.upload()
            t0 = time.perf_counter()
            s.reset_dm()
            s.run()
            dt += time.perf_counter() - t0
        for s in sims.values():
            s.close()
        out[f"one_by_one_{mname}_gate_updates_per_s"] = gates / dt
    out["batched_vs_plain"] = (out["batched_gate_updates_per_s"]
                               / max(out["one_by_one_plain_gate_updates_per_s"],
                                     out["one_by_one_plain_launch_gate_updates_per_s"]))
    out["batched_vs_fused"] = (out["batched_gate_updates_per_s"]
                               / out["one_by_one_fused_gate_updates_per_s"])
    return out


def vqe_published(warmup=1, steps=5):
    """The circuit of the reference's only published single-GPU timing:
    VQE-UCCSD on 8 qubits, 10 808 gates, 249.3 ms on one V100 (README.md:47),
    read from OpenQASM by the engine's front-end.  Median device time of a run
    from |0><0| after a warm-up run, per method."""
    import dmsim_amd

    path = os.path.join(ROOT, "tests", "golden", "qasm", "vqe_uccsd_n8.qasm")
    with open(path) as f:
        src = f.read()
    out = {"workload": "vqe_uccsd_n8.qasm: 8 qubits, 10808 gates",
           "reference_ms": 249.3, "reference_hw": "1x NVIDIA V100 (README.md:47)"}
    import costdata
    model = costdata.default_model()
    for mname in ("fused", "fused_learned", "fused_learned5", "plain"):
        if mname.startswith("fused_learned") and model is None:
            continue
        s = dmsim_amd.Sim(8)
        s.set_method(METHODS[mname])
        if mname.startswith("fused_learned"):
            s.set_cost_model(model)
        if mname == "fused_learned5":
            s.set_fusion_max_k(5)
        s.append_qasm(src)
        s.upload()
        ms = []
        for i in range(warmup + steps):
            s.reset_dm()
            s.run()
            if i >= warmup:
                ms.append(s.last_run_stats()["ms"])
        s.close()
        ms.sort()
        out[f"{mname}_ms"] = ms[len(ms) // 2]
    if model is not None:
        model.close()
    out["speedup_vs_reference"] = out["reference_ms"] / min(
        v for k, v in out.items() if k.endswith("_ms") and k != "reference_ms")
    return out


def calibrate_copy():
    """A plain device-to-device copy of 2 GiB per plane pair (4 GiB read +
    4 GiB written in total) for scaling the FETCH_SIZE/WRITE_SIZE counters."""
    import torch
    a = torch.empty(2 ** 28, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    # the streaming rates a kernel that reads and writes HBM once can reach on
    # this GPU: the best of five timed copies and of five full reductions
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def best_ms(fn):
        ms = []
        for _ in range(5):
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            ms.append(e0.elapsed_time(e1))
        return min(ms)

    nbytes = a.numel() * 8
    copy_ms = best_ms(lambda: b.copy_(a))
    read_ms = best_ms(lambda: a.sum())
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                      "copy_GBs": 2 * nbytes / (copy_ms * 1e-3) / 1e9,
                      "read_GBs": nbytes / (read_ms * 1e-3) / 1e9}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=14)
    ap.add_argument("--methods", default="fused,fused_learned,fused_learned5,plain,plain_launch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-published", action="store_true",
                    help="skip the VQE-UCCSD comparison with the reference's published time")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the n = 16 learned-vs-greedy and n = 15 structured x16 lines")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--batch-trace", action="store_true",
                    help="only the batched configuration, autotuned, launched eagerly "
                         "(for a kernel-trace profiler); prints its JSON line")
    ap.add_argument("--gates", default="",
                    help="LO:HI -- run only gates LO..HI-1 of the workload (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.batch_trace:
        print(json.dumps(batched(args.steps, args.warmup, trace=True)), flush=True)
        return

    import dmsim_amd
    import workloads as W

    n = args.n
    table = W.brickwall_su4(n=n, layers=20, tail=2000, seed=14)
    if args.gates:
        lo, hi = (int(v) for v in args.gates.split(":"))
        table = table.slice(lo, hi)
    gates = len(table)
    num = 4 ** n
    results = {}
    profiles = {}
    for mname in args.methods.split(","):
        sim = dmsim_amd.Sim(n)
        sim.set_method(METHODS[mname])
        model = None
        if mname.startswith("fused_learned"):
            import costdata
            model = costdata.default_model()
            if model is None:
                log(f"[bench] {mname}: no shipped cost model, skipped")
                sim.close()
                continue
            sim.set_cost_model(model)
        if mname == "fused_learned5":  # blocks of up to five qubits
            sim.set_fusion_max_k(5)
        if mname == "fused_inorder":  # greedy plan without Hermitian sweeps
            sim.set_fusion_hermitian(0)
        sim.append_table(table)
        sim.upload()
        if mname.startswith("fused"):
            log(f"[bench] {mname}: plan {sim.plan_info()}")
        log(f"[bench] {mname}: n={n} gates={gates} warmup={args.warmup} steps={args.steps}")
        wall, dev_ms, stats, prof = timed_runs(sim, args.steps, args.warmup)
        ms_step = wall * 1e3 / args.steps
        dev_avg = sum(dev_ms) / len(dev_ms)
        results[mname] = {
            "gate_updates_per_s": gates * args.steps / wall,
            "ms_per_step": ms_step,
            "device_ms_per_step": dev_avg,
            "launches_per_step": stats["launches"],
            "gates_per_launch": gates / max(stats["launches"], 1),
            # 64 * 4^n bytes per gate-update (two full sweeps in the plain method)
            "effective_GBs": 64.0 * num * gates / (dev_avg * 1e-3) / 1e9,
            # the bytes the launches must move (a CX or diagonal sweep touches
            # half of rho; a Hermitian sweep reads the upper tiles), per step,
            # over the step's device time
            "algorithmic_GBs": (sum(v["bytes"] for v in prof.values()) / args.steps
                                / (dev_avg * 1e-3) / 1e9),
        }
        results[mname]["algorithmic_frac"] = results[mname]["algorithmic_GBs"] / HBM_PEAK_GBS
        profiles[mname] = prof
        log(f"[bench] {mname}: {ms_step:.1f} ms/step, device {dev_avg:.1f} ms, "
            f"{results[mname]['gate_updates_per_s']:.1f} gate-updates/s")
        for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"]):
            log(f"[bench]   {k:24s} calls={v['calls']:6d} mean={v['ms'] / v['calls'] * 1e3:9.1f} us "
                f"{v['bytes'] / (v['ms'] * 1e-3) / 1e9:8.1f} GB/s")
        sim.close()
        if model is not None:
            model.close()

    best = max(results, key=lambda k: results[k]["gate_updates_per_s"])
    r = results[best]
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
        # against the faster of the two plain methods (persistent loop, one launch per primitive)
        "speedup_vs_plain": (r["gate_updates_per_s"] / max(results[k]["gate_updates_per_s"]
                                                           for k in ("plain", "plain_launch")
                                                           if k in results)
                             if ("plain" in results or "plain_launch" in results) else None),
        "roofline": roofline_of(profiles[best], n),
    }
    if not args.no_batch:
        log("[bench] batched configuration ...")
        line["batched"] = batched(args.steps, args.warmup)
        log(f"[bench] batched: {line['batched']['batched_gate_updates_per_s']:.0f} "
            f"gate-updates/s, {line['batched']['batched_vs_plain']:.1f}x one-by-one plain")
    if not args.no_large:
        import costdata
        model = costdata.default_model()
        if model is None:
            log("[bench] no shipped cost model: n = 16 and n = 15 lines skipped")
        else:
            log("[bench] learned fusion n=16 (learned vs greedy, blocks of up to 5 qubits) ...")
            line["fusion_n16"] = fusion_n16(model)
            log("[bench] structured n=15 x16 (learned-fused, batched, plain seed 0) ...")
            line["structured_n15"] = structured_n15(model)
            model.close()
    if not args.no_published:
        log("[bench] published comparison (vqe_uccsd_n8) ...")
        line["published_comparison"] = vqe_published()
        pc = line["published_comparison"]
        log(f"[bench] vqe_uccsd_n8: fused {pc['fused_ms']:.2f} ms, learned "
            f"{pc.get('fused_learned_ms', float('nan')):.2f} ms, learned (5-qubit blocks) "
            f"{pc.get('fused_learned5_ms', float('nan')):.2f} ms, plain {pc['plain_ms']:.2f} ms "
            f"(V100: 249.3 ms)")
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(table, n)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
