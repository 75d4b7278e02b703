"""Run an OpenQASM 2.0 circuit on the MI355X engine.

The reference translates a .qasm file into a Python script of Simulation
calls with tool/dmsim_qasm.py (Python 2) and then runs that script; here the
C++ front-end reads the file directly:

    python dm-sim_amd/qasm_run.py circuit.qasm [--method fused] [--shots 10] [--seed 1]

Prints the qubit and gate counts, the device time of the run and the shots
(the reference's script ends with upload(), run(), measure(10)).
"""
import argparse
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import dmsim_amd as dm  # noqa: E402

METHODS = {"plain": dm.METHOD_PLAIN, "plain_launch": dm.METHOD_PLAIN_LAUNCH,
           "fused": dm.METHOD_FUSED}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("qasm")
    ap.add_argument("--method", choices=sorted(METHODS), default="fused")
    ap.add_argument("--shots", type=int, default=10)
    ap.add_argument("--seed", type=int, default=None, help="default: time(0), as the reference")
    args = ap.parse_args()
    with open(args.qasm) as f:
        src = f.read()
    n, ops, _, _ = dm.qasm_parse(src)
    with dm.Sim(n) as sim:
        sim.set_method(METHODS[args.method])
        sim.append_qasm(src)
        sim.upload()
        t0 = time.perf_counter()
        sim.run()
        wall = time.perf_counter() - t0
        shots, _ = sim.measure(args.shots, int(time.time()) if args.seed is None else args.seed)
        print(f"== {os.path.basename(args.qasm)}: {n} qubits, {len(ops)} gates, "
              f"{sim.last_run_stats()['ms']:.3f} ms device, {wall * 1e3:.3f} ms wall ==")
        print(" ".join(str(int(x)) for x in shots))


if __name__ == "__main__":
    main()
