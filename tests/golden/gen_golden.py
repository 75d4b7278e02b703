"""Writes the golden vectors under tests/golden/ from the reference's own CPU
path (oracle/_ref/libdmsim_ref.so, compiled from src/dmsim_cpu_omp.hpp by
oracle/Makefile).  Run in a container that has the reference tree:

    make -C oracle && python tests/golden/gen_golden.py

Every fixture is data: gate tables in, reference outputs out.  All runs are a
single sim() from |0><0| (the only case the reference's serial CPU transpose
gets right) with n <= 10.
"""
from __future__ import annotations

import hashlib
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dm-sim_amd"))

import workloads as W  # noqa: E402
from oracle.oracle import Reference  # noqa: E402

PARAMS = (0.7, 1.3, -0.4)  # theta, phi, lambda used for every parametrised gate matrix


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def table_dict(prefix, t):
    ops, qb, prm, coef = t.arrays()
    return {f"{prefix}ops": ops, f"{prefix}qb": qb, f"{prefix}prm": prm, f"{prefix}coef": coef}


def gate_matrices(ref):
    out = {}
    rng = np.random.default_rng(20)
    for name in W.OP_NAMES:
        k = W.ARITY[name]
        op = W.OP[name]
        for tag, qubits in (("fwd", list(range(k))), ("rev", list(reversed(range(k))))):
            coef = np.zeros(32)
            if name in ("C1", "C2"):
                u = W.haar_unitary(rng, 1 << k)
                coef[0: 2 * u.size: 2] = u.real.ravel()
                coef[1: 2 * u.size: 2] = u.imag.ravel()
                out[f"{name}_{tag}_input"] = u
            m = ref.gate_matrix(k, op, qubits, PARAMS, coef)
            out[f"{name}_{tag}"] = m
            out[f"{name}_{tag}_qubits"] = np.asarray(qubits, dtype=np.uint32)
    out["params"] = np.asarray(PARAMS)
    return out


def main():
    if not Reference.available():
        raise SystemExit("oracle/_ref/libdmsim_ref.so is missing: run `make -C oracle` "
                         "where the reference tree exists")
    ref = Reference()

    np.savez_compressed(os.path.join(HERE, "gate_matrices.npz"), **gate_matrices(ref))

    # all ops incl. the C1/C2 matrix extensions, n = 6, four circuits
    allops = list(W.OP_NAMES)
    d = {}
    for s in range(4):
        t = W.random_mixed(6, 200, 600 + s, gate_set=allops)
        re, im, _ = ref.sim(6, t.arrays())
        d.update(table_dict(f"c{s}_", t))
        d[f"c{s}_re"], d[f"c{s}_im"] = re, im
    np.savez_compressed(os.path.join(HERE, "circuits_n6.npz"), **d)

    # the n = 8, 1000-gate mixed workload, seeds 1..8, with shots
    d = {}
    pos = np.random.default_rng(8).integers(0, 1 << 16, size=1024)
    d["sample_pos"] = pos
    for s in range(1, 9):
        t = W.random_mixed(8, 1000, s)
        re, im, shots = ref.sim(8, t.arrays(), threads=4, seed=1000 + s, reps=64)
        d.update(table_dict(f"s{s}_", t))
        if s == 1:
            d["s1_re"], d["s1_im"] = re, im
        d[f"s{s}_sha_re"] = np.asarray(sha(re))
        d[f"s{s}_sha_im"] = np.asarray(sha(im))
        diag = re.reshape(256, 256).diagonal().copy()
        d[f"s{s}_diag"] = diag
        d[f"s{s}_zexp"] = np.array([sum(diag[i] * (-1 if (i >> q) & 1 else 1)
                                        for i in range(256)) for q in range(8)])
        d[f"s{s}_sample_re"], d[f"s{s}_sample_im"] = re[pos], im[pos]
        d[f"s{s}_fro"] = np.asarray(math.sqrt(float(np.sum(re * re + im * im))))
        d[f"s{s}_shots"] = shots
        d[f"s{s}_shot_seed"] = np.asarray(1000 + s)
    np.savez_compressed(os.path.join(HERE, "mixed_n8.npz"), **d)

    # the reference's adder example as one run (example/adder_n10_amdgpu_omp.cpp)
    first, second = W.adder_n10()
    one = W.GateTable(10)
    for part in (first, second):
        one.ops += part.ops
        one.qbs += part.qbs
        one.prms += part.prms
        one.coefs += part.coefs
    re, im, shots = ref.sim(10, one.arrays(), seed=7, reps=10)
    d = table_dict("", one)
    d["diag"] = re.reshape(1024, 1024).diagonal().copy()
    d["sha_re"], d["sha_im"] = np.asarray(sha(re)), np.asarray(sha(im))
    d["shots"], d["shot_seed"] = shots, np.asarray(7)
    np.savez_compressed(os.path.join(HERE, "adder_n10.npz"), **d)

    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
