"""ctypes face of the checker libraries (TEST INFRASTRUCTURE ONLY).

Only tests/, the package's smoke() and bench.py's cpu_baseline leg import
this module, and only to check or time against it; the product path never
does.

  Oracle     -> oracle/liboracle.so (the repo's C restatement of DM-Sim's
               CPU path, oracle/dmsim_oracle.c)
  Reference  -> oracle/_ref/libdmsim_ref.so (the reference's own
               src/dmsim_cpu_omp.hpp compiled in place; present only where
               the reference tree was available at build time)

Gate tables are the arrays of ``GateTable.arrays()``:
ops int32[ng], qb uint32[ng,5], prm float64[ng,3], coef float64[ng,32].
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = np.ctypeslib.ndpointer


def _dp():
    return _P(dtype=np.float64, flags="C_CONTIGUOUS")


def _ip():
    return _P(dtype=np.int32, flags="C_CONTIGUOUS")


def _up():
    return _P(dtype=np.uint32, flags="C_CONTIGUOUS")


def build(quiet=True):
    """Compile liboracle.so (and _ref where the reference tree exists)."""
    out = subprocess.run(["make", "-B", "-C", HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


class Oracle:
    def __init__(self, path=None):
        path = path or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        lib.orc_set_threads.argtypes = [C.c_int]
        lib.orc_forward.argtypes = [C.c_int, C.c_int64, _dp(), _dp(), C.c_int, _ip(), _up(),
                                    _dp(), _dp()]
        lib.orc_forward.restype = C.c_int
        lib.orc_sim.argtypes = [C.c_int, _dp(), _dp(), C.c_int, _ip(), _up(), _dp(), _dp()]
        lib.orc_sim.restype = C.c_int
        lib.orc_sim_rowwise.argtypes = lib.orc_sim.argtypes
        lib.orc_sim_rowwise.restype = C.c_int
        lib.orc_conj_transpose.argtypes = [C.c_int, _dp(), _dp()]
        lib.orc_init.argtypes = [C.c_int, _dp(), _dp()]
        lib.orc_measure.argtypes = [C.c_int, _dp(), C.c_uint32, C.c_uint32, _up()]
        lib.orc_measure.restype = C.c_double
        lib.orc_measure_diag.argtypes = [C.c_int, _dp(), C.c_uint32, C.c_uint32, _up()]
        lib.orc_measure_diag.restype = C.c_double
        lib.orc_sv_run.argtypes = [C.c_int, _dp(), _dp(), C.c_int, _ip(), _up(), _dp(), _dp()]
        lib.orc_sv_run.restype = C.c_int
        self.lib = lib

    def set_threads(self, t: int):
        self.lib.orc_set_threads(int(t))

    @staticmethod
    def ground(n):
        dim = 1 << n
        re = np.zeros(dim * dim)
        im = np.zeros(dim * dim)
        re[0] = 1.0
        return re, im

    def sim(self, n, table, state=None, rowwise=False):
        """One sim() run on ``state`` (default |0><0|); returns (re, im)
        flattened dim*dim arrays in the reference's result orientation.
        rowwise: the forward passes row by row (bit-identical, cache-friendly
        at large n)."""
        ops, qb, prm, coef = table
        re, im = (self.ground(n) if state is None else
                  (np.array(state[0], dtype=np.float64), np.array(state[1], dtype=np.float64)))
        fn = self.lib.orc_sim_rowwise if rowwise else self.lib.orc_sim
        rc = fn(n, re, im, len(ops), ops, qb.ravel(), prm.ravel(), coef.ravel())
        if rc:
            raise ValueError(f"oracle rejected gate {-rc - 1}")
        return re, im

    def forward(self, n, rows, re, im, table):
        ops, qb, prm, coef = table
        rc = self.lib.orc_forward(n, rows, re, im, len(ops), ops, qb.ravel(), prm.ravel(),
                                  coef.ravel())
        if rc:
            raise ValueError(f"oracle rejected gate {-rc - 1}")

    def measure(self, n, re, reps, seed):
        out = np.zeros(reps, dtype=np.uint32)
        total = self.lib.orc_measure(n, np.ascontiguousarray(re), reps, seed, out)
        return out, total

    def measure_diag(self, n, diag, reps, seed):
        out = np.zeros(reps, dtype=np.uint32)
        total = self.lib.orc_measure_diag(n, np.ascontiguousarray(diag, dtype=np.float64), reps,
                                          seed, out)
        return out, total

    def sv(self, n, table):
        ops, qb, prm, coef = table
        pr = np.zeros(1 << n)
        pi = np.zeros(1 << n)
        pr[0] = 1.0
        rc = self.lib.orc_sv_run(n, pr, pi, len(ops), ops, qb.ravel(), prm.ravel(), coef.ravel())
        if rc:
            raise ValueError(f"oracle rejected gate {-rc - 1}")
        return pr + 1j * pi


class Reference:
    """The reference's own CPU path (only where oracle/_ref was built)."""

    PATH = os.path.join(HERE, "_ref", "libdmsim_ref.so")

    @classmethod
    def available(cls):
        return os.path.exists(cls.PATH)

    def __init__(self):
        lib = C.CDLL(self.PATH)
        lib.ref_sim.argtypes = [C.c_int, C.c_int, C.c_int, _ip(), _up(), _dp(), _dp(), _dp(),
                                C.c_uint, C.c_uint, _up()]
        lib.ref_sim.restype = C.c_int
        lib.ref_sim_ext.argtypes = [C.c_int, C.c_int, _ip(), _up(), _dp(), _dp(), _dp(), _dp()]
        lib.ref_sim_ext.restype = C.c_int
        lib.ref_gate_matrix.argtypes = [C.c_int, C.c_int, _up(), _dp(), _dp(), _dp(), _dp()]
        lib.ref_gate_matrix.restype = C.c_int
        self.lib = lib

    def sim(self, n, table, threads=1, seed=0, reps=0):
        ops, qb, prm, coef = table
        dim = 1 << n
        re = np.zeros(dim * dim)
        im = np.zeros(dim * dim)
        shots = np.zeros(max(reps, 1), dtype=np.uint32)
        if np.any((ops < 0) | (ops > 37)):
            rc = self.lib.ref_sim_ext(n, len(ops), ops, qb.ravel(), prm.ravel(), coef.ravel(),
                                      re, im)
            if reps:
                raise ValueError("shots only through ref_sim (ops 0..37)")
        else:
            rc = self.lib.ref_sim(n, threads, len(ops), ops, qb.ravel(), prm.ravel(), re, im,
                                  seed, reps, shots)
        if rc:
            raise RuntimeError(f"reference harness returned {rc}")
        return re, im, shots[:reps]

    def gate_matrix(self, n, op, qubits, params=(0.0, 0.0, 0.0), coef=None):
        dim = 1 << n
        qb = np.zeros(5, dtype=np.uint32)
        qb[: len(qubits)] = qubits
        prm = np.asarray(params, dtype=np.float64)
        c = np.zeros(32) if coef is None else np.ascontiguousarray(coef, dtype=np.float64)
        re = np.zeros(dim * dim)
        im = np.zeros(dim * dim)
        rc = self.lib.ref_gate_matrix(n, op, qb, prm, c, re, im)
        if rc:
            raise RuntimeError(f"reference harness returned {rc}")
        return (re + 1j * im).reshape(dim, dim)
