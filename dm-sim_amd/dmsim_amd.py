"""ctypes binding of the C ABI (include/dmsim_amd.h) of libdmsim_amd.so.

This is the reference-side binding a Python maintainer would add for the
C-ABI library (the pybind11 module ``libdmsim_py_amdgpu_omp`` next to it is
the drop-in for the reference's own module).  It raises if the library has
not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdmsim_amd.so")

DMX_OK, DMX_EINVAL, DMX_ESTATE, DMX_EHIP, DMX_ENOMEM = 0, -1, -2, -3, -4
NOALLOC_MAX_ENTRIES = 65536  # dmx_get_entries_noalloc limit (include/dmsim_amd.h)
METHOD_PLAIN, METHOD_PLAIN_LAUNCH, METHOD_FUSED = 0, 1, 2
COST_FUSED, COST_DENSE, COST_DIAG, COST_CX, COST_QUAD, COST_FUSED5, COST_HERM = 0, 1, 2, 3, 4, 5, 6
COST_FEATURES, COST_PARAMS = 14, 5185

# every symbol include/dmsim_amd.h declares
EXPORTS = [
    "dmx_create", "dmx_destroy", "dmx_append", "dmx_append_matrix", "dmx_append_kraus",
    "dmx_upload", "dmx_run",
    "dmx_clear_circuit", "dmx_reset", "dmx_reset_dm", "dmx_measure", "dmx_get_dm",
    "dmx_get_entries", "dmx_get_entries_noalloc",
    "dmx_get_diag", "dmx_expect_z", "dmx_info", "dmx_dump", "dmx_set_method",
    "dmx_set_fusion_max_k", "dmx_set_fusion_permute", "dmx_set_fusion_hermitian",
    "dmx_set_fusion_autotune", "dmx_tuned_tile_bits", "dmx_last_run_stats", "dmx_last_run_fallbacks", "dmx_set_report",
    "dmx_set_profiling", "dmx_profile_get",
    "dmx_qasm_parse", "dmx_append_qasm",
    "dmx_cost_create", "dmx_cost_destroy", "dmx_cost_features", "dmx_cost_train",
    "dmx_cost_predict", "dmx_cost_gradient", "dmx_cost_get_params", "dmx_cost_set_params",
    "dmx_set_cost_model", "dmx_set_tile_bits", "dmx_plan_info", "dmx_plan_sweeps", "dmx_time_launch",
    "dmx_batch_create", "dmx_batch_destroy", "dmx_batch_add_circuit", "dmx_batch_add_pauli",
    "dmx_batch_upload", "dmx_batch_set_graph", "dmx_batch_set_autotune", "dmx_batch_tile_bits",
    "dmx_batch_run", "dmx_batch_energies", "dmx_batch_get_dm", "dmx_batch_stats",
    "dmx_batch_sweep_bytes", "dmx_batch_get_diag", "dmx_batch_set_pool_limit",
    "dmx_batch_waves", "dmx_batch_set_cost_model", "dmx_batch_schedule",
    "dmx_batch_set_group_bytes", "dmx_batch_set_hermitian",
    "dmx_last_error",
]

_lib = None


class DmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dmx error {code}: {msg}")
        self.code = code


def load():
    """Load libdmsim_amd.so (built by ``make -C dm-sim_amd``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C dm-sim_amd`")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    lib.dmx_create.argtypes = [C.c_uint32, C.c_uint32, P(vp)]
    lib.dmx_destroy.argtypes = [vp]
    lib.dmx_append.argtypes = [vp, C.c_int32, P(C.c_uint32), C.c_double, C.c_double, C.c_double]
    lib.dmx_append_matrix.argtypes = [vp, C.c_uint32, P(C.c_uint32), P(C.c_double),
                                      P(C.c_double)]
    lib.dmx_append_kraus.argtypes = [vp, C.c_uint32, C.c_uint32, P(C.c_double), P(C.c_double)]
    for name in ("dmx_upload", "dmx_run", "dmx_clear_circuit", "dmx_reset", "dmx_reset_dm"):
        getattr(lib, name).argtypes = [vp]
    lib.dmx_measure.argtypes = [vp, C.c_uint32, C.c_uint32, P(C.c_uint32), P(C.c_double)]
    lib.dmx_get_dm.argtypes = [vp, P(C.c_double), P(C.c_double)]
    lib.dmx_get_diag.argtypes = [vp, P(C.c_double)]
    lib.dmx_get_entries.argtypes = [vp, C.c_uint64, P(C.c_uint64), P(C.c_double),
                                    P(C.c_double)]
    lib.dmx_get_entries_noalloc.argtypes = lib.dmx_get_entries.argtypes
    lib.dmx_expect_z.argtypes = [vp, P(C.c_double)]
    lib.dmx_info.argtypes = [vp, P(C.c_uint32), P(C.c_uint64), P(C.c_uint32)]
    lib.dmx_dump.argtypes = [vp, C.c_char_p, C.c_size_t, P(C.c_size_t)]
    lib.dmx_set_method.argtypes = [vp, C.c_int]
    lib.dmx_set_fusion_max_k.argtypes = [vp, C.c_uint32]
    lib.dmx_set_fusion_permute.argtypes = [vp, C.c_uint32]
    lib.dmx_set_fusion_hermitian.argtypes = [vp, C.c_uint32]
    lib.dmx_set_fusion_autotune.argtypes = [vp, C.c_uint32]
    lib.dmx_tuned_tile_bits.argtypes = [vp, P(C.c_uint32)]
    lib.dmx_last_run_stats.argtypes = [vp, P(C.c_double), P(C.c_uint32), P(C.c_uint32)]
    lib.dmx_last_run_fallbacks.argtypes = [vp, P(C.c_uint32)]
    lib.dmx_set_report.argtypes = [vp, C.c_int]
    lib.dmx_set_profiling.argtypes = [vp, C.c_int]
    lib.dmx_profile_get.argtypes = [vp, C.c_uint32, C.c_char_p, C.c_size_t, P(C.c_uint32),
                                    P(C.c_double), P(C.c_double), P(C.c_uint32)]
    lib.dmx_qasm_parse.argtypes = [C.c_char_p, C.c_uint32, P(C.c_int32), P(C.c_uint32),
                                   P(C.c_double), P(C.c_uint32), P(C.c_uint32)]
    lib.dmx_append_qasm.argtypes = [vp, C.c_char_p, P(C.c_uint32)]
    f32p = P(C.c_float)
    lib.dmx_cost_create.argtypes = [C.c_uint32, P(vp)]
    lib.dmx_cost_destroy.argtypes = [vp]
    lib.dmx_cost_features.argtypes = [C.c_uint32, C.c_int32, C.c_int32, C.c_uint32, C.c_int32,
                                      C.c_int32, C.c_int32, f32p]
    lib.dmx_cost_train.argtypes = [vp, C.c_uint32, f32p, f32p, C.c_uint32, C.c_uint32,
                                   C.c_float, C.c_uint32, f32p]
    lib.dmx_cost_predict.argtypes = [vp, C.c_uint32, f32p, f32p]
    lib.dmx_cost_gradient.argtypes = [vp, C.c_uint32, f32p, f32p, f32p]
    lib.dmx_cost_get_params.argtypes = [vp, f32p, f32p, f32p]
    lib.dmx_cost_set_params.argtypes = [vp, f32p, C.c_float, C.c_float]
    lib.dmx_set_cost_model.argtypes = [vp, vp]
    lib.dmx_set_tile_bits.argtypes = [vp, C.c_uint32]
    lib.dmx_plan_sweeps.argtypes = [vp, P(C.c_uint32), P(C.c_uint32), P(C.c_uint32),
                                    P(C.c_uint32)]
    lib.dmx_plan_info.argtypes = [vp, P(C.c_uint32), P(C.c_uint32), P(C.c_uint32),
                                  P(C.c_double), P(C.c_double), P(C.c_uint32)]
    lib.dmx_time_launch.argtypes = [vp, C.c_int32, C.c_int32, C.c_uint32, C.c_int32, C.c_int32,
                                    C.c_int32, C.c_uint32, P(C.c_double)]
    lib.dmx_batch_create.argtypes = [P(vp)]
    lib.dmx_batch_destroy.argtypes = [vp]
    lib.dmx_batch_add_circuit.argtypes = [vp, C.c_uint32, C.c_uint32, P(C.c_int32),
                                          P(C.c_uint32), P(C.c_double), P(C.c_uint32)]
    lib.dmx_batch_add_pauli.argtypes = [vp, C.c_uint32, C.c_char_p, C.c_double]
    for name in ("dmx_batch_upload", "dmx_batch_run"):
        getattr(lib, name).argtypes = [vp]
    lib.dmx_batch_set_graph.argtypes = [vp, C.c_int]
    lib.dmx_batch_set_autotune.argtypes = [vp, C.c_int]
    lib.dmx_batch_tile_bits.argtypes = [vp, C.c_uint32, P(C.c_uint32)]
    lib.dmx_batch_energies.argtypes = [vp, P(C.c_double)]
    lib.dmx_batch_get_dm.argtypes = [vp, C.c_uint32, P(C.c_double), P(C.c_double)]
    lib.dmx_batch_sweep_bytes.argtypes = [vp, P(C.c_double)]
    lib.dmx_batch_get_diag.argtypes = [vp, C.c_uint32, P(C.c_double)]
    lib.dmx_batch_set_pool_limit.argtypes = [vp, C.c_uint64]
    lib.dmx_batch_waves.argtypes = [vp, P(C.c_uint32)]
    lib.dmx_batch_set_cost_model.argtypes = [vp, vp]
    lib.dmx_batch_schedule.argtypes = [vp, P(C.c_uint64), P(C.c_uint32)]
    lib.dmx_batch_set_group_bytes.argtypes = [vp, C.c_uint64]
    lib.dmx_batch_set_hermitian.argtypes = [vp, C.c_int]
    lib.dmx_batch_stats.argtypes = [vp, P(C.c_double), P(C.c_uint32), P(C.c_uint32),
                                    P(C.c_uint64)]
    lib.dmx_last_error.restype = C.c_char_p
    for name in EXPORTS:
        if name != "dmx_last_error":
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _check(rc):
    if rc != DMX_OK:
        msg = load().dmx_last_error().decode()
        if rc == DMX_EINVAL:
            raise ValueError(msg)
        raise DmxError(rc, msg)


class Sim:
    """One rho on the current GPU (the reference's ``Simulation``)."""

    def __init__(self, n_qubits: int, n_gpus: int = 1):
        lib = load()
        h = C.c_void_p()
        _check(lib.dmx_create(n_qubits, n_gpus, C.byref(h)))
        self._h = h
        self.n = n_qubits
        self.dim = 1 << n_qubits

    def close(self):
        if getattr(self, "_h", None):
            load().dmx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def append(self, op: int, qubits=(), theta=0.0, phi=0.0, lam=0.0):
        qb = np.zeros(5, dtype=np.uint32)
        qb[: len(qubits)] = qubits
        _check(load().dmx_append(self._h, int(op), _ptr(qb, C.c_uint32), theta, phi, lam))

    def append_matrix(self, qubits, mat):
        mat = np.asarray(mat, dtype=complex)
        re = np.ascontiguousarray(mat.real, dtype=np.float64)
        im = np.ascontiguousarray(mat.imag, dtype=np.float64)
        q = np.asarray(qubits, dtype=np.uint32)
        _check(load().dmx_append_matrix(self._h, len(q), _ptr(q, C.c_uint32),
                                        _ptr(re, C.c_double), _ptr(im, C.c_double)))

    def append_kraus(self, qubit: int, ops):
        """Noise channel rho -> sum_i K_i rho K_i^dagger on one qubit (2x2 operators)."""
        ks = np.ascontiguousarray(np.asarray(ops, dtype=complex).reshape(-1, 2, 2))
        re = np.ascontiguousarray(ks.real.ravel())
        im = np.ascontiguousarray(ks.imag.ravel())
        _check(load().dmx_append_kraus(self._h, qubit, len(ks), _ptr(re, C.c_double),
                                       _ptr(im, C.c_double)))

    def append_table(self, table):
        """Append every gate of a workloads.GateTable."""
        ops, qb, prm, coef = table.arrays()
        for g in range(len(ops)):
            op = int(ops[g])
            if op >= 38:  # matrix extensions C1 (38) / C2 (39)
                k = op - 37
                d = 1 << k
                flat = coef[g, : 2 * d * d]
                self.append_matrix(list(qb[g, :k]), (flat[0::2] + 1j * flat[1::2]).reshape(d, d))
            else:
                self.append(op, list(qb[g]), *prm[g])

    def upload(self):
        _check(load().dmx_upload(self._h))
        return self

    def run(self):
        _check(load().dmx_run(self._h))

    def clear_circuit(self):
        _check(load().dmx_clear_circuit(self._h))

    def reset(self):
        _check(load().dmx_reset(self._h))

    def reset_dm(self):
        _check(load().dmx_reset_dm(self._h))

    def measure(self, reps: int, seed: int):
        out = np.zeros(max(reps, 1), dtype=np.uint32)
        tot = C.c_double()
        _check(load().dmx_measure(self._h, reps, seed, _ptr(out, C.c_uint32), C.byref(tot)))
        return out[:reps], tot.value

    def get_dm(self):
        re = np.empty(self.dim * self.dim)
        im = np.empty(self.dim * self.dim)
        _check(load().dmx_get_dm(self._h, _ptr(re, C.c_double), _ptr(im, C.c_double)))
        return re, im

    def get_entries(self, idx):
        """Entries idx (row-major linear indices, get_dm's convention) as
        complex numbers, gathered on the device."""
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        re = np.empty(idx.size)
        im = np.empty(idx.size)
        _check(load().dmx_get_entries(self._h, idx.size, _ptr(idx, C.c_uint64),
                                      _ptr(re, C.c_double), _ptr(im, C.c_double)))
        return re + 1j * im

    def get_entries_noalloc(self, idx):
        """get_entries through the allocation-free entry point (at most
        NOALLOC_MAX_ENTRIES indices per call)."""
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        re = np.empty(idx.size)
        im = np.empty(idx.size)
        _check(load().dmx_get_entries_noalloc(self._h, idx.size, _ptr(idx, C.c_uint64),
                                              _ptr(re, C.c_double), _ptr(im, C.c_double)))
        return re + 1j * im

    def get_diag(self):
        d = np.empty(self.dim)
        _check(load().dmx_get_diag(self._h, _ptr(d, C.c_double)))
        return d

    def expect_z(self):
        z = np.empty(self.n)
        _check(load().dmx_expect_z(self._h, _ptr(z, C.c_double)))
        return z

    def info(self):
        n, dim, g = C.c_uint32(), C.c_uint64(), C.c_uint32()
        _check(load().dmx_info(self._h, C.byref(n), C.byref(dim), C.byref(g)))
        return n.value, dim.value, g.value

    def dump(self):
        need = C.c_size_t()
        _check(load().dmx_dump(self._h, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        _check(load().dmx_dump(self._h, buf, need.value, None))
        return buf.value.decode()

    def set_method(self, m: int):
        _check(load().dmx_set_method(self._h, m))

    def set_fusion_max_k(self, k: int):
        _check(load().dmx_set_fusion_max_k(self._h, k))

    def set_fusion_permute(self, on: bool):
        """Let fused blocks take trailing X/CX/SWAP gates as an output
        permutation (default on; needs a second copy of rho)."""
        _check(load().dmx_set_fusion_permute(self._h, 1 if on else 0))

    def set_fusion_hermitian(self, mode: int):
        """Four-qubit blocks as Hermitian sweeps: 0 off, 1 from 12 qubits
        (default), 2 from 8 qubits (needs permuting blocks on)."""
        _check(load().dmx_set_fusion_hermitian(self._h, int(mode)))

    def set_fusion_autotune(self, on: bool):
        """Greedy plans (no cost model): measure the tile width at upload
        (default on)."""
        _check(load().dmx_set_fusion_autotune(self._h, 1 if on else 0))

    def tuned_tile_bits(self):
        m = C.c_uint32()
        _check(load().dmx_tuned_tile_bits(self._h, C.byref(m)))
        return m.value

    def set_profiling(self, on: bool):
        _check(load().dmx_set_profiling(self._h, 1 if on else 0))

    def set_report(self, on: bool):
        """Print the reference's per-run report after every run (off here)."""
        _check(load().dmx_set_report(self._h, 1 if on else 0))

    def profile(self):
        """{class name: {calls, ms, bytes}} of the last profiled run."""
        lib = load()
        cnt = C.c_uint32()
        _check(lib.dmx_profile_get(self._h, 0, None, 0, None, None, None, C.byref(cnt)))
        out = {}
        for i in range(cnt.value):
            name = C.create_string_buffer(64)
            calls, ms, by = C.c_uint32(), C.c_double(), C.c_double()
            _check(lib.dmx_profile_get(self._h, i, name, 64, C.byref(calls), C.byref(ms),
                                       C.byref(by), None))
            if calls.value:
                out[name.value.decode()] = {"calls": calls.value, "ms": ms.value,
                                            "bytes": by.value}
        return out

    def last_run_stats(self):
        ms, la, sw = C.c_double(), C.c_uint32(), C.c_uint32()
        _check(load().dmx_last_run_stats(self._h, C.byref(ms), C.byref(la), C.byref(sw)))
        fb = C.c_uint32()
        _check(load().dmx_last_run_fallbacks(self._h, C.byref(fb)))
        # fallbacks: persistent-loop segments run one launch per primitive
        # because the cooperative launch was refused
        return {"ms": ms.value, "launches": la.value, "sweeps": sw.value,
                "fallbacks": fb.value}

    def append_qasm(self, source: str):
        """Append an OpenQASM 2.0 program's gates; returns how many."""
        ng = C.c_uint32()
        _check(load().dmx_append_qasm(self._h, source.encode(), C.byref(ng)))
        return ng.value

    def set_cost_model(self, model):
        """Planner model for the next fused upload (None = greedy plan)."""
        self._model = model  # keep it alive while this simulation may plan with it
        _check(load().dmx_set_cost_model(self._h, model._h if model is not None else None))

    def set_tile_bits(self, m: int):
        """Fixed fused tile width 4..8 for the next upload (0 = automatic)."""
        _check(load().dmx_set_tile_bits(self._h, m))

    def plan_info(self):
        fb, pb, wg = C.c_uint32(), C.c_uint32(), C.c_uint32()
        pred, greedy = C.c_double(), C.c_double()
        tiles = (C.c_uint32 * 5)()
        _check(load().dmx_plan_info(self._h, C.byref(fb), C.byref(pb), C.byref(wg),
                                    C.byref(pred), C.byref(greedy), tiles))
        return {"fused_blocks": fb.value, "plain_blocks": pb.value, "fused5_blocks": wg.value,
                "predicted_us": pred.value, "greedy_us": greedy.value,
                "tiles": {m: tiles[m - 4] for m in range(4, 9)}, **self.plan_sweeps()}

    def plan_sweeps(self):
        pb, hb, rl, gr = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(load().dmx_plan_sweeps(self._h, C.byref(pb), C.byref(hb), C.byref(rl),
                                      C.byref(gr)))
        return {"perm_blocks": pb.value, "herm_blocks": hb.value, "relabels": rl.value,
                "grouping": ("least_growth", "frontier", "frontier_timed", "frontier_timed_multi",
                             "frontier_timed_sqrt", "frontier_1q_heavy")[gr.value]}

    def time_launch(self, kind, m=0, qmask=0, target=0, control=-1, slow=0, reps=10):
        """Average device microseconds of one launch configuration (rho is left
        transformed; reset it before simulating)."""
        us = C.c_double()
        _check(load().dmx_time_launch(self._h, kind, m, qmask, target, control, slow, reps,
                                      C.byref(us)))
        return us.value


def depolarizing(p):
    """Kraus operators of the single-qubit depolarizing channel with error
    probability p: (1 - p) rho + p/3 (X rho X + Y rho Y + Z rho Z)."""
    i2 = np.eye(2)
    x = np.array([[0, 1], [1, 0]])
    y = np.array([[0, -1j], [1j, 0]])
    z = np.diag([1, -1])
    return [math.sqrt(1 - p) * i2] + [math.sqrt(p / 3) * m for m in (x, y, z)]


def amplitude_damping(gamma):
    """Kraus operators of amplitude damping with decay probability gamma."""
    return [np.array([[1, 0], [0, math.sqrt(1 - gamma)]]),
            np.array([[0, math.sqrt(gamma)], [0, 0]])]


def phase_damping(lam):
    """Kraus operators of phase damping with parameter lam."""
    return [np.array([[1, 0], [0, math.sqrt(1 - lam)]]), np.array([[0, 0], [0, math.sqrt(lam)]])]


def qasm_parse(source: str):
    """OpenQASM 2.0 text -> (n_qubits, ops int32[g], qb uint32[g, 5], params float64[g, 3])."""
    lib = load()
    src = source.encode()
    nq, ng = C.c_uint32(), C.c_uint32()
    _check(lib.dmx_qasm_parse(src, 0, None, None, None, C.byref(nq), C.byref(ng)))
    g = ng.value
    ops = np.zeros(max(g, 1), dtype=np.int32)
    qb = np.zeros((max(g, 1), 5), dtype=np.uint32)
    prm = np.zeros((max(g, 1), 3))
    _check(lib.dmx_qasm_parse(src, g, _ptr(ops, C.c_int32), _ptr(qb, C.c_uint32),
                              _ptr(prm, C.c_double), C.byref(nq), C.byref(ng)))
    return nq.value, ops[:g], qb[:g], prm[:g]


def cost_features(n, kind, m=0, qmask=0, target=0, control=-1, slow=0):
    """Feature vector (float32[COST_FEATURES]) of one launch configuration."""
    out = np.empty(COST_FEATURES, dtype=np.float32)
    _check(load().dmx_cost_features(n, kind, m, qmask, target, control, slow,
                                    _ptr(out, C.c_float)))
    return out


class CostModel:
    """The fused planner's launch-time model (``dmx_cost_*``): FP32 MLP,
    COST_FEATURES -> 64 -> 64 -> 1, trained and evaluated on the GPU."""

    def __init__(self, seed: int = 16):
        h = C.c_void_p()
        _check(load().dmx_cost_create(seed, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load().dmx_cost_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _f32(a):
        return np.ascontiguousarray(a, dtype=np.float32)

    def train(self, X, y, epochs=200, batch=64, lr=1e-3, seed=0):
        X, y = self._f32(X), self._f32(y)
        loss = C.c_float()
        _check(load().dmx_cost_train(self._h, len(y), _ptr(X, C.c_float), _ptr(y, C.c_float),
                                     epochs, batch, lr, seed, C.byref(loss)))
        return loss.value

    def predict(self, X):
        X = self._f32(X)
        out = np.empty(len(X), dtype=np.float32)
        _check(load().dmx_cost_predict(self._h, len(X), _ptr(X, C.c_float),
                                       _ptr(out, C.c_float)))
        return out

    def gradient(self, X, y):
        X, y = self._f32(X), self._f32(y)
        out = np.empty(COST_PARAMS, dtype=np.float32)
        _check(load().dmx_cost_gradient(self._h, len(y), _ptr(X, C.c_float),
                                        _ptr(y, C.c_float), _ptr(out, C.c_float)))
        return out

    def get_params(self):
        p = np.empty(COST_PARAMS, dtype=np.float32)
        mu, dev = C.c_float(), C.c_float()
        _check(load().dmx_cost_get_params(self._h, _ptr(p, C.c_float), C.byref(mu),
                                          C.byref(dev)))
        return p, mu.value, dev.value

    def set_params(self, params, mean=0.0, dev=1.0):
        p = self._f32(params)
        if p.size != COST_PARAMS:
            raise ValueError(f"expected {COST_PARAMS} parameters")
        _check(load().dmx_cost_set_params(self._h, _ptr(p, C.c_float), mean, dev))

    def save(self, path):
        p, mu, dev = self.get_params()
        np.savez(path, params=p, mean=np.float32(mu), dev=np.float32(dev))

    def load(self, path):
        with np.load(path) as z:
            self.set_params(z["params"], float(z["mean"]), float(z["dev"]))


class Batch:
    """Many independent circuits run together (``dmx_batch_*``)."""

    def __init__(self):
        lib = load()
        h = C.c_void_p()
        _check(lib.dmx_batch_create(C.byref(h)))
        self._h = h
        self.sizes = []

    def close(self):
        if getattr(self, "_h", None):
            load().dmx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_circuit(self, table, paulis=(), weights=()):
        """A workloads.GateTable (ops 0..37) plus its Pauli terms; returns the id."""
        ops, qb, prm, _ = table.arrays()
        cid = C.c_uint32()
        _check(load().dmx_batch_add_circuit(self._h, table.n_qubits, len(ops),
                                            _ptr(ops, C.c_int32), _ptr(qb, C.c_uint32),
                                            _ptr(prm, C.c_double), C.byref(cid)))
        for p, w in zip(paulis, weights):
            _check(load().dmx_batch_add_pauli(self._h, cid.value, p.encode(), float(w)))
        self.sizes.append(table.n_qubits)
        return cid.value

    def upload(self):
        _check(load().dmx_batch_upload(self._h))

    def set_graph(self, on: bool):
        _check(load().dmx_batch_set_graph(self._h, 1 if on else 0))

    def set_autotune(self, on: bool):
        _check(load().dmx_batch_set_autotune(self._h, 1 if on else 0))

    def tile_bits(self, n):
        m = C.c_uint32()
        _check(load().dmx_batch_tile_bits(self._h, n, C.byref(m)))
        return m.value

    def run(self):
        _check(load().dmx_batch_run(self._h))

    def energies(self):
        out = np.empty(len(self.sizes))
        _check(load().dmx_batch_energies(self._h, _ptr(out, C.c_double)))
        return out

    def get_dm(self, cid):
        d = 1 << self.sizes[cid]
        re = np.empty(d * d)
        im = np.empty(d * d)
        _check(load().dmx_batch_get_dm(self._h, cid, _ptr(re, C.c_double), _ptr(im, C.c_double)))
        return re, im

    def get_diag(self, cid):
        out = np.empty(1 << self.sizes[cid])
        _check(load().dmx_batch_get_diag(self._h, cid, _ptr(out, C.c_double)))
        return out

    def set_pool_limit(self, nbytes: int):
        """Bytes one pool of circuit states may take (0: automatic); circuits
        that do not fit together run in waves."""
        _check(load().dmx_batch_set_pool_limit(self._h, int(nbytes)))

    def waves(self):
        w = C.c_uint32()
        _check(load().dmx_batch_waves(self._h, C.byref(w)))
        return w.value

    def set_hermitian(self, on: bool):
        """Hermitian block sweeps for the circuits of 10 and 11 qubits (default on)."""
        _check(load().dmx_batch_set_hermitian(self._h, 1 if on else 0))

    def set_group_bytes(self, nbytes: int):
        """Fixes the schedule's group budget (0 = whole waves) instead of
        timing it at upload."""
        _check(load().dmx_batch_set_group_bytes(self._h, int(nbytes)))

    def schedule(self):
        """The schedule chosen at upload: rho bytes per group of batched
        circuits run step after step together (0: whole waves) and the number
        of groups."""
        gb, ng = C.c_uint64(), C.c_uint32()
        _check(load().dmx_batch_schedule(self._h, C.byref(gb), C.byref(ng)))
        return {"group_bytes": gb.value, "groups": ng.value}

    def set_cost_model(self, model):
        """Planner model for the circuits of 12 or more qubits, which run on
        their own (None = greedy plan); used at upload."""
        self._model = model  # keep it alive while this batch may plan with it
        _check(load().dmx_batch_set_cost_model(self._h, model._h if model is not None else None))

    def stats(self):
        ms, la, nc, ng = C.c_double(), C.c_uint32(), C.c_uint32(), C.c_uint64()
        _check(load().dmx_batch_stats(self._h, C.byref(ms), C.byref(la), C.byref(nc),
                                      C.byref(ng)))
        by = C.c_double()
        _check(load().dmx_batch_sweep_bytes(self._h, C.byref(by)))
        return {"ms": ms.value, "launches": la.value, "circuits": nc.value, "gates": ng.value,
                "sweep_bytes": by.value}
