"""CPU-side checks of the product library: it loads, exports every entry point
include/dmsim_amd.h declares, rejects bad arguments before touching a GPU,
and the drop-in pybind11 module mirrors the reference module's surface
(src/py_amdgpu_omp_wrapper.cpp:29-88) with the factories' parameter placement
(src/dmsim_amdgpu_omp.hpp:594-782).  No kernel runs here."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "dmsim_amd.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(dmx_[a-z_]+)\s*\(", text)))


def test_header_and_library_agree():
    import dmsim_amd
    lib = dmsim_amd.load()
    names = declared()
    assert set(names) == set(dmsim_amd.EXPORTS)
    for name in names:
        assert hasattr(lib, name), name


def test_exports_visible_in_dynamic_symbol_table():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "dm-sim_amd",
                                                                     "libdmsim_amd.so")],
                         capture_output=True, text=True, check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name in declared():
        assert name in syms, name


def test_argument_errors_need_no_gpu():
    import dmsim_amd
    lib = dmsim_amd.load()
    h = C.c_void_p()
    assert lib.dmx_create(10, 2, C.byref(h)) == dmsim_amd.DMX_EINVAL
    assert b"n_gpus" in lib.dmx_last_error()
    assert lib.dmx_create(0, 1, C.byref(h)) == dmsim_amd.DMX_EINVAL
    assert lib.dmx_create(17, 1, C.byref(h)) == dmsim_amd.DMX_EINVAL
    assert lib.dmx_run(None) == dmsim_amd.DMX_EINVAL
    assert lib.dmx_destroy(None) == dmsim_amd.DMX_OK
    with pytest.raises(ValueError):
        dmsim_amd.Sim(8, 4)


FACTORIES = {
    # name: (args, expected (op, qubits, theta, phi, lambda))
    "U3": ((0.1, 0.2, 0.3, 4), (0, [4], 0.1, 0.2, 0.3)),
    "U2": ((0.2, 0.3, 4), (1, [4], 0.0, 0.2, 0.3)),
    "U1": ((0.3, 4), (2, [4], 0.0, 0.0, 0.3)),
    "CX": ((1, 2), (3, [1, 2], 0, 0, 0)),
    "RX": ((0.5, 3), (13, [3], 0.5, 0, 0)),
    "RZ": ((0.5, 3), (15, [3], 0.0, 0.5, 0.0)),
    "CRX": ((0.5, 1, 2), (22, [1, 2], 0, 0, 0.5)),
    "CU1": ((0.5, 1, 2), (25, [1, 2], 0, 0, 0.5)),
    "CU3": ((0.1, 0.2, 0.3, 1, 2), (26, [1, 2], 0.1, 0.2, 0.3)),
    "RXX": ((0.5, 1, 2), (27, [1, 2], 0.5, 0, 0)),
    "C4X": ((0, 1, 2, 3, 4), (33, [0, 1, 2, 3, 4], 0, 0, 0)),
    "R": ((0.5, 3), (34, [3], 0.5, 0, 0)),
    "RYY": ((0.5, 1, 2), (37, [1, 2], 0.5, 0, 0)),
}


def test_pybind_module_surface():
    import libdmsim_py_amdgpu_omp as m
    import workloads as W
    names = [n for n in W.OP_NAMES if n not in ("C1", "C2")]
    for n in names:
        assert hasattr(m.Simulation, n), n
    for meth in ("append", "upload", "clear_circuit", "run", "reset", "measure"):
        assert hasattr(m.Simulation, meth)
    for name, (args, (op, qubits, th, ph, la)) in FACTORIES.items():
        g = getattr(m.Simulation, name)(*args)
        assert g.op == op == W.OP[name]
        assert list(g.qubits[: len(qubits)]) == qubits
        assert (g.theta, g.phi, g.lambda_) == (th, ph, la), name
    g = m.Gate(5, 1, 0, 0, 0, 0, 0.0, 0.0, 0.0)
    assert g.op == 5


def test_pybind_rejects_multi_gpu_without_touching_gpu():
    import libdmsim_py_amdgpu_omp as m
    with pytest.raises(ValueError):
        m.Simulation(4, 2)


def build_cpp_facade_example(out_dir, name="adder_facade"):
    """Compile tests/cpp/<name>.cpp against include/ and the library; returns
    the executable's path."""
    import subprocess
    exe = os.path.join(str(out_dir), name)
    pkg = os.path.join(ROOT, "dm-sim_amd")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", name + ".cpp"), "-L", pkg, "-ldmsim_amd",
           f"-Wl,-rpath,{pkg}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("name", ["adder_facade", "xacc_backend"])
def test_cpp_faces_build(tmp_path, name):
    """The reference-style C++ program (namespace DMSim, Simulation::CX(...),
    sim.sim(), measure, print_measurement, dm_real_res) and the XACC backend
    interface (namespace DmSim, DmSimBackend) compile and link against the
    header-only faces of the C ABI."""
    assert os.path.exists(build_cpp_facade_example(tmp_path, name))


def test_planner_grouping_is_valid(tmp_path):
    """The fused planner's grouping (dmx::plan_blocks, an exported C++ symbol of
    the library) on 108 seeded random circuits of 4..12 qubits with one- to
    five-qubit gates, identities and channels, blocks of at most four and five
    qubits, with and without trailing permutations: every gate in exactly one
    block, matrix members within their block, trailing index maps as stored, no
    gate across a channel, and execution order equal to program order as an
    operator on a random state (tests/cpp/plan_groups.cpp).
    The internal headers include HIP's, so hipcc compiles it (host code only)."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    exe = os.path.join(str(tmp_path), "plan_groups")
    pkg = os.path.join(ROOT, "dm-sim_amd")
    cmd = [hipcc, "-std=c++17", "-O2", "-I", os.path.join(pkg, "csrc"), "-I",
           os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "plan_groups.cpp"),
           "-L", pkg, "-ldmsim_amd", f"-Wl,-rpath,{pkg}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "324 cases OK", r.stdout


def test_hermitian_sweep_maps(tmp_path):
    """The Hermitian fused executor's index maps, executed on the CPU thread by
    thread with the kernels' own map functions (dm-sim_amd/csrc/herm.hpp) on
    plans from dmx::build_fused_plan at n = 8..10, one circuit per width with a
    C4X and two channels with random Kraus operators (in-order steps on
    whatever index bits hold their qubits): every load and store inside the
    planes, every entry written once per sweep, one relabelling per plan, and
    the result equal to the circuit applied gate by gate, channels as their
    Kraus sums (tests/cpp/herm_maps.cpp)."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    exe = os.path.join(str(tmp_path), "herm_maps")
    pkg = os.path.join(ROOT, "dm-sim_amd")
    cmd = [hipcc, "-std=c++17", "-O2", "-I", os.path.join(pkg, "csrc"), "-I",
           os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "herm_maps.cpp"),
           "-L", pkg, "-ldmsim_amd", f"-Wl,-rpath,{pkg}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "10 cases OK", r.stdout
