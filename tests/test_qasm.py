"""CPU checks of the OpenQASM 2.0 front-end (dm-sim_amd/csrc/qasm.cpp): the
C++ parser against the restated reference translator (oracle/qasm_oracle.py)
on the committed circuits and, where the reference tree is present, on all of
its benchmark circuits; hand-written programs with known gate tables; errors."""
import glob
import math
import os

import numpy as np
import pytest

from oracle import qasm_oracle as qo

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "qasm", "*.qasm")))
REF_BENCH = sorted(glob.glob("/root/reference/benchmark/*.qasm"))


@pytest.fixture(scope="module")
def dm():
    return pytest.importorskip("dmsim_amd")


def table(dm, src):
    nq, ops, qb, prm = dm.qasm_parse(src)
    return nq, [(int(o), tuple(int(x) for x in q), tuple(float(v) for v in p))
                for o, q, p in zip(ops, qb, prm)]


def same_as_oracle(dm, src):
    nq, gates = table(dm, src)
    onq, ogates = qo.translate(src)
    assert nq == onq and len(gates) == len(ogates)
    for (op, qb, prm), (oop, oqb, oprm) in zip(gates, ogates):
        assert op == oop and qb[:len(oqb)] == oqb and all(v == 0 for v in qb[len(oqb):])
        assert prm == oprm


@pytest.mark.parametrize("path", FIXTURES + REF_BENCH, ids=os.path.basename)
def test_parser_matches_restated_translator(dm, path):
    same_as_oracle(dm, open(path).read())


def test_registers_broadcast_and_slots(dm):
    src = """OPENQASM 2.0;
    include "qelib1.inc";
    qreg a[2]; qreg b[2];
    creg c[4];
    h a;                 // broadcast over a
    cx a, b;             // pairwise a[i], b[i]
    rz(pi/4) b[1];
    cu1(-pi/2) a[0], b[0];
    u3(0.1, 0.2, 0.3) b[0];
    U(1, 2, 3) a[1]; CX a[1], a[0];
    measure a -> c[0];
    barrier a, b;
    """
    nq, gates = table(dm, src)
    assert nq == 4
    OP = qo.OP
    want = [(OP["H"], (0,), (0, 0, 0)), (OP["H"], (1,), (0, 0, 0)),
            (OP["CX"], (0, 2), (0, 0, 0)), (OP["CX"], (1, 3), (0, 0, 0)),
            (OP["RZ"], (3,), (0, math.pi / 4, 0)),
            (OP["CU1"], (0, 2), (0, 0, -math.pi / 2)),
            (OP["U3"], (2,), (0.1, 0.2, 0.3)), (OP["U3"], (1,), (1, 2, 3)),
            (OP["CX"], (1, 0), (0, 0, 0))]
    assert len(gates) == len(want)
    for (op, qb, prm), (wop, wq, wp) in zip(gates, want):
        assert op == wop and qb[:len(wq)] == wq and prm == tuple(float(v) for v in wp)


def test_user_gates_with_parameters_and_nesting(dm):
    # parametrised user gates (the reference translator cannot read these)
    src = """qreg q[3];
    gate rot(t, p) x { rz(p) x; ry(t/2) x; }
    gate pair(t) x, y { rot(t, 2*t) x; cx x, y; rot(-t, t^2) y; }
    pair(0.5) q[2], q[0];
    """
    _, gates = table(dm, src)
    OP = qo.OP
    want = [(OP["RZ"], (2,), (0, 1.0, 0)), (OP["RY"], (2,), (0.25, 0, 0)),
            (OP["CX"], (2, 0), (0, 0, 0)),
            (OP["RZ"], (0,), (0, 0.25, 0)), (OP["RY"], (0,), (-0.25, 0, 0))]
    assert [(g[0], g[1][:len(w[1])], g[2]) for g, w in zip(gates, want)] == \
        [(w[0], w[1], tuple(float(v) for v in w[2])) for w in want]


def test_expressions(dm):
    src = "qreg q[1]; rx(-pi/4 + 2*(3 - 1)^2 / 8) q[0]; ry(sin(pi/6) * sqrt(16) + ln(exp(1))) q[0];"
    _, gates = table(dm, src)
    assert gates[0][2][0] == -math.pi / 4 + 2 * (3 - 1) ** 2 / 8
    assert abs(gates[1][2][0] - 3.0) < 1e-15


@pytest.mark.parametrize("src,msg", [
    ("qreg q[2]; foo q[0];", "unknown gate"),
    ("qreg q[2]; h q[2];", "out of range"),
    ("qreg q[2]; cx q[0], q[0];", "repeated qubit"),
    ("qreg q[2]; cx q[0];", "takes"),
    ("qreg q[2]; qreg r[3]; cx q, r;", "different sizes"),
    ("qreg q[2]; gate g a { h a; ", "unterminated"),
    ("qreg q[2]; rx(bad) q[0];", "unknown identifier"),
    ("qreg q[2];\n\nh p[0];", "line 3"),
])
def test_errors_name_the_line(dm, src, msg):
    with pytest.raises(ValueError, match=msg):
        dm.qasm_parse(src)


def test_adder_file_is_the_reference_example():
    """Pins both readers to an independent source: the reference's own Python
    example (example/adder_n10_omp.py:38-82, restated in workloads.adder_n10)
    builds the Cuccaro adder that benchmark/adder_n9.qasm describes, gate for
    gate (cin -> 0, a -> 1..4, b -> 5..8, cout -> 9)."""
    import workloads as W
    first, second = W.adder_n10()
    want = []
    for t in (first, second):
        ops, qb, _, _ = t.arrays()
        want += [(int(o), tuple(int(x) for x in q[:W.ARITY[W.OP_NAMES[o]]])) for o, q in zip(ops, qb)]
    src = open(os.path.join(HERE, "golden", "qasm", "adder_n9.qasm")).read()
    nq, gates = qo.translate(src)
    assert nq == 10 and [(g[0], g[1]) for g in gates] == want
    dmx = pytest.importorskip("dmsim_amd")
    _, cgates = table(dmx, src)
    assert [(g[0], g[1][:len(w[1])]) for g, w in zip(cgates, want)] == want


def test_oracle_broadcast_and_slots_known_answer():
    """The restated translator on a program with a known gate table: register
    offsets in declaration order, whole-register broadcast, parameter slots of
    the factories (RZ -> phi, CU1 -> lambda, U2 -> phi and lambda)."""
    src = """OPENQASM 2.0;
include "qelib1.inc";
qreg a[2];
qreg b[3];
h b;
cx a[1],b[2];
rz(pi/2) a[0];
cu1(-pi/4) b[0],a[1];
u2(0.25,0.5) b[1];
"""
    nq, gates = qo.translate(src)
    OP = qo.OP
    assert nq == 5
    assert gates == [(OP["H"], (2,), (0.0, 0.0, 0.0)), (OP["H"], (3,), (0.0, 0.0, 0.0)),
                     (OP["H"], (4,), (0.0, 0.0, 0.0)), (OP["CX"], (1, 4), (0.0, 0.0, 0.0)),
                     (OP["RZ"], (0,), (0.0, math.pi / 2, 0.0)),
                     (OP["CU1"], (2, 1), (0.0, 0.0, -math.pi / 4)),
                     (OP["U2"], (3,), (0.0, 0.25, 0.5))]


def test_expansion_limit(dm):
    """Nested definitions that double at every level stop with an error
    instead of exhausting memory (2^30 gates here)."""
    lines = ["qreg q[1];", "gate g0 a { x a; }"]
    for k in range(1, 31):
        lines.append(f"gate g{k} a {{ g{k - 1} a; g{k - 1} a; }}")
    lines.append("g30 q[0];")
    with pytest.raises(ValueError, match="expands to more than"):
        dm.qasm_parse("\n".join(lines))


def test_application_limit_without_emitted_gates(dm):
    """Nested definitions whose bodies emit nothing (here a barrier only)
    still double the work at every level: the parser stops after a bounded
    number of applications instead of running for 2^60 calls."""
    lines = ["qreg q[1];", "gate g0 a { barrier a; }"]
    for k in range(1, 61):
        lines.append(f"gate g{k} a {{ g{k - 1} a; g{k - 1} a; }}")
    lines.append("g60 q[0];")
    with pytest.raises(ValueError, match="applies more than"):
        dm.qasm_parse("\n".join(lines))


@pytest.mark.parametrize("prefix,suffix", [("(" * 100000, "1" + ")" * 100000),
                                           ("-" * 100000, "1"),
                                           ("2^" * 100000, "1")])
def test_expression_depth_limit(dm, prefix, suffix):
    """Deeply nested parentheses, signs or powers end in a parse error that
    names the line, not in a stack overflow of the host process."""
    with pytest.raises(ValueError, match="line 2.*nests more than"):
        dm.qasm_parse("qreg q[1];\nrx(" + prefix + suffix + ") q[0];")


def test_expression_depth_below_limit(dm):
    nq, gates = table(dm, "qreg q[1];\nrx(" + "(" * 200 + "0.5" + ")" * 200 + ") q[0];")
    assert gates[0][2][0] == 0.5
