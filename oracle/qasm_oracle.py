"""Restatement of the reference's OpenQASM translator (TEST INFRASTRUCTURE ONLY).

Only tests/ import this module, to check the C++ front-end
(dm-sim_amd/csrc/qasm.cpp) against it; the product never does.

The reference, tool/dmsim_qasm.py (Python 2), turns an OpenQASM file into a
Python script of ``sim.append(sim.OP(params..., qubits...))`` calls.  This
module follows the same algorithm but returns the gate table those calls
would build:

* line by line (:240-301); a line starting with // is a comment, a trailing
  // comment is cut (:245-252);
* ``qreg name[n]`` places the register after the previous ones on one qubit
  index space (:286-292); creg, measure, barrier, OPENQASM, include, if and
  reset lines are ignored (:105, :299-300);
* a built-in gate (the GATE_TABLE names, :29-69) applied to whole registers
  is repeated once per bit of the register (paramlist_to_ga, :133-160);
  its parameters are evaluated with pi (:141-143) and passed first to the
  factory of the same upper-case name, so they land in that factory's slots
  (src/dmsim_amdgpu_omp.hpp:595-782, e.g. RZ(phi, q) :670, CU1(lambda, c, t)
  :721);
* ``gate name a,b { ... }`` defines a function whose body statements are
  expanded at each call (function_gate :165-191, parse :257-281).

Two deliberate departures, where the reference's line handling loses
statements that OpenQASM 2.0 defines: a body whose '{' ends the header line
(the reference skips the next line as if it were the brace, :262-263), and a
top-level line holding several statements (the reference reads only the
first).  Both are read as the OpenQASM grammar says.

Parameter expressions are evaluated by a small arithmetic evaluator (numbers,
pi, + - * / ** and unary minus, parentheses, sin/cos/tan/exp/ln/sqrt) in
place of Python's eval, so no text from a circuit file is executed.
"""
from __future__ import annotations

import ast
import math
import operator

# op codes of the reference's enum OP (src/dmsim_amdgpu_omp.hpp:51-57)
OPS = ["U3", "U2", "U1", "CX", "ID", "X", "Y", "Z", "H", "S", "SDG", "T", "TDG", "RX", "RY", "RZ",
       "CZ", "CY", "SWAP", "CH", "CCX", "CSWAP", "CRX", "CRY", "CRZ", "CU1", "CU3", "RXX", "RZZ",
       "RCCX", "RC3X", "C3X", "C3SQRTX", "C4X", "R", "SRN", "W", "RYY"]
OP = {name: i for i, name in enumerate(OPS)}

# GATE_TABLE names (tool/dmsim_qasm.py:29-69) without c1/c2, which take
# matrices rather than angles.
BUILTIN = ["u3", "u2", "u1", "cx", "id", "x", "y", "z", "h", "s", "sdg", "t", "tdg", "rx", "ry",
           "rz", "cz", "cy", "swap", "ch", "ccx", "cswap", "crx", "cry", "crz", "cu1", "cu3",
           "rxx", "rzz", "rccx", "rc3x", "c3x", "c3sqrtx", "c4x"]

# factory parameter slots (0 theta, 1 phi, 2 lambda), src/dmsim_amdgpu_omp.hpp:595-782
SLOTS = {"U3": (0, 1, 2), "U2": (1, 2), "U1": (2,), "RX": (0,), "RY": (0,), "RZ": (1,),
         "CRX": (2,), "CRY": (2,), "CRZ": (2,), "CU1": (2,), "CU3": (0, 1, 2), "RXX": (0,),
         "RZZ": (0,), "RYY": (0,), "R": (0,)}

IGNORED = ("measure", "barrier", "OPENQASM", "include", "creg", "if", "reset")

_BIN = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
        ast.Div: operator.truediv, ast.Pow: operator.pow}
_FN = {"sin": math.sin, "cos": math.cos, "tan": math.tan, "exp": math.exp, "ln": math.log,
       "sqrt": math.sqrt}


def _num(expr: str) -> float:
    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return float(node.value)
        if isinstance(node, ast.Name) and node.id == "pi":
            return math.pi
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            v = ev(node.operand)
            return -v if isinstance(node.op, ast.USub) else v
        if isinstance(node, ast.BinOp) and type(node.op) in _BIN:
            return _BIN[type(node.op)](ev(node.left), ev(node.right))
        if (isinstance(node, ast.Call) and isinstance(node.func, ast.Name)
                and node.func.id in _FN and len(node.args) == 1):
            return _FN[node.func.id](ev(node.args[0]))
        raise ValueError(f"unsupported expression: {expr}")
    return ev(ast.parse(expr.replace("^", "**"), mode="eval"))


def _op(line: str) -> str:  # get_op, :107-111
    if "(" in line:
        line = line[:line.find("(")].strip()
    return line.split(" ")[0].strip()


class Translation:
    def __init__(self):
        self.base = {}    # register -> first global qubit (global_array)
        self.length = {}  # register -> size (field_length)
        self.n_qubits = 0
        self.funcs = {}   # gate name -> (argument names, body statements)
        self.gates = []   # (op, qubits, (theta, phi, lambda))

    def _qubit(self, arg, local):  # qreg_to_ga, :114-131
        arg = arg.strip()
        if "[" in arg:
            reg = arg[:arg.find("[")].strip()
            return self.base[reg] + int(arg[arg.find("[") + 1:arg.find("]")])
        return local[arg] if local is not None else self.base[arg]

    def _emit(self, op, params, qubits):
        name = op.upper()
        prm = [0.0, 0.0, 0.0]
        for slot, v in zip(SLOTS.get(name, ()), params):
            prm[slot] = v
        self.gates.append((OP[name], tuple(qubits), tuple(prm)))

    def _call(self, stmt, local):
        """One gate statement (top level: local is None)."""
        op = _op(stmt)
        rest = stmt[len(op):].strip()
        params = []
        if rest.startswith("("):
            close = rest.find(")")
            params = [_num(p.strip()) for p in rest[1:close].split(",")]
            rest = rest[close + 1:]
        args = [a.strip() for a in rest.strip().strip(";").split(",") if a.strip()]
        reps = 1
        if local is None:  # broadcast over whole registers (paramlist_to_ga)
            for a in args:
                if "[" not in a and self.length.get(a, 1) > 1:
                    reps = self.length[a]
                    break
        for b in range(reps):
            qs = []
            for a in args:
                if local is None and "[" not in a and self.length.get(a, 1) > 1:
                    a = f"{a}[{b}]"
                qs.append(self._qubit(a, local))
            if op in self.funcs:
                names, body = self.funcs[op]
                sub = dict(zip(names, qs))
                for s in body:
                    self._call(s, sub)
            elif op in BUILTIN:
                self._emit(op, params, qs)
            else:
                raise ValueError(f"unknown gate {op}")

    def run(self, text: str):
        lines = text.splitlines()
        i = 0
        while i < len(lines):
            line = lines[i].strip()
            i += 1
            if not line or line.startswith("//"):
                continue
            if "//" in line:
                line = line[:line.find("//")]
            op = _op(line)
            if op == "gate":  # :257-281
                text = line
                while "}" not in text:  # the body, up to its closing brace
                    text += " " + lines[i].split("//")[0].strip()
                    i += 1
                head, body = text[:text.find("{")], text[text.find("{") + 1:text.find("}")]
                head = head.split()
                names = [q.strip() for q in "".join(head[2:]).split(",")]
                self.funcs[head[1].strip()] = (names, [b.strip() for b in body.split(";")
                                                       if b.strip()])
            elif op == "qreg":  # :286-292
                spec = line.split(" ")[1]
                reg = spec[:spec.find("[")]
                size = int(spec[spec.find("[") + 1:spec.find("]")])
                self.base[reg] = self.n_qubits
                self.length[reg] = size
                self.n_qubits += size
            elif op in IGNORED:
                continue
            else:
                for stmt in line.split(";"):
                    if stmt.strip() and _op(stmt.strip()) not in IGNORED:
                        self._call(stmt.strip(), None)
        return self


def translate(text: str):
    """(n_qubits, [(op, qubits, (theta, phi, lambda)), ...])"""
    t = Translation().run(text)
    return t.n_qubits, t.gates
