// qasm.hpp -- OpenQASM 2.0 front-end: source text -> the engine's gate records.
//
// Replaces the reference's translator tool/dmsim_qasm.py (Python 2, which
// writes a Python script of Simulation factory calls): registers are laid out
// in declaration order on one qubit index space (:266-272), qelib1 gates map
// to the factories of the same name with their parameter placement, a gate
// applied to whole registers is broadcast over their bits (:133-160), and
// user-defined gates are expanded at every call (:240-266).  measure, barrier,
// creg, reset and if statements are accepted and ignored, as there (:105).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "lower.hpp"

namespace dmx {

struct QasmProgram {
    int n_qubits = 0;
    std::vector<Gate> gates;
};

// Parses `src`; on failure returns DMX_EINVAL with "line L: message" in the
// error slot.
int parse_qasm(const std::string& src, QasmProgram& out);

}  // namespace dmx
