// pymodule.cpp -- pybind11 module `libdmsim_py_amdgpu_omp`, the drop-in for
// the reference's module of the same name (src/py_amdgpu_omp_wrapper.cpp:29-88):
// classes Gate and Simulation, methods append/upload/clear_circuit/run/reset/
// measure and the 38 static gate factories with the reference's argument order
// (parameters first, then qubits).  It is a thin layer over the C ABI of
// libdmsim_amd.so (include/dmsim_amd.h); errors raise Python exceptions where
// the reference asserts or exits.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <array>
#include <cstdint>
#include <ctime>
#include <stdexcept>
#include <string>
#include <vector>

#include "dmsim_amd.h"

namespace py = pybind11;

namespace {

void check(int rc)
{
    if (rc == DMX_OK) return;
    const std::string msg = dmx_last_error();
    if (rc == DMX_EINVAL) throw py::value_error(msg);
    if (rc == DMX_ENOMEM) throw std::bad_alloc();
    throw std::runtime_error(msg);
}

// The reference's Gate record (src/dmsim_amdgpu_omp.hpp:108-206).
struct Gate {
    int op;
    std::array<uint32_t, 5> qb;
    double theta, phi, lambda;
    Gate(int o, uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t e, double t, double p,
         double l)
        : op(o), qb{a, b, c, d, e}, theta(t), phi(p), lambda(l)
    {
    }
};

Gate* mk(int op, std::array<uint32_t, 5> q, double t = 0, double p = 0, double l = 0)
{
    return new Gate(op, q[0], q[1], q[2], q[3], q[4], t, p, l);
}

class Simulation {
public:
    Simulation(uint32_t n_qubits, uint32_t n_gpus)
    {
        check(dmx_create(n_qubits, n_gpus, &h_));
        check(dmx_set_report(h_, 1));  // the reference prints after every run
    }
    void set_report(bool on) { check(dmx_set_report(h_, on ? 1 : 0)); }
    ~Simulation() { dmx_destroy(h_); }
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    void append(const Gate& g)
    {
        const uint32_t qb[5] = {g.qb[0], g.qb[1], g.qb[2], g.qb[3], g.qb[4]};
        check(dmx_append(h_, g.op, qb, g.theta, g.phi, g.lambda));
    }
    void append_matrix(const std::vector<uint32_t>& qubits,
                       py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> m)
    {
        const size_t k = qubits.size();
        const size_t d = (size_t)1 << k;
        if (m.ndim() != 2 || (size_t)m.shape(0) != d || (size_t)m.shape(1) != d)
            throw py::value_error("matrix must be 2^k x 2^k for k qubits");
        std::vector<double> re(d * d), im(d * d);
        const std::complex<double>* p = m.data();
        for (size_t i = 0; i < d * d; i++) {
            re[i] = p[i].real();
            im[i] = p[i].imag();
        }
        check(dmx_append_matrix(h_, (uint32_t)k, qubits.data(), re.data(), im.data()));
    }
    Simulation* upload()
    {
        check(dmx_upload(h_));
        return this;
    }
    void run() { check(dmx_run(h_)); }
    void clear_circuit() { check(dmx_clear_circuit(h_)); }
    void reset() { check(dmx_reset(h_)); }
    void reset_dm() { check(dmx_reset_dm(h_)); }
    py::list measure(unsigned repetition, py::object seed)
    {
        // src/config.hpp:47 seeds with time(0); an explicit seed is an extension
        const uint32_t s = seed.is_none() ? (uint32_t)time(0) : seed.cast<uint32_t>();
        std::vector<uint32_t> out(repetition);
        check(dmx_measure(h_, repetition, s, out.data(), nullptr));
        py::list rtn;
        for (unsigned i = 0; i < repetition; i++) rtn.append(out[i]);
        return rtn;
    }
    py::tuple get_dm()
    {
        uint32_t n = 0;
        uint64_t dim = 0;
        check(dmx_info(h_, &n, &dim, nullptr));
        py::array_t<double> re({(py::ssize_t)dim, (py::ssize_t)dim});
        py::array_t<double> im({(py::ssize_t)dim, (py::ssize_t)dim});
        check(dmx_get_dm(h_, re.mutable_data(), im.mutable_data()));
        return py::make_tuple(re, im);
    }
    py::array_t<double> dm_real_res()
    {
        uint64_t dim = 0;
        check(dmx_info(h_, nullptr, &dim, nullptr));
        py::array_t<double> re({(py::ssize_t)dim, (py::ssize_t)dim});
        check(dmx_get_dm(h_, re.mutable_data(), nullptr));
        return re;
    }
    py::array_t<double> dm_imag_res()
    {
        uint64_t dim = 0;
        check(dmx_info(h_, nullptr, &dim, nullptr));
        py::array_t<double> im({(py::ssize_t)dim, (py::ssize_t)dim});
        check(dmx_get_dm(h_, nullptr, im.mutable_data()));
        return im;
    }
    py::array_t<double> get_diag()
    {
        uint64_t dim = 0;
        check(dmx_info(h_, nullptr, &dim, nullptr));
        py::array_t<double> d((py::ssize_t)dim);
        check(dmx_get_diag(h_, d.mutable_data()));
        return d;
    }
    py::array_t<double> expect_z()
    {
        uint32_t n = 0;
        check(dmx_info(h_, &n, nullptr, nullptr));
        py::array_t<double> z((py::ssize_t)n);
        check(dmx_expect_z(h_, z.mutable_data()));
        return z;
    }
    std::string dump()
    {
        size_t need = 0;
        check(dmx_dump(h_, nullptr, 0, &need));
        std::string s(need, '\0');
        check(dmx_dump(h_, &s[0], need, nullptr));
        s.resize(need ? need - 1 : 0);
        return s;
    }
    void set_method(int m) { check(dmx_set_method(h_, m)); }
    void set_fusion_max_k(uint32_t k) { check(dmx_set_fusion_max_k(h_, k)); }
    void set_fusion_permute(bool on) { check(dmx_set_fusion_permute(h_, on ? 1u : 0u)); }
    // extension: the gates of an OpenQASM 2.0 program (tool/dmsim_qasm.py's job)
    uint32_t append_qasm(const std::string& source)
    {
        uint32_t n = 0;
        check(dmx_append_qasm(h_, source.c_str(), &n));
        return n;
    }
    py::dict last_run_stats()
    {
        double ms = 0;
        uint32_t launches = 0, sweeps = 0;
        check(dmx_last_run_stats(h_, &ms, &launches, &sweeps));
        uint32_t fallbacks = 0;
        check(dmx_last_run_fallbacks(h_, &fallbacks));
        py::dict d;
        d["ms"] = ms;
        d["launches"] = launches;
        d["sweeps"] = sweeps;
        d["fallbacks"] = fallbacks;
        return d;
    }
    uint32_t n_qubits()
    {
        uint32_t n = 0;
        check(dmx_info(h_, &n, nullptr, nullptr));
        return n;
    }
    uint64_t dim()
    {
        uint64_t d = 0;
        check(dmx_info(h_, nullptr, &d, nullptr));
        return d;
    }
    uint32_t n_gates()
    {
        uint32_t g = 0;
        check(dmx_info(h_, nullptr, nullptr, &g));
        return g;
    }

private:
    dmx_sim* h_ = nullptr;
};

using Q = uint32_t;
using V = double;

}  // namespace

PYBIND11_MODULE(libdmsim_py_amdgpu_omp, m)
{
    m.doc() = "MI355X density-matrix engine (drop-in for DM-Sim's single-GPU module)";
    m.attr("METHOD_PLAIN") = DMX_METHOD_PLAIN;
    m.attr("METHOD_PLAIN_LAUNCH") = DMX_METHOD_PLAIN_LAUNCH;
    m.attr("METHOD_FUSED") = DMX_METHOD_FUSED;

    py::class_<Gate>(m, "Gate")
        .def(py::init<int, Q, Q, Q, Q, Q, V, V, V>())
        .def_readonly("op", &Gate::op)
        .def_readonly("theta", &Gate::theta)
        .def_readonly("phi", &Gate::phi)
        .def_readonly("lambda_", &Gate::lambda)
        .def_property_readonly("qubits", [](const Gate& g) { return g.qb; });

    py::class_<Simulation>(m, "Simulation")
        .def(py::init<Q, Q>())
        .def("append", &Simulation::append)
        .def("upload", &Simulation::upload, py::return_value_policy::reference)
        .def("clear_circuit", &Simulation::clear_circuit)
        .def("run", &Simulation::run)
        .def("sim", &Simulation::run)
        .def("reset", &Simulation::reset)
        .def("reset_dm", &Simulation::reset_dm)
        .def("measure", &Simulation::measure, py::arg("repetition"), py::arg("seed") = py::none())
        // extensions over the reference module
        .def("append_matrix", &Simulation::append_matrix)
        .def("get_dm", &Simulation::get_dm)
        .def_property_readonly("dm_real_res", &Simulation::dm_real_res)
        .def_property_readonly("dm_imag_res", &Simulation::dm_imag_res)
        .def("get_diag", &Simulation::get_diag)
        .def("expect_z", &Simulation::expect_z)
        .def("dump", &Simulation::dump)
        .def("set_method", &Simulation::set_method)
        .def("set_fusion_max_k", &Simulation::set_fusion_max_k)
        .def("set_fusion_permute", &Simulation::set_fusion_permute)
        .def("append_qasm", &Simulation::append_qasm)
        .def("last_run_stats", &Simulation::last_run_stats)
        .def("set_report", &Simulation::set_report)
        .def_property_readonly("n_qubits", &Simulation::n_qubits)
        .def_property_readonly("dim", &Simulation::dim)
        .def_property_readonly("n_gates", &Simulation::n_gates)
        // the 38 factories (src/dmsim_amdgpu_omp.hpp:594-782), same argument order
        .def_static("U3", [](V t, V p, V l, Q a) { return mk(DMX_U3, {a, 0, 0, 0, 0}, t, p, l); })
        .def_static("U2", [](V p, V l, Q a) { return mk(DMX_U2, {a, 0, 0, 0, 0}, 0., p, l); })
        .def_static("U1", [](V l, Q a) { return mk(DMX_U1, {a, 0, 0, 0, 0}, 0., 0., l); })
        .def_static("CX", [](Q a, Q b) { return mk(DMX_CX, {a, b, 0, 0, 0}); })
        .def_static("ID", [](Q a) { return mk(DMX_ID, {a, 0, 0, 0, 0}); })
        .def_static("X", [](Q a) { return mk(DMX_X, {a, 0, 0, 0, 0}); })
        .def_static("Y", [](Q a) { return mk(DMX_Y, {a, 0, 0, 0, 0}); })
        .def_static("Z", [](Q a) { return mk(DMX_Z, {a, 0, 0, 0, 0}); })
        .def_static("H", [](Q a) { return mk(DMX_H, {a, 0, 0, 0, 0}); })
        .def_static("S", [](Q a) { return mk(DMX_S, {a, 0, 0, 0, 0}); })
        .def_static("SDG", [](Q a) { return mk(DMX_SDG, {a, 0, 0, 0, 0}); })
        .def_static("T", [](Q a) { return mk(DMX_T, {a, 0, 0, 0, 0}); })
        .def_static("TDG", [](Q a) { return mk(DMX_TDG, {a, 0, 0, 0, 0}); })
        .def_static("RX", [](V t, Q a) { return mk(DMX_RX, {a, 0, 0, 0, 0}, t); })
        .def_static("RY", [](V t, Q a) { return mk(DMX_RY, {a, 0, 0, 0, 0}, t); })
        .def_static("RZ", [](V p, Q a) { return mk(DMX_RZ, {a, 0, 0, 0, 0}, 0., p); })
        .def_static("CZ", [](Q a, Q b) { return mk(DMX_CZ, {a, b, 0, 0, 0}); })
        .def_static("CY", [](Q a, Q b) { return mk(DMX_CY, {a, b, 0, 0, 0}); })
        .def_static("SWAP", [](Q a, Q b) { return mk(DMX_SWAP, {a, b, 0, 0, 0}); })
        .def_static("CH", [](Q a, Q b) { return mk(DMX_CH, {a, b, 0, 0, 0}); })
        .def_static("CCX", [](Q a, Q b, Q c) { return mk(DMX_CCX, {a, b, c, 0, 0}); })
        .def_static("CSWAP", [](Q a, Q b, Q c) { return mk(DMX_CSWAP, {a, b, c, 0, 0}); })
        .def_static("CRX", [](V l, Q a, Q b) { return mk(DMX_CRX, {a, b, 0, 0, 0}, 0., 0., l); })
        .def_static("CRY", [](V l, Q a, Q b) { return mk(DMX_CRY, {a, b, 0, 0, 0}, 0., 0., l); })
        .def_static("CRZ", [](V l, Q a, Q b) { return mk(DMX_CRZ, {a, b, 0, 0, 0}, 0., 0., l); })
        .def_static("CU1", [](V l, Q a, Q b) { return mk(DMX_CU1, {a, b, 0, 0, 0}, 0., 0., l); })
        .def_static("CU3",
                    [](V t, V p, V l, Q a, Q b) { return mk(DMX_CU3, {a, b, 0, 0, 0}, t, p, l); })
        .def_static("RXX", [](V t, Q a, Q b) { return mk(DMX_RXX, {a, b, 0, 0, 0}, t); })
        .def_static("RZZ", [](V t, Q a, Q b) { return mk(DMX_RZZ, {a, b, 0, 0, 0}, t); })
        .def_static("RCCX", [](Q a, Q b, Q c) { return mk(DMX_RCCX, {a, b, c, 0, 0}); })
        .def_static("RC3X", [](Q a, Q b, Q c, Q d) { return mk(DMX_RC3X, {a, b, c, d, 0}); })
        .def_static("C3X", [](Q a, Q b, Q c, Q d) { return mk(DMX_C3X, {a, b, c, d, 0}); })
        .def_static("C3SQRTX",
                    [](Q a, Q b, Q c, Q d) { return mk(DMX_C3SQRTX, {a, b, c, d, 0}); })
        .def_static("C4X",
                    [](Q a, Q b, Q c, Q d, Q e) { return mk(DMX_C4X, {a, b, c, d, e}); })
        .def_static("R", [](V t, Q a) { return mk(DMX_R, {a, 0, 0, 0, 0}, t); })
        .def_static("SRN", [](Q a) { return mk(DMX_SRN, {a, 0, 0, 0, 0}); })
        .def_static("W", [](Q a) { return mk(DMX_W, {a, 0, 0, 0, 0}); })
        .def_static("RYY", [](V t, Q a, Q b) { return mk(DMX_RYY, {a, b, 0, 0, 0}, t); });
}
