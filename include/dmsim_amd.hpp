// dmsim_amd.hpp -- header-only C++ face of libdmsim_amd.so with the reference's
// C++ API: namespace DMSim, enum OP, class Gate, class Simulation with its 38
// static factories and methods, print_measurement
// (src/dmsim_amdgpu_omp.hpp:51-57, :108-206, :208-829; src/util_amdgpu.hpp:141-156),
// so a program written against the reference's header (e.g.
// example/adder_n10_amdgpu_omp.cpp) builds against this one by changing the
// #include lines.  Everything goes through the C ABI of dmsim_amd.h.
//
// Differences: failures throw std::runtime_error carrying dmx_last_error()
// (the reference asserts or exits); n_gpus must be 1; dm_real_res and
// dm_imag_res copy rho from the device on first use after each run (the
// reference copies after every run); measure() returns new[] storage, free it
// with delete[].
#ifndef DMSIM_AMD_HPP
#define DMSIM_AMD_HPP

#include <cstdio>
#include <ctime>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "dmsim_amd.h"

namespace DMSim {

using IdxType = unsigned;
using ValType = double;

// src/dmsim_amdgpu_omp.hpp:51-57
enum OP {
    U3, U2, U1, CX, ID, X, Y, Z, H, S, SDG, T, TDG, RX, RY, RZ, CZ, CY, SWAP, CH, CCX, CSWAP,
    CRX, CRY, CRZ, CU1, CU3, RXX, RZZ, RCCX, RC3X, C3X, C3SQRTX, C4X, R, SRN, W, RYY
};

namespace detail {
inline void check(int rc)
{
    if (rc != DMX_OK) throw std::runtime_error(std::string("DM-Sim: ") + dmx_last_error());
}
}  // namespace detail

// The gate record (src/dmsim_amdgpu_omp.hpp:108-206).
class Gate {
public:
    Gate(enum OP op, IdxType q0, IdxType q1, IdxType q2, IdxType q3, IdxType q4, ValType t,
         ValType p, ValType l)
        : op_name(op), qb0(q0), qb1(q1), qb2(q2), qb3(q3), qb4(q4), theta(t), phi(p), lambda(l)
    {
    }
    enum OP op_name;
    IdxType qb0, qb1, qb2, qb3, qb4;
    ValType theta, phi, lambda;
};

class Simulation;

// One result plane: rho is copied from the device on first use after a run.
class ResultPlane {
public:
    ValType operator[](size_t i) const { return data()[i]; }
    operator ValType*() const { return data(); }
    ValType* data() const;

private:
    friend class Simulation;
    ResultPlane(Simulation* s, int plane) : sim_(s), plane_(plane) {}
    Simulation* sim_;
    int plane_;
};

class Simulation {
public:
    Simulation(IdxType n, IdxType g)
        : n_qubits(n), n_gpus(g), dim((IdxType)1 << n), n_gates(0), dm_real_res(this, 0),
          dm_imag_res(this, 1)
    {
        detail::check(dmx_create(n, g, &h_));
        detail::check(dmx_set_report(h_, 1));  // PRINT_MEA_PER_CIRCUIT, src/config.hpp:25
    }
    ~Simulation() { dmx_destroy(h_); }
    // Extension: turns the per-run report of sim() off (on by default).
    void set_report(bool on) { detail::check(dmx_set_report(h_, on ? 1 : 0)); }
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    // append(Gate*), :346-359: the record is copied; the caller keeps ownership
    void append(Gate* g)
    {
        const uint32_t qb[5] = {g->qb0, g->qb1, g->qb2, g->qb3, g->qb4};
        detail::check(dmx_append(h_, (int32_t)g->op_name, qb, g->theta, g->phi, g->lambda));
        n_gates++;
    }
    Simulation* upload()
    {
        detail::check(dmx_upload(h_));
        return this;
    }
    void sim()
    {
        detail::check(dmx_run(h_));
        stale_ = true;
    }
    void run() { sim(); }
    void clear_circuit()
    {
        detail::check(dmx_clear_circuit(h_));
        n_gates = 0;
    }
    void reset()
    {
        detail::check(dmx_reset(h_));
        n_gates = 0;
        stale_ = true;
    }
    void reset_dm()
    {
        detail::check(dmx_reset_dm(h_));
        stale_ = true;
    }
    // measure(repetition), :536-564, seeded with time(0) like the reference
    IdxType* measure(unsigned repetition = 10)
    {
        return measure(repetition, (unsigned)time(0));
    }
    IdxType* measure(unsigned repetition, unsigned seed)
    {
        IdxType* out = new IdxType[repetition > 0 ? repetition : 1];
        const int rc = dmx_measure(h_, repetition, seed, out, nullptr);
        if (rc != DMX_OK) {
            delete[] out;
            detail::check(rc);
        }
        return out;
    }
    std::string dump()
    {
        size_t need = 0;
        detail::check(dmx_dump(h_, nullptr, 0, &need));
        std::vector<char> buf(need);
        detail::check(dmx_dump(h_, buf.data(), need, nullptr));
        return std::string(buf.data());
    }
    void print_res_sv()
    {
        printf("----- Real SV ------\n");
        for (IdxType i = 0; i < dim; i++) printf("%lf ", dm_real_res[(size_t)i * dim + i]);
        printf("\n----- Imag SV ------\n");
        for (IdxType i = 0; i < dim; i++) printf("%lf ", dm_imag_res[(size_t)i * dim + i]);
        printf("\n");
    }
    void print_res_dm()
    {
        printf("----- Real DM------\n");
        for (IdxType i = 0; i < dim; i++) {
            for (IdxType j = 0; j < dim; j++) printf("%lf ", dm_real_res[(size_t)i * dim + j]);
            printf("\n");
        }
        printf("----- Imag DM------\n");
        for (IdxType i = 0; i < dim; i++) {
            for (IdxType j = 0; j < dim; j++) printf("%lf ", dm_imag_res[(size_t)i * dim + j]);
            printf("\n");
        }
    }

    // Static factories with the reference's parameter placement (:594-782).
    static Gate* U3(ValType t, ValType p, ValType l, IdxType m) { return mk(OP::U3, m, 0, 0, 0, 0, t, p, l); }
    static Gate* U2(ValType p, ValType l, IdxType m) { return mk(OP::U2, m, 0, 0, 0, 0, 0, p, l); }
    static Gate* U1(ValType l, IdxType m) { return mk(OP::U1, m, 0, 0, 0, 0, 0, 0, l); }
    static Gate* CX(IdxType m, IdxType n) { return mk(OP::CX, m, n); }
    static Gate* ID(IdxType m) { return mk(OP::ID, m); }
    static Gate* X(IdxType m) { return mk(OP::X, m); }
    static Gate* Y(IdxType m) { return mk(OP::Y, m); }
    static Gate* Z(IdxType m) { return mk(OP::Z, m); }
    static Gate* H(IdxType m) { return mk(OP::H, m); }
    static Gate* S(IdxType m) { return mk(OP::S, m); }
    static Gate* SDG(IdxType m) { return mk(OP::SDG, m); }
    static Gate* T(IdxType m) { return mk(OP::T, m); }
    static Gate* TDG(IdxType m) { return mk(OP::TDG, m); }
    static Gate* RX(ValType t, IdxType m) { return mk(OP::RX, m, 0, 0, 0, 0, t); }
    static Gate* RY(ValType t, IdxType m) { return mk(OP::RY, m, 0, 0, 0, 0, t); }
    static Gate* RZ(ValType p, IdxType m) { return mk(OP::RZ, m, 0, 0, 0, 0, 0, p); }
    static Gate* CZ(IdxType m, IdxType n) { return mk(OP::CZ, m, n); }
    static Gate* CY(IdxType m, IdxType n) { return mk(OP::CY, m, n); }
    static Gate* SWAP(IdxType m, IdxType n) { return mk(OP::SWAP, m, n); }
    static Gate* CH(IdxType m, IdxType n) { return mk(OP::CH, m, n); }
    static Gate* CCX(IdxType l, IdxType m, IdxType n) { return mk(OP::CCX, l, m, n); }
    static Gate* CSWAP(IdxType l, IdxType m, IdxType n) { return mk(OP::CSWAP, l, m, n); }
    static Gate* CRX(ValType l, IdxType m, IdxType n) { return mk(OP::CRX, m, n, 0, 0, 0, 0, 0, l); }
    static Gate* CRY(ValType l, IdxType m, IdxType n) { return mk(OP::CRY, m, n, 0, 0, 0, 0, 0, l); }
    static Gate* CRZ(ValType l, IdxType m, IdxType n) { return mk(OP::CRZ, m, n, 0, 0, 0, 0, 0, l); }
    static Gate* CU1(ValType l, IdxType m, IdxType n) { return mk(OP::CU1, m, n, 0, 0, 0, 0, 0, l); }
    static Gate* CU3(ValType t, ValType p, ValType l, IdxType m, IdxType n)
    {
        return mk(OP::CU3, m, n, 0, 0, 0, t, p, l);
    }
    static Gate* RXX(ValType t, IdxType m, IdxType n) { return mk(OP::RXX, m, n, 0, 0, 0, t); }
    static Gate* RZZ(ValType t, IdxType m, IdxType n) { return mk(OP::RZZ, m, n, 0, 0, 0, t); }
    static Gate* RCCX(IdxType l, IdxType m, IdxType n) { return mk(OP::RCCX, l, m, n); }
    static Gate* RC3X(IdxType a, IdxType b, IdxType c, IdxType d) { return mk(OP::RC3X, a, b, c, d); }
    static Gate* C3X(IdxType a, IdxType b, IdxType c, IdxType d) { return mk(OP::C3X, a, b, c, d); }
    static Gate* C3SQRTX(IdxType a, IdxType b, IdxType c, IdxType d)
    {
        return mk(OP::C3SQRTX, a, b, c, d);
    }
    static Gate* C4X(IdxType a, IdxType b, IdxType c, IdxType d, IdxType e)
    {
        return mk(OP::C4X, a, b, c, d, e);
    }
    static Gate* R(ValType t, IdxType m) { return mk(OP::R, m, 0, 0, 0, 0, t); }
    static Gate* SRN(IdxType m) { return mk(OP::SRN, m); }
    static Gate* W(IdxType m) { return mk(OP::W, m); }
    static Gate* RYY(ValType t, IdxType m, IdxType n) { return mk(OP::RYY, m, n, 0, 0, 0, t); }

    // Public members of the reference's object (:786-800).
    const IdxType n_qubits;
    const IdxType n_gpus;
    const IdxType dim;
    IdxType n_gates;
    ResultPlane dm_real_res;
    ResultPlane dm_imag_res;

    // Extension: the C ABI handle (fused method, profiling, ...).
    dmx_sim* handle() const { return h_; }

private:
    friend class ResultPlane;
    static Gate* mk(enum OP op, IdxType a, IdxType b = 0, IdxType c = 0, IdxType d = 0,
                    IdxType e = 0, ValType t = 0, ValType p = 0, ValType l = 0)
    {
        return new Gate(op, a, b, c, d, e, t, p, l);
    }
    void download()
    {
        const size_t num = (size_t)dim * dim;
        if (!re_) {
            re_.reset(new ValType[num]);
            im_.reset(new ValType[num]);
        }
        detail::check(dmx_get_dm(h_, re_.get(), im_.get()));
        stale_ = false;
    }
    dmx_sim* h_ = nullptr;
    bool stale_ = true;
    std::unique_ptr<ValType[]> re_, im_;
};

inline ValType* ResultPlane::data() const
{
    if (sim_->stale_ || !sim_->re_) sim_->download();
    return plane_ == 0 ? sim_->re_.get() : sim_->im_.get();
}

// src/util_amdgpu.hpp:141-156
inline void print_binary(IdxType v, int width)
{
    for (int i = width - 1; i >= 0; i--) putchar('0' + ((v >> i) & 1));
}

inline void print_measurement(IdxType* res, IdxType n_qubits, int repetition)
{
    printf("\n===============  Measurement (tests=%d) ================\n", repetition);
    for (int i = 0; i < repetition; i++) {
        printf("Test-%d: ", i);
        print_binary(res[i], (int)n_qubits);
        printf("\n");
    }
}

}  // namespace DMSim

#endif  // DMSIM_AMD_HPP
