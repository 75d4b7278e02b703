// dmsim_amd_xacc.hpp -- the backend interface of the reference's XACC plugin
// (xacc/DmSimApi.hpp:5-63: namespace DmSim, enum class OP, DmSimBackend with
// init / addGate / measure / finalize, getGpuDmSim()) implemented over the C
// ABI of libdmsim_amd.so, header-only.  An XACC accelerator written against
// DmSimApi.hpp (xacc/DmSimAccelerator.cpp) links this instead of the CUDA
// runner.
//
// Against the reference's runners (xacc/cpu_omp/dm_sim_omp_runner.cpp,
// xacc/nvidia_omp/NvidiaOmpRunner.cu): CH applies CH (the runners append
// SWAP, dm_sim_omp_runner.cpp:108-111); every op of the enum is accepted (the
// runners stop on ID, CCX, CSWAP, CU3, RXX, RZZ, ...); wrong operand or
// parameter counts throw std::invalid_argument instead of asserting.  Circuits
// of four or more qubits run the fused method, smaller ones the plain method
// with one launch per primitive.
#ifndef DMSIM_AMD_XACC_HPP
#define DMSIM_AMD_XACC_HPP

#include <cstdint>
#include <ctime>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "dmsim_amd.h"

namespace DmSim {

enum class OP {
    U3, U2, U1, CX, ID, X, Y, Z, H, S, SDG, T, TDG, RX, RY, RZ, CZ, CY, SWAP, CH, CCX, CSWAP,
    CRX, CRY, CRZ, CU1, CU3, RXX, RZZ, RCCX, RC3X, C3X, C3SQRTX, C4X, R, SRN, W, RYY
};

class DmSimBackend {
public:
    virtual ~DmSimBackend() = default;
    virtual void init(int n_qubits, int n_gpus = 1) = 0;
    virtual void addGate(OP op, const std::vector<int>& qubits,
                         const std::vector<double>& params = {}) = 0;
    virtual std::vector<int64_t> measure(int shots) = 0;
    virtual void finalize() = 0;
};

class AmdGpuBackend : public DmSimBackend {
public:
    ~AmdGpuBackend() override { finalize(); }

    void init(int n_qubits, int n_gpus = 1) override
    {
        finalize();
        check(dmx_create((uint32_t)n_qubits, (uint32_t)n_gpus, &h_));
        check(dmx_set_method(h_, n_qubits >= 4 ? DMX_METHOD_FUSED : DMX_METHOD_PLAIN_LAUNCH));
    }

    void addGate(OP op, const std::vector<int>& qubits,
                 const std::vector<double>& params = {}) override
    {
        if (!h_) throw std::logic_error("DmSim backend: addGate before init");
        const int code = (int)op;
        // parameters as the factories place them: slot 0 theta, 1 phi, 2 lambda
        // (src/dmsim_amdgpu_omp.hpp:594-782)
        static const int kArity[38] = {1, 1, 1, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2,
                                       2, 3, 3, 2, 2, 2, 2, 2, 2, 2, 3, 4, 4, 4, 5, 1, 1, 1, 2};
        std::vector<int> slots;
        switch (op) {
        case OP::U3: case OP::CU3: slots = {0, 1, 2}; break;
        case OP::U2: slots = {1, 2}; break;
        case OP::U1: case OP::CRX: case OP::CRY: case OP::CRZ: case OP::CU1: slots = {2}; break;
        case OP::RZ: slots = {1}; break;
        case OP::RX: case OP::RY: case OP::RXX: case OP::RZZ: case OP::RYY: case OP::R:
            slots = {0};
            break;
        default: break;
        }
        if ((int)qubits.size() != kArity[code] || params.size() != slots.size())
            throw std::invalid_argument("DmSim backend: op " + std::to_string(code) + " takes " +
                                        std::to_string(kArity[code]) + " qubits and " +
                                        std::to_string(slots.size()) + " parameters");
        uint32_t qb[5] = {0, 0, 0, 0, 0};
        for (size_t i = 0; i < qubits.size(); i++) qb[i] = (uint32_t)qubits[i];
        double p[3] = {0, 0, 0};
        for (size_t i = 0; i < params.size(); i++) p[slots[i]] = params[i];
        check(dmx_append(h_, code, qb, p[0], p[1], p[2]));
    }

    // upload, run, sample (the runners' measure, dm_sim_omp_runner.cpp:150-164)
    std::vector<int64_t> measure(int shots) override
    {
        if (!h_) throw std::logic_error("DmSim backend: measure before init");
        check(dmx_upload(h_));
        check(dmx_run(h_));
        std::vector<uint32_t> raw(shots > 0 ? shots : 1);
        check(dmx_measure(h_, (uint32_t)shots, seed_ ? seed_ : (uint32_t)time(0), raw.data(),
                          nullptr));
        return std::vector<int64_t>(raw.begin(), raw.begin() + (shots > 0 ? shots : 0));
    }

    void finalize() override
    {
        if (h_) dmx_destroy(h_);
        h_ = nullptr;
    }

    // Extensions: a fixed sampling seed (0 = time(0), as the reference), the handle.
    void set_seed(uint32_t seed) { seed_ = seed; }
    dmx_sim* handle() const { return h_; }

private:
    static void check(int rc)
    {
        if (rc != DMX_OK) throw std::runtime_error(std::string("DmSim backend: ") + dmx_last_error());
    }
    dmx_sim* h_ = nullptr;
    uint32_t seed_ = 0;
};

inline std::shared_ptr<DmSimBackend> getGpuDmSim() { return std::make_shared<AmdGpuBackend>(); }

}  // namespace DmSim

#endif  // DMSIM_AMD_XACC_HPP
