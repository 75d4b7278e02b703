// The reference's XACC deuteron check (xacc/tests/DmSimCpuAcceleratorTester.cpp:20-47)
// against the DmSimBackend interface of include/dmsim_amd_xacc.hpp: for 20 angles t
// in [-pi, pi], X(q0) Ry(q1, t) CX(q1, q0) H(q0) H(q1), 100000 shots, <Z0 Z1> within
// 0.01 of the reference's expected values; then CH acting as CH (the reference's
// runners append SWAP) and ops the runners reject.  Exits non-zero on a mismatch.
#include <cmath>
#include <cstdio>

#include "dmsim_amd_xacc.hpp"

using DmSim::OP;

int main()
{
    const double expected[20] = {0.0,       -0.324699, -0.614213, -0.837166, -0.9694,
                                 -0.996584, -0.915773, -0.735724, -0.475947, -0.164595,
                                 0.164595,  0.475947,  0.735724,  0.915773,  0.996584,
                                 0.9694,    0.837166,  0.614213,  0.324699,  0.0};
    const double pi = 3.14159265358979323846;
    auto be = DmSim::getGpuDmSim();
    static_cast<DmSim::AmdGpuBackend*>(be.get())->set_seed(2024);
    int bad = 0;
    for (int i = 0; i < 20; i++) {
        const double t = -pi + 2 * pi * i / 19;
        be->init(2);
        be->addGate(OP::X, {0});
        be->addGate(OP::RY, {1}, {t});
        be->addGate(OP::CX, {1, 0});
        be->addGate(OP::H, {0});
        be->addGate(OP::H, {1});
        const auto shots = be->measure(100000);
        double z = 0;
        for (int64_t s : shots) z += (__builtin_popcountll(s) & 1) ? -1.0 : 1.0;
        z /= (double)shots.size();
        const bool ok = std::fabs(z - expected[i]) <= 0.01;
        bad += !ok;
        printf("t=%+.4f <Z0Z1>=%+.5f expected %+.6f %s\n", t, z, expected[i], ok ? "ok" : "FAIL");
        be->finalize();
    }
    // CH on |q0 = 1>: target in |+>, outcomes 1 and 3 about equally
    be->init(2);
    be->addGate(OP::X, {0});
    be->addGate(OP::CH, {0, 1});
    const auto ch = be->measure(4000);
    int c1 = 0, c3 = 0;
    for (int64_t s : ch) c1 += s == 1, c3 += s == 3;
    printf("CH: %d x 01, %d x 11\n", c1, c3);
    bad += (c1 + c3 != 4000) || c1 < 1700 || c3 < 1700;
    // ops the reference runners stop on: CCX(0,1,2) on |110> -> |111>, CU3, RXX
    be->init(5);
    be->addGate(OP::X, {0});
    be->addGate(OP::X, {1});
    be->addGate(OP::CCX, {0, 1, 2});
    be->addGate(OP::CU3, {3, 4}, {0.3, 0.2, 0.1});
    be->addGate(OP::RXX, {3, 4}, {0.0});
    const auto ccx = be->measure(10);
    for (int64_t s : ccx) bad += s != 7;
    printf("CCX: %lld\n", (long long)ccx[0]);
    try {
        be->addGate(OP::U3, {0}, {0.1});
        bad++;
    } catch (const std::invalid_argument&) {
    }
    return bad ? 1 : 0;
}
