// The reference's 10-qubit adder example written against the C++ face of
// this engine (include/dmsim_amd.hpp): cin = 0, a = 0001, b = 1111; every shot
// reads 1000000010 (= 514).  Exits non-zero if the state or the shots differ.
// Optional argument: execution method (0 persistent plain, the default;
// 1 plain, one launch per primitive; 2 fused).
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "dmsim_amd.hpp"

using namespace DMSim;

static void majority(Simulation& sim, IdxType a, IdxType b, IdxType c)
{
    sim.append(Simulation::CX(c, b));
    sim.append(Simulation::CX(c, a));
    sim.append(Simulation::CCX(a, b, c));
}

static void unmaj(Simulation& sim, IdxType a, IdxType b, IdxType c)
{
    sim.append(Simulation::CCX(a, b, c));
    sim.append(Simulation::CX(c, a));
    sim.append(Simulation::CX(a, b));
}

int main(int argc, char** argv)
{
    Simulation sim(10, 1);
    if (argc > 1) detail::check(dmx_set_method(sim.handle(), std::atoi(argv[1])));
    for (IdxType q : {1u, 5u, 6u, 7u, 8u}) {
        Gate* g = Simulation::X(q);
        sim.append(g);
        delete g;
    }
    majority(sim, 0, 5, 1);
    majority(sim, 1, 6, 2);
    majority(sim, 2, 7, 3);
    majority(sim, 3, 8, 4);
    sim.append(Simulation::CX(4, 9));
    unmaj(sim, 3, 8, 4);
    unmaj(sim, 2, 7, 3);
    unmaj(sim, 1, 6, 2);
    unmaj(sim, 0, 5, 1);
    sim.upload();
    sim.sim();
    uint32_t fallbacks = 0;
    detail::check(dmx_last_run_fallbacks(sim.handle(), &fallbacks));
    printf("persistent segments run as single launches: %u\n", fallbacks);
    IdxType* res = sim.measure(5, 1234u);
    print_measurement(res, 10, 5);
    int bad = 0;
    for (int i = 0; i < 5; i++) bad += res[i] != 514;
    delete[] res;
    const size_t p = (size_t)514 * sim.dim + 514;
    printf("rho[514][514] = %.17g, gates = %u\n", sim.dm_real_res[p], sim.n_gates);
    if (std::fabs(sim.dm_real_res[p] - 1.0) > 1e-12 || sim.n_gates != 30) bad++;
    return bad ? 1 : 0;
}
