// fuse.cpp -- fusion planner and fused executor.
#include "fuse.hpp"

#include "dmsim_amd.h"
#include "engine.hpp"

namespace dmx {

int build_fused_plan(const std::vector<Gate>&, int, int, FusedPlan** out)
{
    *out = nullptr;
    set_error("fused method not available in this build");
    return DMX_ESTATE;
}

}  // namespace dmx
