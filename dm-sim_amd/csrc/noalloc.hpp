// noalloc.hpp -- the entry gather that allocates nothing (noalloc.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dmx {

// Most entries one allocation-free gather takes: its device staging (the
// indices and both output planes, 24 bytes an entry, 1.5 MiB in all) is a
// fixed array that each GPU the library runs on holds until the process ends.
constexpr uint64_t kNoAllocMaxEntries = 65536;

// Looks up, once per device, the device address of that array on the current
// device (hipGetSymbolAddress) and keeps it; *out receives it.
hipError_t noalloc_scratch(void** out);

// The gather of launch_gather with its staging in the fixed array: d_idx = the
// first count uint64 slots, outputs after them; copies idx in and the entries
// out on `stream` and waits for them.  count <= kNoAllocMaxEntries (checked
// by the caller).  Calls on one device share the array: overlapping calls on
// one GPU race.
hipError_t gather_noalloc(const double* re, const double* im, int n, const uint64_t* idx,
                          uint64_t count, bool transposed, double* out_re, double* out_im,
                          hipStream_t stream);

}  // namespace dmx
