// dmd_noalloc.h -- device scratch of the allocation-free entry point.
#pragma once
#include <hip/hip_runtime.h>

namespace dmd {

// Sites whose diagonal blocks the scratch array holds (9 doubles each:
// 4.5 MiB per device).
constexpr int kScratchSites = 65536;

// Takes the device address of this device's copy of the scratch array
// (hipGetSymbolAddress); call once with `device` current before any use.
hipError_t noalloc_init(int device);
// That address, or nullptr before noalloc_init(device).
double* noalloc_scratch(int device);
// out[9 perm[base + i] + t] = blocks[9 i + t], i < n, t < 9.
hipError_t launch_scatter_blocks(hipStream_t st, int base, int n, const int* perm,
                                 const double* blocks, double* out);

}  // namespace dmd
