// dmd_noalloc.hip -- the device scratch of the allocation-free entry point.
//
// dmdgpu_diag_blocks_noalloc stages the diagonal Hessian blocks of the
// context's instance in this file-scope array instead of a buffer the library
// allocates, then writes them in the caller's site order into the caller's
// device buffer.  Each device holds its own copy of the array (at its own
// address) until the process ends; every call on one device shares it, so
// calls that overlap on one device race.
//
// In host code `scratch` names a host-side stand-in, not device memory: the
// device address of each device's copy comes from hipGetSymbolAddress, taken
// once per device in noalloc_init (called by dmdgpu_create with the device
// current) and used everywhere the host passes the array to a kernel.
#include <mutex>

#include "dmd_device.h"
#include "dmd_noalloc.h"

namespace dmd {

static __device__ __align__(256) double scratch[kScratchSites * 9];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};
std::mutex g_mu;
}  // namespace

hipError_t noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = p;
  return hipSuccess;
}

double* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return static_cast<double*>(g_scratch[device]);
}

// out[9 perm[base + i] + t] = blocks[9 i + t] for the n sites of the instance.
__global__ __launch_bounds__(256) void k_scatter_blocks(int base, int n, const int* perm,
                                                        const double* blocks, double* out) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= 9LL * n) return;
  const int i = (int)(e / 9), t = (int)(e % 9);
  out[9LL * perm[base + i] + t] = blocks[e];
}

hipError_t launch_scatter_blocks(hipStream_t st, int base, int n, const int* perm,
                                 const double* blocks, double* out) {
  if (n <= 0) return hipSuccess;
  const long long cnt = 9LL * n;
  k_scatter_blocks<<<(int)((cnt + 255) / 256), 256, 0, st>>>(base, n, perm, blocks, out);
  return hipGetLastError();
}

}  // namespace dmd
