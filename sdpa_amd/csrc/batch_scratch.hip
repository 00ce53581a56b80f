// batch_scratch.hip -- the fixed device array of sdpa_batch_solve_noalloc
// (include/sdpa_batch.h).  It sits in a code object of its own: the array
// is reserved on a device when this code object is loaded there, which
// happens only for callers of the no-allocation entry point.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

namespace sdpa_amd {

namespace {

constexpr size_t kScratchBytes = (size_t)128 << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

// touches the array so that the code object holds a kernel (lazy loading
// loads code objects by their kernels); never launched
__global__ void k_scratch_anchor(int i) { scratch[i] = 0; }

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};

}  // namespace

size_t noalloc_scratch_bytes() { return kScratchBytes; }

// The array's device address on `device`, looked up once per device with
// that device current (the host-side name is only a stand-in, never a
// device address).  The caller has made `device` current.
int noalloc_scratch_address(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return -100;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_addr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -100;
    g_addr[device] = p;
  }
  *out = g_addr[device];
  return 0;
}

}  // namespace sdpa_amd
