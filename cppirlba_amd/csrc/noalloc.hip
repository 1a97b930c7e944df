// Fixed device scratch for the allocation-free entry point
// (irlba_hip_sddmm_noalloc): callers that may not allocate device memory
// (code under stream capture, tight training or serving loops) get the SDDMM's
// temporaries from this file-scope array instead of the context's grow-only
// workspace.
//
// Every GPU the library runs on holds its own copy of the array (its size is
// kNoAllocScratchBytes) from the time the code object is loaded on that device
// until the process ends.  Host code never passes the array's name to a
// kernel: the name is a host-side stand-in, so its device address is looked up
// once per device with hipGetSymbolAddress (with that device current) and kept
// in a per-device table.  Calls that use the scratch on one GPU share it: they
// must not overlap in time (two contexts on one device, or two threads on two
// contexts of one device, race on it).
#include <mutex>
#include <vector>

#include "common.hpp"

namespace irlba_hip {

static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_addr;  // per device, null until resolved

__global__ void k_clear_flag(int* flag) {
    if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = 0;
}

}  // namespace

void* noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0) throw std::runtime_error("invalid device for the fixed scratch");
    if (static_cast<size_t>(device) >= g_scratch_addr.size()) g_scratch_addr.resize(device + 1, nullptr);
    if (!g_scratch_addr[device]) {
        int current = -1;
        IRLBA_HIP_CHECK(hipGetDevice(&current));
        if (current != device) throw std::runtime_error("the fixed scratch must be resolved with its device current");
        void* p = nullptr;
        IRLBA_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the fixed scratch");
        g_scratch_addr[device] = p;
    }
    return g_scratch_addr[device];
}

// A kernel, not hipMemsetAsync: the call may be recorded into a graph by its
// caller, and a kernel node clears the flag on every replay.
void launch_clear_flag(int* flag, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, s, flag);
    IRLBA_HIP_LAUNCH_CHECK();
}

}  // namespace irlba_hip
