// Scratch memory of the entry point that allocates nothing
// (fi_icgn_compute_noalloc): the per-POI reference records (PoiPrep, 176 B)
// of one call, in a file-scope device array that every GPU the library runs
// on holds from the code object's load until the process ends.  Its device
// address is taken once per device with hipGetSymbolAddress; host code uses
// that address only (the array's name in host code is not a device address).
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// race.  Each call writes the record of every POI it processes before the
// IC-GN kernel reads it, so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "fi_host.h"

namespace fi {

static __device__ __align__(256) PoiPrep scratch[kNoAllocMaxPois];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
PoiPrep* g_addr[kMaxDevices] = {};
}  // namespace

PoiPrep* noalloc_prep_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) fail(FI_E_INVALID_ARGUMENT, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_addr[device] == nullptr) {
        // the current device is `device` (the caller's DeviceGuard)
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            fail(FI_E_DEVICE, std::string("hipGetSymbolAddress of the no-allocation scratch failed: ") +
                                  hipGetErrorString(e));
        g_addr[device] = static_cast<PoiPrep*>(p);
    }
    return g_addr[device];
}

}  // namespace fi
