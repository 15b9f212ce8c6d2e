// noalloc.hip -- the static device buffer of the allocation-free entry point
// (ela3d_run_noalloc in capi.cpp): a whole shot's device state, placed here
// instead of in memory obtained from hipMalloc.
#include <mutex>

#include "shot.hpp"

namespace ela3d {

// One copy per GPU the library runs on, held until the process ends.  In host
// code `scratch` is a host-side stand-in, not a device address: only the address
// hipGetSymbolAddress returns for the current device may be handed to kernels or
// copies.
static __device__ __align__(256) unsigned char scratch[ELA3D_NOALLOC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on `device`, looked up once per device with that
// device current, and cached.
void* noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) throw Error{ELA3D_E_INVALID, "device ordinal out of range"};
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        int prev = 0;
        HIP_CHECK(hipGetDevice(&prev));
        HIP_CHECK(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(prev);
        if (e != hipSuccess || !p)
            throw Error{ELA3D_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
        if (reinterpret_cast<uintptr_t>(p) % 256 != 0) throw Error{ELA3D_E_HIP, "static device buffer misaligned"};
        g_scratch[device] = p;
    }
    return g_scratch[device];
}

}  // namespace ela3d
