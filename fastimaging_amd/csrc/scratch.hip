// Device scratch region of the no-allocation entry point
// (stp_image_visibilities_device_noalloc): a file-scope device array that every
// device holds from the first such call until the process ends.  Host code never
// passes the array's name to a kernel: it resolves the array's address on each
// device once (hipGetSymbolAddress, with that device current) and hands that
// address to the workspace carving of the context (Arena).
#include <mutex>

#include "common.h"

static __device__ __align__(256) unsigned char scratch[STP_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mutex;
void* g_ptr[MAX_DEVICES] = {};

} // namespace

namespace stp_hip {

// Address of `scratch` on `device`, which must be the current device.  Resolved
// on the first call per device and kept.
void* scratch_region(int device)
{
    if (device < 0 || device >= MAX_DEVICES)
        throw Error(STP_ERR_INVALID_ARGUMENT, "device index out of range for the scratch region");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(STP_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[device] = p;
    }
    return g_ptr[device];
}

} // namespace stp_hip
