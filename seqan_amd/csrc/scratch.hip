// scratch.hip -- the device array behind sqmi_align_score_noalloc.
//
// One file-scope array in device memory: every GPU the library runs on holds
// its own copy, at its own address, for the life of the process.  Host code
// never uses the array's name as an address (on the host it is only a
// stand-in for the symbol); the entry point's setup looks the device address
// up once per device with hipGetSymbolAddress, with that device current, and
// every kernel argument, fill and copy then uses that address.
#include <mutex>

#include "kernels.hpp"

namespace sqmi {

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
void *gScratchPtr[kMaxDevices] = {};
}  // namespace

hipError_t scratchAddress(int device, void **out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratchPtr[device]) {
        void *ptr = nullptr;
        const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (err != hipSuccess)
            return err;
        gScratchPtr[device] = ptr;
    }
    *out = gScratchPtr[device];
    return hipSuccess;
}

}  // namespace sqmi
