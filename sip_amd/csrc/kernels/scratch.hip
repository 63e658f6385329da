// Scratch memory of the no-allocation exec entry point: one file-scope __device__ array, of which
// every GPU the library runs on holds its own copy at its own address. In host code the array's
// name is only a host-side stand-in, never a device address, so host code uses the address that
// hipGetSymbolAddress returns for the current device, looked up once per device and kept.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.h"
#include "scratch.h"

namespace sip {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void *scratch_device_ptr(int device)
{
    if (device < 0 || device >= kMaxDevices)
        throw SipError(AsdSip::ACL_ERROR_INTERNAL_ERROR, "scratch: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return g_scratch_ptr[device];
    int prev = -1;
    SIP_HIP_CHECK(hipGetDevice(&prev));
    if (prev != device) SIP_HIP_CHECK(hipSetDevice(device));
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (prev != device) SIP_HIP_CHECK(hipSetDevice(prev));
    SIP_HIP_CHECK(e);
    if (ptr == nullptr) throw SipError(AsdSip::ACL_ERROR_INTERNAL_ERROR, "scratch: no device address");
    g_scratch_ptr[device] = ptr;
    return ptr;
}

}  // namespace sip
