// Per-device scratch memory of the no-allocation exec entry point (sipFftExecC2CNoAlloc).
#pragma once
#include <cstddef>

namespace sip {

// Size of the scratch array each GPU holds from the first no-allocation call until the process ends.
constexpr size_t kScratchBytes = size_t(64) << 20;

// Device address of `device`'s copy of the scratch array, looked up with hipGetSymbolAddress on the
// first call for that device (with that device current) and kept; later calls return the kept
// address. Throws SipError when the lookup fails.
void *scratch_device_ptr(int device);

}  // namespace sip
