// Fixed device workspace of ckh_key_switch_batch_static: the batched key switch for callers that
// cannot allocate device memory (code under stream capture, serving loops that forbid hipMalloc).
//
// The array below is a file-scope device variable: every GPU the library is loaded on holds its
// own copy (CKH_STATIC_SCRATCH_BYTES) until the process ends. In host code its name is only a
// stand-in; the device address comes from hipGetSymbolAddress, looked up once per device with that
// device current (at context creation) and used wherever host code hands the memory to a kernel
// or a copy.
#include <mutex>
#include <vector>

#include "runtime.h"

namespace ckh {

static __device__ __align__(256) unsigned char scratch[CKH_STATIC_SCRATCH_BYTES];

uint8_t* static_scratch_address(int device) {
    static std::mutex mu;
    static std::vector<uint8_t*> addr;  // by device ordinal
    std::lock_guard<std::mutex> lk(mu);
    if ((int)addr.size() <= device) addr.resize(device + 1, nullptr);
    if (!addr[device]) {
        void* p = nullptr;
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(static scratch)");
        if (!p) fail(CKH_ERR_DEVICE, "static scratch has no device address");
        addr[device] = (uint8_t*)p;
    }
    return addr[device];
}

}  // namespace ckh
