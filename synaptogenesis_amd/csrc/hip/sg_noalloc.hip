// sg_hip_activate_base_noalloc: the base activator (sg_hip_activate_base)
// for callers that cannot allocate device memory.
//
// The plane path of sg_hip_activate_base packs the source's spike bits into
// a scratch plane (sg_conn.plane, one bit per source) before summing the
// weights of the spiking visits.  Here that plane lives in a file-scope
// device array of fixed size instead: every GPU the library runs on holds
// one copy of it from the library's load until the process ends, and the
// entry point calls no hipMalloc, hipFree or allocating library routine.
//
// On the host the array's name is not a device address.  Its address on a
// device comes from hipGetSymbolAddress, called once per device (with that
// device current) on the entry point's first call there, and kept per
// device; kernels only ever get that address.
//
// Calls on one GPU share its array: calls that overlap on one GPU (on
// different streams, or from different threads) race on it.  Each call
// packs every plane word it reads before reading it, so nothing needs
// clearing between calls.

#include <hip/hip_runtime.h>

#include <mutex>

#include "sg_hip.h"

constexpr long kScratchWords = SG_NOALLOC_PLANE_WORDS;

static __device__ __align__(256) uint64_t scratch[kScratchWords];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_lock;
uint64_t *g_scratch[kMaxDevices] = {nullptr};

// The array's address on the current device (looked up on the first call
// there).
int device_scratch(uint64_t **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_lock);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch[dev] = static_cast<uint64_t *>(p);
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

}  // namespace

extern "C" int sg_hip_activate_base_noalloc(const sg_conn *c, void *stream) {
    if (!c) return hipErrorInvalidValue;
    if (c->num_dest <= 0) return 0;
    sg_conn d = *c;
    d.plane_ready = 0;
    // Without the plane path (a FLOAT source, weights that may be
    // infinite, one-to-one) the kernels need no scratch at all.
    if (!(d.skip_silent && d.out_type == SG_OUT_BIT &&
          d.type != SG_ONE_TO_ONE)) {
        d.plane = nullptr;
        return sg_hip_activate_base(&d, stream);
    }
    if (d.from_size < 0 ||
        ((long)d.from_size + 63) / 64 > kScratchWords)
        return SG_HIP_ERROR_NOALLOC_LIMIT;
    uint64_t *plane = nullptr;
    const int e = device_scratch(&plane);
    if (e != hipSuccess) return e;
    d.plane = plane;
    return sg_hip_activate_base(&d, stream);
}
