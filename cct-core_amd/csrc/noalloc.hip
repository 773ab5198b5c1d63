// cogconv_launch_noalloc: cogconv_launch for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  The plan's workspace (split-reduction partials,
// packed low-precision filter banks) comes from `scratch`, a file-scope device
// array of NOALLOC_SCRATCH_BYTES, instead of a caller buffer.
//
// Every device holds its own copy of the array from the first use of this
// library's code on it until the process ends.  On the host, `scratch` names a
// host-side stand-in, not device memory: its device address comes from
// hipGetSymbolAddress, taken once per device (with that device current) by
// cogconv_init, or by the entry point's first call on a device that was not
// initialised, and kept in scratch_addr[].  Only that address is ever handed to
// a kernel.
//
// All calls on one device share its array, and the array keeps whatever the
// last call left in it: every kernel that reads a workspace writes all of it
// first in the same launch sequence (partials, packed banks), so nothing has to
// start at zero.  Calls that overlap on one device (several host threads, or
// several streams with work in flight) race on the array: the caller
// serialises them.
#include <mutex>

#include "internal.h"

namespace cogconv {
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
}

static __device__ __align__(256) unsigned char scratch[cogconv::NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex addr_mu;
void* scratch_addr[MAX_DEVICES];  // device address of `scratch` on each device (null until set up)

}  // namespace

// Takes the device address of `scratch` on `device`, which must be current.
int cogconv::noalloc_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES)
        return fail(COGCONV_ERR_INVALID_ARG, "device %d beyond the %d the library tracks", device, MAX_DEVICES);
    std::lock_guard<std::mutex> lk(addr_mu);
    if (scratch_addr[device]) return COGCONV_OK;
    void* p = nullptr;
    const int e = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch symbol address");
    if (e) return e;
    if (!p) return fail(COGCONV_ERR_RUNTIME, "scratch symbol address is null");
    scratch_addr[device] = p;
    return COGCONV_OK;
}

extern "C" int cogconv_noalloc_scratch_bytes(size_t* bytes) {
    cogconv::clear_error();
    if (!bytes) return cogconv::fail(COGCONV_ERR_INVALID_ARG, "null argument");
    *bytes = cogconv::NOALLOC_SCRATCH_BYTES;
    return COGCONV_OK;
}

extern "C" int cogconv_launch_noalloc(cogconv_plan plan, const float* in0, const float* in1, float* out,
                                      cogconv_stream stream) {
    cogconv::clear_error();
    if (!plan || !in0 || !in1 || !out) return cogconv::fail(COGCONV_ERR_INVALID_ARG, "null plan or buffer");
    if (plan->kind == cogconv::KK_FFT)
        return cogconv::fail(COGCONV_ERR_UNSUPPORTED,
                             "the FFT route's transform library allocates its work area: not offered without allocation");
    if (plan->workspace > cogconv::NOALLOC_SCRATCH_BYTES)
        return cogconv::fail(COGCONV_ERR_UNSUPPORTED,
                             "plan needs %zu bytes of workspace; the no-allocation entry point holds at most %zu",
                             plan->workspace, cogconv::NOALLOC_SCRATCH_BYTES);
    void* ws = nullptr;
    if (plan->workspace) {
        int dev = 0;
        int e = cogconv::check_hip(hipGetDevice(&dev), "hipGetDevice");
        if (e) return e;
        if (dev < 0 || dev >= MAX_DEVICES)
            return cogconv::fail(COGCONV_ERR_INVALID_ARG, "device %d beyond the %d the library tracks", dev,
                                 MAX_DEVICES);
        {
            std::lock_guard<std::mutex> lk(addr_mu);
            ws = scratch_addr[dev];
        }
        if (!ws) {
            e = cogconv::noalloc_setup(dev);
            if (e) return e;
            std::lock_guard<std::mutex> lk(addr_mu);
            ws = scratch_addr[dev];
        }
    }
    return cogconv_launch(plan, in0, in1, out, ws, stream);
}
