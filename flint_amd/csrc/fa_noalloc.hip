// fa_noalloc.hip -- an integer product entry point that allocates nothing.
//
// For callers that cannot call hipMalloc (code under graph capture, tight
// serving loops): fa_mpn_mul_device_noalloc runs the same kernels as
// fa_mpn_mul_device, with their scratch (residue rows and carry flags) in
// the file-scope device array below instead of the thread's grow-only
// buffer.  Every device that runs the library holds its own copy of the
// array until the process ends.
#include <mutex>

#include "fa_internal.h"

#define FA_NOALLOC_SCRATCH_BYTES (64u << 20)

static __device__ __align__(256) unsigned char scratch[FA_NOALLOC_SCRATCH_BYTES];

// the device address of `scratch` on each device (the host-side symbol is a
// stand-in, not a device address)
static void* scratch_ptr[16];
static std::mutex scratch_mutex;

// once per device, with that device current (called by fa_init)
int fa_noalloc_setup()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16)
        return fa_set_error(FA_E_NODEV, "no HIP device");
    std::lock_guard<std::mutex> g(scratch_mutex);
    if (scratch_ptr[dev]) return FA_OK;
    void* p = NULL;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fa_set_error(FA_E_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                            hipGetErrorString(e));
    scratch_ptr[dev] = p;
    return FA_OK;
}

extern "C" uint64_t fa_noalloc_scratch_bytes(void) { return FA_NOALLOC_SCRATCH_BYTES; }

extern "C" int fa_mpn_mul_device_noalloc(uint64_t* z, const uint64_t* a, int64_t an,
                                         const uint64_t* b, int64_t bn, void* stream)
{
    if (an <= 0 || bn <= 0 || !z || !a || !b)
        return fa_set_error(FA_E_ARG, "fa_mpn_mul_device_noalloc: bad arguments");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16)
        return fa_set_error(FA_E_NODEV, "no HIP device");
    void* p;
    {
        std::lock_guard<std::mutex> g(scratch_mutex);
        p = scratch_ptr[dev];
    }
    if (!p)
        return fa_set_error(FA_E_ARG,
                            "fa_mpn_mul_device_noalloc: call fa_init() on device %d first", dev);
    return fa_mpn_mul_prealloc(z, a, (uint64_t)an, b, (uint64_t)bn, (hipStream_t)stream, p,
                               FA_NOALLOC_SCRATCH_BYTES);
}
