// axspmv_dev_build_axt_noalloc: the device AXT builder for callers that must
// not allocate device memory.
//
// axspmv_dev_build_axt allocates its scratch (the CSR check flag, the per-row
// counts, their 64-bit exclusive scan and rocPRIM's temporary storage) and
// its output arrays with hipMalloc on every call.  This variant takes the
// output arrays from the caller and keeps the scratch in `scratch`, a
// file-scope device array of AXSPMV_DEV_BUILD_NOALLOC_SCRATCH bytes that each
// GPU running the library holds until the process ends.  The host reaches
// it through the device address hipGetSymbolAddress returns, looked up once
// per device (with that device current) and kept per device: the symbol's
// name in host code is not a device address.
//
// Every call on a GPU shares that GPU's array, and the array keeps the last
// call's contents; each call clears what must start at zero (the check
// flag, the scan's last input) with fill kernels.  Calls that overlap on one
// GPU (several host threads, or several streams) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../host/common.hpp"
#include "launch.hpp"

namespace axspmv {
// devbuild.cpp: the build with the scratch at `scratch` (cap bytes) and the
// caller's output arrays in f
int dev_build_axt_into(const axspmv_dcsr *a, const double *d_x, uint32_t bs, uint32_t thw, uint32_t th, uint32_t mode,
                       hipStream_t s, char *scratch, size_t cap, axspmv_dfmt *f);
}  // namespace axspmv

static __device__ __align__(256) unsigned char scratch[AXSPMV_DEV_BUILD_NOALLOC_SCRATCH];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
char *g_ptr[MAX_DEVICES] = {};  // device address of `scratch` on each device, once looked up

// *out = the current device's copy of `scratch`.
int scratch_of_current_device(char **out) {
    *out = nullptr;
    int dev = 0;
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return axspmv::fail(AXSPMV_E_HIP, "dev_build_axt_noalloc: no HIP device (%s)", hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES)
        return axspmv::fail(AXSPMV_E_UNSUPPORTED, "dev_build_axt_noalloc: device %d beyond the %d this build tracks",
                            dev, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (r != hipSuccess || !p)
            return axspmv::fail(AXSPMV_E_HIP,
                                "dev_build_axt_noalloc: hipGetSymbolAddress of the scratch array failed (%s)",
                                hipGetErrorString(r));
        g_ptr[dev] = static_cast<char *>(p);
    }
    *out = g_ptr[dev];
    return AXSPMV_OK;
}

}  // namespace

extern "C" int axspmv_dev_build_axt_noalloc(const axspmv_dcsr *a, const double *d_x, uint32_t bs, uint32_t thw,
                                            uint32_t th, uint32_t mode, void *stream, axspmv_dfmt *f) {
    char *w = nullptr;
    AX_TRY(scratch_of_current_device(&w));
    return axspmv::dev_build_axt_into(a, d_x, bs, thw, th, mode, static_cast<hipStream_t>(stream), w,
                                      AXSPMV_DEV_BUILD_NOALLOC_SCRATCH, f);
}
