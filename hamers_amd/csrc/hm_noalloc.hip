// hm_noalloc.hip -- the allocation-free per-patch RHS entry point
// (hm_compute_convective_flux_and_source_on_patch_noalloc): the pipeline's
// scratch lives in a file-scope device array instead of the context's arena,
// so a call makes no device (or host) allocation.
#include <atomic>
#include <mutex>

#include "hm_ctx.hpp"

namespace {

// One copy per device, held until the process ends.  Every call on a device
// shares it and leaves its contents behind; the pipeline writes every scratch
// value before it reads it (V, c over the 4-ghost box, theta and Omega over
// the 2-ghost box, every midpoint flux and u*), so nothing needs clearing.
static __device__ __align__(256) double scratch[HM_NOALLOC_SCRATCH_DOUBLES];

constexpr int kMaxDevices = 64;
// Device address of `scratch` on each device (the host-side name is only a
// stand-in for it), resolved once per device.
std::atomic<double*> g_scratch[kMaxDevices];
std::mutex g_setup;

}  // namespace

namespace hm {

int noalloc_setup(int device, std::string& err) {
    if (device < 0 || device >= kMaxDevices) {
        err = "device ordinal out of range for the allocation-free entry point";
        return HM_ERR_INVALID_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(g_setup);
    if (g_scratch[device].load()) return HM_OK;
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) {
        err = "hipGetDevice failed";
        return HM_ERR_DEVICE;
    }
    if (prev != device && hipSetDevice(device) != hipSuccess) {
        err = "hipSetDevice failed";
        return HM_ERR_DEVICE;
    }
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (prev >= 0 && prev != device) (void)hipSetDevice(prev);
    if (e != hipSuccess || !ptr) {
        err = std::string("hipGetSymbolAddress of the scratch array failed: ") + hipGetErrorString(e);
        return HM_ERR_DEVICE;
    }
    g_scratch[device].store(static_cast<double*>(ptr));
    return HM_OK;
}

}  // namespace hm

extern "C" int hm_compute_convective_flux_and_source_on_patch_noalloc(hm_ctx* ctx, const hm_patch_view* patch,
                                                                       double time, double dt, int rk_step,
                                                                       void* stream) {
    (void)time;
    (void)rk_step;
    if (!ctx || !ctx->rec) return HM_ERR_INVALID_ARGUMENT;
    if (!patch) return ctx->rec->fail(HM_ERR_INVALID_ARGUMENT, "NULL patch view");
    const int dev = ctx->rec->device();
    double* p = (dev >= 0 && dev < kMaxDevices) ? g_scratch[dev].load() : nullptr;
    if (!p) return ctx->rec->fail(HM_ERR_DEVICE, "the scratch array of this device was not set up");
    return ctx->rec->computeOnPatchInScratch(*patch, dt, static_cast<hipStream_t>(stream), p,
                                             size_t(HM_NOALLOC_SCRATCH_DOUBLES));
}
