// mpsx_two_site_static: the batched two-site update on a fixed, library-owned
// device workspace, for callers that cannot allocate device memory.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "driver.h"

// The workspace: one copy per device, held until the process ends. The
// kernels initialise every part of it they read before reading it
// (k_init_rows resets the header counters, the per-problem control blocks
// and the identity rows; the theta kernel writes the working rows), so what
// an earlier call left here is never used, and nothing needs clearing
// between calls. In host code `scratch` names a host-side stand-in, not the
// device array: only the address hipGetSymbolAddress returns for the current
// device is passed to the kernels.
static __device__ __align__(256) char scratch[MPSX_STATIC_WS_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {nullptr};

// Device address of `scratch` on the current device, looked up on the first
// call on that device (with it current) and kept.
int scratch_on_current_device(hipStream_t stream, void **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return mpsx_internal_fail(MPSX_EHIP, "hipGetDevice failed or device index out of range");
    if (stream) {
        int sdev = -1;
        if (hipStreamGetDevice(stream, &sdev) != hipSuccess)
            return mpsx_internal_fail(MPSX_EHIP, "hipStreamGetDevice failed");
        if (sdev != dev)
            return mpsx_internal_fail(MPSX_EINVAL,
                                      "the stream belongs to another device than the current one");
    }
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            char msg[256];
            std::snprintf(msg, sizeof msg, "hipGetSymbolAddress(scratch): %s",
                          hipGetErrorString(e));
            return mpsx_internal_fail(MPSX_EHIP, msg);
        }
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return MPSX_OK;
}

}  // namespace

extern "C" int mpsx_two_site_static(mpsx_stream stream_, const mpsx_two_site_op *h,
                                    const mpsx_two_site_op *d, int count,
                                    const mpsx_svd_opts *opts, int *sweeps_out) {
    hipStream_t stream = (hipStream_t)stream_;
    if (sweeps_out) *sweeps_out = 0;
    if (count == 0) return MPSX_OK;
    void *ws = nullptr;
    if (int e = scratch_on_current_device(stream, &ws)) return e;
    size_t need = 0;
    const int rc = mpsx_internal_two_site(stream, h, d, count, opts, static_cast<char *>(ws),
                                          MPSX_STATIC_WS_BYTES, sweeps_out, false, &need);
    if (rc == MPSX_EWORKSPACE) {
        char msg[256];
        std::snprintf(msg, sizeof msg,
                      "batch needs %zu bytes of workspace, more than the %zu of "
                      "mpsx_two_site_static's fixed workspace; split the batch or use "
                      "mpsx_two_site",
                      need, (size_t)MPSX_STATIC_WS_BYTES);
        return mpsx_internal_fail(MPSX_EWORKSPACE, msg);
    }
    return rc;
}
