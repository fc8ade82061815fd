// Patch standard deviations without any device allocation: the sample list
// and the per-sample results are staged through a fixed device array of this
// file, so the call makes no hipMalloc / hipFree and needs no handle.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "inlm_capi.h"
#include "inlm_internal.hpp"
#include "inlm_kernels.hpp"

namespace {

constexpr int64_t kMaxSamples = INLM_NOALLOC_MAX_SAMPLES;
constexpr size_t kScratchBytes = (size_t)kMaxSamples * (sizeof(int64_t) + sizeof(double));  // 4 MiB
constexpr int kMaxDevices = 64;

// Each GPU holds its own copy until the process ends.  Host code never uses
// this name as an address: the per-device address comes from
// hipGetSymbolAddress below.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mutex;
unsigned char *g_scratch_dev[kMaxDevices] = {};

// The address of this device's copy of `scratch`, looked up once per device
// with that device current.
hipError_t scratch_of(int device, unsigned char **out)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch_dev[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_dev[device] = static_cast<unsigned char *>(ptr);
    }
    *out = g_scratch_dev[device];
    return hipSuccess;
}

inlm_status bad(const std::string &msg)
{
    inlm::set_global_error(msg);
    return INLM_BAD_ARG;
}

inlm_status hip_bad(hipError_t e, const char *what)
{
    inlm::set_global_error(std::string(what) + ": " + hipGetErrorString(e));
    return INLM_HIP_ERROR;
}

}  // namespace

inlm_status inlm_patchstd_noalloc(const float *vol, const int32_t shape[3], int32_t patchsize, const int64_t *samples,
                                  int64_t n, double *out, int32_t device)
{
    if (!vol || !shape || (n > 0 && (!samples || !out))) return bad("null argument");
    const int nx = shape[0], ny = shape[1], nz = shape[2];
    if (nx < 1 || ny < 1 || nz < 1) return bad("shape must be positive");
    if (n < 0) return bad("negative sample count");
    if (n > kMaxSamples)
        return bad("inlm_patchstd_noalloc takes at most " + std::to_string(kMaxSamples) + " samples per call (got " +
                   std::to_string(n) + ")");
    if (patchsize < 1 || patchsize > nx || patchsize > ny) return bad("patch size does not fit the slice");
    if (device < 0 || device >= kMaxDevices) return bad("device index out of range");
    const int64_t nslice = (int64_t)nx * ny, nvox = nslice * nz;
    for (int64_t i = 0; i < n; i++) {  // every patch must lie inside its slice
        const int64_t idx = samples[i];
        if (idx < 0 || idx >= nvox) return bad("sample index out of range");
        const int64_t r = idx % nslice;
        if (r % nx + patchsize > nx || r / nx + patchsize > ny) return bad("sample patch leaves the slice");
    }
    if (n == 0) return INLM_OK;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_bad(e, "hipSetDevice");
    unsigned char *base = nullptr;
    if ((e = scratch_of(device, &base)) != hipSuccess) return hip_bad(e, "hipGetSymbolAddress(scratch)");
    int64_t *samples_dev = reinterpret_cast<int64_t *>(base);
    double *std_dev = reinterpret_cast<double *>(base + (size_t)kMaxSamples * sizeof(int64_t));
    // the null stream: the copies and the kernel run in order, and the last
    // copy returns when the results are on the host
    if ((e = hipMemcpy(samples_dev, samples, sizeof(int64_t) * n, hipMemcpyHostToDevice)) != hipSuccess)
        return hip_bad(e, "hipMemcpy(samples)");
    if ((e = inlm::launch_patchstd(vol, nx, nslice, samples_dev, n, nx, nslice, patchsize, std_dev, nullptr)) !=
        hipSuccess)
        return hip_bad(e, "patchstd kernel");
    if ((e = hipMemcpy(out, std_dev, sizeof(double) * n, hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_bad(e, "hipMemcpy(stds)");
    return INLM_OK;
}
