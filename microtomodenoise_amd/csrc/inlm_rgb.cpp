// Colour path (Color/color_denoising.h:19-150): 2D RGB images as three
// channels, Anscombe variance stabilisation, a noise level per channel, the
// three-channel filter of iternlm_cpu.cpp:932-1075 on the GPU, and the
// reference's iteration schedule.
//
// The schedule differs from the volume path in two ways the reference has:
// the blended noise level replaces sigma0 (so each re-estimate blends with
// the previous blend), and the program writes every iteration's image after
// an inverse and a forward Anscombe transform, whose rounding then feeds the
// next iteration (the iteration hook below lets the file program do that).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "inlm_capi.h"
#include "inlm_internal.hpp"
#include "inlm_kernels.hpp"
#include "inlm_tables.hpp"

namespace {

struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

inlm_status rgb_fail(inlm_status s, const std::string &m)
{
    inlm::set_global_error(m);
    return s;
}

}  // namespace

namespace inlm {

void anscombe(float *v, size_t n)
{
    for (size_t i = 0; i < n; i++) v[i] = 2.f * sqrtf(v[i] + 3.f / 8.f);  // colorspace.h:99-101
}

void inverse_anscombe(float *v, size_t n)
{
    for (size_t i = 0; i < n; i++) v[i] = (v[i] / 2.f) * (v[i] / 2.f) - 3.f / 8.f;  // colorspace.h:118-120
}

inlm_status denoise_rgb(const inlm_params *pin, const inlm_rgb_options *c, const int32_t shape2[2], int32_t device,
                        const float *in, float *out, float *sigma_log,
                        const std::function<bool(int, float *)> &on_iteration)
{
    if (!pin || !c || !shape2 || !in || !out) return rgb_fail(INLM_BAD_ARG, "null argument");
    const int nx = shape2[0], ny = shape2[1];
    if (nx < 2 || ny < 2) return rgb_fail(INLM_BAD_ARG, "image smaller than 2x2");
    // color_denoising.h:25-29: 2D images, one slice, 2048 noise samples
    inlm_params p = *pin;
    p.radius_patchspace[2] = 0;
    p.radius_searchspace[2] = 0;
    p.nslices = 1;
    p.n_samples = 2048;
    if (p.maxiterations < 1) return rgb_fail(INLM_BAD_ARG, "maxiterations must be >= 1");
    if (p.patchsize < 1 || p.patchsize >= nx || p.patchsize >= ny)
        return rgb_fail(INLM_BAD_ARG, "noise patch does not fit the image");
    const size_t npix = (size_t)nx * ny;

    int sh[3] = {nx, ny, 1};
    const Geometry geo = make_geometry(sh, p.radius_searchspace, p.radius_patchspace, p.nslices);
    if (geo.px >= nx || geo.py >= ny) return rgb_fail(INLM_BAD_ARG, "search window wider than the image");
    const std::vector<Offset3> S = search_offsets(p.radius_searchspace, p.nslices);
    const std::vector<Offset3> PT = patch_offsets(p.radius_patchspace, p.nslices);
    if (S.empty()) return rgb_fail(INLM_BAD_ARG, "empty search space");
    if ((int)PT.size() > kMaxTaps) return rgb_fail(INLM_BAD_ARG, "patch has more than 128 taps");
    // the colour filter uses the computed weights (iternlm_cpu.cpp:985, setup_distweight)
    const std::vector<float> W = computed_weights(p.radius_patchspace, p.nslices, p.z_anisotropy, geo.nzp);
    FilterGeom g{geo.nx, geo.ny, geo.nz, geo.nxp, geo.nyp, geo.nzp, geo.px, geo.py, geo.pz};
    const long sp = (long)geo.nxp * geo.nyp;
    TapTable taps;
    std::memset(&taps, 0, sizeof(taps));
    taps.n = (int)PT.size();
    for (int i = 0; i < taps.n; i++) {
        taps.w[i] = W[i];
        taps.lin[i] = (int)(PT[i].dz * sp + (long)PT[i].dy * geo.nxp + PT[i].dx);
        taps.dxyz[3 * i] = PT[i].dx;
        taps.dxyz[3 * i + 1] = PT[i].dy;
        taps.dxyz[3 * i + 2] = PT[i].dz;
    }
    std::vector<SearchEntry> se(S.size());
    for (size_t i = 0; i < S.size(); i++)
        se[i] = SearchEntry{S[i].dx, S[i].dy, S[i].dz, (int)(S[i].dz * sp + (long)S[i].dy * geo.nxp + S[i].dx)};

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return rgb_fail(INLM_NO_DEVICE, "no GPU visible");
    if (device < 0 || device >= ndev) return rgb_fail(INLM_BAD_ARG, "device index out of range");
#define RGB_HIP(x)                                                                                     \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) return rgb_fail(e_ == hipErrorOutOfMemory ? INLM_OOM : INLM_HIP_ERROR, \
                                              std::string("HIP: ") + hipGetErrorString(e_));           \
    } while (0)
    RGB_HIP(hipSetDevice(device));
    hipStream_t st;
    RGB_HIP(hipStreamCreate(&st));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};

    const size_t padn = (size_t)geo.voxels_p();
    DevBuf bR, bP, bO, bU, bS, bSm, bSd;
    RGB_HIP(hipMalloc(&bR.p, 3 * padn * sizeof(float)));
    RGB_HIP(hipMalloc(&bP.p, 3 * padn * sizeof(float)));
    RGB_HIP(hipMalloc(&bO.p, 3 * npix * sizeof(float)));
    RGB_HIP(hipMalloc(&bU.p, 3 * npix * sizeof(float)));
    RGB_HIP(hipMalloc(&bS.p, se.size() * sizeof(SearchEntry)));
    float *Rd = static_cast<float *>(bR.p), *Pd = static_cast<float *>(bP.p), *Od = static_cast<float *>(bO.p);
    float *Ud = static_cast<float *>(bU.p);
    RGB_HIP(hipMemcpyAsync(bS.p, se.data(), se.size() * sizeof(SearchEntry), hipMemcpyHostToDevice, st));

    // the channels, Anscombe-transformed on the host as the reference does
    std::vector<float> host(in, in + 3 * npix);
    if (c->stabilize_variance) anscombe(host.data(), host.size());
    RGB_HIP(hipMemcpyAsync(Ud, host.data(), 3 * npix * sizeof(float), hipMemcpyHostToDevice, st));
    for (int ch = 0; ch < 3; ch++) RGB_HIP(launch_pad_reflect(Ud + ch * npix, Rd + ch * padn, g, st));

    // noise level per channel (noiselevel.cpp:96-184): one sample list for
    // every estimate, two clusters over the patch deviations of each channel
    // (the z-adaptive mode is global for a 2D image), or the semimanual mode
    // with a reference sample per channel
    std::vector<int64_t> samples;
    const int64_t ns = p.noise_mode == INLM_NOISE_MANUAL ? 0 : p.n_samples;
    if (ns > 0) {
        samples.resize((size_t)ns);
        if (!draw_samples(sh, ns, p.patchsize, p.sample_seed, samples.data(), p.circular_mask_diameter))
            return rgb_fail(INLM_BAD_ARG, "no noise sample fits the image (or the circular mask)");
        RGB_HIP(hipMalloc(&bSm.p, (size_t)ns * sizeof(int64_t)));
        RGB_HIP(hipMalloc(&bSd.p, (size_t)ns * sizeof(double)));
        RGB_HIP(hipMemcpyAsync(bSm.p, samples.data(), (size_t)ns * sizeof(int64_t), hipMemcpyHostToDevice, st));
    }
    std::vector<double> sd((size_t)std::max<int64_t>(ns, 1));
    inlm_params pc = p;
    if (pc.noise_mode == INLM_NOISE_ZADAPTIVE) pc.noise_mode = INLM_NOISE_GLOBAL;
    int64_t semi_ref[3] = {-1, -1, -1};
    auto estimate = [&](const float *chans_dev, int which, float sig[3]) -> inlm_status {
        const int one_slice[3] = {nx, ny, 1};
        for (int ch = 0; ch < 3; ch++) {
            if (ns > 0) {
                RGB_HIP(launch_patchstd(chans_dev + ch * npix, nx, (int64_t)npix, static_cast<int64_t *>(bSm.p), ns,
                                        nx, (int64_t)npix, p.patchsize, static_cast<double *>(bSd.p), st));
                RGB_HIP(hipMemcpyAsync(sd.data(), bSd.p, (size_t)ns * sizeof(double), hipMemcpyDeviceToHost, st));
                RGB_HIP(hipStreamSynchronize(st));
            }
            std::string err;
            const inlm_status r = sigma_from_std(pc, one_slice, samples.data(), sd.data(), ns, which, semi_ref[ch],
                                                 &sig[ch], err);
            if (r != INLM_OK) return rgb_fail(r, err);
        }
        if (c->average_channelsigma) sig[0] = sig[1] = sig[2] = (sig[0] + sig[1] + sig[2]) / 3.f;
        return INLM_OK;
    };

    float s0[3], s1[3];
    inlm_status s = estimate(Ud, 0, s0);
    if (s != INLM_OK) return s;
    std::vector<float> res(3 * npix);
    for (int iter = 1; iter <= p.maxiterations; iter++) {
        const float *R3[3] = {Rd, Rd + padn, Rd + 2 * padn};
        const float *P3[3];
        float *O3[3] = {Od, Od + npix, Od + 2 * npix};
        if (iter == 1) {
            for (int ch = 0; ch < 3; ch++) P3[ch] = R3[ch];
        } else {
            for (int ch = 0; ch < 3; ch++) {
                RGB_HIP(launch_pad_reflect(Od + ch * npix, Pd + ch * padn, g, st));
                P3[ch] = Pd + ch * padn;
            }
        }
        float mult[3];
        for (int ch = 0; ch < 3; ch++) mult[ch] = -1.f / (s0[ch] * s0[ch] * p.beta);  // iternlm_cpu.cpp:384-386
        if (sigma_log) std::memcpy(sigma_log + 3 * (iter - 1), s0, sizeof(s0));
        RGB_HIP(launch_filter_rgb(p.weight_fn, R3, P3, O3, mult, c->independent_channels, static_cast<SearchEntry *>(bS.p),
                                  (int)se.size(), taps, g, st));
        if (on_iteration) {
            RGB_HIP(hipMemcpyAsync(res.data(), Od, 3 * npix * sizeof(float), hipMemcpyDeviceToHost, st));
            RGB_HIP(hipStreamSynchronize(st));
            if (!on_iteration(iter, res.data())) return rgb_fail(INLM_IO_ERROR, "writing an iteration image failed");
            RGB_HIP(hipMemcpyAsync(Od, res.data(), 3 * npix * sizeof(float), hipMemcpyHostToDevice, st));
        }
        // color_denoising.h:85-96 and :116-126: re-estimate on the result and
        // blend into sigma0
        const bool again = iter == 1 ? p.maxiterations > 1 : p.continuous_estimate != 0;
        if (again) {
            if ((s = estimate(Od, 1, s1)) != INLM_OK) return s;
            for (int ch = 0; ch < 3; ch++) s0[ch] = p.alpha * s0[ch] + (1.f - p.alpha) * s1[ch];
        }
    }
    RGB_HIP(hipMemcpyAsync(res.data(), Od, 3 * npix * sizeof(float), hipMemcpyDeviceToHost, st));
    RGB_HIP(hipStreamSynchronize(st));
    if (c->stabilize_variance) inverse_anscombe(res.data(), res.size());
    std::memcpy(out, res.data(), 3 * npix * sizeof(float));
    return INLM_OK;
#undef RGB_HIP
}

}  // namespace inlm

extern "C" {

void inlm_default_rgb_options(inlm_rgb_options *c)
{
    std::memset(c, 0, sizeof(*c));
    c->independent_channels = 0;
    c->stabilize_variance = 1;
    c->average_channelsigma = 1;
}

inlm_status inlm_denoise_rgb(const inlm_params *p, const inlm_rgb_options *c, const int32_t shape[2], int32_t device,
                             const float *in, float *out, float *sigma_log)
{
    inlm_rgb_options d;
    if (!c) {
        inlm_default_rgb_options(&d);
        c = &d;
    }
    return inlm::denoise_rgb(p, c, shape, device, in, out, sigma_log, nullptr);
}

}  // extern "C"
