/*
 * iternlm_gpu.h -- header-only C++ drop-in for denoise::IterativeNLM_GPU.
 *
 * Keeps the class name, public members and method signatures of the
 * reference (iterNLM_v0.3/Denoiser/iternlm_gpu.h:20-56) and maps each method
 * onto the C ABI of include/inlm_capi.h (link with -linlm).  Differences a
 * caller can observe:
 *   - one GPU: configure_device holds the whole volume on deviceID and
 *     returns shape[2] (slices it can process at once), -1 without a GPU (the
 *     reference's signal to take the CPU path, main.cpp:317) and -2 on a HIP
 *     error (iternlm_gpu.cu:1608-1618);
 *   - failures of the other methods are printed like the reference's, and
 *     also kept: last_status() returns the inlm_status of the last call;
 *   - Run_GaussianNoise_GPUBlocks (the multi-GPU / host-blocked drivers) is
 *     not provided: the volume stays resident in HBM.
 */
#ifndef INLM_ITERNLM_GPU_H
#define INLM_ITERNLM_GPU_H

#include <cstdio>
#include <string>

#include "denoise_parameters.h"
#include "inlm_capi.h"

namespace denoise
{
class IterativeNLM_GPU
{
public:
    int n_parallelslices = 0;
    bool resumed = false;

    int configure_device(int shape[3], protocol::DenoiseParameters *params)
    {
        free_device();
        if (inlm_device_count() < 1) {
            params->gpu.n_gpus = 0;
            return -1;
        }
        inlm_params p = to_c(*params);
        const int32_t sh[3] = {shape[0], shape[1], shape[2]};
        status_ = inlm_create(&p, sh, params->gpu.deviceID, &h_);
        if (status_ == INLM_NO_DEVICE) {
            params->gpu.n_gpus = 0;
            return -1;
        }
        if (status_ != INLM_OK) {
            std::fprintf(stderr, "Device Configuration: %s\n", inlm_last_error(nullptr));
            return -2;
        }
        params->gpu.n_gpus = 1;
        n_parallelslices = shape[2];
        return n_parallelslices;
    }

    void free_device()
    {
        if (h_) inlm_destroy(h_);
        h_ = nullptr;
    }

    void Run_GaussianNoise(int iter, float *instack, int shape[3], protocol::DenoiseParameters *params)
    {
        (void)shape;
        (void)params;
        report(inlm_run_iteration(h_, iter, iter == 1 ? instack : nullptr), "Run_GaussianNoise");
    }

    void set_sigma(float *sigmalist, int shape[3]) { report(inlm_set_sigma(h_, sigmalist, shape[2]), "set_sigma"); }

    void get_result(float *output, int shape[3])
    {
        (void)shape;
        report(inlm_get_result(h_, output, 0), "get_result");
    }

    // build-specific: status of the last call and the handle for the C ABI
    inlm_status last_status() const { return status_; }
    inlm_handle *handle() const { return h_; }

    ~IterativeNLM_GPU() { free_device(); }

    static inlm_params to_c(const protocol::DenoiseParameters &d)
    {
        inlm_params p;
        inlm_default_params(&p);
        p.alpha = d.alpha;
        p.maxiterations = d.maxiterations;
        p.nslices = d.nslices;
        for (int i = 0; i < 3; i++) {
            p.radius_searchspace[i] = d.radius_searchspace[i];
            p.radius_patchspace[i] = d.radius_patchspace[i];
        }
        p.beta = d.beta;
        p.z_anisotropy = d.z_anisotropy;
        const std::string &m = d.noiselevel.mode;
        p.noise_mode = (m == "global" || m == "auto") ? INLM_NOISE_GLOBAL
                       : (m == "manual")              ? INLM_NOISE_MANUAL
                                                      : INLM_NOISE_ZADAPTIVE;
        p.zwindow = d.noiselevel.zwindow;
        p.sigma[0] = d.noiselevel.sigma[0];
        p.sigma[1] = d.noiselevel.sigma[1];
        p.n_samples = (int64_t)d.noiselevel.n_samples;
        p.patchsize = d.noiselevel.patchsize;
        p.stds_from_mean = d.noiselevel.stds_from_mean;
        p.continuous_estimate = d.noiselevel.continuous_estimate ? 1 : 0;
        p.sample_seed = d.gpu.sample_seed;
        p.cpu_blockwise = d.cpu.blockwise ? 1 : 0;
        const std::string &pr = d.gpu.precision;
        p.precision = pr == "fp32_fast" ? INLM_PREC_FP32_FAST
                      : pr == "bf16"    ? INLM_PREC_BF16
                      : pr == "fp16"    ? INLM_PREC_FP16
                      : pr == "fp16_fp8" ? INLM_PREC_FP16_FP8
                                        : INLM_PREC_FP32;
        p.weight_fn = d.gpu.weight_fn == "exp" ? INLM_WEIGHT_EXP : INLM_WEIGHT_PADE;
        return p;
    }

private:
    inlm_handle *h_ = nullptr;
    inlm_status status_ = INLM_OK;

    void report(inlm_status st, const char *where)
    {
        status_ = st;
        if (st != INLM_OK) std::fprintf(stderr, "%s: %s\n", where, inlm_last_error(h_));
    }
};
}  // namespace denoise

#endif  // INLM_ITERNLM_GPU_H
