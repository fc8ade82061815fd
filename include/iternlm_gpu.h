/*
 * iternlm_gpu.h -- header-only C++ drop-in for denoise::IterativeNLM_GPU.
 *
 * Keeps the class name, public members and method signatures of the
 * reference (iterNLM_v0.3/Denoiser/iternlm_gpu.h:20-56) and maps each method
 * onto the C ABI of include/inlm_capi.h (link with -linlm), so a caller
 * written against the reference, such as its main.cpp:311-563, compiles
 * against this header unchanged.  Differences a caller can observe:
 *   - one GPU: configure_device returns the number of slices one slab may
 *     hold (shape[2] when the whole volume fits, as the reference sizes it
 *     against the free device memory, iternlm_gpu.cu:1528-1566), -1 without
 *     a GPU (the reference's signal to take the CPU path, main.cpp:317) and
 *     -2 on a HIP error or when not even one slice fits
 *     (iternlm_gpu.cu:1608-1618).  gpu.memory_limit, a field of this
 *     build, caps the device bytes it may plan for;
 *   - both Run_GaussianNoise_GPUBlocks overloads run z-slabs whose halo comes
 *     from the previous iterate and whose filter strength is indexed by
 *     global z (the reference reads a block's lower halo from the buffer it
 *     is overwriting and indexes sigma by block-local z,
 *     iternlm_gpu.cu:121 and :1792-1819), so their FP32 result equals the
 *     whole-volume iteration;
 *   - resuming (io.resume) an iteration found complete on disk sets
 *     `resumed` and leaves that iteration as the result (the reference
 *     returns before setting it and, at iteration 1, then overwrites the
 *     read-in result with the input, iternlm_gpu.cu:1683-1697, :2117-2131);
 *   - failures of the other methods are printed like the reference's, and
 *     also kept: last_status() returns the inlm_status of the last call.
 */
#ifndef INLM_ITERNLM_GPU_H
#define INLM_ITERNLM_GPU_H

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

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
        deviceID_ = params->gpu.deviceID;
        int32_t n = 0;
        status_ = inlm_capacity_slices(&p, sh, deviceID_, params->gpu.memory_limit, &n);
        if (status_ == INLM_NO_DEVICE) {
            params->gpu.n_gpus = 0;
            return -1;
        }
        if (status_ != INLM_OK) {
            std::fprintf(stderr, "Device Configuration: %s\n", inlm_last_error(nullptr));
            return -2;
        }
        if (n >= shape[2]) {
            status_ = inlm_create(&p, sh, deviceID_, &h_);
            if (status_ == INLM_NO_DEVICE) {
                params->gpu.n_gpus = 0;
                return -1;
            }
            if (status_ != INLM_OK) {
                std::fprintf(stderr, "Device Configuration: %s\n", inlm_last_error(nullptr));
                return -2;
            }
        }
        params->gpu.n_gpus = 1;
        n_parallelslices = n;
        return n_parallelslices;
    }

    void free_device()
    {
        if (h_) inlm_destroy(h_);
        h_ = nullptr;
    }

    void Run_GaussianNoise(int iter, float *instack, int shape[3], protocol::DenoiseParameters *params)
    {
        if (!h_) {
            report(INLM_BAD_ARG, "Run_GaussianNoise", "the volume exceeds the device; use Run_GaussianNoise_GPUBlocks");
            return;
        }
        // a finished iteration on disk becomes the result (iternlm_gpu.cu:1665-1697)
        if (params->io.resume) {
            std::vector<float> vol;
            if (iteration_on_disk(*params, shape, vol)) {
                if (iter == 1) report(inlm_load_raw(h_, instack, 0), "Run_GaussianNoise");
                report(inlm_set_result(h_, iter, vol.data(), 0), "Run_GaussianNoise");
                resumed = true;
                std::printf("iteration %d read in from disk\n", iter);
                return;
            }
            resumed = false;
        }
        report(inlm_run_iteration(h_, iter, iter == 1 ? instack : nullptr), "Run_GaussianNoise");
    }

    // Host-RAM block driver (iternlm_gpu.cu:1729-1882): one iteration over
    // slabs of n_parallelslices slices; `previous` holds the previous iterate
    // (unused at iteration 1) and receives the result.
    void Run_GaussianNoise_GPUBlocks(int iter, float *instack, float *&previous, int shape[3],
                                     protocol::DenoiseParameters *params)
    {
        if (params->io.resume) {
            std::vector<float> vol;
            if (iteration_on_disk(*params, shape, vol)) {
                float *fresh = static_cast<float *>(std::malloc(sizeof(float) * vol.size()));
                if (!fresh) {
                    report(INLM_OOM, "Run_GaussianNoise_GPUBlocks", "out of host memory");
                    return;
                }
                std::memcpy(fresh, vol.data(), sizeof(float) * vol.size());
                std::free(previous);
                previous = fresh;
                resumed = true;
                std::printf("iteration %d read in from disk\n", iter);
                return;
            }
            resumed = false;
        }
        if (!previous || sigma_.size() != (size_t)shape[2]) {
            report(INLM_BAD_ARG, "Run_GaussianNoise_GPUBlocks", "no output buffer or sigma not set for this shape");
            return;
        }
        inlm_params p = to_c(*params);
        const int32_t sh[3] = {shape[0], shape[1], shape[2]};
        report_g(inlm_iterate_slabs(&p, sh, deviceID_, iter, instack, previous, sigma_.data(), previous,
                                    std::max(1, n_parallelslices)),
                 "Run_GaussianNoise_GPUBlocks");
    }

    // File-list block driver (iternlm_gpu.cu:1883-2115): one iteration over
    // slabs of n_parallelslices slices read from the raw and previous
    // iterations' TIFF files (each slab with the filter's z reach on both
    // sides), written as denoised<slice>.tif into io.active_outpath (16-bit
    // with io.save_type "16bit"); with io.resume a slab whose files exist is
    // skipped.
    void Run_GaussianNoise_GPUBlocks(int iter, std::vector<std::string> &filelist_raw,
                                     std::vector<std::string> &filelist_prev, int shape[3],
                                     protocol::DenoiseParameters *params)
    {
        const char *who = "Run_GaussianNoise_GPUBlocks";
        const int nx = shape[0], ny = shape[1], nz = shape[2];
        const int fs = params->io.firstslice > 0 ? params->io.firstslice : 0;
        // sub-stacks: the raw list starts at the first slice, the previous
        // iteration's list at its file of that slice (iternlm_gpu.cu:1897-1908)
        const int off_raw = fs;
        int off_prev = 0;
        if (fs != 0) {
            const std::string sub = zfill4(fs) + ".tif";
            for (size_t i = 0; i < filelist_prev.size(); i++)
                if (ends_with(filelist_prev[i], sub)) {
                    off_prev = (int)i;
                    break;
                }
        }
        if (sigma_.size() != (size_t)nz || n_parallelslices < 1) {
            report(INLM_BAD_ARG, who, "sigma not set for this shape or device not configured");
            return;
        }
        if ((int)filelist_raw.size() < off_raw + nz || (iter > 1 && (int)filelist_prev.size() < off_prev + nz)) {
            report(INLM_BAD_ARG, who, "file lists shorter than the stack");
            return;
        }
        inlm_params p = to_c(*params);
        const int halo = nz > 1 ? std::min(p.nslices / 2, p.radius_searchspace[2]) + p.radius_patchspace[2] : 0;
        const size_t slice = (size_t)nx * ny;
        const int block = n_parallelslices;
        const bool as16 = params->io.save_type == "16bit";
        std::vector<float> raw, prev, out;
        for (int z0 = 0; z0 < nz; z0 += block) {
            const int z1 = std::min(nz, z0 + block);
            if (params->io.resume && inlm_iteration_on_disk(params->io.active_outpath.c_str(), z0 + fs, z1 - 1 + fs))
                continue;
            const int lo = std::max(0, z0 - halo), hi = std::min(nz, z1 + halo), d = hi - lo;
            raw.resize(slice * d);
            prev.resize(slice * d);
            out.resize(slice * d);
            for (int z = lo; z < hi; z++) {
                if (!read_slice(filelist_raw[off_raw + z], raw.data() + slice * (z - lo), nx, ny, who)) return;
                if (iter > 1 && !read_slice(filelist_prev[off_prev + z], prev.data() + slice * (z - lo), nx, ny, who))
                    return;
            }
            const int32_t sh[3] = {nx, ny, d};
            if (!report_g(inlm_iterate_slabs(&p, sh, deviceID_, iter, raw.data(), prev.data(), sigma_.data() + lo,
                                             out.data(), d),
                          who))
                return;
            const int32_t bs[3] = {nx, ny, z1 - z0};
            if (!report_g(inlm_tiff_write_stack(params->io.active_outpath.c_str(), "denoised",
                                                out.data() + slice * (z0 - lo), bs, z0 + fs, as16 ? 1 : 0),
                          who))
                return;
        }
    }

    void set_sigma(float *sigmalist, int shape[3])
    {
        sigma_.assign(sigmalist, sigmalist + shape[2]);
        if (h_) report(inlm_set_sigma(h_, sigmalist, shape[2]), "set_sigma");
    }

    void get_result(float *output, int shape[3])
    {
        (void)shape;
        if (!h_) {
            report(INLM_BAD_ARG, "get_result", "no resident result (block mode)");
            return;
        }
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
                       : (m == "semimanual")          ? INLM_NOISE_SEMIMANUAL
                                                      : INLM_NOISE_ZADAPTIVE;
        p.zwindow = d.noiselevel.zwindow;
        p.sigma[0] = d.noiselevel.sigma[0];
        p.sigma[1] = d.noiselevel.sigma[1];
        p.n_samples = (int64_t)d.noiselevel.n_samples;
        p.patchsize = d.noiselevel.patchsize;
        p.stds_from_mean = d.noiselevel.stds_from_mean;
        p.continuous_estimate = d.noiselevel.continuous_estimate ? 1 : 0;
        p.circular_mask_diameter = d.noiselevel.circular_mask_diameter;
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
    int deviceID_ = 0;
    std::vector<float> sigma_;  // the filter strength per slice, for the block drivers

    static std::string zfill4(int v)
    {
        char b[32];
        std::snprintf(b, sizeof b, "%04d", v);
        return b;
    }
    static bool ends_with(const std::string &s, const std::string &t)
    {
        return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
    }
    bool read_slice(const std::string &file, float *dst, int nx, int ny, const char *who)
    {
        int32_t sh[3] = {0, 0, 0};
        inlm_status st = inlm_tiff_stack_shape(file.c_str(), sh);
        if (st == INLM_OK && (sh[0] != nx || sh[1] != ny)) st = INLM_BAD_ARG;
        if (st == INLM_BAD_ARG) return report(st, who, "slice file of another size");
        if (st == INLM_OK) st = inlm_tiff_read_stack(file.c_str(), 0, 0, dst);
        return report_g(st, who);
    }
    // io.firstslice..io.lastslice of io.active_outpath when every file is
    // there and complete (iternlm_gpu.cu:1665-1681), read into vol
    bool iteration_on_disk(const protocol::DenoiseParameters &pr, const int shape[3], std::vector<float> &vol)
    {
        const int fs = pr.io.firstslice > 0 ? pr.io.firstslice : 0;
        const int ls = pr.io.lastslice >= fs ? pr.io.lastslice : fs + shape[2] - 1;
        if (ls - fs + 1 != shape[2] || !inlm_iteration_on_disk(pr.io.active_outpath.c_str(), fs, ls)) return false;
        const size_t slice = (size_t)shape[0] * shape[1];
        vol.resize(slice * shape[2]);
        for (int z = fs; z <= ls; z++)
            if (!read_slice(pr.io.active_outpath + "/denoised" + zfill4(z) + ".tif", vol.data() + slice * (z - fs),
                            shape[0], shape[1], "resume"))
                return false;
        return true;
    }
    // a call whose message is the library's global one (no handle involved)
    bool report_g(inlm_status st, const char *where) { return report(st, where, inlm_last_error(nullptr)); }
    bool report(inlm_status st, const char *where, const char *msg = nullptr)
    {
        status_ = st;
        if (st != INLM_OK)
            std::fprintf(stderr, "%s: %s\n", where, msg ? msg : h_ ? inlm_last_error(h_) : inlm_last_error(nullptr));
        return st == INLM_OK;
    }
};
}  // namespace denoise

#endif  // INLM_ITERNLM_GPU_H
