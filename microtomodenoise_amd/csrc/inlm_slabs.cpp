// Out-of-core driver: the whole schedule of inlm_denoise over z-slabs, for
// volumes whose padded buffers do not fit one GPU together.
//
// The reference streams blocks of slices through the GPU from host memory
// or from the file list (Run_GaussianNoise_GPUBlocks, iternlm_gpu.cu:1729-
// 2115; the CPU block mode, iternlm_cpu.cpp:403-577).  Here every iteration
// visits the slabs in z order; a slab of output slices [z0, z1) is uploaded
// with the filter's z reach (search |dz| plus patch |dz|) on both sides, so
// every output reads real neighbours, and the volume's own ends are
// reflect-padded exactly as in the whole-volume run.  sigma is indexed by
// global z (the reference's block driver indexes it by block-local z,
// iternlm_gpu.cu:121).  The noise estimate uses the whole volume's sample
// list: samples are sorted by linear index, each sample's patch lies in one
// slice, so slab by slab gives the same standard deviations.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "inlm_capi.h"
#include "inlm_internal.hpp"
#include "inlm_tables.hpp"

namespace {

struct Handle {
    inlm_handle *h = nullptr;
    ~Handle()
    {
        if (h) inlm_destroy(h);
    }
};

}  // namespace

extern "C" inlm_status inlm_denoise_slabs(const inlm_params *p, const int32_t shape[3], int32_t device,
                                          const float *raw, float *out, int32_t slab_slices, float *sigma_log)
{
    if (!p || !shape || !raw || !out) {
        inlm::set_global_error("inlm_denoise_slabs: null argument");
        return INLM_BAD_ARG;
    }
    if (slab_slices < 1) {
        inlm::set_global_error("inlm_denoise_slabs: slab_slices must be >= 1");
        return INLM_BAD_ARG;
    }
    const int nx = shape[0], ny = shape[1], nz = shape[2];
    if (nx < 1 || ny < 1 || nz < 1) {
        inlm::set_global_error("inlm_denoise_slabs: empty shape");
        return INLM_BAD_ARG;
    }
    const int64_t slice = (int64_t)nx * ny;
    inlm_status s;

    // one slab holds the volume: the in-memory schedule itself
    if (slab_slices >= nz || nz == 1) {
        Handle H;
        if ((s = inlm_create(p, shape, device, &H.h)) != INLM_OK) return s;
        return inlm_denoise(H.h, raw, out, sigma_log);
    }
    // z reach of one output: search |dz| <= min(nslices/2, rs_z), patch |dz| <= rp_z
    const int halo = std::min(p->nslices / 2, p->radius_searchspace[2]) + p->radius_patchspace[2];

    std::vector<int64_t> samples;
    if (p->noise_mode != INLM_NOISE_MANUAL) {
        samples.resize((size_t)p->n_samples);
        int sh[3] = {nx, ny, nz};
        if (!inlm::draw_samples(sh, p->n_samples, p->patchsize, p->sample_seed, samples.data())) {
            inlm::set_global_error("inlm_denoise_slabs: volume smaller than the noise patch");
            return INLM_BAD_ARG;
        }
    }
    const int64_t ns = (int64_t)samples.size();

    // noise level of a whole host volume (which: 0 first estimate, 1 later)
    auto estimate = [&](const float *vol, int which, float *sig) -> inlm_status {
        if (p->noise_mode == INLM_NOISE_MANUAL) {
            for (int i = 0; i < nz; i++) sig[i] = p->sigma[which == 0 ? 0 : 1];
            return INLM_OK;
        }
        std::vector<double> sd((size_t)std::max<int64_t>(ns, 1));
        std::vector<int64_t> local;
        int64_t i0 = 0;
        for (int z0 = 0; z0 < nz; z0 += slab_slices) {
            const int z1 = std::min(nz, z0 + slab_slices);
            int64_t i1 = i0;
            while (i1 < ns && samples[(size_t)i1] < (int64_t)z1 * slice) i1++;
            if (i1 > i0) {
                local.resize((size_t)(i1 - i0));
                for (int64_t i = i0; i < i1; i++) local[(size_t)(i - i0)] = samples[(size_t)i] - (int64_t)z0 * slice;
                const int32_t sh[3] = {nx, ny, z1 - z0};
                inlm_params q = *p;
                q.precision = INLM_PREC_FP32;  // the sample statistics read the volume as given
                Handle H;
                if ((s = inlm_create(&q, sh, device, &H.h)) != INLM_OK) return s;
                if ((s = inlm_load_raw(H.h, vol + (int64_t)z0 * slice, 0)) != INLM_OK) return s;
                if ((s = inlm_patchstd(H.h, 0, local.data(), i1 - i0, sd.data() + i0)) != INLM_OK) return s;
            }
            i0 = i1;
        }
        const double k = p->stds_from_mean;
        if (p->noise_mode == INLM_NOISE_ZADAPTIVE) {
            std::vector<double> mean, stdev;
            int sh[3] = {nx, ny, nz};
            inlm::zwindow_clusters(samples.data(), sd.data(), ns, sh, p->zwindow, mean, stdev);
            for (int i = 0; i < nz; i++) sig[i] = (float)(mean[i] + k * stdev[i]);
        } else {
            double mean = 0, stdev = 0;
            inlm::create2clusters(sd.data(), (int)ns, 0, (int)ns - 1, mean, stdev);
            for (int i = 0; i < nz; i++) sig[i] = (float)(mean + k * stdev);
        }
        return INLM_OK;
    };

    std::vector<float> prev((size_t)(slice * nz)), next((size_t)(slice * nz)), part;
    // one iteration over the slabs: previous iterate in prev (unused at 1)
    auto iterate = [&](int iter, const float *sig) -> inlm_status {
        for (int z0 = 0; z0 < nz; z0 += slab_slices) {
            const int z1 = std::min(nz, z0 + slab_slices);
            const int lo = std::max(0, z0 - halo), hi = std::min(nz, z1 + halo);
            const int32_t sh[3] = {nx, ny, hi - lo};
            Handle H;
            if ((s = inlm_create(p, sh, device, &H.h)) != INLM_OK) return s;
            if ((s = inlm_load_raw(H.h, raw + (int64_t)lo * slice, 0)) != INLM_OK) return s;
            if ((s = inlm_set_sigma(H.h, sig + lo, hi - lo)) != INLM_OK) return s;
            if (iter > 1 && (s = inlm_set_result(H.h, iter - 1, prev.data() + (int64_t)lo * slice, 0)) != INLM_OK)
                return s;
            if ((s = inlm_run_iteration(H.h, iter, nullptr)) != INLM_OK) return s;
            part.resize((size_t)(slice * (hi - lo)));
            if ((s = inlm_get_result(H.h, part.data(), 0)) != INLM_OK) return s;
            std::memcpy(next.data() + (int64_t)z0 * slice, part.data() + (int64_t)(z0 - lo) * slice,
                        sizeof(float) * (size_t)(slice * (z1 - z0)));
        }
        prev.swap(next);
        return INLM_OK;
    };

    // the schedule of inlm_denoise (main.cpp:325-448, :517-544)
    std::vector<float> s0(nz), s1(nz), tmp(nz);
    if ((s = estimate(raw, 0, s0.data())) != INLM_OK) return s;
    if ((s = iterate(1, s0.data())) != INLM_OK) return s;
    if (sigma_log) std::memcpy(sigma_log, s0.data(), sizeof(float) * nz);
    auto reestimate = [&]() -> inlm_status {
        inlm_status r = estimate(prev.data(), 1, tmp.data());
        if (r != INLM_OK) return r;
        const float a = p->alpha;
        for (int i = 0; i < nz; i++) s1[i] = a * s0[i] + (1.f - a) * tmp[i];
        return INLM_OK;
    };
    if (p->maxiterations > 1 && (s = reestimate()) != INLM_OK) return s;
    for (int iter = 2; iter <= p->maxiterations; iter++) {
        if ((s = iterate(iter, s1.data())) != INLM_OK) return s;
        if (sigma_log) std::memcpy(sigma_log + (size_t)(iter - 1) * nz, s1.data(), sizeof(float) * nz);
        if (p->continuous_estimate && p->maxiterations > iter + 1 && (s = reestimate()) != INLM_OK) return s;
    }
    std::memcpy(out, prev.data(), sizeof(float) * (size_t)(slice * nz));
    return INLM_OK;
}
