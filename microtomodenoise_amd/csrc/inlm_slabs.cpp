// Slab drivers: the filter and the whole schedule over z-slabs, for volumes
// whose padded buffers do not fit one GPU together.
//
// The reference streams blocks of slices through the GPU from host memory
// or from the file list (Run_GaussianNoise_GPUBlocks, iternlm_gpu.cu:1729-
// 2115; the CPU block mode, iternlm_cpu.cpp:403-577).  Here an iteration
// visits the slabs in z order; a slab of output slices [z0, z1) is uploaded
// as a region that holds the filter's z reach (search |dz| plus patch |dz|)
// on both sides, so every output reads real neighbours, and the volume's
// own ends are reflect-padded exactly as in the whole-volume run.  Every
// region of one call has the same depth (regions at the volume's ends are
// shifted inward), so one handle serves all slabs.  sigma is indexed by
// global z (the reference's block driver indexes it by block-local z,
// iternlm_gpu.cu:121), and the halo comes from the previous iterate (the
// reference's host-RAM block driver reads it from the buffer it is
// overwriting, so a block's lower halo holds the new iterate).  The noise
// estimate uses the whole volume's sample list: samples are sorted by linear
// index, each sample's patch lies in one slice, so slab by slab gives the
// same standard deviations.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

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

// z reach of one output: search |dz| <= min(nslices/2, rs_z), patch |dz| <= rp_z
int z_reach(const inlm_params &p, int nz)
{
    if (nz == 1) return 0;
    return std::min(p.nslices / 2, p.radius_searchspace[2]) + p.radius_patchspace[2];
}

// Depth of the region a slab of `slab` output slices is uploaded as: the
// outputs and the reach on both sides, and at least two slices (a one-slice
// handle drops the z radii and, at patch radius 1, switches the weight
// table), at most the volume.
int region_depth(int slab, int halo, int nz) { return std::min(nz, std::max(2, slab + 2 * halo)); }

inlm_status check_common(const inlm_params *p, const int32_t shape[3], const char *who)
{
    if (!p || !shape) {
        inlm::set_global_error(std::string(who) + ": null argument");
        return INLM_BAD_ARG;
    }
    if (shape[0] < 1 || shape[1] < 1 || shape[2] < 1) {
        inlm::set_global_error(std::string(who) + ": empty shape");
        return INLM_BAD_ARG;
    }
    return INLM_OK;
}

}  // namespace

extern "C" inlm_status inlm_iterate_slabs(const inlm_params *p, const int32_t shape[3], int32_t device, int32_t iter,
                                          const float *raw, const float *prev, const float *sigma, float *out,
                                          int32_t slab_slices)
{
    inlm_status s = check_common(p, shape, "inlm_iterate_slabs");
    if (s != INLM_OK) return s;
    if (!raw || !sigma || !out || (iter > 1 && !prev)) {
        inlm::set_global_error("inlm_iterate_slabs: null argument");
        return INLM_BAD_ARG;
    }
    if (iter < 1 || slab_slices < 1) {
        inlm::set_global_error("inlm_iterate_slabs: iter and slab_slices must be >= 1");
        return INLM_BAD_ARG;
    }
    const int nx = shape[0], ny = shape[1], nz = shape[2];
    const int64_t slice = (int64_t)nx * ny;
    const size_t nvox = (size_t)(slice * nz);
    const int halo = z_reach(*p, nz);
    const int slab = std::min<int>(slab_slices, nz);
    const int D = region_depth(slab, halo, nz);

    // the halo of every slab comes from the previous iterate, also when the
    // caller lets the result overwrite it
    const float *pv = prev;
    std::vector<float> prev_copy;
    if (iter > 1 && out < prev + nvox && prev < out + nvox) {
        prev_copy.assign(prev, prev + nvox);
        pv = prev_copy.data();
    }
    Handle H;
    const int32_t sh[3] = {nx, ny, D};
    if ((s = inlm_create(p, sh, device, &H.h)) != INLM_OK) return s;
    auto bail = [&](inlm_status st) {
        inlm::set_global_error(std::string("inlm_iterate_slabs: ") + inlm_last_error(H.h));
        return st;
    };
    std::vector<float> part((size_t)(slice * D));
    for (int z0 = 0; z0 < nz; z0 += slab) {
        const int z1 = std::min(nz, z0 + slab);
        const int lo = std::min(std::max(0, z0 - halo), nz - D);
        if ((s = inlm_load_raw(H.h, raw + (int64_t)lo * slice, 0)) != INLM_OK) return bail(s);
        if ((s = inlm_set_sigma(H.h, sigma + lo, D)) != INLM_OK) return bail(s);
        if (iter > 1 && (s = inlm_set_result(H.h, iter - 1, pv + (int64_t)lo * slice, 0)) != INLM_OK) return bail(s);
        if ((s = inlm_run_iteration(H.h, iter, nullptr)) != INLM_OK) return bail(s);
        if ((s = inlm_get_result(H.h, part.data(), 0)) != INLM_OK) return bail(s);
        std::memcpy(out + (int64_t)z0 * slice, part.data() + (int64_t)(z0 - lo) * slice,
                    sizeof(float) * (size_t)(slice * (z1 - z0)));
    }
    return INLM_OK;
}

extern "C" inlm_status inlm_capacity_slices(const inlm_params *p, const int32_t shape[3], int32_t device,
                                            int64_t memory_limit, int32_t *n_slices)
{
    inlm_status s = check_common(p, shape, "inlm_capacity_slices");
    if (s != INLM_OK) return s;
    if (!n_slices) {
        inlm::set_global_error("inlm_capacity_slices: null argument");
        return INLM_BAD_ARG;
    }
    int64_t budget = memory_limit;
    if (budget <= 0) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
            inlm::set_global_error("no GPU visible");
            return INLM_NO_DEVICE;
        }
        if (device < 0 || device >= ndev || hipSetDevice(device) != hipSuccess) {
            inlm::set_global_error("inlm_capacity_slices: device index out of range");
            return INLM_BAD_ARG;
        }
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            inlm::set_global_error("inlm_capacity_slices: hipMemGetInfo failed");
            return INLM_HIP_ERROR;
        }
        // the runtime's own allocations (streams, events, kernel arguments)
        budget = (int64_t)fr - ((int64_t)256 << 20);
    }
    const int nz = shape[2];
    const int halo = z_reach(*p, nz);
    auto fits = [&](int slab) {
        const int32_t sh[3] = {shape[0], shape[1], slab >= nz ? nz : region_depth(slab, halo, nz)};
        return (int64_t)inlm::handle_bytes(*p, sh) <= budget;
    };
    if (fits(nz)) {
        *n_slices = nz;
        return INLM_OK;
    }
    if (!fits(1)) {
        inlm::set_global_error("inlm_capacity_slices: not even one slice fits the device memory");
        return INLM_OOM;
    }
    int lo = 1, hi = nz;  // fits(lo), !fits(hi)
    while (hi - lo > 1) {
        const int mid = lo + (hi - lo) / 2;
        (fits(mid) ? lo : hi) = mid;
    }
    *n_slices = lo;
    return INLM_OK;
}

extern "C" inlm_status inlm_denoise_slabs(const inlm_params *p, const int32_t shape[3], int32_t device,
                                          const float *raw, float *out, int32_t slab_slices, float *sigma_log)
{
    inlm_status s = check_common(p, shape, "inlm_denoise_slabs");
    if (s != INLM_OK) return s;
    if (!raw || !out) {
        inlm::set_global_error("inlm_denoise_slabs: null argument");
        return INLM_BAD_ARG;
    }
    if (slab_slices < 1) {
        inlm::set_global_error("inlm_denoise_slabs: slab_slices must be >= 1");
        return INLM_BAD_ARG;
    }
    const int nx = shape[0], ny = shape[1], nz = shape[2];
    const int64_t slice = (int64_t)nx * ny;

    // one slab holds the volume: the in-memory schedule itself
    if (slab_slices >= nz || nz == 1) {
        Handle H;
        if ((s = inlm_create(p, shape, device, &H.h)) != INLM_OK) return s;
        s = inlm_denoise(H.h, raw, out, sigma_log);
        if (s != INLM_OK) inlm::set_global_error(inlm_last_error(H.h));
        return s;
    }

    std::vector<int64_t> samples;
    if (p->noise_mode != INLM_NOISE_MANUAL) {
        samples.resize((size_t)p->n_samples);
        int sh[3] = {nx, ny, nz};
        if (!inlm::draw_samples(sh, p->n_samples, p->patchsize, p->sample_seed, samples.data(),
                                p->circular_mask_diameter)) {
            inlm::set_global_error("inlm_denoise_slabs: no noise sample fits the volume (or the circular mask)");
            return INLM_BAD_ARG;
        }
    }
    const int64_t ns = (int64_t)samples.size();
    int64_t semi_ref = -1;

    // noise level of a whole host volume (which: 0 first estimate, 1 later)
    auto estimate = [&](const float *vol, int which, float *sig) -> inlm_status {
        std::vector<double> sd((size_t)std::max<int64_t>(ns, 1));
        std::vector<int64_t> local;
        int64_t i0 = 0;
        for (int z0 = 0; z0 < nz && ns > 0; z0 += slab_slices) {
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
                if ((s = inlm_load_raw(H.h, vol + (int64_t)z0 * slice, 0)) != INLM_OK ||
                    (s = inlm_patchstd(H.h, 0, local.data(), i1 - i0, sd.data() + i0)) != INLM_OK) {
                    inlm::set_global_error(inlm_last_error(H.h));
                    return s;
                }
            }
            i0 = i1;
        }
        const int sh[3] = {nx, ny, nz};
        std::string err;
        s = inlm::sigma_from_std(*p, sh, samples.data(), sd.data(), ns, which, semi_ref, sig, err);
        if (s != INLM_OK) inlm::set_global_error("inlm_denoise_slabs: " + err);
        return s;
    };

    std::vector<float> cur((size_t)(slice * nz));
    // the schedule of inlm_denoise (main.cpp:325-448, :517-544)
    std::vector<float> s0(nz), s1(nz), tmp(nz);
    if ((s = estimate(raw, 0, s0.data())) != INLM_OK) return s;
    if ((s = inlm_iterate_slabs(p, shape, device, 1, raw, nullptr, s0.data(), cur.data(), slab_slices)) != INLM_OK)
        return s;
    if (sigma_log) std::memcpy(sigma_log, s0.data(), sizeof(float) * nz);
    auto reestimate = [&]() -> inlm_status {
        inlm_status r = estimate(cur.data(), 1, tmp.data());
        if (r != INLM_OK) return r;
        const float a = p->alpha;
        for (int i = 0; i < nz; i++) s1[i] = a * s0[i] + (1.f - a) * tmp[i];
        return INLM_OK;
    };
    if (p->maxiterations > 1 && (s = reestimate()) != INLM_OK) return s;
    for (int iter = 2; iter <= p->maxiterations; iter++) {
        // the result overwrites the previous iterate; the slabs read their
        // halo from a copy of it
        if ((s = inlm_iterate_slabs(p, shape, device, iter, raw, cur.data(), s1.data(), cur.data(), slab_slices)) !=
            INLM_OK)
            return s;
        if (sigma_log) std::memcpy(sigma_log + (size_t)(iter - 1) * nz, s1.data(), sizeof(float) * nz);
        if (p->continuous_estimate && p->maxiterations > iter + 1 && (s = reestimate()) != INLM_OK) return s;
    }
    std::memcpy(out, cur.data(), sizeof(float) * (size_t)(slice * nz));
    return INLM_OK;
}
