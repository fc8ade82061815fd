// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Drives the reference's own CPU code, compiled from its sources where they
// lie (Denoiser/iternlm_prepare.cpp and Denoiser/iternlm_cpu_kernels.cpp of
// iterNLM_v0.3), so the oracle restatement can be pinned bit for bit.  The
// result is oracle/_ref/libinlm_ref.so; oracle/Makefile builds it only when
// the reference tree is present.  Nothing in the shipped library uses it.
//
// The specialised kernels filterslice_p111...p333 and the table pointers are
// private members of denoise::IterativeNLM_CPU (Denoiser/iternlm_cpu.h:118-154).
// They are reached through explicit template instantiation, for which the
// language waives access checking; no reference header is edited or replaced.
// Denoiser/iternlm_cpu.cpp (the slice loop and the generic kernel) includes
// libtiff headers that this image lacks and is therefore not built; the loop
// below restates its per-slice driver (iternlm_cpu.cpp:91-114): sigma per
// slice, m = -1/(sigma^2*beta), one OpenMP thread per slice.
#include <cstdlib>
#include <cstring>
#include "Denoiser/iternlm_cpu.h"

using denoise::IterativeNLM_CPU;
typedef void (IterativeNLM_CPU::*SliceFn)(int, float, float *, float *, float *, int *, protocol::DenoiseParameters *);

template <typename Tag, typename Tag::type M>
struct Expose { friend typename Tag::type get(Tag) { return M; } };

#define EXPOSE(TAG, TYPE, MEMBER)                                    \
    struct TAG { typedef TYPE type; friend type get(TAG); };          \
    template struct Expose<TAG, &IterativeNLM_CPU::MEMBER>;

EXPOSE(TagNS, int IterativeNLM_CPU::*, nsize_search)
EXPOSE(TagNP, int IterativeNLM_CPU::*, nsize_patch)
EXPOSE(TagSP, long long int *IterativeNLM_CPU::*, search_positions)
EXPOSE(TagPP, long long int *IterativeNLM_CPU::*, patch_positions)
EXPOSE(TagDW, float *IterativeNLM_CPU::*, distweight)
EXPOSE(TagP111, SliceFn, filterslice_p111)
EXPOSE(TagP112, SliceFn, filterslice_p112)
EXPOSE(TagP113, SliceFn, filterslice_p113)
EXPOSE(TagP221, SliceFn, filterslice_p221)
EXPOSE(TagP222, SliceFn, filterslice_p222)
EXPOSE(TagP331, SliceFn, filterslice_p331)
EXPOSE(TagP332, SliceFn, filterslice_p332)
EXPOSE(TagP333, SliceFn, filterslice_p333)

static void fill_params(protocol::DenoiseParameters &p, const int rs[3], const int rp[3], int nslices, float beta,
                        float z_aniso)
{
    for (int i = 0; i < 3; i++) {
        p.radius_searchspace[i] = rs[i];
        p.radius_patchspace[i] = rp[i];
    }
    p.nslices = nslices;
    p.beta = beta;
    p.z_anisotropy = z_aniso;
}

extern "C" {

// Tables from setup_searchspace / setup_patchspace / setup_distweight
// (iternlm_prepare.cpp:102-263) on a padded shape.  Returns counts through
// nS/nP and copies up to cap entries.
__attribute__((visibility("default"))) void ref_tables(const int shape_padded[3], const int rs[3], const int rp[3],
                                                        int nslices, float z_aniso, int *nS, long long *soff,
                                                        int *nP, long long *poff, float *w, int cap)
{
    protocol::DenoiseParameters p;
    fill_params(p, rs, rp, nslices, 1.f, z_aniso);
    int sh[3] = {shape_padded[0], shape_padded[1], shape_padded[2]};
    long long *s = denoise::setup_searchspace(sh, &p, *nS);
    long long *pp = denoise::setup_patchspace(sh, &p, *nP);
    float *ww = denoise::setup_distweight(sh, &p);
    for (int i = 0; i < *nS && i < cap; i++) soff[i] = s[i];
    for (int i = 0; i < *nP && i < cap; i++) {
        poff[i] = pp[i];
        w[i] = ww[i];
    }
    free(s);
    free(pp);
    free(ww);
}

// pad_reflective (iternlm_prepare.cpp:6-48).
__attribute__((visibility("default"))) void ref_pad_reflective(const float *in, const int shape[3], const int pad[6],
                                                                float *out)
{
    int pd[6];
    for (int i = 0; i < 6; i++) pd[i] = pad[i];
    int outshape[3];
    float *o = denoise::pad_reflective(const_cast<float *>(in), pd, shape, outshape);
    memcpy(out, o, sizeof(float) * (size_t)outshape[0] * outshape[1] * outshape[2]);
    free(o);
}

// One filter iteration through the reference's specialised kernel `kernel`
// (111,112,113,221,222,331,332,333) over all slices of an unpadded volume of
// shape[3].  Rpad/Ppad are padded as the reference pads them.
__attribute__((visibility("default"))) int ref_filter(int kernel, float *Rpad, float *Ppad, const int shape[3],
                                                       const int rs[3], const int rp[3], int nslices, float beta,
                                                       const float *sigma, float *out, int nthreads)
{
    SliceFn fn;
    switch (kernel) {
    case 111: fn = get(TagP111()); break;
    case 112: fn = get(TagP112()); break;
    case 113: fn = get(TagP113()); break;
    case 221: fn = get(TagP221()); break;
    case 222: fn = get(TagP222()); break;
    case 331: fn = get(TagP331()); break;
    case 332: fn = get(TagP332()); break;
    case 333: fn = get(TagP333()); break;
    default: return -1;
    }
    protocol::DenoiseParameters p;
    fill_params(p, rs, rp, nslices, beta, 1.f);
    int pz = std::min(nslices / 2, rs[2]) + rp[2];
    int sh[3] = {shape[0], shape[1], shape[2]};
    int shp[3] = {shape[0] + 2 * (rs[0] + rp[0]), shape[1] + 2 * (rs[1] + rp[1]), shape[2] + 2 * pz};
    IterativeNLM_CPU obj;
    obj.*get(TagSP()) = denoise::setup_searchspace(shp, &p, obj.*get(TagNS()));
    obj.*get(TagPP()) = denoise::setup_patchspace(shp, &p, obj.*get(TagNP()));
    obj.*get(TagDW()) = denoise::setup_distweight(shp, &p);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < shape[2]; i++) {
        float s = sigma[i];
        float multiplier = -1.f / (s * s * p.beta);
        (obj.*fn)(i, multiplier, Rpad, Ppad, out, sh, &p);
    }
    free(obj.*get(TagSP()));
    free(obj.*get(TagPP()));
    free(obj.*get(TagDW()));
    return 0;
}

}  // extern "C"
