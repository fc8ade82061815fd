/*
 * inlm_oracle.c -- CPU restatement of the iterNLM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline use it, always as the checker.
 *
 * Two families live here:
 *   - *_f32: a faithful FP32 restatement of the reference CPU (OpenMP) path,
 *     operation for operation (same loop order, same tap order, unfused
 *     multiply/add, Pade [3/3] weight with the -3.56648 cutoff, float
 *     normalisation).  Built with -ffp-contract=off so that every
 *     intermediate rounds to float exactly as the reference's x86-64 SSE
 *     build does.
 *   - *_f64: the same algorithm carried in double precision; it is the
 *     yardstick for the reduced-precision forms (BF16/FP16 storage, FMA
 *     fast mode, pruned offsets).
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to iterNLM_v0.3/ in iternlm/microtomodenoise.
 *
 * How each part is pinned (details in tests/):
 *   tables, padding, FP32 filter: bit-exact against oracle/_ref, which is
 *     compiled from the reference's own iternlm_prepare.cpp and
 *     iternlm_cpu_kernels.cpp, plus closed-form counts and invariants;
 *   generic filter with computed weights: same code path as the pinned
 *     specialised form, weights pinned against oracle/_ref;
 *   noise estimator: the reference's Noise/noiselevel.cpp needs libtiff
 *     headers that this image lacks, so it cannot be compiled; pinned by
 *     closed forms (constant and ramp patches), a planted bimodal sample set
 *     and fixed-point invariance of the clustering;
 *   FP64 filter: closed forms (sigma -> infinity box mean, sigma -> 0
 *     identity, constant volume, shift equivariance) and agreement with the
 *     FP32 restatement within a stated budget.
 *   exp weight mode: parity unpinned (the reference GPU uses CUDA's expf,
 *     which cannot be run here); checked against libm exp only.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

static int imax(int a, int b) { return a > b ? a : b; }
static int iabs(int a) { return a < 0 ? -a : a; }

/* ------------------------------------------------------------------------ */
/* Tables                                                                    */
/* ------------------------------------------------------------------------ */

/* Search offsets as (dx,dy,dz) triples in the reference's z->y->x scan order.
 * Restates Denoiser/iternlm_prepare.cpp:102-151: keep |dz| <= nslices/2
 * (integer division, :137), drop the centre (:138), keep the float32
 * ellipsoid (x/rx)^2+(y/ry)^2+(z/rz)^2 <= 1 with r = max(radius,1)
 * (:119-121, :139).  Returns the count; writes at most cap triples. */
ORC_API int orc_search_offsets(const int rs[3], int nslices, int *dxyz, int cap)
{
    const float rxs = (float)imax(rs[0], 1);
    const float rys = (float)imax(rs[1], 1);
    const float rzs = (float)imax(rs[2], 1);
    int n = 0;
    for (int zs = -rs[2]; zs <= rs[2]; zs++)
        for (int ys = -rs[1]; ys <= rs[1]; ys++)
            for (int xs = -rs[0]; xs <= rs[0]; xs++) {
                if (iabs(zs) > nslices / 2) continue;
                if (zs == 0 && ys == 0 && xs == 0) continue;
                float fx = (float)xs / rxs, fy = (float)ys / rys, fz = (float)zs / rzs;
                float e = fx * fx;
                e = e + fy * fy;
                e = e + fz * fz;
                if (e <= 1.f) {
                    if (n < cap) { dxyz[3 * n] = xs; dxyz[3 * n + 1] = ys; dxyz[3 * n + 2] = zs; }
                    n++;
                }
            }
    return n;
}

/* Patch taps, centre first.  Restates iternlm_prepare.cpp:152-203. */
ORC_API int orc_patch_offsets(const int rp[3], int nslices, int *dxyz, int cap)
{
    const float rxp = (float)imax(rp[0], 1);
    const float ryp = (float)imax(rp[1], 1);
    const float rzp = (float)imax(rp[2], 1);
    int n = 0;
    if (cap > 0) { dxyz[0] = dxyz[1] = dxyz[2] = 0; }
    n = 1;
    for (int zp = -rp[2]; zp <= rp[2]; zp++)
        for (int yp = -rp[1]; yp <= rp[1]; yp++)
            for (int xp = -rp[0]; xp <= rp[0]; xp++) {
                if (iabs(zp) > nslices / 2) continue;
                if (zp == 0 && yp == 0 && xp == 0) continue;
                float fx = (float)xp / rxp, fy = (float)yp / ryp, fz = (float)zp / rzp;
                float e = fx * fx;
                e = e + fy * fy;
                e = e + fz * fz;
                if (e <= 1.f) {
                    if (n < cap) { dxyz[3 * n] = xp; dxyz[3 * n + 1] = yp; dxyz[3 * n + 2] = zp; }
                    n++;
                }
            }
    return n;
}

/* Computed patch weights 1/(2d+1)^2 with d = sqrtf(x^2+y^2+z^2*aniso^2); the
 * centre gets the largest weight; all 1/7 when every max(r,1) == 1 and the
 * padded depth exceeds 1.  Restates iternlm_prepare.cpp:204-263. */
ORC_API int orc_distweight(const int rp[3], int nslices, float z_aniso, int nz_padded, float *w, int cap)
{
    const float rxp = (float)imax(rp[0], 1);
    const float ryp = (float)imax(rp[1], 1);
    const float rzp = (float)imax(rp[2], 1);
    const float sq_an = z_aniso * z_aniso;
    float maxw = 0.f;
    int n = 1;
    float *tmp = (float *)malloc(sizeof(float) * (size_t)(2 * rp[0] + 1) * (2 * rp[1] + 1) * (2 * rp[2] + 1) + 16);
    tmp[0] = 1.f;
    for (int zp = -rp[2]; zp <= rp[2]; zp++)
        for (int yp = -rp[1]; yp <= rp[1]; yp++)
            for (int xp = -rp[0]; xp <= rp[0]; xp++) {
                if (iabs(zp) > nslices / 2) continue;
                if (zp == 0 && yp == 0 && xp == 0) continue;
                float fx = (float)xp / rxp, fy = (float)yp / ryp, fz = (float)zp / rzp;
                float e = fx * fx;
                e = e + fy * fy;
                e = e + fz * fz;
                if (e <= 1.f) {
                    /* sqrtf((xp*xp)+(yp*yp)+(zp*zp)*sq_anisotropy): the integer
                     * sum xp*xp+yp*yp converts to float before adding the
                     * float z term (iternlm_prepare.cpp:242). */
                    int ixy = xp * xp + yp * yp;
                    float zz = (float)(zp * zp) * sq_an;
                    float ed = sqrtf((float)ixy + zz);
                    float t = 2.f * ed + 1.f;
                    float wd = 1.f / (t * t);
                    if (wd > maxw) maxw = wd;
                    tmp[n++] = wd;
                }
            }
    tmp[0] = maxw;
    int special = (rxp == 1.f && ryp == 1.f && rzp == 1.f && nz_padded > 1);
    for (int i = 0; i < n && i < cap; i++) w[i] = special ? 1.f / 7.f : tmp[i];
    free(tmp);
    return n;
}

/* The reference's specialised CPU kernels hard-code the weights printed by
 * its kernel-code generator (iternlm_gpu.cu:1516-1517), i.e. the computed
 * float weight streamed with the default 6 significant digits and parsed back
 * as a float literal.  The r=1 kernel is the exception: its literal
 * 0.142857143f is 1/7 itself (iternlm_cpu_kernels.cpp:59-65). */
ORC_API void orc_literal_weights(const float *w, int n, float *out)
{
    char buf[64];
    for (int i = 0; i < n; i++) {
        snprintf(buf, sizeof buf, "%g", (double)w[i]);
        out[i] = strtof(buf, NULL);
    }
}

/* Which weight table the reference CPU path applies for patch radius rp.
 * blockwise selects the dispatch of Run_GaussianNoise_blocks
 * (iternlm_cpu.cpp:529-548), otherwise Run_GaussianNoise
 * (iternlm_cpu.cpp:97-114).  Returns 1 for the literal table of a
 * specialised kernel, 2 for the r=1 kernel (1/7), 0 for the generic kernel
 * (computed weights), -1 for a dispatch that reaches a specialised kernel
 * whose hard-coded tap count does not match the patch (undefined behaviour
 * in the reference: block mode with radius (2,2,3) lands in
 * filterslice_p333). */
ORC_API int orc_cpu_weight_kind(const int rp[3], int blockwise)
{
    int a = rp[0], b = rp[1], c = rp[2];
    if (a == 1 && b == 1 && c == 1) return 2;
    if (a == 1 && b == 1 && (c == 2 || c == 3)) return 1;
    if (a == 2 && b == 2 && (c == 1 || c == 2)) return 1;
    if (!blockwise) {
        if (a == 3 && b == 3 && (c == 1 || c == 2 || c == 3)) return 1;
        return 0;
    }
    if (a == 2 && b == 2 && c == 3) return -1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Reflective padding: iternlm_prepare.cpp:6-48 (mirror without repeating the */
/* edge, iterated until in range).  pad = {x0,y0,z0,x1,y1,z1}.               */
/* ------------------------------------------------------------------------ */
ORC_API int orc_pad_reflect(const float *in, const int shape[3], const int pad[6], float *out)
{
    const int nx0 = shape[0], ny0 = shape[1], nz0 = shape[2];
    for (int a = 0; a < 3; a++)
        if (shape[a] < 2 && (pad[a] > 0 || pad[a + 3] > 0)) return -1; /* reference loops forever */
    const int nx1 = nx0 + pad[0] + pad[3], ny1 = ny0 + pad[1] + pad[4], nz1 = nz0 + pad[2] + pad[5];
    const long long nslice0 = (long long)nx0 * ny0, nslice1 = (long long)nx1 * ny1;
#pragma omp parallel for schedule(static)
    for (int z1 = 0; z1 < nz1; z1++)
        for (int y1 = 0; y1 < ny1; y1++)
            for (int x1 = 0; x1 < nx1; x1++) {
                int z0 = z1 - pad[2], y0 = y1 - pad[1], x0 = x1 - pad[0];
                while (z0 < 0 || z0 >= nz0 || y0 < 0 || y0 >= ny0 || x0 < 0 || x0 >= nx0) {
                    if (z0 < 0) z0 = -z0;
                    if (y0 < 0) y0 = -y0;
                    if (x0 < 0) x0 = -x0;
                    if (z0 >= nz0) z0 = 2 * nz0 - z0 - 2;
                    if (y0 >= ny0) y0 = 2 * ny0 - y0 - 2;
                    if (x0 >= nx0) x0 = 2 * nx0 - x0 - 2;
                }
                out[z1 * nslice1 + (long long)y1 * nx1 + x1] = in[z0 * nslice0 + (long long)y0 * nx0 + x0];
            }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Weight function                                                           */
/* ------------------------------------------------------------------------ */

/* Pade [3/3] of exp, evaluated left to right as written at
 * Denoiser/iternlm_cpu.h:125, with the cutoff of iternlm_cpu.h:126 applied as
 * at iternlm_cpu.cpp:815. */
static float pade_cut_f32(float x)
{
    if (!(x > -3.56648f)) return 0.0f;
    float num = 120.f + 60.f * x;
    num = num + (12.f * x) * x;
    num = num + (x * x) * x;
    float den = 120.f - 60.f * x;
    den = den + (12.f * x) * x;
    den = den - (x * x) * x;
    return num / den;
}
static double pade_cut_f64(double x)
{
    if (!(x > (double)-3.56648f)) return 0.0;
    double num = 120.0 + 60.0 * x + 12.0 * x * x + x * x * x;
    double den = 120.0 - 60.0 * x + 12.0 * x * x - x * x * x;
    return num / den;
}
ORC_API float orc_pade_f32(float x) { return pade_cut_f32(x); }
ORC_API double orc_pade_f64(double x) { return pade_cut_f64(x); }

/* ------------------------------------------------------------------------ */
/* Filter core                                                               */
/* ------------------------------------------------------------------------ */

/* One iteration over output slices [z_begin, z_end) of an unpadded volume of
 * shape nx,ny,nz.  Rpad/Ppad are the reflect-padded original and previous
 * iterate with padding pad = {px,py,pz} on both sides; soff/poff are linear
 * shifts in the padded layout (x fastest).  Restates
 * iternlm_cpu.cpp:758-843 (generic) and the specialised
 * iternlm_cpu_kernels.cpp:7-1108, which differ only in the weight table:
 *   d = sum_p ((P[x+s+p]-P[x+p])^2 * W_p) in tap order, unfused;
 *   a = d*m with m = -1/(sigma^2*beta) per slice (iternlm_cpu.cpp:94-95);
 *   w = pade(a) if a > -3.56648 else 0 (weight_mode 0), or expf(a)
 *       (weight_mode 1, the reference GPU's choice, iternlm_gpu.cu:150);
 *   num += w*R[x+s]; den += w; wmax = max;
 *   out = (num + wmax*R[x]) / (den + wmax) if wmax > 0 else R[x]. */
ORC_API void orc_filter_f32(const float *Rpad, const float *Ppad, const int shape[3], const int pad[3],
                            const float *sigma, float beta, int nS, const long long *soff, int nP,
                            const long long *poff, const float *w, int weight_mode, float *out,
                            int z_begin, int z_end, int nthreads)
{
    const int nx = shape[0], ny = shape[1];
    const long long nxp = nx + 2 * pad[0], nyp = ny + 2 * pad[1];
    const long long nslp = nxp * nyp;
    float *orig_all = (float *)malloc(sizeof(float) * (size_t)nP * (size_t)(nthreads > 0 ? nthreads : 256));
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int z = z_begin; z < z_end; z++) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        float *orig = orig_all + (size_t)tid * nP;
        const float s = sigma[z];
        const float m = -1.f / (s * s * beta);
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const long long i0 = (z + pad[2]) * nslp + (y + pad[1]) * nxp + (x + pad[0]);
                const float r0 = Rpad[i0];
                for (int p = 0; p < nP; p++) orig[p] = Ppad[i0 + poff[p]];
                float fv = 0.f, fw = 0.f, wmax = 0.f;
                for (int k = 0; k < nS; k++) {
                    const long long i1 = i0 + soff[k];
                    const float rs = Rpad[i1];
                    float d = 0.f;
                    for (int p = 0; p < nP; p++) {
                        float t = Ppad[i1 + poff[p]] - orig[p];
                        float t2 = t * t;
                        d = d + t2 * w[p];
                    }
                    d = d * m;
                    float wt = weight_mode == 0 ? pade_cut_f32(d) : expf(d);
                    fv = fv + wt * rs;
                    fw = fw + wt;
                    if (wt > wmax) wmax = wt;
                }
                float res;
                if (wmax > 0.f) {
                    fv = fv + wmax * r0;
                    fw = fw + wmax;
                    res = fv / fw;
                } else
                    res = r0;
                out[((long long)z * ny + y) * nx + x] = res;
            }
    }
    free(orig_all);
}

/* Three-channel filter of the colour path, Denoiser/iternlm_cpu.cpp:932-1075
 * (filterslice_3channels), operation for operation.  Rpad/Ppad hold the three
 * padded channels one after the other; mult[c] = -1/(sigma_c^2 beta)
 * (iternlm_cpu.cpp:384-386).  Joint weights (independent == 0): the mean of
 * the three scaled distances, (d0+d1+d2)/3, weights every channel; the
 * centre enters once with the largest weight.  Independent channels: one
 * weight per channel and no centre term (the reference's else-branch). */
ORC_API void orc_filter_rgb_f32(const float *Rpad, const float *Ppad, const int shape[3], const int pad[3],
                                const float mult[3], int independent, int nS, const long long *soff, int nP,
                                const long long *poff, const float *w, float *out)
{
    const int nx = shape[0], ny = shape[1], nz = shape[2];
    const long long nxp = nx + 2 * pad[0], nyp = ny + 2 * pad[1];
    const long long nslp = nxp * nyp, nzp = nz + 2 * pad[2], npad = nslp * nzp;
    const long long nvox = (long long)nx * ny * nz;
    const float *R0 = Rpad, *R1 = Rpad + npad, *R2 = Rpad + 2 * npad;
    const float *P0 = Ppad, *P1 = Ppad + npad, *P2 = Ppad + 2 * npad;
#pragma omp parallel for schedule(dynamic, 1)
    for (int y = 0; y < ny; y++)
        for (int z = 0; z < nz; z++)
            for (int x = 0; x < nx; x++) {
                const long long i0 = (z + pad[2]) * nslp + (y + pad[1]) * nxp + (x + pad[0]);
                float fv0 = 0.f, fv1 = 0.f, fv2 = 0.f, fw0 = 0.f, fw1 = 0.f, fw2 = 0.f;
                float wm0 = 0.f, wm1 = 0.f, wm2 = 0.f;
                for (int k = 0; k < nS; k++) {
                    const long long i1 = i0 + soff[k];
                    float d0 = 0.f, d1 = 0.f, d2 = 0.f;
                    for (int p = 0; p < nP; p++) {
                        const long long pp = poff[p];
                        const float dw = w[p];
                        float t = P0[i1 + pp] - P0[i0 + pp];
                        d0 += (t * t) * dw;
                        t = P1[i1 + pp] - P1[i0 + pp];
                        d1 += (t * t) * dw;
                        t = P2[i1 + pp] - P2[i0 + pp];
                        d2 += (t * t) * dw;
                    }
                    d0 = d0 * mult[0];
                    d1 = d1 * mult[1];
                    d2 = d2 * mult[2];
                    if (!independent) {
                        const float ds = (d0 + d1 + d2) / 3.f;
                        const float wt = pade_cut_f32(ds);
                        fv0 += wt * R0[i1];
                        fv1 += wt * R1[i1];
                        fv2 += wt * R2[i1];
                        fw0 += wt;
                        if (wt > wm0) wm0 = wt;
                    } else {
                        const float w0 = pade_cut_f32(d0), w1 = pade_cut_f32(d1), w2 = pade_cut_f32(d2);
                        fv0 += w0 * R0[i1];
                        fv1 += w1 * R1[i1];
                        fv2 += w2 * R2[i1];
                        fw0 += w0;
                        fw1 += w1;
                        fw2 += w2;
                        if (w0 > wm0) wm0 = w0;
                        if (w1 > wm1) wm1 = w1;
                        if (w2 > wm2) wm2 = w2;
                    }
                }
                const long long o = ((long long)z * ny + y) * nx + x;
                const float r0 = R0[i0], r1 = R1[i0], r2 = R2[i0];
                if (!independent && wm0 > 0.f) {
                    fv0 += wm0 * r0;
                    fv1 += wm0 * r1;
                    fv2 += wm0 * r2;
                    fw0 += wm0;
                    out[o] = fv0 / fw0;
                    out[nvox + o] = fv1 / fw0;
                    out[2 * nvox + o] = fv2 / fw0;
                } else if (independent) {
                    out[o] = wm0 > 0.f ? fv0 / fw0 : r0;
                    out[nvox + o] = wm1 > 0.f ? fv1 / fw1 : r1;
                    out[2 * nvox + o] = wm2 > 0.f ? fv2 / fw2 : r2;
                } else {
                    out[o] = r0;
                    out[nvox + o] = r1;
                    out[2 * nvox + o] = r2;
                }
            }
}

/* Weight mass per filter tile and search offset: the weights of
 * orc_filter_f32 (same FP32 operations, so the same values), summed in
 * double over the voxels of each tile of tile[0] x tile[1] x tile[2] voxels
 * (tiles numbered x fastest, then y, then z).  mass has ntiles * nS entries
 * and is overwritten.  The reference has no such output; it labels the
 * learned offset proposal. */
ORC_API void orc_weight_mass_f32(const float *Ppad, const int shape[3], const int pad[3], const float *sigma,
                                 float beta, int nS, const long long *soff, int nP, const long long *poff,
                                 const float *w, int weight_mode, const int tile[3], double *mass)
{
    const int nx = shape[0], ny = shape[1], nz = shape[2];
    const long long nxp = nx + 2 * pad[0], nyp = ny + 2 * pad[1];
    const long long nslp = nxp * nyp;
    const int ntx = (nx + tile[0] - 1) / tile[0], nty = (ny + tile[1] - 1) / tile[1];
    const int ntz = (nz + tile[2] - 1) / tile[2];
    memset(mass, 0, sizeof(double) * (size_t)ntx * nty * ntz * nS);
    for (int z = 0; z < nz; z++) {
        const float s = sigma[z];
        const float m = -1.f / (s * s * beta);
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const long long i0 = (z + pad[2]) * nslp + (y + pad[1]) * nxp + (x + pad[0]);
                const long long t = ((long long)(z / tile[2]) * nty + y / tile[1]) * ntx + x / tile[0];
                double *mt = mass + t * nS;
                for (int k = 0; k < nS; k++) {
                    const long long i1 = i0 + soff[k];
                    float d = 0.f;
                    for (int p = 0; p < nP; p++) {
                        float dd = Ppad[i1 + poff[p]] - Ppad[i0 + poff[p]];
                        float t2 = dd * dd;
                        d = d + t2 * w[p];
                    }
                    d = d * m;
                    mt[k] += (double)(weight_mode == 0 ? pade_cut_f32(d) : expf(d));
                }
            }
    }
}

/* FP64 twin of orc_filter_f32: identical algorithm, every intermediate in
 * double (inputs and weights are the same float values). */
ORC_API void orc_filter_f64(const float *Rpad, const float *Ppad, const int shape[3], const int pad[3],
                            const double *sigma, double beta, int nS, const long long *soff, int nP,
                            const long long *poff, const float *w, int weight_mode, double *out,
                            int z_begin, int z_end, int nthreads)
{
    const int nx = shape[0], ny = shape[1];
    const long long nxp = nx + 2 * pad[0], nyp = ny + 2 * pad[1];
    const long long nslp = nxp * nyp;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int z = z_begin; z < z_end; z++) {
        const double s = sigma[z];
        const double m = -1.0 / (s * s * beta);
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const long long i0 = (z + pad[2]) * nslp + (y + pad[1]) * nxp + (x + pad[0]);
                const double r0 = Rpad[i0];
                double fv = 0, fw = 0, wmax = 0;
                for (int k = 0; k < nS; k++) {
                    const long long i1 = i0 + soff[k];
                    double d = 0;
                    for (int p = 0; p < nP; p++) {
                        double t = (double)Ppad[i1 + poff[p]] - (double)Ppad[i0 + poff[p]];
                        d += t * t * (double)w[p];
                    }
                    d *= m;
                    double wt = weight_mode == 0 ? pade_cut_f64(d) : exp(d);
                    fv += wt * (double)Rpad[i1];
                    fw += wt;
                    if (wt > wmax) wmax = wt;
                }
                out[((long long)z * ny + y) * nx + x] = wmax > 0 ? (fv + wmax * r0) / (fw + wmax) : r0;
            }
    }
}

/* uint16 re-quantisation of Geometry/hdcommunication.cpp:1473:
 * min(max(0, round(x)), 65535) with C round (half away from zero). */
ORC_API void orc_requant_u16(const float *in, long long n, uint16_t *out)
{
    for (long long i = 0; i < n; i++) {
        float v = roundf(in[i]);
        if (v < 0.f) v = 0.f;
        if (v > 65535.f) v = 65535.f;
        out[i] = (uint16_t)v;
    }
}

/* ------------------------------------------------------------------------ */
/* Noise estimator                                                           */
/* ------------------------------------------------------------------------ */

/* Sample positions.  The reference draws them with rand() seeded from the
 * clock inside an OpenMP loop (Noise/noiselevel.h:24,
 * noiselevel.cpp:192-234), which no test can reproduce.  This restatement
 * keeps the distribution -- x in [0,nx-ps), y in [0,ny-ps), z in [0,nz),
 * sorted linear indices -- and draws from SplitMix64(seed) so that a sample
 * list is reproducible; the shipped library uses the same generator so both
 * sides see the same positions. */
static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static int cmp_ll(const void *a, const void *b)
{
    long long x = *(const long long *)a, y = *(const long long *)b;
    return (x > y) - (x < y);
}
ORC_API int orc_draw_samples(const int shape[3], long long n, int patchsize, uint64_t seed, long long *out)
{
    const int nx = shape[0], ny = shape[1], nz = shape[2];
    if (nx <= patchsize || ny <= patchsize || nz < 1) return -1;
    uint64_t st = seed;
    for (long long i = 0; i < n; i++) {
        int x = (int)(splitmix64(&st) % (uint64_t)(nx - patchsize));
        int y = (int)(splitmix64(&st) % (uint64_t)(ny - patchsize));
        int z = (int)(splitmix64(&st) % (uint64_t)nz);
        out[i] = (long long)z * nx * ny + (long long)y * nx + x;
    }
    qsort(out, (size_t)n, sizeof(long long), cmp_ll);
    return 0;
}

/* Circular field of view (noiselevel.cpp:204-224): a sample corner (x, y)
 * is redrawn while (x - nx/2)^2 + (y - ny/2)^2 > r^2, r = diameter/2 -
 * sqrt(2 ps^2), the square taken in float.  The reference would loop for
 * ever when no corner of the sampling range is inside; here a brute-force
 * scan of the range returns -1 for that case. */
ORC_API int orc_draw_samples_mask(const int shape[3], long long n, int patchsize, int mask_diameter, uint64_t seed,
                                  long long *out)
{
    if (mask_diameter <= 0) return orc_draw_samples(shape, n, patchsize, seed, out);
    const int nx = shape[0], ny = shape[1], nz = shape[2];
    if (nx <= patchsize || ny <= patchsize || nz < 1) return -1;
    float maxr = (float)((double)((float)mask_diameter / 2.f) - sqrt((double)(2 * patchsize * patchsize)));
    maxr = maxr * maxr;
    const int cx = nx / 2, cy = ny / 2;
    int feasible = 0;
    for (int y = 0; y < ny - patchsize && !feasible; y++)
        for (int x = 0; x < nx - patchsize; x++)
            if ((float)((x - cx) * (x - cx) + (y - cy) * (y - cy)) <= maxr) {
                feasible = 1;
                break;
            }
    if (!feasible) return -1;
    uint64_t st = seed;
    for (long long i = 0; i < n; i++) {
        int x = (int)(splitmix64(&st) % (uint64_t)(nx - patchsize));
        int y = (int)(splitmix64(&st) % (uint64_t)(ny - patchsize));
        int z = (int)(splitmix64(&st) % (uint64_t)nz);
        float r = (float)((x - cx) * (x - cx) + (y - cy) * (y - cy));
        while (r > maxr) {
            x = (int)(splitmix64(&st) % (uint64_t)(nx - patchsize));
            y = (int)(splitmix64(&st) % (uint64_t)(ny - patchsize));
            r = (float)((x - cx) * (x - cx) + (y - cy) * (y - cy));
        }
        out[i] = (long long)z * nx * ny + (long long)y * nx + x;
    }
    qsort(out, (size_t)n, sizeof(long long), cmp_ll);
    return 0;
}

/* Per-sample 2D patch standard deviation: float mean, then float population
 * variance, then sqrtf (noiselevel.cpp:235-283). */
ORC_API void orc_patchstd_f32(const float *vol, const int shape[3], const long long *samples, long long n,
                              int patchsize, float *out)
{
    const int nx = shape[0];
    const long long nslice = (long long)shape[0] * shape[1];
    const int plen = patchsize * patchsize;
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < n; k++) {
        long long idx = samples[k];
        int z0 = (int)(idx / nslice);
        int y0 = (int)((idx - z0 * nslice) / nx);
        int x0 = (int)(idx - z0 * nslice - (long long)y0 * nx);
        float sum = 0;
        for (int y = y0; y < y0 + patchsize; y++) {
            long long p0 = z0 * nslice + (long long)y * nx + x0;
            for (int x = 0; x < patchsize; x++) sum = sum + vol[p0 + x];
        }
        float mean = sum / (float)plen;
        float sq = 0.f;
        for (int y = y0; y < y0 + patchsize; y++) {
            long long p0 = z0 * nslice + (long long)y * nx + x0;
            for (int x = 0; x < patchsize; x++) {
                float v = vol[p0 + x];
                float dv = v - mean;
                sq = sq + dv * dv;
            }
        }
        float var = sq / (float)plen;
        out[k] = sqrtf(var);
    }
}
/* FP64 twin: two-pass mean/variance in double. */
ORC_API void orc_patchstd_f64(const float *vol, const int shape[3], const long long *samples, long long n,
                              int patchsize, double *out)
{
    const int nx = shape[0];
    const long long nslice = (long long)shape[0] * shape[1];
    const int plen = patchsize * patchsize;
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < n; k++) {
        long long idx = samples[k];
        int z0 = (int)(idx / nslice);
        int y0 = (int)((idx - z0 * nslice) / nx);
        int x0 = (int)(idx - z0 * nslice - (long long)y0 * nx);
        double sum = 0;
        for (int y = y0; y < y0 + patchsize; y++)
            for (int x = 0; x < patchsize; x++) sum += vol[z0 * nslice + (long long)y * nx + x0 + x];
        double mean = sum / plen, sq = 0;
        for (int y = y0; y < y0 + patchsize; y++)
            for (int x = 0; x < patchsize; x++) {
                double dv = vol[z0 * nslice + (long long)y * nx + x0 + x] - mean;
                sq += dv * dv;
            }
        out[k] = sqrt(sq / plen);
    }
}

/* Two-cluster hard EM over values[fs..ls] (inclusive), zeros excluded.
 * Restates noiselevel.cpp:366-528 with its float/double mix: initial means
 * at min+1/4 range and max-1/4 range, nearest-mean assignment, then up to
 * 1000 rounds of Gaussian-likelihood reassignment.  Returns the lower
 * cluster's (mean, std). */
ORC_API void orc_create2clusters_f32(const float *values, int n_samples, int fs, int ls, float *mean_out,
                                     float *std_out)
{
    if (fs < 0) fs = 0;
    if (ls < 0) ls = n_samples - 1;
    float minimum = 1e20f, maximum = -1e20f;
    for (int i = fs; i <= ls; i++) {
        float v = values[i];
        if (v < minimum) minimum = v;
        if (v > maximum) maximum = v;
    }
    float mean1 = (float)((double)minimum + 0.25 * (double)(maximum - minimum));
    float mean2 = (float)((double)maximum - 0.25 * (double)(maximum - minimum));
    int *asg = (int *)malloc(sizeof(int) * (size_t)(n_samples > 0 ? n_samples : 1));
    float cs1 = 0.f, cs2 = 0.f, s1 = 0.f, s2 = 0.f;
    for (int i = fs; i <= ls; i++) {
        float v = values[i];
        if (v == 0.0f)
            asg[i] = -1;
        else if (fabsf(v - mean1) < fabsf(v - mean2)) {
            cs1 = cs1 + v; s1 = s1 + 1.f; asg[i] = 0;
        } else {
            cs2 = cs2 + v; s2 = s2 + 1.f; asg[i] = 1;
        }
    }
    if (s1 > 0.f) mean1 = cs1 / s1;
    if (s2 > 0.f) mean2 = cs2 / s2;
    if (s1 == 0 && s2 == 0) {
        free(asg);
        *mean_out = 0.f; *std_out = 0.f;
        return;
    }
    float var1 = 0.f, var2 = 0.f;
    for (int i = fs; i <= ls; i++) {
        float v = values[i];
        float a = (float)asg[i];
        if (a == 0) var1 = var1 + (v - mean1) * (v - mean1);
        else if (a == 1) var2 = var2 + (v - mean2) * (v - mean2);
    }
    var1 = var1 / (s1 > 1.f ? s1 : 1.f);
    var2 = var2 / (s2 > 1.f ? s2 : 1.f);
    const float PI = 3.141592654f;
    int change = 1;
    uint64_t iters = 0;
    while (change && var1 > 0.0f && var2 > 0.0f) {
        change = 0;
        iters++;
        cs1 = cs2 = s1 = s2 = 0.f;
        for (int i = fs; i <= ls; i++) {
            float v = values[i];
            if (v == 0.0f) continue;
            float e1 = -(((v - mean1) * (v - mean1)) / (2.f * var1));
            float e2 = -(((v - mean2) * (v - mean2)) / (2.f * var2));
            float p1 = (float)(1. / (double)sqrtf(var1 * 2.f * PI) * (double)expf(e1));
            float p2 = (float)(1. / (double)sqrtf(var2 * 2.f * PI) * (double)expf(e2));
            if (p2 > p1) {
                if (asg[i] != 1) { change = 1; asg[i] = 1; }
                cs2 = cs2 + v; s2 = s2 + 1.f;
            } else {
                if (asg[i] != 0) { change = 1; asg[i] = 0; }
                cs1 = cs1 + v; s1 = s1 + 1.f;
            }
        }
        if (s1 > 0.f) mean1 = cs1 / s1;
        if (s2 > 0.f) mean2 = cs2 / s2;
        var1 = 0.f; var2 = 0.f;
        for (int i = fs; i <= ls; i++) {
            float v = values[i];
            float a = (float)asg[i];
            if (a == 0) var1 = var1 + (v - mean1) * (v - mean1);
            else if (a == 1) var2 = var2 + (v - mean2) * (v - mean2);
        }
        var1 = var1 / (s1 > 1.f ? s1 : 1.f);
        var2 = var2 / (s2 > 1.f ? s2 : 1.f);
        if (iters == 1000) break;
    }
    free(asg);
    if (mean1 <= mean2 || var2 <= 0.0) {
        *mean_out = mean1; *std_out = (float)sqrt((double)var1);
    } else {
        *mean_out = mean2; *std_out = (float)sqrt((double)var2);
    }
}

/* FP64 twin of orc_create2clusters_f32 (same decisions, double arithmetic). */
ORC_API void orc_create2clusters_f64(const double *values, int n_samples, int fs, int ls, double *mean_out,
                                     double *std_out)
{
    if (fs < 0) fs = 0;
    if (ls < 0) ls = n_samples - 1;
    double mn = 1e300, mx = -1e300;
    for (int i = fs; i <= ls; i++) {
        if (values[i] < mn) mn = values[i];
        if (values[i] > mx) mx = values[i];
    }
    double m1 = mn + 0.25 * (mx - mn), m2 = mx - 0.25 * (mx - mn);
    int *asg = (int *)malloc(sizeof(int) * (size_t)(n_samples > 0 ? n_samples : 1));
    double c1 = 0, c2 = 0, s1 = 0, s2 = 0;
    for (int i = fs; i <= ls; i++) {
        double v = values[i];
        if (v == 0.0) asg[i] = -1;
        else if (fabs(v - m1) < fabs(v - m2)) { c1 += v; s1++; asg[i] = 0; }
        else { c2 += v; s2++; asg[i] = 1; }
    }
    if (s1 > 0) m1 = c1 / s1;
    if (s2 > 0) m2 = c2 / s2;
    if (s1 == 0 && s2 == 0) { free(asg); *mean_out = 0; *std_out = 0; return; }
    double v1 = 0, v2 = 0;
    for (int i = fs; i <= ls; i++) {
        if (asg[i] == 0) v1 += (values[i] - m1) * (values[i] - m1);
        else if (asg[i] == 1) v2 += (values[i] - m2) * (values[i] - m2);
    }
    v1 /= (s1 > 1 ? s1 : 1);
    v2 /= (s2 > 1 ? s2 : 1);
    const double PI = 3.14159265358979323846;
    int change = 1, iters = 0;
    while (change && v1 > 0 && v2 > 0) {
        change = 0;
        iters++;
        c1 = c2 = s1 = s2 = 0;
        for (int i = fs; i <= ls; i++) {
            double v = values[i];
            if (v == 0.0) continue;
            double p1 = 1. / sqrt(v1 * 2 * PI) * exp(-((v - m1) * (v - m1) / (2 * v1)));
            double p2 = 1. / sqrt(v2 * 2 * PI) * exp(-((v - m2) * (v - m2) / (2 * v2)));
            if (p2 > p1) { if (asg[i] != 1) { change = 1; asg[i] = 1; } c2 += v; s2++; }
            else { if (asg[i] != 0) { change = 1; asg[i] = 0; } c1 += v; s1++; }
        }
        if (s1 > 0) m1 = c1 / s1;
        if (s2 > 0) m2 = c2 / s2;
        v1 = v2 = 0;
        for (int i = fs; i <= ls; i++) {
            if (asg[i] == 0) v1 += (values[i] - m1) * (values[i] - m1);
            else if (asg[i] == 1) v2 += (values[i] - m2) * (values[i] - m2);
        }
        v1 /= (s1 > 1 ? s1 : 1);
        v2 /= (s2 > 1 ? s2 : 1);
        if (iters == 1000) break;
    }
    free(asg);
    if (m1 <= m2 || v2 <= 0.0) { *mean_out = m1; *std_out = sqrt(v1); }
    else { *mean_out = m2; *std_out = sqrt(v2); }
}

/* z-adaptive windows (noiselevel.cpp:529-606): per slice pos, samples with
 * slice index in (pos - w/2, pos + w/2] plus the first sample past the window
 * (the reference's endpos stops on it, :546); empty slices (mean 0) are
 * filled from their neighbours until none is left.  out[3*z] = {z, mean,
 * std}.  use_f64 selects the FP64 clustering on double copies. */
ORC_API void orc_zwindow(const long long *samples, const float *values_f32, const double *values_f64, long long n,
                         const int shape[3], int windowsize, int use_f64, double *out)
{
    const long long nslice = (long long)shape[0] * shape[1];
    const int nz = shape[2];
    int *st = (int *)malloc(sizeof(int) * nz), *en = (int *)malloc(sizeof(int) * nz);
    int startpos = 0, endpos = 0;
    for (int pos = 0; pos < nz; pos++) {
        while (startpos < n - 1 && samples[startpos] / nslice <= (pos - windowsize / 2)) startpos++;
        while ((endpos < (int)(n - 1)) && (samples[endpos] / nslice <= (pos + windowsize / 2))) endpos++;
        st[pos] = startpos;
        en[pos] = endpos;
    }
    double max_std = -1.0;
    int zero_slice = 0;
    for (int pos = 0; pos < nz; pos++) {
        double mean, sd;
        if (use_f64) orc_create2clusters_f64(values_f64, (int)n, st[pos], en[pos], &mean, &sd);
        else {
            float mf, sf;
            orc_create2clusters_f32(values_f32, (int)n, st[pos], en[pos], &mf, &sf);
            mean = mf; sd = sf;
        }
        out[3 * pos] = pos; out[3 * pos + 1] = mean; out[3 * pos + 2] = sd;
        if (mean > max_std) max_std = mean;
        if (mean == 0.0) zero_slice = 1;
    }
    if (zero_slice && max_std > 0.0) {
        while (zero_slice) {
            zero_slice = 0;
            for (int i = 0; i < nz; i++) {
                double tv = out[3 * i + 1];
                if (tv == 0.0) {
                    double nv = 0, ns = 0;
                    int c = 0;
                    if (i > 0 && out[3 * (i - 1) + 1] > 0) { nv += out[3 * (i - 1) + 1]; ns += out[3 * (i - 1) + 2]; c++; }
                    if (i < nz - 1 && out[3 * (i + 1) + 1] > 0) { nv += out[3 * (i + 1) + 1]; ns += out[3 * (i + 1) + 2]; c++; }
                    if (c > 0) {
                        if (use_f64) { tv = nv / c; out[3 * i + 1] = tv; out[3 * i + 2] = ns / c; }
                        else {
                            /* float accumulation as in the reference */
                            float fnv = 0.f, fns = 0.f;
                            if (i > 0 && out[3 * (i - 1) + 1] > 0) { fnv = fnv + (float)out[3 * (i - 1) + 1]; fns = fns + (float)out[3 * (i - 1) + 2]; }
                            if (i < nz - 1 && out[3 * (i + 1) + 1] > 0) { fnv = fnv + (float)out[3 * (i + 1) + 1]; fns = fns + (float)out[3 * (i + 1) + 2]; }
                            tv = fnv / (float)c; out[3 * i + 1] = tv; out[3 * i + 2] = fns / (float)c;
                        }
                    }
                }
                if (tv == 0.0) zero_slice = 1;
            }
        }
    }
    free(st);
    free(en);
}
