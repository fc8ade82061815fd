// Launch wrappers of the iterNLM HIP kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace inlm {

// Offset entry of the search table: shift in voxels plus the linear shift in
// the padded layout.
struct SearchEntry {
    int dx, dy, dz;
    int lin;  // dz*nxp*nyp + dy*nxp + dx (fits in 32 bits for every supported shape)
};

struct FilterGeom {
    int nx, ny, nz;
    int nxp, nyp, nzp;
    int px, py, pz;
};

enum FilterForm { FORM_FP32 = 0, FORM_FP32_FAST = 1, FORM_BF16 = 2, FORM_FP16 = 3, FORM_FP16_FP8 = 4 };

constexpr int kMaxTaps = 128;
constexpr int kMfmaMaxPatterns = 16;

struct TapTable {
    float w[kMaxTaps];
    int lin[kMaxTaps];       // linear patch shifts in the padded layout
    int dxyz[3 * kMaxTaps];  // (dx,dy,dz) per tap
    int n;
};

// Reflect-pad an unpadded float volume into a padded float volume
// (mirror without repeating the edge, iternlm_prepare.cpp:6-48).
hipError_t launch_pad_reflect(const float *in, float *out, const FilterGeom &g, hipStream_t st);

// One filter pass.  R/P are padded volumes; out is unpadded; sigma has nz
// entries.  Returns hipErrorInvalidValue when no kernel covers the patch.
// rp is the (isotropic) patch radius served by a tiled kernel, or 0 to force
// the per-voxel kernel, which runs any tap table.  With mass != nullptr the
// tiled kernel also writes, per 32x8x8 tile and search offset, the sum of
// the weights over the tile's voxels ([tile][offset], tiles x-fastest).
hipError_t launch_filter(int form, int rp, int weight_fn, const float *R, const float *P, float *out,
                         const float *sigma, float beta, const SearchEntry *search, int n_search,
                         const TapTable &taps, const FilterGeom &g, hipStream_t st, float *mass = nullptr);

// Three-channel filter of the colour path (iternlm_cpu.cpp:932-1075) on
// padded FP32 channels; mult[c] = -1 / (sigma_c^2 beta).
hipError_t launch_filter_rgb(int weight_fn, const float *const R[3], const float *const P[3], float *const out[3],
                             const float mult[3], int independent, const SearchEntry *search, int n_search,
                             const TapTable &taps, const FilterGeom &g, hipStream_t st);

// Whether the tiled kernel for isotropic radius rp uses exactly the given tap
// order (dx,dy,dz per tap, centre first) and whether all taps it groups into
// one weight class carry the same weight.
bool tiled_taps_match(int rp, const int *dxyz, const float *w, int n);

// Exhaustive check of the parity kernel's division: counts weights that
// differ from IEEE division over the float bit patterns 0x80000000+[first,
// first+count).
hipError_t launch_selftest_div(unsigned first, unsigned count, unsigned long long *mismatch, unsigned *example,
                               hipStream_t st);

// Patch standard deviation per sample (one wavefront per sample).  Sample
// indices are linear in the unpadded volume (nx, nslice = nx*ny); the volume
// itself is addressed through `origin` (voxel 0,0,0) with the given row and
// slice pitches, so the padded original can be read in place.
hipError_t launch_patchstd(const float *origin, int64_t row_pitch, int64_t slice_pitch, const int64_t *samples,
                           int64_t n, int nx, int64_t nslice, int patchsize, double *out, hipStream_t st);

// Matrix-core forms (FP32_FAST, FP16, BF16).  Rs/Ps are normalised padded
// volumes in the form's storage type; raw is the unpadded original in float
// (returned unchanged where every weight vanishes).
bool mfma_taps_supported(int rp, const int *dxyz, const float *w, int n);
// elems: allocated elements of each padded volume from Rs/Ps on (bounds the
// buffer loads).  shifted: for the 16-bit forms, the distance in elements
// from each volume to its copy shifted by one element (copy[j] =
// volume[j + 1], written by launch_pad_normalise); those forms need an even
// row pitch g.nxp.  0 for FP32_FAST.
// min_lin: the smallest linear offset of the search table (<= 0).
// Per-tile offset lists for the matrix-core filter (32x8x8 tiles, x fastest):
// tile t visits the search entries lists[t*K .. t*K + counts[t]) when
// counts[t] < n_search, else the whole search.  Device pointers.
struct OffsetLists {
    const int *lists;
    const int *counts;
    int K;
    long ntiles;
};
hipError_t launch_filter_mfma(int form, int rp, int weight_fn, const void *Rs, const void *Ps, size_t elems,
                              long shifted, const float *raw, float *out, const float *sigma, float beta, float mscale,
                              double shift, double inv_scale, const SearchEntry *search,
                              const SearchEntry *search_full, int n_search, int min_lin, const TapTable &taps,
                              const FilterGeom &g, hipStream_t st, const OffsetLists *ol = nullptr,
                              bool row_form = false);
hipError_t launch_pad_normalise(int form, const float *in, void *out, long shifted, const FilterGeom &g, float shift,
                                float scale, hipStream_t st);
// FP64 sum and sum of squares of n floats into acc[0], acc[1]; acc holds
// kMomentDoubles doubles (the rest are per-workgroup partials), needs no
// clearing, and the result does not depend on scheduling.
constexpr int kMomentBlocks = 1024;
constexpr int kMomentDoubles = 2 + 2 * kMomentBlocks;

// The full-search order of the matrix-core forms is followed on the device by
// n_search + kStageMovesPad ints, one per offset and zero past the last: 0
// where the block kernel loads the staging chunks of a wavefront's own slot,
// s = 1 .. kStageMovesMax where the offset is the one before it moved by
// dx + 2 within one dz and dy (one dword per region row), s such moves after
// the last load.
constexpr int kStageMovesPad = 64;
constexpr int kStageMovesMax = 2;
hipError_t launch_moments(const float *in, long n, double *acc, hipStream_t st);
inline size_t form_elem_size(int form) { return form == FORM_FP32 || form == FORM_FP32_FAST ? 4 : 2; }

}  // namespace inlm
