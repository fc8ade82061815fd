// Host-side tables of the iterNLM filter: search offsets, patch taps and
// patch weights, plus the reflect-padding geometry.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace inlm {

struct Offset3 {
    int dx, dy, dz;
};

struct Geometry {
    int nx, ny, nz;    // unpadded
    int px, py, pz;    // symmetric padding per axis
    int nxp, nyp, nzp; // padded
    int64_t slice_p() const { return (int64_t)nxp * nyp; }
    int64_t voxels() const { return (int64_t)nx * ny * nz; }
    int64_t voxels_p() const { return slice_p() * nzp; }
};

// Padding of iternlm_gpu.cu:1454-1459: rs+rp in x/y, min(nslices/2, rs_z)+rp in z.
Geometry make_geometry(const int shape[3], const int rs[3], const int rp[3], int nslices);

// Search offsets in the reference z->y->x order (iternlm_prepare.cpp:102-151).
std::vector<Offset3> search_offsets(const int rs[3], int nslices);
// Patch taps, centre first (iternlm_prepare.cpp:152-203).
std::vector<Offset3> patch_offsets(const int rp[3], int nslices);
// Computed weights 1/(2d+1)^2 (iternlm_prepare.cpp:204-263).
std::vector<float> computed_weights(const int rp[3], int nslices, float z_aniso, int nz_padded);
// The weight table the reference CPU path applies (literal tables of the
// specialised kernels, iternlm_cpu_kernels.cpp, or computed weights of the
// generic kernel).  Returns false and sets err for dispatches whose
// reference behaviour is undefined.
bool cpu_weight_table(const int rp[3], int nslices, float z_aniso, int nz_padded, bool blockwise,
                      std::vector<float> &out, std::string &err);

// SplitMix64 sample positions with the distribution of
// NoiseLevel::drawsamples (noiselevel.cpp:192-234), sorted; mask_diameter > 0
// restricts the sample corners to the circular field of view of
// noiselevel.cpp:204-224.  False when the volume is smaller than the patch or
// no corner of the sampling range lies inside the circle.
bool draw_samples(const int shape[3], int64_t n, int patchsize, uint64_t seed, int64_t *out, int mask_diameter = 0);

// Two-cluster hard EM (noiselevel.cpp:366-528) in FP64 over values[fs..ls].
void create2clusters(const double *values, int n, int fs, int ls, double &mean, double &std);
// z-adaptive windows (noiselevel.cpp:529-606) in FP64: out[z] = {mean, std}.
void zwindow_clusters(const int64_t *samples, const double *values, int64_t n, const int shape[3], int windowsize,
                      std::vector<double> &mean, std::vector<double> &std);

}  // namespace inlm
