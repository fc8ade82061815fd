// Drives denoise::IterativeNLM_GPU (include/iternlm_gpu.h) through the GPU
// branch of the reference's main.cpp:311-545: configure_device, set_sigma,
// then per iteration Run_GaussianNoise when the volume fits
// (n_parallelslices >= shape[2]), the host-RAM Run_GaussianNoise_GPUBlocks
// when it does not, or the file-list Run_GaussianNoise_GPUBlocks with
// gpu.blockwise_host; `resumed` decides whether a result is written, as in
// main.cpp:379-389 and :462-482.  The noise estimate is not part of the
// class, so the two sigma lists come from files.  TIFF reading and writing
// stand in for hdcom (inlm_tiff_*).
//
//   facade_main <dir> nx ny nz rs rp blockwise blockwise_host memory_limit resume [precision]
//
// <dir>/raw.bin (float32 volume), <dir>/sigma1.bin, <dir>/sigma2.bin (nz
// floats each), <dir>/raw/raw<z>.tif (one file per slice, for the file-list
// driver).  Writes <dir>/out/denoise-iteration0N/denoised<z>.tif for N = 1,
// 2 and prints "n_parallelslices <n>" and "iteration <i> resumed <0|1>".
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "iternlm_gpu.h"

static std::vector<float> load(const std::string &path, size_t n)
{
    std::vector<float> v(n);
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(float), n, f) != n) {
        std::fprintf(stderr, "cannot read %s\n", path.c_str());
        std::exit(3);
    }
    std::fclose(f);
    return v;
}
static std::string zfill(int v, int w)
{
    char b[32];
    std::snprintf(b, sizeof b, "%0*d", w, v);
    return b;
}
static void save(const std::string &dir, const float *vol, int shape[3], int first)
{
    const int32_t sh[3] = {shape[0], shape[1], shape[2]};
    if (inlm_tiff_write_stack(dir.c_str(), "denoised", vol, sh, first, 0) != INLM_OK) {
        std::fprintf(stderr, "cannot write %s: %s\n", dir.c_str(), inlm_last_error(nullptr));
        std::exit(4);
    }
}

int main(int argc, char **argv)
{
    if (argc != 11 && argc != 12) return 2;
    const std::string dir = argv[1];
    int shape[3] = {std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4])};
    const int rs = std::atoi(argv[5]), rp = std::atoi(argv[6]);
    protocol::DenoiseParameters params;
    for (int i = 0; i < 3; i++) {
        params.radius_searchspace[i] = rs;
        params.radius_patchspace[i] = rp;
    }
    params.cpu.blockwise = std::atoi(argv[7]) != 0;
    params.gpu.blockwise_host = std::atoi(argv[8]) != 0;
    params.gpu.memory_limit = std::atoll(argv[9]);
    params.io.resume = std::atoi(argv[10]) != 0;
    if (argc == 12) params.gpu.precision = argv[11];
    params.io.firstslice = 0;
    params.io.lastslice = shape[2] - 1;
    const int nz = shape[2];
    const size_t n = (size_t)shape[0] * shape[1] * nz;
    float *instack = nullptr;
    if (!params.gpu.blockwise_host) {
        std::vector<float> raw = load(dir + "/raw.bin", n);
        instack = static_cast<float *>(std::malloc(n * sizeof(float)));
        std::copy(raw.begin(), raw.end(), instack);
    }
    std::vector<float> s0 = load(dir + "/sigma1.bin", nz);
    std::vector<float> s1 = load(dir + "/sigma2.bin", nz);
    std::vector<std::string> filelist, filelist_prev;
    for (int z = 0; z < nz; z++) filelist.push_back(dir + "/raw/raw" + zfill(z, 4) + ".tif");
    const std::string outpath = dir + "/out";
    float *output = nullptr;

    denoise::IterativeNLM_GPU gpu;
    const int maxpar = gpu.configure_device(shape, &params);
    if (maxpar < 0) return 5;
    std::printf("n_parallelslices %d\n", gpu.n_parallelslices);
    if (params.gpu.n_gpus > 0 && params.gpu.blockwise_host == false) params.cpu.blockwise = false;
    gpu.set_sigma(s0.data(), shape);
    for (int iter = 1; iter <= 2; iter++) {
        params.io.active_outpath = outpath + "/denoise-iteration" + zfill(iter, 2) + "/";
        if (iter == 2) gpu.set_sigma(s1.data(), shape);
        if (!params.gpu.blockwise_host && gpu.n_parallelslices >= shape[2]) {
            if (!output) output = static_cast<float *>(std::malloc(n * sizeof(float)));
            gpu.Run_GaussianNoise(iter, instack, shape, &params);
            if (!gpu.resumed) {
                gpu.get_result(output, shape);
                save(params.io.active_outpath, output, shape, params.io.firstslice);
            }
        } else if (!params.gpu.blockwise_host) {
            if (!output) output = static_cast<float *>(std::malloc(n * sizeof(float)));
            gpu.Run_GaussianNoise_GPUBlocks(iter, instack, output, shape, &params);
            if (!gpu.resumed) save(params.io.active_outpath, output, shape, params.io.firstslice);
        } else {
            gpu.Run_GaussianNoise_GPUBlocks(iter, filelist, iter == 1 ? filelist : filelist_prev, shape, &params);
            filelist_prev.clear();
            for (int z = 0; z < nz; z++) filelist_prev.push_back(params.io.active_outpath + "denoised" + zfill(z, 4) + ".tif");
        }
        std::printf("iteration %d resumed %d\n", iter, gpu.resumed ? 1 : 0);
        if (gpu.last_status() != INLM_OK) return 6;
    }
    gpu.free_device();
    std::free(instack);
    std::free(output);
    return 0;
}
