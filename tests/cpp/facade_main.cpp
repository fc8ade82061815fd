// Drives denoise::IterativeNLM_GPU (include/iternlm_gpu.h) the way the
// reference's main.cpp does (configure_device, set_sigma, Run_GaussianNoise,
// get_result), reading raw float32 files written by the test and writing the
// two iterations' results.
//   facade_main <dir> nx ny nz rs rp blockwise
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
static void save(const std::string &path, const std::vector<float> &v)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(float), v.size(), f) != v.size()) std::exit(4);
    std::fclose(f);
}

int main(int argc, char **argv)
{
    if (argc != 8) return 2;
    const std::string dir = argv[1];
    int shape[3] = {std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4])};
    const int rs = std::atoi(argv[5]), rp = std::atoi(argv[6]);
    protocol::DenoiseParameters params;
    for (int i = 0; i < 3; i++) {
        params.radius_searchspace[i] = rs;
        params.radius_patchspace[i] = rp;
    }
    params.cpu.blockwise = std::atoi(argv[7]) != 0;
    const size_t n = (size_t)shape[0] * shape[1] * shape[2];
    std::vector<float> raw = load(dir + "/raw.bin", n);
    std::vector<float> s1 = load(dir + "/sigma1.bin", shape[2]);
    std::vector<float> s2 = load(dir + "/sigma2.bin", shape[2]);
    std::vector<float> out(n);

    denoise::IterativeNLM_GPU gpu;
    const int slices = gpu.configure_device(shape, &params);
    if (slices < shape[2]) return 5;
    gpu.set_sigma(s1.data(), shape);
    gpu.Run_GaussianNoise(1, raw.data(), shape, &params);
    gpu.get_result(out.data(), shape);
    save(dir + "/out1.bin", out);
    gpu.set_sigma(s2.data(), shape);
    gpu.Run_GaussianNoise(2, raw.data(), shape, &params);
    gpu.get_result(out.data(), shape);
    save(dir + "/out2.bin", out);
    gpu.free_device();
    return gpu.last_status() == INLM_OK ? 0 : 6;
}
