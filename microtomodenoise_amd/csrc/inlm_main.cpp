// inlm: the reference's command-line program (main.cpp:41-219) over the C ABI.
//
//   inlm <input folder or TIFF> [-o <output folder>] [options]
//
// Flags keep the reference's names and meanings; the run itself is
// inlm_denoise_files (noise estimate, iterations, output folders, resume,
// cleanup, log file, as main.cpp:231-591), or inlm_denoise_rgb_files with
// --rgb (the colour path, main.cpp:254-257).  Flags that select parts this
// build leaves out (several GPUs, CPU threading and unrolling) are accepted
// where the reference accepts them and reported.  Added flags: -precision, -weight, -seed, -device (same as
// -gpu0) and --print-params, which prints the parsed settings and exits
// without touching a GPU.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "inlm_capi.h"

namespace {

const char *kUsage =
    "usage: inlm <input folder or TIFF stack> [options]\n"
    "  data:     -i/-input PATH  -o/-output PATH  -r/-resume true|false  --resume\n"
    "            -fs/-firstslice N  -ls/-lastslice N (inclusive)  --16bit  --cleanup  -aniso/-anisotropy F\n"
    "            --rgb/--color (2D RGB images)  --independent  --nopoisson  --noaverage\n"
    "  denoise:  -maxiter/-iter/-iterations N  -slices/-nslices N  -a/-alpha F\n"
    "            -search N  -search0/1/2 N  -patch N  -patch0/1/2 N\n"
    "  noise:    -noise/-noisemode global|auto|z-adaptive|manual|semimanual  -sigma/-s F  -sigma0/-s0 F  -sigma1/-s1 F\n"
    "            -nsamples/-noise_samples N  -noisepatch N  -noiseshift F  -mask/-mask_diameter N\n"
    "            --continuous  --preserve_detail\n"
    "  hardware: -gpu0/-device N  -gpus N (one GPU is used)  --blocks  --noblocks  --verbose/--v\n"
    "  additions:  -precision fp32|fp32_fast|fp16|bf16|fp16_fp8  -weight pade|exp  -seed N  --print-params\n";

bool is(const char *a, const char *b) { return std::strcmp(a, b) == 0; }

int fail(const char *what, const char *arg)
{
    std::fprintf(stderr, "inlm: %s%s%s\n", what, arg ? ": " : "", arg ? arg : "");
    return 2;
}

const char *noise_name(int m)
{
    return m == INLM_NOISE_GLOBAL       ? "global"
           : m == INLM_NOISE_MANUAL     ? "manual"
           : m == INLM_NOISE_SEMIMANUAL ? "semimanual"
                                        : "z-adaptive";
}
const char *precision_name(int p)
{
    return p == INLM_PREC_FP32_FAST ? "fp32_fast"
           : p == INLM_PREC_FP16    ? "fp16"
           : p == INLM_PREC_BF16    ? "bf16"
           : p == INLM_PREC_FP16_FP8 ? "fp16_fp8"
                                     : "fp32";
}

}  // namespace

int main(int argc, char **argv)
{
    inlm_params p;
    inlm_default_params(&p);
    inlm_io_options o;
    inlm_default_io_options(&o);
    std::string inpath, outpath;
    int device = 0, ngpus = 1;
    bool print_params = false, rgb = false;
    inlm_rgb_options c;
    inlm_default_rgb_options(&c);

    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        // flags that take a value read it from the next argument
        auto next = [&](const char *flag) -> const char * {
            if (i + 1 >= argc) return nullptr;
            (void)flag;
            return argv[++i];
        };
        const char *v = nullptr;
#define NEED(flag)                                              \
    do {                                                        \
        v = next(flag);                                         \
        if (!v) return fail("missing value after", flag);       \
    } while (0)
        // data
        if (is(a, "-i") || is(a, "-input")) { NEED(a); inpath = v; }
        else if (is(a, "-o") || is(a, "-output")) { NEED(a); outpath = v; }
        else if (is(a, "-r") || is(a, "-resume")) {
            NEED(a);
            if (is(v, "false") || is(v, "False")) o.resume = 0;
            else if (is(v, "true") || is(v, "True")) o.resume = 1;
        }
        else if (is(a, "--resume")) o.resume = 1;
        else if (is(a, "-fs") || is(a, "-firstslice")) { NEED(a); o.firstslice = std::atoi(v); }
        else if (is(a, "-ls") || is(a, "-lastslice")) { NEED(a); o.lastslice = std::atoi(v); }
        else if (is(a, "--16bit")) o.save_16bit = 1;
        else if (is(a, "--cleanup")) o.cleanup = 1;
        else if (is(a, "--color") || is(a, "--rgb")) rgb = true;
        else if (is(a, "-aniso") || is(a, "-anisotropy")) { NEED(a); p.z_anisotropy = (float)std::atof(v); }
        // denoising
        else if (is(a, "-maxiter") || is(a, "-iter") || is(a, "-iterations")) { NEED(a); p.maxiterations = std::atoi(v); }
        else if (is(a, "-slices") || is(a, "-nslices")) { NEED(a); p.nslices = std::atoi(v); }
        else if (is(a, "-a") || is(a, "-alpha")) { NEED(a); p.alpha = (float)std::atof(v); }
        else if (is(a, "-search")) {
            NEED(a);
            p.radius_searchspace[0] = p.radius_searchspace[1] = p.radius_searchspace[2] = std::atoi(v);
        }
        else if (is(a, "-search0")) { NEED(a); p.radius_searchspace[0] = std::atoi(v); }
        else if (is(a, "-search1")) { NEED(a); p.radius_searchspace[1] = std::atoi(v); }
        else if (is(a, "-search2")) { NEED(a); p.radius_searchspace[2] = std::atoi(v); }
        else if (is(a, "-patch")) {
            NEED(a);
            p.radius_patchspace[0] = p.radius_patchspace[1] = p.radius_patchspace[2] = std::atoi(v);
        }
        else if (is(a, "-patch0")) { NEED(a); p.radius_patchspace[0] = std::atoi(v); }
        else if (is(a, "-patch1")) { NEED(a); p.radius_patchspace[1] = std::atoi(v); }
        else if (is(a, "-patch2")) { NEED(a); p.radius_patchspace[2] = std::atoi(v); }
        else if (is(a, "--independent")) c.independent_channels = 1;
        else if (is(a, "--nopoisson")) c.stabilize_variance = 0;
        else if (is(a, "--noaverage")) c.average_channelsigma = 0;
        // hardware
        else if (is(a, "-cpus") || is(a, "-n_cpus") || is(a, "-threads")) { NEED(a); }
        else if (is(a, "-gpus") || is(a, "-n_gpus") || is(a, "-ngpus")) { NEED(a); ngpus = std::atoi(v); }
        else if (is(a, "-gpu0") || is(a, "-device")) { NEED(a); device = std::atoi(v); }
        else if (is(a, "--noblocks")) p.cpu_blockwise = 0;
        else if (is(a, "--blocks")) p.cpu_blockwise = 1;
        else if (is(a, "--unroll")) {
        }
        else if (is(a, "--verbose") || is(a, "--v")) o.verbose = 1;
        // noise estimation
        else if (is(a, "-noise") || is(a, "-noisemode")) {
            NEED(a);
            if (is(v, "global") || is(v, "auto")) p.noise_mode = INLM_NOISE_GLOBAL;
            else if (is(v, "z-adaptive")) p.noise_mode = INLM_NOISE_ZADAPTIVE;
            else if (is(v, "manual")) p.noise_mode = INLM_NOISE_MANUAL;
            else if (is(v, "semimanual")) p.noise_mode = INLM_NOISE_SEMIMANUAL;
            // the reference prints an error and leaves sigma at zero (noiselevel.cpp:89-90)
            else return fail("unknown noise estimation mode", v);
        }
        else if (is(a, "-sigma") || is(a, "-s")) { NEED(a); p.sigma[0] = p.sigma[1] = (float)std::atof(v); }
        else if (is(a, "-sigma0") || is(a, "-s0")) { NEED(a); p.sigma[0] = (float)std::atof(v); }
        else if (is(a, "-sigma1") || is(a, "-s1")) { NEED(a); p.sigma[1] = (float)std::atof(v); }
        else if (is(a, "-nsamples") || is(a, "-noise_samples")) { NEED(a); p.n_samples = std::atoll(v); }
        else if (is(a, "-noisepatch")) { NEED(a); p.patchsize = std::atoi(v); }
        else if (is(a, "-noiseshift")) { NEED(a); p.stds_from_mean = (float)std::atof(v); }
        else if (is(a, "-mask") || is(a, "-mask_diameter")) { NEED(a); p.circular_mask_diameter = std::atoi(v); }
        else if (is(a, "--continuous")) p.continuous_estimate = 1;
        else if (is(a, "--preserve_detail")) {
            // the reference's suggested setting (main.cpp:195-202)
            p.continuous_estimate = 1;
            p.alpha = 0.f;
            p.stds_from_mean = 4.f;
            p.maxiterations = 6;
        }
        // flags of this library, not in the reference
        else if (is(a, "-precision")) {
            NEED(a);
            if (is(v, "fp32")) p.precision = INLM_PREC_FP32;
            else if (is(v, "fp32_fast")) p.precision = INLM_PREC_FP32_FAST;
            else if (is(v, "fp16")) p.precision = INLM_PREC_FP16;
            else if (is(v, "fp16_fp8")) p.precision = INLM_PREC_FP16_FP8;
            else if (is(v, "bf16")) p.precision = INLM_PREC_BF16;
            else return fail("unknown precision", v);
        }
        else if (is(a, "-weight")) {
            NEED(a);
            if (is(v, "pade")) p.weight_fn = INLM_WEIGHT_PADE;
            else if (is(v, "exp")) p.weight_fn = INLM_WEIGHT_EXP;
            else return fail("unknown weight function", v);
        }
        else if (is(a, "-seed")) { NEED(a); p.sample_seed = std::strtoull(v, nullptr, 10); }
        else if (is(a, "--print-params")) print_params = true;
        else if (is(a, "-h") || is(a, "-help") || is(a, "--help")) {
            std::fputs(kUsage, stdout);
            return 0;
        }
        // the first other argument is the input, as in the reference
        else if (inpath.empty()) inpath = a;
#undef NEED
    }

    if (print_params) {
        std::printf("input=%s\noutput=%s\n", inpath.c_str(), outpath.c_str());
        std::printf("maxiterations=%d\nnslices=%d\nalpha=%g\nbeta=%g\nz_anisotropy=%g\n", p.maxiterations,
                    p.nslices, p.alpha, p.beta, p.z_anisotropy);
        std::printf("search=%d,%d,%d\npatch=%d,%d,%d\n", p.radius_searchspace[0], p.radius_searchspace[1],
                    p.radius_searchspace[2], p.radius_patchspace[0], p.radius_patchspace[1],
                    p.radius_patchspace[2]);
        std::printf("noise=%s\nsigma=%g,%g\nn_samples=%lld\npatchsize=%d\nstds_from_mean=%g\ncontinuous=%d\n"
                    "mask_diameter=%d\n",
                    noise_name(p.noise_mode), p.sigma[0], p.sigma[1], (long long)p.n_samples, p.patchsize,
                    p.stds_from_mean, p.continuous_estimate, p.circular_mask_diameter);
        std::printf("cpu_blockwise=%d\nprecision=%s\nweight=%s\nseed=%llu\n", p.cpu_blockwise,
                    precision_name(p.precision), p.weight_fn == INLM_WEIGHT_EXP ? "exp" : "pade",
                    (unsigned long long)p.sample_seed);
        std::printf("resume=%d\nfirstslice=%d\nlastslice=%d\nsave_16bit=%d\ncleanup=%d\nverbose=%d\ndevice=%d\n",
                    o.resume, o.firstslice, o.lastslice, o.save_16bit, o.cleanup, o.verbose, device);
        std::printf("rgb=%d\nindependent_channels=%d\nstabilize_variance=%d\naverage_channelsigma=%d\n", rgb ? 1 : 0,
                    c.independent_channels, c.stabilize_variance, c.average_channelsigma);
        return 0;
    }
    if (inpath.empty()) {
        std::fputs(kUsage, stderr);
        return fail("no input given", nullptr);
    }
    if (ngpus > 1) std::fprintf(stderr, "inlm: %d GPUs requested; one GPU holds the volume, using device %d\n", ngpus, device);

    const inlm_status st = rgb ? inlm_denoise_rgb_files(&p, &c, device, inpath.c_str(), outpath.c_str(), &o)
                               : inlm_denoise_files(&p, device, inpath.c_str(), outpath.c_str(), &o);
    if (st != INLM_OK) {
        std::fprintf(stderr, "inlm: %s (status %d)\n", inlm_last_error(nullptr), (int)st);
        return st == INLM_NO_DEVICE ? 3 : 1;
    }
    return 0;
}
