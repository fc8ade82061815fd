/*
 * denoise_parameters.h -- parameter surface of the iterNLM filter.
 *
 * Restates protocol::DenoiseParameters of the reference
 * (iterNLM_v0.3/denoise_parameters.h:9-74) with the same namespace, field
 * names, nesting and defaults, so code written against the reference compiles
 * unchanged against these headers.  Four fields are additions of this
 * library and sit in the GPU block: the arithmetic form, the weight
 * function, the sample seed of the noise estimator and a device memory
 * limit.
 */
#ifndef INLM_DENOISE_PARAMETERS_H
#define INLM_DENOISE_PARAMETERS_H

#include <cstdint>
#include <string>

namespace protocol
{
struct DenoiseParameters {
    float alpha = 0.5f;
    int maxiterations = 4;

    int nslices = 11;                          // adjacent slices the search space may reach
    int radius_searchspace[3] = {10, 10, 10};  // ellipsoidal search space
    int radius_patchspace[3] = {1, 1, 1};      // ellipsoidal patch
    float beta = 1.f;                          // scales the filter strength

    float z_anisotropy = 1.f;

    struct NoiseLevel {
        std::string mode = "z-adaptive";  // "z-adaptive", "global" ("auto"), "manual", "semimanual"
        int zwindow = 100;
        float sigma[2] = {0.0f, 0.0f};    // manual mode
        uint64_t n_samples = 100000;
        int patchsize = 15;
        float stds_from_mean = 1;
        bool continuous_estimate = false;
        int circular_mask_diameter = 0;
    };
    struct CPU {
        int max_threads = 128;
        bool unrolled_patchspace = false;
        bool blockwise = true;  // also selects the reference CPU weight table the GPU reproduces
    };
    struct GPU {
        int n_gpus = 2;
        int deviceID = 0;
        int threadsPerBlock = 128;
        int memory_buffer = 0;
        bool blockwise_host = false;
        // build-specific
        std::string precision = "fp32";  // "fp32" (parity), "fp32_fast", "bf16", "fp16", "fp16_fp8"
        std::string weight_fn = "pade";  // "pade" (reference CPU) or "exp" (reference GPU)
        uint64_t sample_seed = 0;
        // device bytes configure_device may plan for (0: the free device
        // memory); below the volume's footprint it returns fewer slices than
        // shape[2] and the caller takes the block path
        int64_t memory_limit = 0;
    };
    struct IO {
        bool resume = false;
        int firstslice = -1;
        int lastslice = -1;
        std::string save_type = "32bit";
        bool cleanup = false;
        bool rgb = false;
        std::string active_outpath = "";
    };
    struct Color {
        bool stabilize_variance = true;
        bool independent_channels = false;
        bool average_channelsigma = true;
    };

    NoiseLevel noiselevel;
    CPU cpu;
    GPU gpu;
    IO io;
    Color color;
};
}  // namespace protocol

#endif  // INLM_DENOISE_PARAMETERS_H
