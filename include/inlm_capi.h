/*
 * inlm_capi.h -- C ABI of the MI355X-native iterNLM hot path.
 *
 * This is the drop-in boundary for the reference's GPU denoiser class
 * denoise::IterativeNLM_GPU (Denoiser/iternlm_gpu.h:20-56 in
 * iternlm/microtomodenoise iterNLM_v0.3).  Each entry point names the
 * reference member it replaces.  Plain C types only: no C++ and no torch
 * types cross this boundary.
 *
 * Conventions
 *   - Volumes are float32, x fastest: idx = z*nx*ny + y*nx + x, with
 *     shape = {nx, ny, nz} (the reference's shape[0] = TIFF width).
 *   - Host buffers belong to the caller; a handle owns all device state.
 *   - A handle is single-threaded.  Different handles are independent: there
 *     is no global mutable state, all constants travel as kernel arguments
 *     (the reference keeps them in __constant__ globals, which made it
 *     non-reentrant, gpu_constants.cuh:13-22).
 *   - Every call is synchronous at return; work runs on one HIP stream of
 *     default priority owned by the handle.
 *   - Errors are status codes; inlm_last_error() gives the message.
 */
#ifndef INLM_CAPI_H
#define INLM_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define INLM_API __attribute__((visibility("default")))

/* Status codes.  NO_DEVICE mirrors configure_device's -1 (no GPU, the caller
 * falls back to the CPU path, main.cpp:317) and HIP_ERROR its -2
 * (iternlm_gpu.cu:1608-1618). */
typedef enum {
    INLM_OK = 0,
    INLM_NO_DEVICE = -1,
    INLM_HIP_ERROR = -2,
    INLM_BAD_ARG = -3,
    INLM_OOM = -4,
    INLM_IO_ERROR = -5 /* a file could not be read, parsed or written */
} inlm_status;

/* NoiseLevel.mode of denoise_parameters.h:23 ("global"/"auto",
 * "z-adaptive", "manual", "semimanual"; noiselevel.cpp:14-83).  Semimanual:
 * the first estimate is sigma[0] on every slice and picks the sample whose
 * patch std lies closest to it; later estimates are that sample's patch std
 * on the current result (noiselevel.cpp:56-79). */
typedef enum {
    INLM_NOISE_GLOBAL = 0,
    INLM_NOISE_ZADAPTIVE = 1,
    INLM_NOISE_MANUAL = 2,
    INLM_NOISE_SEMIMANUAL = 3
} inlm_noise_mode;

/* Arithmetic form of the filter kernel.
 *   FP32      -- parity form: operation-faithful to the reference CPU path
 *                (tap order, unfused multiply/add, IEEE division), bit-exact
 *                against it.
 *   FP32_FAST -- FP32 storage, fused multiply-add stencil and reciprocal-based
 *                Pade; differs from FP32 by rounding only.
 *   BF16/FP16 -- R and P stored in 16 bits after the affine map
 *                (v - shift) * scale; squared differences in FP16 (BF16
 *                values are widened to FP16 first), patch sums, weight sums
 *                and accumulators in FP32, normalisation in FP64.
 *   FP16_FP8  -- FP16 storage as FP16, squared-difference planes in FP8
 *                e4m3 scaled per tile by a power of two, patch weights in
 *                FP8 times a common factor; distance sums in FP32. */
typedef enum {
    INLM_PREC_FP32 = 0,
    INLM_PREC_FP32_FAST = 1,
    INLM_PREC_BF16 = 2,
    INLM_PREC_FP16 = 3,
    INLM_PREC_FP16_FP8 = 4
} inlm_precision;

/* Weight function: the reference CPU's Pade [3/3] with the -3.56648 cutoff
 * (Denoiser/iternlm_cpu.h:125-126) or the reference GPU's expf
 * (iternlm_gpu.cu:150). */
typedef enum { INLM_WEIGHT_PADE = 0, INLM_WEIGHT_EXP = 1 } inlm_weight_fn;

/* POD mirror of protocol::DenoiseParameters (denoise_parameters.h:9-74),
 * restricted to the fields the hot path reads, plus this library's own
 * precision / weight-function / sampler-seed choices. */
typedef struct inlm_params {
    float alpha;                   /* 0.5 */
    int32_t maxiterations;         /* 4 */
    int32_t nslices;               /* 11 */
    int32_t radius_searchspace[3]; /* 10,10,10 */
    int32_t radius_patchspace[3];  /* 1,1,1 */
    float beta;                    /* 1 */
    float z_anisotropy;            /* 1 */
    /* NoiseLevel */
    int32_t noise_mode;            /* inlm_noise_mode, z-adaptive */
    int32_t zwindow;               /* 100 */
    float sigma[2];                /* manual mode */
    int64_t n_samples;             /* 1e5 */
    int32_t patchsize;             /* 15 */
    float stds_from_mean;          /* 1 */
    int32_t continuous_estimate;   /* 0 */
    uint64_t sample_seed;          /* seeds the sample positions (the reference seeds from the clock) */
    /* CPU.blockwise: selects which reference CPU dispatch the weight table
     * follows (iternlm_cpu.cpp:97-114 vs :529-548). */
    int32_t cpu_blockwise;         /* 1 */
    /* build-specific */
    int32_t precision;             /* inlm_precision, FP32 */
    int32_t weight_fn;             /* inlm_weight_fn, PADE */
    /* FP32 form: 0 the tiled kernel where one covers the patch, 1 the
     * per-voxel kernel.  Matrix-core forms: 0 the fastest kernel for the
     * case (at patch radius 2 with 16-bit storage and FP16 planes, the
     * 4x4-block kernel), 2 the row kernel. */
    int32_t kernel_select;
    /* NoiseLevel.circular_mask_diameter (denoise_parameters.h:33): when > 0
     * noise samples are drawn only inside that circle around the slice
     * centre (noiselevel.cpp:204-224) */
    int32_t circular_mask_diameter; /* 0 */
    int32_t reserved[6];
} inlm_params;

typedef struct inlm_handle inlm_handle;

/* Fills the reference defaults (denoise_parameters.h:11-48). */
INLM_API void inlm_default_params(inlm_params *p);

/* Number of visible GPUs (0 when none). */
INLM_API int32_t inlm_device_count(void);

/* Replaces IterativeNLM_GPU::configure_device (iternlm_gpu.cu:1440-1623):
 * builds the search/patch tables on the padded shape, allocates the padded
 * original, the padded previous iterate and the result on `device`.
 * Returns NO_DEVICE when no GPU is visible. */
INLM_API inlm_status inlm_create(const inlm_params *p, const int32_t shape[3], int32_t device, inlm_handle **out);

/* Replaces IterativeNLM_GPU::free_device (iternlm_gpu.cu:1625-1649). */
INLM_API inlm_status inlm_destroy(inlm_handle *h);

/* Last error message of a handle (or of the last failed inlm_create when h
 * is NULL). */
INLM_API const char *inlm_last_error(const inlm_handle *h);

/* Replaces IterativeNLM_GPU::set_sigma (iternlm_gpu.cu:2341-2355): one
 * filter strength per slice (nz values). */
INLM_API inlm_status inlm_set_sigma(inlm_handle *h, const float *sigma, int32_t nz);

/* Loads the original noisy volume (nx*ny*nz floats) from host memory or, with
 * is_device != 0, from device memory on the handle's GPU; reflect-pads it
 * into the resident original and makes it the previous iterate.  This is
 * the first half of prepare_iteration1 (iternlm_gpu.cu:2117-2152). */
INLM_API inlm_status inlm_load_raw(inlm_handle *h, const float *raw, int32_t is_device);

/* Replaces IterativeNLM_GPU::Run_GaussianNoise (iternlm_gpu.cu:1651-1728):
 * iteration `iter` (1-based).  At iter == 1 a non-NULL host `raw` is loaded
 * first (as the reference uploads instack); at iter > 1 the previous result
 * is reflect-padded into the previous iterate (prepare_nextiteration,
 * iternlm_gpu.cu:2230-2243).  Then one filter pass writes the result. */
INLM_API inlm_status inlm_run_iteration(inlm_handle *h, int32_t iter, const float *raw);

/* Replaces IterativeNLM_GPU::get_result (iternlm_gpu.cu:2356-2368): copies
 * the last result (nx*ny*nz floats) to host memory, or to device memory when
 * is_device != 0. */
INLM_API inlm_status inlm_get_result(inlm_handle *h, float *out, int32_t is_device);

/* Sample positions for the noise estimator: the distribution of
 * NoiseLevel::drawsamples (Noise/noiselevel.cpp:192-234) drawn from a
 * SplitMix64 stream seeded with `seed`, sorted.  Host-only; needs no GPU. */
INLM_API inlm_status inlm_draw_samples(const int32_t shape[3], int64_t n, int32_t patchsize, uint64_t seed,
                                       int64_t *out);

/* inlm_draw_samples restricted to a circular field of view
 * (noiselevel.cpp:204-224): with mask_diameter > 0, a sample's corner (x, y)
 * is redrawn until (x - nx/2)^2 + (y - ny/2)^2 <= (mask_diameter/2 -
 * sqrt(2 patchsize^2))^2 (integer square, float threshold, as the
 * reference).  BAD_ARG when no position of the sampling range satisfies it.
 * mask_diameter <= 0 is inlm_draw_samples.  Host-only. */
INLM_API inlm_status inlm_draw_samples_mask(const int32_t shape[3], int64_t n, int32_t patchsize,
                                            int32_t mask_diameter, uint64_t seed, int64_t *out);

/* Per-sample 15x15 (patchsize^2) in-slice standard deviation on the GPU
 * (NoiseLevel::get_patchstd, noiselevel.cpp:235-283), one wavefront per
 * sample, FP64 accumulation.  which = 0: the original volume; 1: the last
 * result. */
INLM_API inlm_status inlm_patchstd(inlm_handle *h, int32_t which, const int64_t *samples, int64_t n, double *out);

/* No counterpart in the reference.  inlm_patchstd for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc): no handle, no hipMalloc/hipFree.  `vol`
 * is a device pointer to an unpadded float volume of `shape` (x fastest);
 * `samples` (host, n patch corners as for inlm_patchstd) and `out` (host, n
 * doubles) are the caller's.  The sample list and the per-sample results
 * are staged in a fixed device array of this library of
 * INLM_NOALLOC_MAX_SAMPLES * 16 bytes (4 MiB), which each GPU the call runs
 * on holds until the process ends; larger n is rejected with INLM_BAD_ARG.
 * Every call on one device shares that array (each call overwrites it, so
 * nothing needs clearing): calls that overlap on one GPU race.  Runs on the
 * null stream and returns when `out` is written.  Same kernel, same values
 * as inlm_patchstd. */
#define INLM_NOALLOC_MAX_SAMPLES 262144
INLM_API inlm_status inlm_patchstd_noalloc(const float *vol, const int32_t shape[3], int32_t patchsize,
                                           const int64_t *samples, int64_t n, double *out, int32_t device);

/* Noise level per slice (NoiseLevel::get_noiselevel, noiselevel.cpp:9-95):
 * patch std on the GPU, two-cluster EM in FP64 on the host, global or
 * z-adaptive, sigma = mean + stds_from_mean * std.  Writes nz floats. */
INLM_API inlm_status inlm_estimate_noise(inlm_handle *h, int32_t which, const int64_t *samples, int64_t n,
                                         float *sigma_out);

/* The host half of get_noiselevel (noiselevel.cpp:9-95) on per-sample patch
 * standard deviations `stds` (n values for the sorted sample list): FP64
 * two-cluster EM, global or z-adaptive (with the reference's window and its
 * extrapolation into slices without valid samples, noiselevel.cpp:529-606),
 * manual, or semimanual.  which = 0 is the first estimate of a run, 1 a
 * later one.  Semimanual keeps its reference sample in *semimanual_ref:
 * which = 0 writes it (the sample closest to sigma[0]), which = 1 reads it.
 * Writes nz floats.  Host-only. */
INLM_API inlm_status inlm_sigma_from_patchstd(const inlm_params *p, const int32_t shape[3], const int64_t *samples,
                                              const double *stds, int64_t n, int32_t which, int64_t *semimanual_ref,
                                              float *sigma_out);

/* The whole schedule of main.cpp:325-545 on a resident volume: sigma0 from
 * the original, iteration 1, sigma1 re-estimated on its output with the same
 * samples and blended alpha*s0 + (1-alpha)*s1, iterations 2..maxiterations
 * (re-estimating each iteration with continuous_estimate).  raw/out are host
 * pointers (raw may be NULL when inlm_load_raw was called).  sigma_log, when
 * not NULL, receives the sigma list used by each iteration
 * (maxiterations*nz floats). */
INLM_API inlm_status inlm_denoise(inlm_handle *h, const float *raw, float *out, float *sigma_log);

/* Out-of-core schedule: inlm_denoise over z-slabs of at most slab_slices
 * output slices, each uploaded with the filter's z reach on both sides, for
 * volumes whose buffers do not fit one GPU together.  Replaces the host- and
 * file-staged block drivers Run_GaussianNoise_GPUBlocks
 * (iternlm_gpu.cu:1729-2115) and the CPU block mode (iternlm_cpu.cpp:403-577),
 * with sigma indexed by global z (the reference indexes it by block-local
 * z, iternlm_gpu.cu:121).  raw, out: host, nx*ny*nz floats; sigma_log NULL
 * or maxiterations*nz floats as in inlm_denoise.  The FP32 form gives the
 * whole-volume result bit for bit; the 16-bit forms normalise each slab
 * by its own moments. */
INLM_API inlm_status inlm_denoise_slabs(const inlm_params *p, const int32_t shape[3], int32_t device,
                                        const float *raw, float *out, int32_t slab_slices, float *sigma_log);

/* Capacity of one GPU for the slab drivers, as configure_device sizes it
 * (iternlm_gpu.cu:1528-1566): the largest number of output slices whose slab
 * (with the filter's z reach on both sides) a handle of these parameters
 * holds within `memory_limit` bytes of device memory, or within the
 * device's free memory when memory_limit <= 0.  nz when the whole volume
 * fits; OOM when not even one slice does. */
INLM_API inlm_status inlm_capacity_slices(const inlm_params *p, const int32_t shape[3], int32_t device,
                                          int64_t memory_limit, int32_t *n_slices);

/* One filter iteration over z-slabs of at most slab_slices output slices
 * on host buffers: the GPU half of Run_GaussianNoise_GPUBlocks
 * (iternlm_gpu.cu:1729-1882).  raw: the original volume; prev: the previous
 * iterate (ignored at iter 1, where the original is the previous iterate);
 * sigma: nz floats indexed by global z; out: nx*ny*nz floats (may alias
 * prev: the slabs read their halo from a copy of the previous iterate, not
 * from slices the same call has already overwritten, unlike the
 * reference's in-place update).  Every slab is uploaded with the filter's z
 * reach (search |dz| plus patch |dz|) on both sides, so the FP32 form gives
 * the whole-volume iteration bit for bit. */
INLM_API inlm_status inlm_iterate_slabs(const inlm_params *p, const int32_t shape[3], int32_t device, int32_t iter,
                                        const float *raw, const float *prev, const float *sigma, float *out,
                                        int32_t slab_slices);

/* Host-only query of the tables a handle with these parameters would use on
 * a volume of `shape`: search offsets and patch taps as (dx,dy,dz) triples in
 * the reference order (setup_searchspace / setup_patchspace,
 * iternlm_prepare.cpp:102-203) and the patch weights the reference CPU path
 * applies (literal tables of iternlm_cpu_kernels.cpp or setup_distweight,
 * iternlm_prepare.cpp:204-263).  Writes at most `cap` entries per table and
 * returns the true sizes.  Needs no GPU. */
INLM_API inlm_status inlm_tables(const inlm_params *p, const int32_t shape[3], int32_t cap, int32_t *n_search,
                                 int32_t *search_dxyz, int32_t *n_patch, int32_t *patch_dxyz, float *weights);

/* Self-test of the division inside the FP32 parity kernel's Pade weight: on
 * `device`, evaluates the weight for every float exponent from -0 down to the
 * cutoff -3.56648 (about 1.08e9 values) with the kernel's division and with
 * IEEE division and counts the results that differ.  It has no counterpart
 * in the reference; it backs the bit-exactness claim of INLM_PREC_FP32. */
INLM_API inlm_status inlm_selftest_division(int32_t device, uint64_t *mismatches, uint64_t *checked,
                                            uint32_t *example);

/* Kernel time of the last filter pass in milliseconds (HIP events recorded
 * on the handle's stream around the filter launch). */
INLM_API float inlm_last_filter_ms(const inlm_handle *h);

/* Table sizes of a handle: offsets visited and patch taps. */
INLM_API inlm_status inlm_table_sizes(const inlm_handle *h, int32_t *n_search, int32_t *n_patch);

/* Makes `vol` (nx*ny*nz floats, host or device) the result of iteration
 * `iter` without filtering: the resume path of Run_GaussianNoise
 * (iternlm_gpu.cu:1665-1697), which loads a finished iteration from disk
 * into the device result.  The original must be loaded. */
INLM_API inlm_status inlm_set_result(inlm_handle *h, int32_t iter, const float *vol, int32_t is_device);

/* ---------------------------------------------------------------------------
 * Learned offset proposal.  None of these has a counterpart in the reference:
 * they serve the pruned matrix-core path, in which each 32x8x8 filter tile
 * (tiles numbered x fastest, then y, then z) visits only a subset of the
 * search offsets.
 * ------------------------------------------------------------------------- */
typedef struct inlm_proposal inlm_proposal;

/* Number of 32x8x8 filter tiles of a handle's volume. */
INLM_API int64_t inlm_tile_count(const inlm_handle *h);

/* Training labels: runs iteration `iter` exactly as inlm_run_iteration does
 * with the FP32 parity form (same result, bit for bit) and also writes, for
 * every tile and search offset, the sum over the tile's voxels of the weight
 * that offset received ([tiles][n_search] floats, to device memory when
 * is_device != 0).  FP32 handles whose patch runs on the tiled kernel only. */
INLM_API inlm_status inlm_weight_mass(inlm_handle *h, int32_t iter, float *mass, int32_t is_device);

/* A 3D CNN (four 3x3x3 convolutions of 16 channels, ReLU, per-tile mean
 * pool, linear head over the n_search offsets, softmax) that predicts each
 * tile's distribution of weight mass over the offsets.  Its convolutions run
 * on BF16 matrix cores with FP32 accumulation. */
INLM_API inlm_status inlm_proposal_create(int32_t n_search, int32_t device, uint64_t seed, inlm_proposal **out);
INLM_API inlm_status inlm_proposal_destroy(inlm_proposal *p);
INLM_API const char *inlm_proposal_last_error(const inlm_proposal *p);
INLM_API int64_t inlm_proposal_num_params(const inlm_proposal *p);
/* Parameter vector (host floats): w1[27][16] b1[16]; three times
 * w[27 tap][16 in][16 out] b[16]; Wh[16][n_search] bh[n_search].  Taps are
 * numbered (dz+1)*9 + (dy+1)*3 + (dx+1).  set != 0 loads `params` (and resets
 * the optimiser state), set == 0 reads them out. */
INLM_API inlm_status inlm_proposal_params(inlm_proposal *p, float *params, int32_t set);
/* Gradient of the last training step's loss (host floats, same layout). */
INLM_API inlm_status inlm_proposal_grads(const inlm_proposal *p, float *grads);
/* One training step on a device volume (nx*ny*nz floats) and its device
 * labels from inlm_weight_mass: forward, cross-entropy of the predicted
 * distribution against mass/sum(mass) averaged over tiles, backward; with
 * lr > 0 also one Adam update.  *loss receives the loss before the update. */
INLM_API inlm_status inlm_proposal_train_step(inlm_proposal *p, const float *vol, const int32_t shape[3],
                                              const float *mass, float lr, float *loss);
/* Predicted distributions [tiles][n_search] for a device volume, written to
 * device memory when is_device != 0, else to host memory. */
INLM_API inlm_status inlm_proposal_predict(inlm_proposal *p, const float *vol, const int32_t shape[3], float *probs,
                                           int32_t is_device);

/* Attaches a proposal to a matrix-core handle whose original is loaded: the
 * proposal predicts once from the original; every later iteration visits,
 * per tile, its K most probable offsets in search order, except in tiles
 * whose K offsets carry less than `threshold` of the predicted mass
 * (fallback) and in a fraction `audit_frac` of tiles drawn afresh each
 * iteration (audit), which visit the whole search.  p == NULL detaches (and
 * clears fixed lists). */
INLM_API inlm_status inlm_set_proposal(inlm_handle *h, inlm_proposal *p, int32_t K, float threshold,
                                       float audit_frac, uint64_t seed);
/* Fixed per-tile offset lists for a matrix-core handle (host arrays):
 * lists[t*K + j] are indices into the search table, counts[t] in [1, n_search]
 * entries of tile t are used, counts[t] == n_search means the whole search.
 * Replaces an attached proposal. */
INLM_API inlm_status inlm_set_offset_lists(inlm_handle *h, const int32_t *lists, const int32_t *counts, int32_t K);
/* Statistics of the last pruned iteration: tiles, fallback tiles, audited
 * tiles, mean predicted mass of the kept offsets, mean offsets visited per
 * tile. */
INLM_API inlm_status inlm_proposal_stats(const inlm_handle *h, float out[5]);

/* ---------------------------------------------------------------------------
 * Volume I/O and the command-line program's output contract
 * (Geometry/hdcommunication.cpp, main.cpp:231-563).  Host-only except
 * inlm_denoise_files; baseline uncompressed TIFF without libtiff.
 * ------------------------------------------------------------------------- */

/* Shape {nx, ny, nz} of a TIFF input: a directory of slice files (*.tif,
 * *.tiff, sorted by name) or one file with one image per IFD or an ImageJ
 * stack (GetFilelist_And_ImageSequenceDimensions, hdcommunication.cpp:51-104). */
INLM_API inlm_status inlm_tiff_stack_shape(const char *path, int32_t shape[3]);

/* Slices first..last of a TIFF input as float32 (8/16-bit unsigned samples
 * converted, 32-bit samples taken as IEEE floats: Get3DTifSequence_32bitPointer,
 * hdcommunication.cpp:975-1060).  out holds nx*ny*(last-first+1) floats. */
INLM_API inlm_status inlm_tiff_read_stack(const char *path, int32_t first, int32_t last, float *out);

/* One TIFF per slice, dir/name<z0+i, 4 digits>.tif, float32 or (as16bit)
 * uint16 = min(max(0, round(x)), 65535) (SaveTifSequence_32bit / _as16bit,
 * hdcommunication.cpp:1550-1627); creates dir. */
INLM_API inlm_status inlm_tiff_write_stack(const char *dir, const char *name, const float *vol,
                                           const int32_t shape[3], int32_t z0, int32_t as16bit);

/* Noise files "i,sigma" per line (SaveNoisefile / ReadNoisefile,
 * hdcommunication.cpp:1736-1786); reading fails unless there are nz lines. */
INLM_API inlm_status inlm_noise_csv_write(const char *dir, const char *name, const float *sigma, int32_t n);
INLM_API inlm_status inlm_noise_csv_read(const char *file, int32_t nz, float *sigma);

/* Resume check of iternlm_cpu.cpp:51-85: 1 when dir holds denoised<first..last>.tif
 * and no file is smaller than one before it. */
INLM_API int32_t inlm_iteration_on_disk(const char *dir, int32_t first, int32_t last);

/* Options of the command-line program (protocol::DenoiseParameters::io,
 * denoise_parameters.h). firstslice/lastslice -1: the whole stack. */
typedef struct inlm_io_options {
    int32_t save_16bit; /* --16bit: uint16 output */
    int32_t resume;     /* --resume: reuse finished iterations and noise files */
    int32_t cleanup;    /* keep only the last iteration, in outpath itself */
    int32_t firstslice, lastslice;
    int32_t verbose;
    int32_t reserved[8];
} inlm_io_options;
INLM_API void inlm_default_io_options(inlm_io_options *o);

/* The whole program of main.cpp:231-563 on one GPU: read the TIFF input,
 * estimate sigma0 (noise_vs_slice00.csv), iteration 1 into
 * outpath/denoise-iteration01/denoised<z>.tif, sigma1 (noise_vs_slice01.csv)
 * blended with alpha, iterations 2..maxiterations, resume and cleanup as the
 * reference, and a line block appended to logfile.txt next to outpath.
 * outpath NULL or empty: <parent of inpath>/denoised/. */
INLM_API inlm_status inlm_denoise_files(const inlm_params *p, int32_t device, const char *inpath,
                                        const char *outpath, const inlm_io_options *o);

/* ---- colour path (Color/color_denoising.h, iternlm_cpu.cpp:932-1075) ----- */
typedef struct inlm_rgb_options {
    int32_t independent_channels; /* --independent: a weight per channel (default: one joint weight) */
    int32_t stabilize_variance;   /* Anscombe transform of the channels (default 1; --nopoisson: 0) */
    int32_t average_channelsigma; /* one noise level for all channels (default 1; --noaverage: 0) */
    int32_t reserved[5];
} inlm_rgb_options;
INLM_API void inlm_default_rgb_options(inlm_rgb_options *c);

/* One 2D RGB image, replacing color_denoise::run_colorimage_denoising for an
 * image in memory: in and out hold the R, G and B planes of nx*ny floats
 * each (shape = {nx, ny}); the z radii are ignored (2D) and the noise level
 * of each channel comes from 2048 samples.  The result is returned after the
 * inverse Anscombe transform, unquantised.  sigma_log: NULL or
 * 3*maxiterations floats, the channel noise levels each iteration used. */
INLM_API inlm_status inlm_denoise_rgb(const inlm_params *p, const inlm_rgb_options *c, const int32_t shape[2],
                                      int32_t device, const float *in, float *out, float *sigma_log);
/* RGB TIFF (8 or 16 bits, 3 or 4 samples per pixel) to three float planes,
 * and three float planes to an 8-bit RGB TIFF (values capped at 255 and
 * truncated, as color_denoising.h:72-77). */
INLM_API inlm_status inlm_rgb_tiff_shape(const char *file, int32_t shape[2]);
INLM_API inlm_status inlm_rgb_tiff_read(const char *file, float *planes);
INLM_API inlm_status inlm_rgb_tiff_write(const char *file, const float *planes, const int32_t shape[2]);
/* The reference program's colour mode (main.cpp:254-257): every RGB TIFF in
 * inpath (or the one file) to outpath/<name>_denoised.tif, and, unless
 * cleanup, every iteration to outpath/denoise_iterations/<name>_iterNN.tif. */
INLM_API inlm_status inlm_denoise_rgb_files(const inlm_params *p, const inlm_rgb_options *c, int32_t device,
                                            const char *inpath, const char *outpath, const inlm_io_options *o);

#ifdef __cplusplus
}
#endif
#endif /* INLM_CAPI_H */
