// C ABI of the iterNLM hot path: one handle per volume, owning the resident
// padded original, the padded previous iterate, the result and the tables.
#include "inlm_capi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "inlm_internal.hpp"
#include "inlm_kernels.hpp"
#include "inlm_learn.hpp"
#include "inlm_tables.hpp"

namespace {

std::string g_create_error;  // message of the last failed call without a handle
constexpr size_t kGuard = 64;

}  // namespace

namespace inlm {
void set_global_error(const std::string &msg) { g_create_error = msg; }

inlm_status sigma_from_std(const inlm_params &p, const int shape[3], const int64_t *samples, const double *sd,
                           int64_t n, int which, int64_t &semi_ref, float *sig, std::string &err)
{
    const int nz = shape[2];
    const double k = p.stds_from_mean;
    switch (p.noise_mode) {
    case INLM_NOISE_MANUAL: {
        // sigma[0] on the first estimate, sigma[1] afterwards (noiselevel.cpp:81-88)
        const float v = p.sigma[which == 0 ? 0 : 1];
        for (int i = 0; i < nz; i++) sig[i] = v;
        return INLM_OK;
    }
    case INLM_NOISE_SEMIMANUAL: {
        // noiselevel.cpp:56-79: the first estimate is sigma[0] and fixes the
        // sample whose patch std lies closest to it (the first of equals);
        // later estimates are that sample's patch std
        if (n < 1) {
            err = "semimanual noise mode needs noise samples";
            return INLM_BAD_ARG;
        }
        float v;
        if (which == 0) {
            double best = 1e9;
            int64_t ref = 0;
            for (int64_t i = 0; i < n; i++) {
                const double d = std::fabs(sd[i] - (double)p.sigma[0]);
                if (d < best) {
                    best = d;
                    ref = i;
                }
            }
            semi_ref = ref;
            v = p.sigma[0];
        } else {
            if (semi_ref < 0 || semi_ref >= n) {
                err = "semimanual noise mode: no reference sample from a first estimate";
                return INLM_BAD_ARG;
            }
            v = (float)sd[semi_ref];
        }
        for (int i = 0; i < nz; i++) sig[i] = v;
        return INLM_OK;
    }
    case INLM_NOISE_ZADAPTIVE: {
        std::vector<double> mean, stdev;
        zwindow_clusters(samples, sd, n, shape, p.zwindow, mean, stdev);
        for (int i = 0; i < nz; i++) sig[i] = (float)(mean[i] + k * stdev[i]);
        return INLM_OK;
    }
    case INLM_NOISE_GLOBAL: {
        double mean = 0, stdev = 0;
        create2clusters(sd, (int)n, 0, (int)n - 1, mean, stdev);
        for (int i = 0; i < nz; i++) sig[i] = (float)(mean + k * stdev);
        return INLM_OK;
    }
    default: err = "unknown noise mode"; return INLM_BAD_ARG;
    }
}
}  // namespace inlm

struct inlm_handle {
    inlm_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    inlm::Geometry geo{};
    inlm::FilterGeom fg{};
    std::vector<inlm::SearchEntry> search, search_fast;
    std::vector<unsigned char> search_fast_blob;  // the fast order and its staging moves, as uploaded
    std::vector<int> tap_dxyz;
    inlm::TapTable taps{};
    int rp_tiled = 0;
    size_t pad_elems = 0;  // allocated elements of each padded volume (>= padded voxels)
    long shifted = 0;      // 16-bit forms: elements from a padded volume to its copy shifted by one
    // FP32 parity form: padded original and previous iterate in float.
    // Matrix-core forms: the same two volumes normalised, in the form's
    // storage type (Rs/Ps), plus the unpadded original in float.
    float *R_pad = nullptr, *P_pad = nullptr, *result = nullptr, *sigma = nullptr;
    void *R_alloc = nullptr, *P_alloc = nullptr;
    void *Rs = nullptr, *Ps = nullptr;
    float *raw = nullptr;
    double *moments = nullptr;
    bool fast = false;
    double shift = 0.0, scale = 1.0;
    inlm::SearchEntry *search_dev = nullptr;
    // the same offsets in the order the matrix-core kernels sweep the whole
    // search (the lists of the pruned forms index search_dev)
    inlm::SearchEntry *search_fast_dev = nullptr;
    int min_lin = 0;  // smallest linear search offset
    int64_t *samples_dev = nullptr;
    double *std_dev = nullptr;
    int64_t samples_cap = 0;
    int64_t semi_ref = -1;  // semimanual noise mode: the reference sample of the first estimate
    bool raw_loaded = false, sigma_set = false;
    int last_iter = 0;
    float last_ms = 0.f;
    // pruned matrix-core path: per-tile offset lists, either fixed by the
    // caller or re-drawn each iteration from an attached proposal's
    // prediction (probs_dev)
    int *lists_dev = nullptr, *counts_dev = nullptr;
    int lists_K = 0;
    long lists_cap = 0;
    bool lists_fixed = false;
    float *probs_dev = nullptr;
    int prop_K = 0;
    float prop_thr = 0.f, prop_audit = 0.f;
    uint64_t prop_seed = 0;
    float *stats_dev = nullptr;
    float stats[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    std::string err;
};

struct inlm_proposal {
    inlm::ProposalNet net;
    int device = 0;
    int ns = 0;
    std::string err;
};

namespace {

inlm_status fail(inlm_handle *h, inlm_status st, const std::string &msg)
{
    if (h) h->err = msg;
    else g_create_error = msg;
    return st;
}

inlm_status hip_fail(inlm_handle *h, hipError_t e, const char *where)
{
    std::string m = std::string(where) + ": " + hipGetErrorString(e);
    return fail(h, e == hipErrorOutOfMemory ? INLM_OOM : INLM_HIP_ERROR, m);
}

#define HIP_TRY(h, expr)                                  \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)

// Highest element index the tiled filter kernel may read in a padded volume.
// The kernel reads whole tiles, so for shapes that are not multiples of the
// tile it reads past the padded volume; those voxels feed only discarded
// outputs, and the allocation covers them.
size_t padded_alloc_elems(const inlm::Geometry &g)
{
    const long sp = (long)g.nxp * g.nyp;
    const long ntz = (g.nz + 7) / 8, nty = (g.ny + 7) / 8, ntx = (g.nx + 31) / 32;
    const long R = 3, pitch = 40;  // largest tile radius and its LDS pitch
    const long zmax = (ntz * 8 - 1) + 2L * g.pz;
    const long ymax = (nty * 8 - 1) + 2L * g.py;
    const long xmax = (ntx - 1) * 32 + pitch - 1 - 2 * 0 + 2L * g.px;
    long maxidx = zmax * sp + ymax * g.nxp + xmax + 2 * R;
    long need = std::max<long>(maxidx + 1, (long)g.voxels_p());
    return (size_t)need + 64;
}

long tile_count(const inlm::Geometry &g)
{
    return (long)((g.nx + 31) / 32) * ((g.ny + 7) / 8) * ((g.nz + 7) / 8);
}

inlm_status ensure_lists(inlm_handle *h, int K)
{
    const long nt = tile_count(h->geo);
    const long need = nt * (long)K;
    if (!h->counts_dev) HIP_TRY(h, hipMalloc(&h->counts_dev, sizeof(int) * nt));
    if (!h->stats_dev) HIP_TRY(h, hipMalloc(&h->stats_dev, sizeof(float) * 4));
    if (need > h->lists_cap) {
        (void)hipFree(h->lists_dev);
        h->lists_dev = nullptr;
        h->lists_cap = 0;
        HIP_TRY(h, hipMalloc(&h->lists_dev, sizeof(int) * need));
        h->lists_cap = need;
    }
    h->lists_K = K;
    return INLM_OK;
}

void clear_pruning(inlm_handle *h)
{
    (void)hipFree(h->probs_dev);
    h->probs_dev = nullptr;
    h->lists_fixed = false;
    h->prop_K = 0;
}

inlm_status ensure_samples(inlm_handle *h, int64_t n)
{
    if (n <= h->samples_cap) return INLM_OK;
    if (h->samples_dev) (void)hipFree(h->samples_dev);
    if (h->std_dev) (void)hipFree(h->std_dev);
    h->samples_dev = nullptr;
    h->std_dev = nullptr;
    h->samples_cap = 0;
    HIP_TRY(h, hipMalloc(&h->samples_dev, sizeof(int64_t) * n));
    HIP_TRY(h, hipMalloc(&h->std_dev, sizeof(double) * n));
    h->samples_cap = n;
    return INLM_OK;
}

}  // namespace

namespace inlm {
// Device bytes a handle of these parameters allocates for `shape` (the sums
// of inlm_create's allocations, plus the sample buffers of one noise
// estimate with n_samples samples).
size_t handle_bytes(const inlm_params &pin, const int32_t shape_in[3])
{
    inlm_params p = pin;
    int shape[3] = {shape_in[0], shape_in[1], shape_in[2]};
    if (shape[2] == 1) {
        p.radius_searchspace[2] = 0;
        p.radius_patchspace[2] = 0;
    }
    Geometry g = make_geometry(shape, p.radius_searchspace, p.radius_patchspace, p.nslices);
    const size_t esz = form_elem_size(p.precision);
    if (esz == 2 && (g.nxp & 1) != 0) g.nxp += 1;
    const size_t pad_elems = padded_alloc_elems(g);
    const bool fast = p.precision != INLM_PREC_FP32;
    const size_t shifted = fast && esz == 2 ? (pad_elems + kGuard + 63) / 64 * 64 : 0;
    const size_t padb = (shifted > 0 ? kGuard + shifted + pad_elems : pad_elems + kGuard) * esz;
    const size_t vox = (size_t)g.voxels();
    size_t b = 2 * padb + sizeof(float) * vox + sizeof(float) * (size_t)g.nz;
    if (fast) b += sizeof(float) * vox + inlm::kMomentDoubles * sizeof(double);
    const size_t ns = search_offsets(p.radius_searchspace, p.nslices).size();
    b += 2 * ns * sizeof(SearchEntry) + (ns + kStageMovesPad) * sizeof(int);
    if (p.noise_mode != INLM_NOISE_MANUAL && p.n_samples > 0) b += (size_t)p.n_samples * (sizeof(int64_t) + sizeof(double));
    return b;
}
}  // namespace inlm

extern "C" {

void inlm_default_params(inlm_params *p)
{
    std::memset(p, 0, sizeof(*p));
    p->alpha = 0.5f;
    p->maxiterations = 4;
    p->nslices = 11;
    for (int i = 0; i < 3; i++) {
        p->radius_searchspace[i] = 10;
        p->radius_patchspace[i] = 1;
    }
    p->beta = 1.f;
    p->z_anisotropy = 1.f;
    p->noise_mode = INLM_NOISE_ZADAPTIVE;
    p->zwindow = 100;
    p->n_samples = 100000;
    p->patchsize = 15;
    p->stds_from_mean = 1.f;
    p->continuous_estimate = 0;
    p->sample_seed = 0;
    p->cpu_blockwise = 1;
    p->precision = INLM_PREC_FP32;
    p->weight_fn = INLM_WEIGHT_PADE;
}

int32_t inlm_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char *inlm_last_error(const inlm_handle *h) { return h ? h->err.c_str() : g_create_error.c_str(); }

inlm_status inlm_create(const inlm_params *pin, const int32_t shape_in[3], int32_t device, inlm_handle **out)
{
    if (!pin || !shape_in || !out) return fail(nullptr, INLM_BAD_ARG, "null argument");
    *out = nullptr;
    inlm_params p = *pin;
    int shape[3] = {shape_in[0], shape_in[1], shape_in[2]};
    for (int i = 0; i < 3; i++) {
        if (shape[i] < 1) return fail(nullptr, INLM_BAD_ARG, "shape must be positive");
        if (p.radius_searchspace[i] < 0 || p.radius_patchspace[i] < 0)
            return fail(nullptr, INLM_BAD_ARG, "radii must be non-negative");
    }
    if (p.nslices < 1) return fail(nullptr, INLM_BAD_ARG, "nslices must be positive");
    // Single-slice input: the reference's main drops the z radii
    // (main.cpp:261-265).
    if (shape[2] == 1) {
        p.radius_searchspace[2] = 0;
        p.radius_patchspace[2] = 0;
    }
    if (p.precision < INLM_PREC_FP32 || p.precision > INLM_PREC_FP16_FP8)
        return fail(nullptr, INLM_BAD_ARG, "unknown precision form");
    if (p.weight_fn != INLM_WEIGHT_PADE && p.weight_fn != INLM_WEIGHT_EXP)
        return fail(nullptr, INLM_BAD_ARG, "unknown weight function");

    inlm::Geometry g = inlm::make_geometry(shape, p.radius_searchspace, p.radius_patchspace, p.nslices);
    // the 16-bit matrix-core forms keep an even row pitch (one unused column
    // when nx + 2 px is odd), so that the parity of a load's element index
    // depends on the search offset alone
    if (inlm::form_elem_size(p.precision) == 2 && (g.nxp & 1) != 0) g.nxp += 1;
    const int padv[3] = {g.px, g.py, g.pz};
    for (int i = 0; i < 3; i++)
        if (shape[i] < 2 && padv[i] > 0)
            return fail(nullptr, INLM_BAD_ARG, "reflective padding needs at least 2 voxels on every padded axis");
    const long sp = (long)g.nxp * g.nyp;
    if ((double)sp * (g.pz + 8) + (double)g.nxp * 16 > 2.0e9)
        return fail(nullptr, INLM_BAD_ARG, "slice too large for 32-bit offsets");

    inlm_handle *h = new inlm_handle();
    h->p = p;
    h->device = device;
    h->geo = g;
    h->fg = inlm::FilterGeom{g.nx, g.ny, g.nz, g.nxp, g.nyp, g.nzp, g.px, g.py, g.pz};

    // Tables on the padded shape (configure_device, iternlm_gpu.cu:1468-1470).
    std::vector<inlm::Offset3> S = inlm::search_offsets(p.radius_searchspace, p.nslices);
    std::vector<inlm::Offset3> PT = inlm::patch_offsets(p.radius_patchspace, p.nslices);
    std::vector<float> W;
    std::string werr;
    if (!inlm::cpu_weight_table(p.radius_patchspace, p.nslices, p.z_anisotropy, g.nzp, p.cpu_blockwise != 0, W,
                                werr)) {
        delete h;
        return fail(nullptr, INLM_BAD_ARG, werr);
    }
    if ((int)PT.size() > inlm::kMaxTaps) {
        delete h;
        return fail(nullptr, INLM_BAD_ARG, "patch has more than 128 taps");
    }
    if (S.empty()) {
        delete h;
        return fail(nullptr, INLM_BAD_ARG, "empty search space");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        delete h;
        return fail(nullptr, INLM_NO_DEVICE, "no GPU visible");
    }
    if (device < 0 || device >= ndev) {
        delete h;
        return fail(nullptr, INLM_BAD_ARG, "device index out of range");
    }
    for (const auto &o : S) h->search.push_back({o.dx, o.dy, o.dz, (int)(o.dz * sp + (long)o.dy * g.nxp + o.dx)});
    for (const auto &e : h->search) h->min_lin = std::min(h->min_lin, e.lin);
    h->taps.n = (int)PT.size();
    for (size_t i = 0; i < PT.size(); i++) {
        h->taps.w[i] = W[i];
        h->taps.lin[i] = (int)(PT[i].dz * sp + (long)PT[i].dy * g.nxp + PT[i].dx);
        h->taps.dxyz[3 * i] = PT[i].dx;
        h->taps.dxyz[3 * i + 1] = PT[i].dy;
        h->taps.dxyz[3 * i + 2] = PT[i].dz;
        h->tap_dxyz.push_back(PT[i].dx);
        h->tap_dxyz.push_back(PT[i].dy);
        h->tap_dxyz.push_back(PT[i].dz);
    }
    const int *rp = p.radius_patchspace;
    const bool iso = rp[0] == rp[1] && rp[1] == rp[2] && rp[0] >= 1 && rp[0] <= 3;
    h->fast = p.precision != INLM_PREC_FP32;
    if (h->fast) {
        // the matrix-core forms cover isotropic patches of radius 1..3 whose
        // weights depend on (dx^2+dy^2, dz^2) only
        if (!iso || !inlm::mfma_taps_supported(rp[0], h->tap_dxyz.data(), h->taps.w, h->taps.n)) {
            delete h;
            return fail(nullptr, INLM_BAD_ARG, "the matrix-core forms need an isotropic patch of radius 1, 2 or 3");
        }
        h->rp_tiled = rp[0];
    } else if (p.kernel_select == 0 && iso && inlm::tiled_taps_match(rp[0], h->tap_dxyz.data(), h->taps.w, h->taps.n))
        h->rp_tiled = rp[0];

    inlm_status st = INLM_OK;
    auto bail = [&](inlm_status s) {
        inlm_destroy(h);
        return s;
    };
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return bail(hip_fail(nullptr, e, "hipSetDevice"));
    if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(hip_fail(nullptr, e, "hipStreamCreate"));
    if ((e = hipEventCreate(&h->ev0)) != hipSuccess || (e = hipEventCreate(&h->ev1)) != hipSuccess)
        return bail(hip_fail(nullptr, e, "hipEventCreate"));
    h->pad_elems = padded_alloc_elems(g);
    // kGuard elements in front of each padded volume: the matrix-core staging
    // reads chunks that start up to 3 elements before the volume.  The
    // 16-bit forms follow each volume with a copy shifted by one element
    // (copy[j] = volume[j + 1]), 64-element aligned, whose own guard is the
    // zeroed gap in front of it.
    const size_t esz = inlm::form_elem_size(h->p.precision);
    if (h->fast && esz == 2) h->shifted = (long)((h->pad_elems + kGuard + 63) / 64 * 64);
    const size_t padb = (h->shifted > 0 ? kGuard + (size_t)h->shifted + h->pad_elems : h->pad_elems + kGuard) * esz;
    if ((e = hipMalloc(&h->R_alloc, padb)) != hipSuccess) return bail(hip_fail(nullptr, e, "hipMalloc(original)"));
    if ((e = hipMalloc(&h->P_alloc, padb)) != hipSuccess) return bail(hip_fail(nullptr, e, "hipMalloc(previous)"));
    if ((e = hipMemsetAsync(h->R_alloc, 0, padb, h->stream)) != hipSuccess ||
        (e = hipMemsetAsync(h->P_alloc, 0, padb, h->stream)) != hipSuccess)
        return bail(hip_fail(nullptr, e, "hipMemset"));
    h->Rs = static_cast<char *>(h->R_alloc) + kGuard * esz;
    h->Ps = static_cast<char *>(h->P_alloc) + kGuard * esz;
    if (h->fast) {
        if ((e = hipMalloc(&h->raw, sizeof(float) * g.voxels())) != hipSuccess)
            return bail(hip_fail(nullptr, e, "hipMalloc(raw)"));
        if ((e = hipMalloc(&h->moments, inlm::kMomentDoubles * sizeof(double))) != hipSuccess)
            return bail(hip_fail(nullptr, e, "hipMalloc(moments)"));
    } else {
        h->R_pad = static_cast<float *>(h->Rs);
        h->P_pad = static_cast<float *>(h->Ps);
    }
    if ((e = hipMalloc(&h->result, sizeof(float) * g.voxels())) != hipSuccess)
        return bail(hip_fail(nullptr, e, "hipMalloc(result)"));
    if ((e = hipMalloc(&h->sigma, sizeof(float) * g.nz)) != hipSuccess)
        return bail(hip_fail(nullptr, e, "hipMalloc(sigma)"));
    if ((e = hipMalloc(&h->search_dev, sizeof(inlm::SearchEntry) * h->search.size())) != hipSuccess)
        return bail(hip_fail(nullptr, e, "hipMalloc(search)"));
    if ((e = hipMemcpyAsync(h->search_dev, h->search.data(), sizeof(inlm::SearchEntry) * h->search.size(),
                            hipMemcpyHostToDevice, h->stream)) != hipSuccess)
        return bail(hip_fail(nullptr, e, "hipMemcpy(search)"));
    // Full-search order of the matrix-core forms: per search slice, the
    // offsets whose 16-bit loads use one volume copy (even dx with an even
    // row pitch) before those that use the other, each group row by row.
    // Consecutive steps then read one copy at regions two columns apart, so
    // the L1 cache keeps serving them; the reference order alternates the
    // two copies at every step.  The sums are order-independent up to FP32
    // rounding.
    h->search_fast = h->search;
    std::stable_sort(h->search_fast.begin(), h->search_fast.end(), [](const inlm::SearchEntry &a, const inlm::SearchEntry &b) {
        if (a.dz != b.dz) return a.dz < b.dz;
        if ((a.lin & 1) != (b.lin & 1)) return (a.lin & 1) < (b.lin & 1);
        if (a.dy != b.dy) return a.dy < b.dy;
        return a.dx < b.dx;
    });
    // followed by the block kernel's staging moves (kStageMovesPad)
    {
        const size_t ns = h->search_fast.size();
        std::vector<int> moves(ns + inlm::kStageMovesPad, 0);
        for (size_t k = 1; k < ns; k++) {
            const inlm::SearchEntry &c = h->search_fast[k], &b = h->search_fast[k - 1];
            if (c.dz == b.dz && c.dy == b.dy && c.dx == b.dx + 2 && moves[k - 1] < inlm::kStageMovesMax)
                moves[k] = moves[k - 1] + 1;
        }
        const size_t tb = sizeof(inlm::SearchEntry) * ns, mb = sizeof(int) * moves.size();
        h->search_fast_blob.resize(tb + mb);
        std::memcpy(h->search_fast_blob.data(), h->search_fast.data(), tb);
        std::memcpy(h->search_fast_blob.data() + tb, moves.data(), mb);
        if ((e = hipMalloc(&h->search_fast_dev, tb + mb)) != hipSuccess)
            return bail(hip_fail(nullptr, e, "hipMalloc(search)"));
        if ((e = hipMemcpyAsync(h->search_fast_dev, h->search_fast_blob.data(), tb + mb, hipMemcpyHostToDevice,
                                h->stream)) != hipSuccess)
            return bail(hip_fail(nullptr, e, "hipMemcpy(search)"));
    }
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return bail(hip_fail(nullptr, e, "sync"));
    (void)st;
    *out = h;
    return INLM_OK;
}

inlm_status inlm_destroy(inlm_handle *h)
{
    if (!h) return INLM_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    (void)hipFree(h->R_alloc);
    (void)hipFree(h->P_alloc);
    (void)hipFree(h->raw);
    (void)hipFree(h->moments);
    (void)hipFree(h->result);
    (void)hipFree(h->sigma);
    (void)hipFree(h->search_dev);
    (void)hipFree(h->search_fast_dev);
    (void)hipFree(h->samples_dev);
    (void)hipFree(h->std_dev);
    (void)hipFree(h->lists_dev);
    (void)hipFree(h->counts_dev);
    (void)hipFree(h->probs_dev);
    (void)hipFree(h->stats_dev);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return INLM_OK;
}

inlm_status inlm_set_sigma(inlm_handle *h, const float *sigma, int32_t nz)
{
    if (!h || !sigma) return fail(h, INLM_BAD_ARG, "null argument");
    if (nz != h->geo.nz) return fail(h, INLM_BAD_ARG, "sigma list length must equal nz");
    HIP_TRY(h, hipSetDevice(h->device));
    HIP_TRY(h, hipMemcpyAsync(h->sigma, sigma, sizeof(float) * nz, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    h->sigma_set = true;
    return INLM_OK;
}

inlm_status inlm_load_raw(inlm_handle *h, const float *raw, int32_t is_device)
{
    if (!h || !raw) return fail(h, INLM_BAD_ARG, "null argument");
    HIP_TRY(h, hipSetDevice(h->device));
    const size_t bytes = sizeof(float) * h->geo.voxels();
    if (!h->fast) {
        HIP_TRY(h, hipMemcpyAsync(h->result, raw, bytes, is_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                  h->stream));
        HIP_TRY(h, inlm::launch_pad_reflect(h->result, h->R_pad, h->fg, h->stream));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
    } else {
        // normalisation: shift = mean, scale = 1/std of the original (FP64)
        HIP_TRY(h, hipMemcpyAsync(h->raw, raw, bytes, is_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                  h->stream));
        HIP_TRY(h, inlm::launch_moments(h->raw, (long)h->geo.voxels(), h->moments, h->stream));
        double mom[2];
        HIP_TRY(h, hipMemcpyAsync(mom, h->moments, sizeof mom, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
        const double nvox = (double)h->geo.voxels();
        const double mean = mom[0] / nvox;
        const double var = std::max(mom[1] / nvox - mean * mean, 0.0);
        h->shift = mean;
        h->scale = var > 0 ? 1.0 / std::sqrt(var) : 1.0;
        HIP_TRY(h, inlm::launch_pad_normalise(h->p.precision, h->raw, h->Rs, h->shifted, h->fg, (float)h->shift,
                                              (float)h->scale, h->stream));
        HIP_TRY(h, hipStreamSynchronize(h->stream));
    }
    h->raw_loaded = true;
    h->last_iter = 0;
    // a proposal predicted from the previous original no longer applies
    if (h->probs_dev) clear_pruning(h);
    return INLM_OK;
}

}  // extern "C"

namespace {

// One iteration (inlm_run_iteration); with `mass` (device) the FP32 tiled
// kernel also writes the per-tile weight mass of every offset.
inlm_status run_iteration_impl(inlm_handle *h, int32_t iter, const float *raw, float *mass)
{
    if (!h) return fail(h, INLM_BAD_ARG, "null handle");
    if (iter < 1) return fail(h, INLM_BAD_ARG, "iterations count from 1");
    HIP_TRY(h, hipSetDevice(h->device));
    if (iter == 1 && raw) {
        inlm_status s = inlm_load_raw(h, raw, 0);
        if (s != INLM_OK) return s;
    }
    if (!h->raw_loaded) return fail(h, INLM_BAD_ARG, "no original volume loaded");
    if (!h->sigma_set) return fail(h, INLM_BAD_ARG, "sigma not set");
    const void *P = h->Rs;  // iteration 1: the previous iterate is the original
    if (iter > 1) {
        if (h->last_iter < 1) return fail(h, INLM_BAD_ARG, "iteration > 1 before iteration 1");
        if (h->fast)
            HIP_TRY(h, inlm::launch_pad_normalise(h->p.precision, h->result, h->Ps, h->shifted, h->fg, (float)h->shift,
                                                  (float)h->scale, h->stream));
        else
            HIP_TRY(h, inlm::launch_pad_reflect(h->result, h->P_pad, h->fg, h->stream));
        P = h->Ps;
    }
    const int ns = (int)h->search.size();
    const long nt = tile_count(h->geo);
    HIP_TRY(h, hipEventRecord(h->ev0, h->stream));
    if (h->fast) {
        inlm::OffsetLists ol{h->lists_dev, h->counts_dev, h->lists_K, nt};
        bool pruned = h->probs_dev || h->lists_fixed;
        if (h->probs_dev) {
            // fresh lists for this iteration: top-K, fallback, audit draw
            HIP_TRY(h, hipMemsetAsync(h->stats_dev, 0, sizeof(float) * 4, h->stream));
            HIP_TRY(h, inlm::launch_topk(h->probs_dev, (int)nt, ns, h->prop_K, h->prop_thr, h->prop_audit,
                                         h->prop_seed + 0x9E3779B97F4A7C15ull * (uint64_t)iter, h->lists_dev,
                                         h->counts_dev, h->stats_dev, h->stream));
            // when every tile falls back or is audited, the full-search
            // kernel runs instead of the pruned one
            float full_tiles = 0.f;
            HIP_TRY(h, hipMemcpyAsync(&full_tiles, h->stats_dev, sizeof(float), hipMemcpyDeviceToHost, h->stream));
            HIP_TRY(h, hipStreamSynchronize(h->stream));
            if (full_tiles >= (float)nt) pruned = false;
        }
        HIP_TRY(h, inlm::launch_filter_mfma(h->p.precision, h->rp_tiled, h->p.weight_fn, h->Rs, P, h->pad_elems, h->shifted,
                                            h->raw, h->result, h->sigma, h->p.beta,
                                            (float)(1.0 / (h->scale * h->scale)), h->shift, 1.0 / h->scale,
                                            h->search_dev, h->search_fast_dev, ns, h->min_lin, h->taps,
                                            h->fg, h->stream, pruned ? &ol : nullptr, h->p.kernel_select == 2));
    } else {
        HIP_TRY(h, inlm::launch_filter(h->p.precision, h->rp_tiled, h->p.weight_fn, h->R_pad,
                                       static_cast<const float *>(P), h->result, h->sigma, h->p.beta, h->search_dev,
                                       ns, h->taps, h->fg, h->stream, mass));
    }
    HIP_TRY(h, hipEventRecord(h->ev1, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    HIP_TRY(h, hipEventElapsedTime(&h->last_ms, h->ev0, h->ev1));
    if (h->fast && h->probs_dev) {
        float st[4];
        HIP_TRY(h, hipMemcpy(st, h->stats_dev, sizeof st, hipMemcpyDeviceToHost));
        h->stats[0] = (float)nt;
        h->stats[1] = st[0];
        h->stats[2] = st[1];
        h->stats[3] = st[2] / (float)nt;
        h->stats[4] = st[3] / (float)nt;
    }
    h->last_iter = iter;
    return INLM_OK;
}

}  // namespace

extern "C" {

inlm_status inlm_run_iteration(inlm_handle *h, int32_t iter, const float *raw)
{
    return run_iteration_impl(h, iter, raw, nullptr);
}

inlm_status inlm_get_result(inlm_handle *h, float *out, int32_t is_device)
{
    if (!h || !out) return fail(h, INLM_BAD_ARG, "null argument");
    HIP_TRY(h, hipSetDevice(h->device));
    HIP_TRY(h, hipMemcpyAsync(out, h->result, sizeof(float) * h->geo.voxels(),
                              is_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return INLM_OK;
}

inlm_status inlm_draw_samples(const int32_t shape[3], int64_t n, int32_t patchsize, uint64_t seed, int64_t *out)
{
    if (!shape || (!out && n > 0)) return fail(nullptr, INLM_BAD_ARG, "null argument");
    int sh[3] = {shape[0], shape[1], shape[2]};
    if (!inlm::draw_samples(sh, n, patchsize, seed, out))
        return fail(nullptr, INLM_BAD_ARG, "volume smaller than the noise patch");
    return INLM_OK;
}

inlm_status inlm_draw_samples_mask(const int32_t shape[3], int64_t n, int32_t patchsize, int32_t mask_diameter,
                                   uint64_t seed, int64_t *out)
{
    if (!shape || (!out && n > 0)) return fail(nullptr, INLM_BAD_ARG, "null argument");
    int sh[3] = {shape[0], shape[1], shape[2]};
    if (!inlm::draw_samples(sh, n, patchsize, seed, out, mask_diameter))
        return fail(nullptr, INLM_BAD_ARG,
                    mask_diameter > 0 ? "no sample position lies inside the circular mask (or the volume is smaller "
                                        "than the noise patch)"
                                      : "volume smaller than the noise patch");
    return INLM_OK;
}

inlm_status inlm_sigma_from_patchstd(const inlm_params *p, const int32_t shape[3], const int64_t *samples,
                                     const double *stds, int64_t n, int32_t which, int64_t *semimanual_ref,
                                     float *sigma_out)
{
    if (!p || !shape || !sigma_out || (n > 0 && (!samples || !stds)) || n < 0)
        return fail(nullptr, INLM_BAD_ARG, "null argument");
    if (shape[0] < 1 || shape[1] < 1 || shape[2] < 1) return fail(nullptr, INLM_BAD_ARG, "shape must be positive");
    if (p->noise_mode == INLM_NOISE_SEMIMANUAL && !semimanual_ref)
        return fail(nullptr, INLM_BAD_ARG, "semimanual noise mode needs semimanual_ref");
    const int64_t nslice = (int64_t)shape[0] * shape[1];
    for (int64_t i = 0; i < n; i++)
        if (samples[i] < 0 || samples[i] >= nslice * shape[2] || (i > 0 && samples[i] < samples[i - 1]))
            return fail(nullptr, INLM_BAD_ARG, "samples must be sorted voxel indices of the volume");
    int sh[3] = {shape[0], shape[1], shape[2]};
    int64_t dummy = -1;
    std::string err;
    const inlm_status s = inlm::sigma_from_std(*p, sh, samples, stds, n, which,
                                               semimanual_ref ? *semimanual_ref : dummy, sigma_out, err);
    if (s != INLM_OK) return fail(nullptr, s, err);
    return INLM_OK;
}

inlm_status inlm_patchstd(inlm_handle *h, int32_t which, const int64_t *samples, int64_t n, double *out)
{
    if (!h || (!samples && n > 0) || (!out && n > 0)) return fail(h, INLM_BAD_ARG, "null argument");
    if (which != 0 && which != 1) return fail(h, INLM_BAD_ARG, "which must be 0 (original) or 1 (result)");
    if (which == 0 && !h->raw_loaded) return fail(h, INLM_BAD_ARG, "no original volume loaded");
    if (which == 1 && h->last_iter < 1) return fail(h, INLM_BAD_ARG, "no result yet");
    const inlm::Geometry &g = h->geo;
    const int ps = h->p.patchsize;
    if (ps < 1 || ps > g.nx || ps > g.ny) return fail(h, INLM_BAD_ARG, "patch size does not fit the slice");
    const int64_t nslice = (int64_t)g.nx * g.ny;
    for (int64_t i = 0; i < n; i++) {  // every patch must lie inside its slice
        const int64_t idx = samples[i];
        if (idx < 0 || idx >= g.voxels()) return fail(h, INLM_BAD_ARG, "sample index out of range");
        const int64_t r = idx % nslice;
        if (r % g.nx + ps > g.nx || r / g.nx + ps > g.ny) return fail(h, INLM_BAD_ARG, "sample patch leaves the slice");
    }
    if (n == 0) return INLM_OK;
    HIP_TRY(h, hipSetDevice(h->device));
    inlm_status s = ensure_samples(h, n);
    if (s != INLM_OK) return s;
    HIP_TRY(h, hipMemcpyAsync(h->samples_dev, samples, sizeof(int64_t) * n, hipMemcpyHostToDevice, h->stream));
    const float *origin;
    int64_t rowp, slp;
    if (which == 0 && h->fast) {
        origin = h->raw;
        rowp = g.nx;
        slp = nslice;
    } else if (which == 0) {
        const long sp = (long)g.nxp * g.nyp;
        origin = h->R_pad + (long)g.pz * sp + (long)g.py * g.nxp + g.px;
        rowp = g.nxp;
        slp = sp;
    } else {
        origin = h->result;
        rowp = g.nx;
        slp = nslice;
    }
    HIP_TRY(h, inlm::launch_patchstd(origin, rowp, slp, h->samples_dev, n, g.nx, nslice, ps, h->std_dev, h->stream));
    HIP_TRY(h, hipMemcpyAsync(out, h->std_dev, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return INLM_OK;
}

inlm_status inlm_estimate_noise(inlm_handle *h, int32_t which, const int64_t *samples, int64_t n, float *sigma_out)
{
    if (!h || !sigma_out) return fail(h, INLM_BAD_ARG, "null argument");
    std::vector<double> sd((size_t)std::max<int64_t>(n, 1));
    if (h->p.noise_mode != INLM_NOISE_MANUAL) {
        inlm_status s = inlm_patchstd(h, which, samples, n, sd.data());
        if (s != INLM_OK) return s;
    }
    int shape[3] = {h->geo.nx, h->geo.ny, h->geo.nz};
    std::string err;
    const inlm_status s = inlm::sigma_from_std(h->p, shape, samples, sd.data(), n, which, h->semi_ref, sigma_out, err);
    if (s != INLM_OK) return fail(h, s, err);
    return INLM_OK;
}

inlm_status inlm_denoise(inlm_handle *h, const float *raw, float *out, float *sigma_log)
{
    if (!h) return fail(h, INLM_BAD_ARG, "null handle");
    inlm_status s;
    if (raw && (s = inlm_load_raw(h, raw, 0)) != INLM_OK) return s;
    if (!h->raw_loaded) return fail(h, INLM_BAD_ARG, "no original volume loaded");
    const int nz = h->geo.nz;
    std::vector<int64_t> samples;
    if (h->p.noise_mode != INLM_NOISE_MANUAL) {
        samples.resize((size_t)h->p.n_samples);
        int sh[3] = {h->geo.nx, h->geo.ny, h->geo.nz};
        if (!inlm::draw_samples(sh, h->p.n_samples, h->p.patchsize, h->p.sample_seed, samples.data(),
                                h->p.circular_mask_diameter))
            return fail(h, INLM_BAD_ARG, "no noise sample fits the volume (or the circular mask)");
    }
    const int64_t ns = (int64_t)samples.size();
    std::vector<float> s0(nz), s1(nz), tmp(nz);
    // main.cpp:325-351: initial noise level on the original.
    if ((s = inlm_estimate_noise(h, 0, samples.data(), ns, s0.data())) != INLM_OK) return s;
    if ((s = inlm_set_sigma(h, s0.data(), nz)) != INLM_OK) return s;
    if ((s = inlm_run_iteration(h, 1, nullptr)) != INLM_OK) return s;
    if (sigma_log) std::memcpy(sigma_log, s0.data(), sizeof(float) * nz);
    auto reestimate = [&]() -> inlm_status {
        // main.cpp:420-448: re-estimate on the current result with the same
        // samples, then sigma = alpha*s0 + (1-alpha)*s1 in float.
        inlm_status r = inlm_estimate_noise(h, 1, samples.data(), ns, tmp.data());
        if (r != INLM_OK) return r;
        const float a = h->p.alpha;
        for (int i = 0; i < nz; i++) s1[i] = a * s0[i] + (1.f - a) * tmp[i];
        return inlm_set_sigma(h, s1.data(), nz);
    };
    if (h->p.maxiterations > 1 && (s = reestimate()) != INLM_OK) return s;
    for (int iter = 2; iter <= h->p.maxiterations; iter++) {
        if ((s = inlm_run_iteration(h, iter, nullptr)) != INLM_OK) return s;
        if (sigma_log) std::memcpy(sigma_log + (size_t)(iter - 1) * nz, s1.data(), sizeof(float) * nz);
        if (h->p.continuous_estimate && h->p.maxiterations > iter + 1 && (s = reestimate()) != INLM_OK) return s;
    }
    if (out) return inlm_get_result(h, out, 0);
    return INLM_OK;
}

inlm_status inlm_tables(const inlm_params *p, const int32_t shape[3], int32_t cap, int32_t *n_search,
                        int32_t *search_dxyz, int32_t *n_patch, int32_t *patch_dxyz, float *weights)
{
    if (!p || !shape || !n_search || !n_patch) return fail(nullptr, INLM_BAD_ARG, "null argument");
    int sh[3] = {shape[0], shape[1], shape[2]};
    inlm::Geometry g = inlm::make_geometry(sh, p->radius_searchspace, p->radius_patchspace, p->nslices);
    std::vector<inlm::Offset3> S = inlm::search_offsets(p->radius_searchspace, p->nslices);
    std::vector<inlm::Offset3> PT = inlm::patch_offsets(p->radius_patchspace, p->nslices);
    std::vector<float> W;
    std::string err;
    if (!inlm::cpu_weight_table(p->radius_patchspace, p->nslices, p->z_anisotropy, g.nzp, p->cpu_blockwise != 0, W,
                                err))
        return fail(nullptr, INLM_BAD_ARG, err);
    *n_search = (int32_t)S.size();
    *n_patch = (int32_t)PT.size();
    for (int i = 0; i < (int)S.size() && i < cap && search_dxyz; i++) {
        search_dxyz[3 * i] = S[i].dx;
        search_dxyz[3 * i + 1] = S[i].dy;
        search_dxyz[3 * i + 2] = S[i].dz;
    }
    for (int i = 0; i < (int)PT.size() && i < cap; i++) {
        if (patch_dxyz) {
            patch_dxyz[3 * i] = PT[i].dx;
            patch_dxyz[3 * i + 1] = PT[i].dy;
            patch_dxyz[3 * i + 2] = PT[i].dz;
        }
        if (weights) weights[i] = W[i];
    }
    return INLM_OK;
}

inlm_status inlm_selftest_division(int32_t device, uint64_t *mismatches, uint64_t *checked, uint32_t *example)
{
    if (!mismatches || !checked) return fail(nullptr, INLM_BAD_ARG, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, INLM_NO_DEVICE, "no GPU visible");
    if (device < 0 || device >= ndev) return fail(nullptr, INLM_BAD_ARG, "device index out of range");
    HIP_TRY(nullptr, hipSetDevice(device));
    unsigned long long *cnt = nullptr;
    HIP_TRY(nullptr, hipMalloc(&cnt, 16));
    // bit patterns of -0 ... the cutoff -3.56648f
    const float cut = -3.56648f;
    uint32_t cb;
    std::memcpy(&cb, &cut, 4);
    const unsigned count = (cb - 0x80000000u) + 1u;
    hipError_t e = hipMemset(cnt, 0, 16);
    if (e == hipSuccess) e = inlm::launch_selftest_div(0u, count, cnt, reinterpret_cast<unsigned *>(cnt + 1), nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    unsigned long long host[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpy(host, cnt, 16, hipMemcpyDeviceToHost);
    (void)hipFree(cnt);
    if (e != hipSuccess) return hip_fail(nullptr, e, "division self-test");
    *mismatches = host[0];
    *checked = count;
    if (example) *example = (uint32_t)host[1];
    return INLM_OK;
}

float inlm_last_filter_ms(const inlm_handle *h) { return h ? h->last_ms : 0.f; }

inlm_status inlm_table_sizes(const inlm_handle *h, int32_t *n_search, int32_t *n_patch)
{
    if (!h) return INLM_BAD_ARG;
    if (n_search) *n_search = (int32_t)h->search.size();
    if (n_patch) *n_patch = h->taps.n;
    return INLM_OK;
}

inlm_status inlm_set_result(inlm_handle *h, int32_t iter, const float *vol, int32_t is_device)
{
    if (!h || !vol) return fail(h, INLM_BAD_ARG, "null argument");
    if (iter < 1) return fail(h, INLM_BAD_ARG, "iterations count from 1");
    if (!h->raw_loaded) return fail(h, INLM_BAD_ARG, "no original volume loaded");
    HIP_TRY(h, hipSetDevice(h->device));
    HIP_TRY(h, hipMemcpyAsync(h->result, vol, sizeof(float) * h->geo.voxels(),
                              is_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    h->last_iter = iter;
    return INLM_OK;
}

// ---------------------------------------------------------------------------
// Learned offset proposal
// ---------------------------------------------------------------------------

int64_t inlm_tile_count(const inlm_handle *h) { return h ? (int64_t)tile_count(h->geo) : 0; }

inlm_status inlm_weight_mass(inlm_handle *h, int32_t iter, float *mass, int32_t is_device)
{
    if (!h || !mass) return fail(h, INLM_BAD_ARG, "null argument");
    if (h->fast || h->rp_tiled == 0)
        return fail(h, INLM_BAD_ARG, "weight mass needs the FP32 form on the tiled kernel (isotropic patch r 1..3)");
    HIP_TRY(h, hipSetDevice(h->device));
    const size_t bytes = sizeof(float) * (size_t)tile_count(h->geo) * h->search.size();
    float *dst = mass;
    if (!is_device) HIP_TRY(h, hipMalloc(&dst, bytes));
    inlm_status s = run_iteration_impl(h, iter, nullptr, dst);
    if (s == INLM_OK && !is_device) {
        hipError_t e = hipMemcpy(mass, dst, bytes, hipMemcpyDeviceToHost);
        if (e != hipSuccess) s = hip_fail(h, e, "hipMemcpy(mass)");
    }
    if (!is_device) (void)hipFree(dst);
    return s;
}

namespace {
inlm_status pfail(inlm_proposal *p, inlm_status st, const std::string &msg)
{
    if (p) p->err = msg;
    else g_create_error = msg;
    return st;
}
inlm_status phip(inlm_proposal *p, hipError_t e, const char *where)
{
    return pfail(p, e == hipErrorOutOfMemory ? INLM_OOM : INLM_HIP_ERROR, std::string(where) + ": " +
                                                                              hipGetErrorString(e));
}
}  // namespace

inlm_status inlm_proposal_create(int32_t n_search, int32_t device, uint64_t seed, inlm_proposal **out)
{
    if (!out) return pfail(nullptr, INLM_BAD_ARG, "null argument");
    *out = nullptr;
    if (n_search < 1 || n_search > inlm::kMaxProposalOffsets)
        return pfail(nullptr, INLM_BAD_ARG, "n_search must be in [1, 2048]");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return pfail(nullptr, INLM_NO_DEVICE, "no GPU visible");
    if (device < 0 || device >= ndev) return pfail(nullptr, INLM_BAD_ARG, "device index out of range");
    inlm_proposal *p = new inlm_proposal();
    p->device = device;
    p->ns = n_search;
    hipError_t e = p->net.init(n_search, device, seed);
    if (e != hipSuccess) {
        inlm_status st = phip(nullptr, e, "proposal init");
        delete p;
        return st;
    }
    *out = p;
    return INLM_OK;
}

inlm_status inlm_proposal_destroy(inlm_proposal *p)
{
    delete p;
    return INLM_OK;
}

const char *inlm_proposal_last_error(const inlm_proposal *p) { return p ? p->err.c_str() : g_create_error.c_str(); }

int64_t inlm_proposal_num_params(const inlm_proposal *p) { return p ? (int64_t)p->net.num_params() : 0; }

inlm_status inlm_proposal_params(inlm_proposal *p, float *params, int32_t set)
{
    if (!p || !params) return pfail(p, INLM_BAD_ARG, "null argument");
    hipError_t e = set ? p->net.set_params(params, p->net.num_params()) : p->net.get_params(params, p->net.num_params());
    return e == hipSuccess ? INLM_OK : phip(p, e, "proposal params");
}

inlm_status inlm_proposal_grads(const inlm_proposal *p, float *grads)
{
    if (!p || !grads) return pfail(const_cast<inlm_proposal *>(p), INLM_BAD_ARG, "null argument");
    hipError_t e = p->net.get_grads(grads, p->net.num_params());
    return e == hipSuccess ? INLM_OK : phip(const_cast<inlm_proposal *>(p), e, "proposal grads");
}

inlm_status inlm_proposal_train_step(inlm_proposal *p, const float *vol, const int32_t shape[3], const float *mass,
                                     float lr, float *loss)
{
    if (!p || !vol || !shape || !mass) return pfail(p, INLM_BAD_ARG, "null argument");
    if (shape[0] < 1 || shape[1] < 1 || shape[2] < 1) return pfail(p, INLM_BAD_ARG, "shape must be positive");
    const int sh[3] = {shape[0], shape[1], shape[2]};
    float l = 0.f;
    hipError_t e = p->net.run(vol, sh, mass, true, &l);
    if (e != hipSuccess) return phip(p, e, "proposal forward/backward");
    if (loss) *loss = l;
    if (lr > 0.f && (e = p->net.adam(lr)) != hipSuccess) return phip(p, e, "proposal update");
    return INLM_OK;
}

inlm_status inlm_proposal_predict(inlm_proposal *p, const float *vol, const int32_t shape[3], float *probs,
                                  int32_t is_device)
{
    if (!p || !vol || !shape || !probs) return pfail(p, INLM_BAD_ARG, "null argument");
    if (shape[0] < 1 || shape[1] < 1 || shape[2] < 1) return pfail(p, INLM_BAD_ARG, "shape must be positive");
    const int sh[3] = {shape[0], shape[1], shape[2]};
    hipError_t e = p->net.run(vol, sh, nullptr, false, nullptr);
    if (e != hipSuccess) return phip(p, e, "proposal forward");
    const long nt = (long)((sh[0] + 31) / 32) * ((sh[1] + 7) / 8) * ((sh[2] + 7) / 8);
    e = hipMemcpy(probs, p->net.probs(), sizeof(float) * nt * p->ns,
                  is_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost);
    return e == hipSuccess ? INLM_OK : phip(p, e, "hipMemcpy(probs)");
}

inlm_status inlm_set_proposal(inlm_handle *h, inlm_proposal *p, int32_t K, float threshold, float audit_frac,
                              uint64_t seed)
{
    if (!h) return fail(h, INLM_BAD_ARG, "null handle");
    HIP_TRY(h, hipSetDevice(h->device));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    if (!p) {
        clear_pruning(h);
        return INLM_OK;
    }
    if (!h->fast) return fail(h, INLM_BAD_ARG, "offset pruning needs a matrix-core form");
    if (!h->raw_loaded) return fail(h, INLM_BAD_ARG, "no original volume loaded");
    const int ns = (int)h->search.size();
    if (p->ns != ns) return fail(h, INLM_BAD_ARG, "proposal was built for a different number of search offsets");
    if (p->device != h->device) return fail(h, INLM_BAD_ARG, "proposal lives on another device");
    if (K < 1 || K > ns) return fail(h, INLM_BAD_ARG, "K must be in [1, n_search]");
    if (!(threshold >= 0.f && threshold <= 1.f) || !(audit_frac >= 0.f && audit_frac <= 1.f))
        return fail(h, INLM_BAD_ARG, "threshold and audit_frac must lie in [0, 1]");
    clear_pruning(h);
    inlm_status s = ensure_lists(h, K);
    if (s != INLM_OK) return s;
    const long nt = tile_count(h->geo);
    HIP_TRY(h, hipMalloc(&h->probs_dev, sizeof(float) * nt * ns));
    const int sh[3] = {h->geo.nx, h->geo.ny, h->geo.nz};
    hipError_t e = p->net.run(h->raw, sh, nullptr, false, nullptr);
    if (e != hipSuccess) {
        clear_pruning(h);
        return hip_fail(h, e, "proposal forward");
    }
    HIP_TRY(h, hipMemcpy(h->probs_dev, p->net.probs(), sizeof(float) * nt * ns, hipMemcpyDeviceToDevice));
    h->prop_K = K;
    h->prop_thr = threshold;
    h->prop_audit = audit_frac;
    h->prop_seed = seed;
    return INLM_OK;
}

inlm_status inlm_set_offset_lists(inlm_handle *h, const int32_t *lists, const int32_t *counts, int32_t K)
{
    if (!h || !lists || !counts) return fail(h, INLM_BAD_ARG, "null argument");
    if (!h->fast) return fail(h, INLM_BAD_ARG, "offset pruning needs a matrix-core form");
    const int ns = (int)h->search.size();
    if (K < 1 || K > ns) return fail(h, INLM_BAD_ARG, "K must be in [1, n_search]");
    const long nt = tile_count(h->geo);
    for (long t = 0; t < nt; t++) {
        if (counts[t] < 1 || counts[t] > ns) return fail(h, INLM_BAD_ARG, "counts must lie in [1, n_search]");
        const int used = counts[t] < ns ? counts[t] : 0;
        if (used > K) return fail(h, INLM_BAD_ARG, "a count exceeds K");
        for (int j = 0; j < used; j++)
            if (lists[t * K + j] < 0 || lists[t * K + j] >= ns)
                return fail(h, INLM_BAD_ARG, "list entry outside the search table");
    }
    HIP_TRY(h, hipSetDevice(h->device));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    clear_pruning(h);
    inlm_status s = ensure_lists(h, K);
    if (s != INLM_OK) return s;
    HIP_TRY(h, hipMemcpy(h->lists_dev, lists, sizeof(int) * nt * K, hipMemcpyHostToDevice));
    HIP_TRY(h, hipMemcpy(h->counts_dev, counts, sizeof(int) * nt, hipMemcpyHostToDevice));
    h->lists_fixed = true;
    return INLM_OK;
}

inlm_status inlm_proposal_stats(const inlm_handle *h, float out[5])
{
    if (!h || !out) return INLM_BAD_ARG;
    for (int i = 0; i < 5; i++) out[i] = h->stats[i];
    return INLM_OK;
}

}  // extern "C"
