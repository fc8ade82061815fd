// Library-internal helpers shared between translation units.
#pragma once
#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>

#include "inlm_capi.h"

namespace inlm {

// message returned by inlm_last_error(NULL)
void set_global_error(const std::string &msg);

// The host half of get_noiselevel (noiselevel.cpp:9-95): sigma per slice
// from the per-sample patch stds sd[0..n) of the sorted sample list, in the
// parameters' noise mode.  which = 0: first estimate; semi_ref keeps the
// semimanual reference sample between estimates.
inlm_status sigma_from_std(const inlm_params &p, const int shape[3], const int64_t *samples, const double *sd,
                           int64_t n, int which, int64_t &semi_ref, float *sig, std::string &err);

// Device bytes a handle of these parameters allocates for `shape`.
size_t handle_bytes(const inlm_params &p, const int32_t shape[3]);

// Anscombe transform and its algebraic inverse (Color/colorspace.h:87-123)
void anscombe(float *v, size_t n);
void inverse_anscombe(float *v, size_t n);

// The colour schedule; on_iteration(iter, channels) sees (and may change)
// each iteration's result on the host before it is used further.
inlm_status denoise_rgb(const inlm_params *p, const inlm_rgb_options *c, const int32_t shape[2], int32_t device,
                        const float *in, float *out, float *sigma_log,
                        const std::function<bool(int, float *)> &on_iteration);

}  // namespace inlm
