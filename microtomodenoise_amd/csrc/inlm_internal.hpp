// Library-internal helpers shared between translation units.
#pragma once
#include <cstddef>
#include <functional>
#include <string>

#include "inlm_capi.h"

namespace inlm {

// message returned by inlm_last_error(NULL)
void set_global_error(const std::string &msg);

// Anscombe transform and its algebraic inverse (Color/colorspace.h:87-123)
void anscombe(float *v, size_t n);
void inverse_anscombe(float *v, size_t n);

// The colour schedule; on_iteration(iter, channels) sees (and may change)
// each iteration's result on the host before it is used further.
inlm_status denoise_rgb(const inlm_params *p, const inlm_rgb_options *c, const int32_t shape[2], int32_t device,
                        const float *in, float *out, float *sigma_log,
                        const std::function<bool(int, float *)> &on_iteration);

}  // namespace inlm
