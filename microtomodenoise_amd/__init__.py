"""MI355X-native iterNLM: the iterative non-local-means filter of
iternlm/microtomodenoise as hand-written gfx950 HIP kernels behind a C ABI
(include/inlm_capi.h), with a Python mirror of the reference's
``denoise::IterativeNLM_GPU`` / ``protocol::DenoiseParameters`` interface."""
from .denoiser import CPU, GPU, DenoiseParameters, IterativeNLM_GPU, NoiseLevel, denoise, denoise_rgb, denoise_slabs  # noqa: F401
from ._capi import InlmError, draw_samples  # noqa: F401
