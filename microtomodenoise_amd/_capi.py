"""ctypes binding of ``libinlm.so`` (the C ABI declared in include/inlm_capi.h).

The library is loaded from this package directory.  There is no fallback:
when the library is missing, importing the binding raises, so no call can
silently run anything other than the HIP path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libinlm.so")

INLM_OK, INLM_NO_DEVICE, INLM_HIP_ERROR, INLM_BAD_ARG, INLM_OOM, INLM_IO_ERROR = 0, -1, -2, -3, -4, -5
NOISE_GLOBAL, NOISE_ZADAPTIVE, NOISE_MANUAL, NOISE_SEMIMANUAL = 0, 1, 2, 3
PREC_FP32, PREC_FP32_FAST, PREC_BF16, PREC_FP16, PREC_FP16_FP8 = 0, 1, 2, 3, 4
WEIGHT_PADE, WEIGHT_EXP = 0, 1

STATUS_NAMES = {0: "OK", -1: "NO_DEVICE", -2: "HIP_ERROR", -3: "BAD_ARG", -4: "OOM", -5: "IO_ERROR"}

# Every symbol include/inlm_capi.h declares.
EXPORTS = [
    "inlm_default_params", "inlm_device_count", "inlm_create", "inlm_destroy", "inlm_last_error",
    "inlm_set_sigma", "inlm_load_raw", "inlm_run_iteration", "inlm_get_result", "inlm_draw_samples",
    "inlm_patchstd", "inlm_estimate_noise", "inlm_denoise", "inlm_tables", "inlm_selftest_division",
    "inlm_last_filter_ms", "inlm_table_sizes",
    # learned offset proposal
    "inlm_tile_count", "inlm_weight_mass", "inlm_proposal_create", "inlm_proposal_destroy",
    "inlm_proposal_last_error", "inlm_proposal_num_params", "inlm_proposal_params", "inlm_proposal_grads",
    "inlm_proposal_train_step", "inlm_proposal_predict", "inlm_set_proposal", "inlm_set_offset_lists",
    "inlm_proposal_stats",
    # resume, volume I/O and the program's output contract
    "inlm_set_result", "inlm_tiff_stack_shape", "inlm_tiff_read_stack", "inlm_tiff_write_stack",
    "inlm_noise_csv_write", "inlm_noise_csv_read", "inlm_iteration_on_disk", "inlm_default_io_options",
    "inlm_denoise_files", "inlm_denoise_slabs", "inlm_default_rgb_options", "inlm_denoise_rgb",
    "inlm_rgb_tiff_shape", "inlm_rgb_tiff_read", "inlm_rgb_tiff_write", "inlm_denoise_rgb_files",
    # noise estimator host half, circular sampling mask, slab drivers
    "inlm_draw_samples_mask", "inlm_sigma_from_patchstd", "inlm_capacity_slices", "inlm_iterate_slabs",
    # patch standard deviations without device allocation
    "inlm_patchstd_noalloc",
]
NOALLOC_MAX_SAMPLES = 262144  # INLM_NOALLOC_MAX_SAMPLES


class InlmRgbOptions(C.Structure):
    """Mirror of ``inlm_rgb_options`` (include/inlm_capi.h)."""
    _fields_ = [
        ("independent_channels", C.c_int32),
        ("stabilize_variance", C.c_int32),
        ("average_channelsigma", C.c_int32),
        ("reserved", C.c_int32 * 5),
    ]


class InlmIoOptions(C.Structure):
    """Mirror of ``inlm_io_options`` (include/inlm_capi.h)."""
    _fields_ = [
        ("save_16bit", C.c_int32),
        ("resume", C.c_int32),
        ("cleanup", C.c_int32),
        ("firstslice", C.c_int32),
        ("lastslice", C.c_int32),
        ("verbose", C.c_int32),
        ("reserved", C.c_int32 * 8),
    ]


class InlmParams(C.Structure):
    """Mirror of ``inlm_params`` (include/inlm_capi.h)."""
    _fields_ = [
        ("alpha", C.c_float),
        ("maxiterations", C.c_int32),
        ("nslices", C.c_int32),
        ("radius_searchspace", C.c_int32 * 3),
        ("radius_patchspace", C.c_int32 * 3),
        ("beta", C.c_float),
        ("z_anisotropy", C.c_float),
        ("noise_mode", C.c_int32),
        ("zwindow", C.c_int32),
        ("sigma", C.c_float * 2),
        ("n_samples", C.c_int64),
        ("patchsize", C.c_int32),
        ("stds_from_mean", C.c_float),
        ("continuous_estimate", C.c_int32),
        ("sample_seed", C.c_uint64),
        ("cpu_blockwise", C.c_int32),
        ("precision", C.c_int32),
        ("weight_fn", C.c_int32),
        ("kernel_select", C.c_int32),
        ("circular_mask_diameter", C.c_int32),
        ("reserved", C.c_int32 * 6),
    ]


class InlmError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m microtomodenoise_amd.build`")
        L = C.CDLL(LIB_PATH)
        H = C.c_void_p
        f32p, f64p, i64p, i32p = (C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int32))
        L.inlm_default_params.argtypes = [C.POINTER(InlmParams)]
        L.inlm_default_params.restype = None
        L.inlm_device_count.restype = C.c_int32
        L.inlm_create.argtypes = [C.POINTER(InlmParams), i32p, C.c_int32, C.POINTER(H)]
        L.inlm_destroy.argtypes = [H]
        L.inlm_last_error.argtypes = [H]
        L.inlm_last_error.restype = C.c_char_p
        L.inlm_set_sigma.argtypes = [H, f32p, C.c_int32]
        L.inlm_load_raw.argtypes = [H, C.c_void_p, C.c_int32]
        L.inlm_run_iteration.argtypes = [H, C.c_int32, f32p]
        L.inlm_get_result.argtypes = [H, C.c_void_p, C.c_int32]
        L.inlm_draw_samples.argtypes = [i32p, C.c_int64, C.c_int32, C.c_uint64, i64p]
        L.inlm_patchstd.argtypes = [H, C.c_int32, i64p, C.c_int64, f64p]
        L.inlm_patchstd_noalloc.argtypes = [C.c_void_p, i32p, C.c_int32, i64p, C.c_int64, f64p, C.c_int32]
        L.inlm_estimate_noise.argtypes = [H, C.c_int32, i64p, C.c_int64, f32p]
        L.inlm_denoise.argtypes = [H, f32p, f32p, f32p]
        L.inlm_last_filter_ms.argtypes = [H]
        L.inlm_last_filter_ms.restype = C.c_float
        L.inlm_table_sizes.argtypes = [H, i32p, i32p]
        L.inlm_tables.argtypes = [C.POINTER(InlmParams), i32p, C.c_int32, i32p, i32p, i32p, i32p, f32p]
        L.inlm_selftest_division.argtypes = [C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                             C.POINTER(C.c_uint32)]
        L.inlm_tile_count.argtypes = [H]
        L.inlm_tile_count.restype = C.c_int64
        L.inlm_weight_mass.argtypes = [H, C.c_int32, C.c_void_p, C.c_int32]
        L.inlm_proposal_create.argtypes = [C.c_int32, C.c_int32, C.c_uint64, C.POINTER(H)]
        L.inlm_proposal_destroy.argtypes = [H]
        L.inlm_proposal_last_error.argtypes = [H]
        L.inlm_proposal_last_error.restype = C.c_char_p
        L.inlm_proposal_num_params.argtypes = [H]
        L.inlm_proposal_num_params.restype = C.c_int64
        L.inlm_proposal_params.argtypes = [H, f32p, C.c_int32]
        L.inlm_proposal_grads.argtypes = [H, f32p]
        L.inlm_proposal_train_step.argtypes = [H, C.c_void_p, i32p, C.c_void_p, C.c_float, f32p]
        L.inlm_proposal_predict.argtypes = [H, C.c_void_p, i32p, C.c_void_p, C.c_int32]
        L.inlm_set_proposal.argtypes = [H, H, C.c_int32, C.c_float, C.c_float, C.c_uint64]
        L.inlm_set_offset_lists.argtypes = [H, i32p, i32p, C.c_int32]
        L.inlm_proposal_stats.argtypes = [H, f32p]
        L.inlm_set_result.argtypes = [H, C.c_int32, C.c_void_p, C.c_int32]
        L.inlm_tiff_stack_shape.argtypes = [C.c_char_p, i32p]
        L.inlm_tiff_read_stack.argtypes = [C.c_char_p, C.c_int32, C.c_int32, f32p]
        L.inlm_tiff_write_stack.argtypes = [C.c_char_p, C.c_char_p, f32p, i32p, C.c_int32, C.c_int32]
        L.inlm_noise_csv_write.argtypes = [C.c_char_p, C.c_char_p, f32p, C.c_int32]
        L.inlm_noise_csv_read.argtypes = [C.c_char_p, C.c_int32, f32p]
        L.inlm_iteration_on_disk.argtypes = [C.c_char_p, C.c_int32, C.c_int32]
        L.inlm_iteration_on_disk.restype = C.c_int32
        L.inlm_default_io_options.argtypes = [C.POINTER(InlmIoOptions)]
        L.inlm_default_io_options.restype = None
        L.inlm_denoise_files.argtypes = [C.POINTER(InlmParams), C.c_int32, C.c_char_p, C.c_char_p,
                                         C.POINTER(InlmIoOptions)]
        L.inlm_default_rgb_options.argtypes = [C.POINTER(InlmRgbOptions)]
        L.inlm_default_rgb_options.restype = None
        L.inlm_denoise_rgb.argtypes = [C.POINTER(InlmParams), C.POINTER(InlmRgbOptions), C.POINTER(C.c_int32),
                                       C.c_int32, f32p, f32p, f32p]
        L.inlm_rgb_tiff_shape.argtypes = [C.c_char_p, C.POINTER(C.c_int32)]
        L.inlm_rgb_tiff_read.argtypes = [C.c_char_p, f32p]
        L.inlm_rgb_tiff_write.argtypes = [C.c_char_p, f32p, C.POINTER(C.c_int32)]
        L.inlm_denoise_rgb_files.argtypes = [C.POINTER(InlmParams), C.POINTER(InlmRgbOptions), C.c_int32, C.c_char_p,
                                             C.c_char_p, C.POINTER(InlmIoOptions)]
        L.inlm_denoise_slabs.argtypes = [C.POINTER(InlmParams), C.POINTER(C.c_int32), C.c_int32, f32p, f32p,
                                         C.c_int32, f32p]
        L.inlm_draw_samples_mask.argtypes = [i32p, C.c_int64, C.c_int32, C.c_int32, C.c_uint64, i64p]
        L.inlm_sigma_from_patchstd.argtypes = [C.POINTER(InlmParams), i32p, i64p, f64p, C.c_int64, C.c_int32, i64p,
                                               f32p]
        L.inlm_capacity_slices.argtypes = [C.POINTER(InlmParams), i32p, C.c_int32, C.c_int64, i32p]
        L.inlm_iterate_slabs.argtypes = [C.POINTER(InlmParams), i32p, C.c_int32, C.c_int32, f32p, f32p, f32p, f32p,
                                         C.c_int32]
        _LIB = L
    return _LIB


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def check(status: int, handle=None):
    if status != INLM_OK:
        msg = lib().inlm_last_error(handle)
        raise InlmError(status, msg.decode() if msg else "")


def check_proposal(status: int, prop=None):
    if status != INLM_OK:
        msg = lib().inlm_proposal_last_error(prop)
        raise InlmError(status, msg.decode() if msg else "")


def default_params() -> InlmParams:
    p = InlmParams()
    lib().inlm_default_params(C.byref(p))
    return p


def patchstd_noalloc(vol_dev_ptr: int, shape_xyz, patchsize: int, samples, device: int = 0) -> np.ndarray:
    """Patch standard deviations of a device volume (unpadded float32, x
    fastest) at the sample corners, through the entry point that allocates
    no device memory (at most NOALLOC_MAX_SAMPLES samples per call)."""
    s = np.ascontiguousarray(samples, np.int64)
    out = np.zeros(len(s), np.float64)
    sh = np.asarray(shape_xyz, np.int32)
    check(lib().inlm_patchstd_noalloc(C.c_void_p(vol_dev_ptr), ptr(sh, C.c_int32), patchsize, ptr(s, C.c_int64),
                                      len(s), ptr(out, C.c_double), device))
    return out


def draw_samples(shape_xyz, n: int, patchsize: int = 15, seed: int = 0, mask_diameter: int = 0) -> np.ndarray:
    """Sample positions of the noise estimator (host-only, no GPU needed);
    mask_diameter > 0 keeps them inside the circular field of view."""
    out = np.zeros(n, np.int64)
    sh = np.asarray(shape_xyz, np.int32)
    if mask_diameter > 0:
        check(lib().inlm_draw_samples_mask(ptr(sh, C.c_int32), n, patchsize, mask_diameter, seed,
                                           ptr(out, C.c_int64)))
    else:
        check(lib().inlm_draw_samples(ptr(sh, C.c_int32), n, patchsize, seed, ptr(out, C.c_int64)))
    return out


def sigma_from_patchstd(params: InlmParams, shape_xyz, samples, stds, which: int = 0, semimanual_ref=None):
    """The host half of the noise estimator (inlm_sigma_from_patchstd):
    sigma per slice from per-sample patch stds.  semimanual_ref is a
    one-element int64 array that carries the semimanual reference sample
    from the first estimate (which=0) to later ones."""
    sh = np.asarray(shape_xyz, np.int32)
    smp = np.ascontiguousarray(samples, np.int64)
    sd = np.ascontiguousarray(stds, np.float64)
    out = np.zeros(int(sh[2]), np.float32)
    ref = semimanual_ref if semimanual_ref is not None else np.full(1, -1, np.int64)
    check(lib().inlm_sigma_from_patchstd(C.byref(params), ptr(sh, C.c_int32), ptr(smp, C.c_int64),
                                         ptr(sd, C.c_double), len(smp), which, ptr(ref, C.c_int64),
                                         ptr(out, C.c_float)))
    return out


def capacity_slices(params: InlmParams, shape_xyz, device: int = 0, memory_limit: int = 0) -> int:
    """Output slices one slab may hold within memory_limit device bytes (the
    free device memory when 0): the sizing of configure_device."""
    sh = np.asarray(shape_xyz, np.int32)
    n = C.c_int32()
    check(lib().inlm_capacity_slices(C.byref(params), ptr(sh, C.c_int32), device, memory_limit, C.byref(n)))
    return n.value


def tables(params: InlmParams, shape_xyz):
    """Host-only table query (inlm_tables): search offsets, patch taps and
    weights as the library builds them."""
    cap = 16384
    sd = np.zeros(3 * cap, np.int32)
    pd = np.zeros(3 * cap, np.int32)
    w = np.zeros(cap, np.float32)
    ns, npt = C.c_int32(), C.c_int32()
    sh = np.asarray(shape_xyz, np.int32)
    check(lib().inlm_tables(C.byref(params), ptr(sh, C.c_int32), cap, C.byref(ns), ptr(sd, C.c_int32),
                            C.byref(npt), ptr(pd, C.c_int32), ptr(w, C.c_float)))
    return (sd[: 3 * ns.value].reshape(-1, 3).copy(), pd[: 3 * npt.value].reshape(-1, 3).copy(),
            w[: npt.value].copy())
