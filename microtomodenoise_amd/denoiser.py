"""Python mirror of the reference's denoiser interface over the C ABI.

``DenoiseParameters`` restates ``protocol::DenoiseParameters``
(iterNLM_v0.3/denoise_parameters.h:9-74) with the same field names and
defaults.  ``IterativeNLM_GPU`` keeps the method names and argument meaning of
``denoise::IterativeNLM_GPU`` (Denoiser/iternlm_gpu.h:20-56):
``configure_device`` returns the number of slices one slab may hold
(shape[2] when the whole volume fits the free device memory, or
``gpu.memory_limit`` bytes), -1 when no GPU is visible and -2 on a HIP
error, like the reference (iternlm_gpu.cu:1443-1444, :1528-1566,
:1608-1618); below shape[2] the caller takes
``Run_GaussianNoise_GPUBlocks``.  Unlike the reference, failures of the other methods raise
``InlmError`` instead of printing and returning.

Volumes are numpy float32 arrays of shape ``(nz, ny, nx)`` (x fastest); the
``shape`` arguments are ``[nx, ny, nz]`` as in the reference.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _capi as capi


@dataclass
class NoiseLevel:
    mode: str = "z-adaptive"  # "z-adaptive", "global" ("auto"), "manual", "semimanual"
    zwindow: int = 100
    sigma: List[float] = field(default_factory=lambda: [0.0, 0.0])
    n_samples: int = 100000
    patchsize: int = 15
    stds_from_mean: float = 1.0
    continuous_estimate: bool = False
    circular_mask_diameter: int = 0


@dataclass
class CPU:
    max_threads: int = 128
    unrolled_patchspace: bool = False
    blockwise: bool = True


@dataclass
class GPU:
    n_gpus: int = 2
    deviceID: int = 0
    threadsPerBlock: int = 128
    memory_buffer: int = 0
    blockwise_host: bool = False
    # device bytes configure_device may plan for (0: the free device memory)
    memory_limit: int = 0


@dataclass
class DenoiseParameters:
    alpha: float = 0.5
    maxiterations: int = 4
    nslices: int = 11
    radius_searchspace: List[int] = field(default_factory=lambda: [10, 10, 10])
    radius_patchspace: List[int] = field(default_factory=lambda: [1, 1, 1])
    beta: float = 1.0
    z_anisotropy: float = 1.0
    noiselevel: NoiseLevel = field(default_factory=NoiseLevel)
    cpu: CPU = field(default_factory=CPU)
    gpu: GPU = field(default_factory=GPU)
    # build-specific choices
    precision: str = "fp32"  # fp32 (parity), fp32_fast, bf16, fp16, fp16_fp8
    weight_fn: str = "pade"  # pade (reference CPU), exp (reference GPU)
    sample_seed: int = 0
    kernel_select: int = 0  # 0 tiled where available, 1 per-voxel

    def to_c(self) -> capi.InlmParams:
        p = capi.default_params()
        p.alpha = self.alpha
        p.maxiterations = self.maxiterations
        p.nslices = self.nslices
        for i in range(3):
            p.radius_searchspace[i] = int(self.radius_searchspace[i])
            p.radius_patchspace[i] = int(self.radius_patchspace[i])
        p.beta = self.beta
        p.z_anisotropy = self.z_anisotropy
        mode = self.noiselevel.mode
        if mode in ("global", "auto"):
            p.noise_mode = capi.NOISE_GLOBAL
        elif mode == "z-adaptive":
            p.noise_mode = capi.NOISE_ZADAPTIVE
        elif mode == "manual":
            p.noise_mode = capi.NOISE_MANUAL
        elif mode == "semimanual":
            p.noise_mode = capi.NOISE_SEMIMANUAL
        else:
            raise ValueError(f"unsupported noise mode {mode!r}")
        p.zwindow = self.noiselevel.zwindow
        p.sigma[0], p.sigma[1] = self.noiselevel.sigma
        p.n_samples = int(self.noiselevel.n_samples)
        p.patchsize = self.noiselevel.patchsize
        p.stds_from_mean = self.noiselevel.stds_from_mean
        p.continuous_estimate = int(self.noiselevel.continuous_estimate)
        p.circular_mask_diameter = int(self.noiselevel.circular_mask_diameter)
        p.sample_seed = int(self.sample_seed)
        p.cpu_blockwise = int(self.cpu.blockwise)
        p.precision = {"fp32": capi.PREC_FP32, "fp32_fast": capi.PREC_FP32_FAST, "bf16": capi.PREC_BF16,
                       "fp16": capi.PREC_FP16, "fp16_fp8": capi.PREC_FP16_FP8}[self.precision]
        p.weight_fn = {"pade": capi.WEIGHT_PADE, "exp": capi.WEIGHT_EXP}[self.weight_fn]
        p.kernel_select = int(self.kernel_select)
        return p


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


class IterativeNLM_GPU:
    """Drop-in for denoise::IterativeNLM_GPU on one MI355X."""

    def __init__(self):
        self.n_parallelslices = 0
        self.resumed = False
        self._h = C.c_void_p()
        self._shape = None
        self._sigma = None

    # -- iternlm_gpu.cu:1440-1623 -------------------------------------------
    def configure_device(self, shape, params: DenoiseParameters) -> int:
        L = capi.lib()
        if L.inlm_device_count() < 1:
            params.gpu.n_gpus = 0
            return -1
        sh = np.asarray(shape, np.int32)
        cp = params.to_c()
        n = C.c_int32()
        st = L.inlm_capacity_slices(C.byref(cp), capi.ptr(sh, C.c_int32), params.gpu.deviceID,
                                    int(params.gpu.memory_limit), C.byref(n))
        if st == capi.INLM_OK and n.value >= int(sh[2]):
            st = L.inlm_create(C.byref(cp), capi.ptr(sh, C.c_int32), params.gpu.deviceID, C.byref(self._h))
        if st == capi.INLM_NO_DEVICE:
            params.gpu.n_gpus = 0
            return -1
        if st != capi.INLM_OK:
            self.error = L.inlm_last_error(None).decode()
            if st == capi.INLM_BAD_ARG:
                raise capi.InlmError(st, self.error)
            return -2
        params.gpu.n_gpus = 1
        self._maxit = int(params.maxiterations)
        self._shape = tuple(int(v) for v in shape)
        self._params = params
        self._device = params.gpu.deviceID
        self.n_parallelslices = n.value
        return self.n_parallelslices

    def free_device(self):
        if self._h:
            capi.lib().inlm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.free_device()
        except Exception:
            pass

    def _check(self, st):
        capi.check(st, self._h)

    # -- iternlm_gpu.cu:2341-2355 -------------------------------------------
    def set_sigma(self, sigmalist, shape=None):
        s = _f32(sigmalist)
        self._sigma = s.copy()
        if self._h:
            self._check(capi.lib().inlm_set_sigma(self._h, capi.ptr(s, C.c_float), len(s)))

    # -- iternlm_gpu.cu:1729-1882 -------------------------------------------
    def Run_GaussianNoise_GPUBlocks(self, iter: int, instack, previous, shape=None, params=None) -> np.ndarray:
        """One iteration over slabs of n_parallelslices slices on host
        arrays; returns the new iterate (the reference writes it into
        `previous`, which this mirror also does when it is a float32 array)."""
        raw = _f32(instack)
        prev = _f32(previous) if previous is not None else raw
        out = iterate_slabs(raw, prev, self._sigma, params or self._params, iter, max(1, self.n_parallelslices),
                            self._device)
        if isinstance(previous, np.ndarray) and previous.dtype == np.float32 and previous.shape == out.shape:
            previous[...] = out
        return out

    # -- iternlm_gpu.cu:1651-1728 -------------------------------------------
    def Run_GaussianNoise(self, iter: int, instack=None, shape=None, params=None):
        raw = None
        if iter == 1 and instack is not None:
            raw = _f32(instack)
        self._check(capi.lib().inlm_run_iteration(self._h, iter, capi.ptr(raw, C.c_float) if raw is not None
                                                  else None))

    # -- iternlm_gpu.cu:2356-2368 -------------------------------------------
    def get_result(self, output: Optional[np.ndarray] = None, shape=None) -> np.ndarray:
        nx, ny, nz = self._shape
        if output is None:
            output = np.empty((nz, ny, nx), np.float32)
        assert output.dtype == np.float32 and output.flags.c_contiguous and output.size == nx * ny * nz
        self._check(capi.lib().inlm_get_result(self._h, output.ctypes.data_as(C.c_void_p), 0))
        return output

    # -- device-resident variants ---------------------------------------------
    def load_raw_device(self, dev_ptr: int):
        self._check(capi.lib().inlm_load_raw(self._h, C.c_void_p(dev_ptr), 1))

    def get_result_device(self, dev_ptr: int):
        self._check(capi.lib().inlm_get_result(self._h, C.c_void_p(dev_ptr), 1))

    def set_result(self, iter: int, vol):
        """Make `vol` the result of iteration `iter` (the resume path of
        Run_GaussianNoise, iternlm_gpu.cu:1665-1697)."""
        v = _f32(vol)
        self._check(capi.lib().inlm_set_result(self._h, iter, v.ctypes.data_as(C.c_void_p), 0))

    def load_raw(self, instack):
        raw = _f32(instack)
        self._check(capi.lib().inlm_load_raw(self._h, raw.ctypes.data_as(C.c_void_p), 0))

    # -- noise estimator ------------------------------------------------------
    def patchstd(self, samples, which: int = 0) -> np.ndarray:
        s = np.ascontiguousarray(samples, np.int64)
        out = np.zeros(len(s), np.float64)
        self._check(capi.lib().inlm_patchstd(self._h, which, capi.ptr(s, C.c_int64), len(s),
                                             capi.ptr(out, C.c_double)))
        return out

    def estimate_noise(self, samples, which: int = 0) -> np.ndarray:
        s = np.ascontiguousarray(samples, np.int64)
        out = np.zeros(self._shape[2], np.float32)
        self._check(capi.lib().inlm_estimate_noise(self._h, which, capi.ptr(s, C.c_int64), len(s),
                                                   capi.ptr(out, C.c_float)))
        return out

    def denoise(self, instack=None):
        """The main.cpp:325-545 schedule; returns (result, sigma per iteration)."""
        nx, ny, nz = self._shape
        out = np.empty((nz, ny, nx), np.float32)
        raw = _f32(instack) if instack is not None else None
        L = capi.lib()
        log = np.zeros((max(self._maxit, 1), nz), np.float32)
        self._check(L.inlm_denoise(self._h, capi.ptr(raw, C.c_float) if raw is not None else None,
                                   capi.ptr(out, C.c_float), capi.ptr(log, C.c_float)))
        return out, log

    def last_filter_ms(self) -> float:
        return float(capi.lib().inlm_last_filter_ms(self._h))

    def table_sizes(self):
        a, b = C.c_int32(), C.c_int32()
        self._check(capi.lib().inlm_table_sizes(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # -- learned offset proposal (no counterpart in the reference) -------------
    def tile_count(self) -> int:
        return int(capi.lib().inlm_tile_count(self._h))

    def weight_mass(self, iter: int, dev_ptr: Optional[int] = None):
        """Runs iteration `iter` (FP32 form) and returns the per-tile weight
        mass of every search offset ([tiles, n_search] floats), or writes it
        to device memory at dev_ptr."""
        if dev_ptr is not None:
            self._check(capi.lib().inlm_weight_mass(self._h, iter, C.c_void_p(dev_ptr), 1))
            return None
        ns, _ = self.table_sizes()
        out = np.empty((self.tile_count(), ns), np.float32)
        self._check(capi.lib().inlm_weight_mass(self._h, iter, out.ctypes.data_as(C.c_void_p), 0))
        return out

    def set_proposal(self, proposal, K: int, threshold: float = 0.95, audit_frac: float = 0.02, seed: int = 0):
        """Attach a trained OffsetProposal (None detaches)."""
        ph = proposal.handle if proposal is not None else None
        self._check(capi.lib().inlm_set_proposal(self._h, ph, int(K), float(threshold), float(audit_frac),
                                                 int(seed)))

    def set_offset_lists(self, lists, counts, K: int):
        lists = np.ascontiguousarray(lists, np.int32)
        counts = np.ascontiguousarray(counts, np.int32)
        self._check(capi.lib().inlm_set_offset_lists(self._h, capi.ptr(lists, C.c_int32),
                                                     capi.ptr(counts, C.c_int32), int(K)))

    def proposal_stats(self) -> dict:
        out = np.zeros(5, np.float32)
        self._check(capi.lib().inlm_proposal_stats(self._h, capi.ptr(out, C.c_float)))
        return {"tiles": int(out[0]), "fallback_tiles": int(out[1]), "audited_tiles": int(out[2]),
                "mean_predicted_mass": float(out[3]), "mean_offsets_visited": float(out[4])}


def denoise(instack: np.ndarray, params: DenoiseParameters, device: int = 0):
    """Full iterNLM on one GPU: returns (result, per-iteration sigma lists)."""
    nz, ny, nx = instack.shape
    params.gpu.deviceID = device
    g = IterativeNLM_GPU()
    rc = g.configure_device([nx, ny, nz], params)
    if rc < 0:
        raise capi.InlmError(rc, getattr(g, "error", "no GPU"))
    try:
        out, log = g.denoise(instack)
    finally:
        g.free_device()
    return out, log[: params.maxiterations]


def denoise_slabs(instack: np.ndarray, params: DenoiseParameters, slab_slices: int, device: int = 0):
    """The same schedule over z-slabs of at most `slab_slices` output slices
    (inlm_denoise_slabs), for volumes too large for one GPU at once: returns
    (result, per-iteration sigma lists).  Replaces the reference's block
    drivers (Run_GaussianNoise_GPUBlocks, iternlm_gpu.cu:1729-2115)."""
    raw = _f32(instack)
    nz, ny, nx = raw.shape
    out = np.empty_like(raw)
    log = np.zeros((max(params.maxiterations, 1), nz), np.float32)
    sh = np.array([nx, ny, nz], np.int32)
    L = capi.lib()
    st = L.inlm_denoise_slabs(C.byref(params.to_c()), capi.ptr(sh, C.c_int32), device, capi.ptr(raw, C.c_float),
                              capi.ptr(out, C.c_float), int(slab_slices), capi.ptr(log, C.c_float))
    capi.check(st)
    return out, log[: params.maxiterations]


def iterate_slabs(raw: np.ndarray, prev: np.ndarray, sigma, params: DenoiseParameters, iter: int,
                  slab_slices: int, device: int = 0) -> np.ndarray:
    """One filter iteration over z-slabs of at most `slab_slices` output
    slices (inlm_iterate_slabs); prev is ignored at iteration 1."""
    raw = _f32(raw)
    prev = _f32(prev)
    sig = _f32(sigma)
    nz, ny, nx = raw.shape
    out = np.empty_like(raw)
    sh = np.array([nx, ny, nz], np.int32)
    capi.check(capi.lib().inlm_iterate_slabs(C.byref(params.to_c()), capi.ptr(sh, C.c_int32), device, int(iter),
                                             capi.ptr(raw, C.c_float), capi.ptr(prev, C.c_float),
                                             capi.ptr(sig, C.c_float), capi.ptr(out, C.c_float), int(slab_slices)))
    return out


def rgb_options(independent_channels=False, stabilize_variance=True, average_channelsigma=True):
    """``inlm_rgb_options``: the reference's Color parameters
    (denoise_parameters.h:63-67; --independent, --nopoisson, --noaverage)."""
    c = capi.InlmRgbOptions()
    capi.lib().inlm_default_rgb_options(C.byref(c))
    c.independent_channels = int(independent_channels)
    c.stabilize_variance = int(stabilize_variance)
    c.average_channelsigma = int(average_channelsigma)
    return c


def denoise_rgb(image: np.ndarray, params: DenoiseParameters, device: int = 0, **color):
    """One 2D RGB image ([3, ny, nx] channels, or [ny, nx, 3]) through the
    colour path (Color/color_denoising.h:19-150; inlm_denoise_rgb): returns
    ([3, ny, nx] float channels after the inverse Anscombe transform, the
    three noise levels each iteration used)."""
    img = np.asarray(image)
    if img.ndim == 3 and img.shape[0] != 3 and img.shape[2] == 3:
        img = np.moveaxis(img, 2, 0)
    chans = np.ascontiguousarray(img, np.float32)
    if chans.ndim != 3 or chans.shape[0] != 3:
        raise ValueError("expected [3, ny, nx] or [ny, nx, 3] channels")
    _, ny, nx = chans.shape
    out = np.empty_like(chans)
    log = np.zeros((max(params.maxiterations, 1), 3), np.float32)
    sh = np.array([nx, ny], np.int32)
    c = rgb_options(**color)
    capi.check(capi.lib().inlm_denoise_rgb(C.byref(params.to_c()), C.byref(c), capi.ptr(sh, C.c_int32), device,
                                           capi.ptr(chans, C.c_float), capi.ptr(out, C.c_float),
                                           capi.ptr(log, C.c_float)))
    return out, log[: params.maxiterations]

