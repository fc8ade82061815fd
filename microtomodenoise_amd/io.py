"""Volume I/O and the command-line program's output contract, through the
library's C ABI (Geometry/hdcommunication.cpp and main.cpp:231-563 of the
reference): TIFF stacks in and out, noise CSV files, the resume check and the
whole on-disk denoising run."""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _capi as capi


def _b(s: str) -> bytes:
    return s.encode()


def stack_shape(path: str):
    """(nx, ny, nz) of a directory of TIFF slices or a single TIFF file."""
    sh = np.zeros(3, np.int32)
    capi.check(capi.lib().inlm_tiff_stack_shape(_b(path), capi.ptr(sh, C.c_int32)))
    return tuple(int(v) for v in sh)


def read_stack(path: str, first: int = 0, last: Optional[int] = None) -> np.ndarray:
    """Slices first..last as float32 [nz, ny, nx]."""
    nx, ny, nz = stack_shape(path)
    last = nz - 1 if last is None else last
    out = np.empty((last - first + 1, ny, nx), np.float32)
    capi.check(capi.lib().inlm_tiff_read_stack(_b(path), first, last, capi.ptr(out, C.c_float)))
    return out


def write_stack(directory: str, name: str, vol: np.ndarray, z0: int = 0, as16bit: bool = False):
    """One TIFF per slice: directory/name<z0+i:04d>.tif."""
    v = np.ascontiguousarray(vol, np.float32)
    nz, ny, nx = v.shape
    sh = np.asarray([nx, ny, nz], np.int32)
    capi.check(capi.lib().inlm_tiff_write_stack(_b(directory), _b(name), capi.ptr(v, C.c_float),
                                                capi.ptr(sh, C.c_int32), z0, int(as16bit)))


def write_noise_csv(directory: str, name: str, sigma):
    s = np.ascontiguousarray(sigma, np.float32)
    capi.check(capi.lib().inlm_noise_csv_write(_b(directory), _b(name), capi.ptr(s, C.c_float), len(s)))


def read_noise_csv(path: str, nz: int) -> Optional[np.ndarray]:
    """The sigma list of a noise CSV, or None when missing or not nz lines."""
    out = np.zeros(nz, np.float32)
    st = capi.lib().inlm_noise_csv_read(_b(path), nz, capi.ptr(out, C.c_float))
    return out if st == capi.INLM_OK else None


def iteration_on_disk(directory: str, first: int, last: int) -> bool:
    return bool(capi.lib().inlm_iteration_on_disk(_b(directory), first, last))


def io_options(save_16bit=False, resume=False, cleanup=False, firstslice=-1, lastslice=-1, verbose=False):
    o = capi.InlmIoOptions()
    capi.lib().inlm_default_io_options(C.byref(o))
    o.save_16bit, o.resume, o.cleanup = int(save_16bit), int(resume), int(cleanup)
    o.firstslice, o.lastslice, o.verbose = int(firstslice), int(lastslice), int(verbose)
    return o


def denoise_files(inpath: str, params, outpath: Optional[str] = None, device: int = 0, **options):
    """The reference program's whole run on files (main.cpp:231-563)."""
    o = io_options(**options)
    capi.check(capi.lib().inlm_denoise_files(C.byref(params.to_c()), device, _b(inpath),
                                             _b(outpath) if outpath else None, C.byref(o)))


def read_rgb(path: str) -> np.ndarray:
    """An RGB TIFF (8 or 16 bits, 3 or 4 samples per pixel) as [3, ny, nx]
    float channels (hdcommunication.cpp:674-760)."""
    sh = np.zeros(2, np.int32)
    L = capi.lib()
    capi.check(L.inlm_rgb_tiff_shape(_b(path), capi.ptr(sh, C.c_int32)))
    out = np.empty((3, int(sh[1]), int(sh[0])), np.float32)
    capi.check(L.inlm_rgb_tiff_read(_b(path), capi.ptr(out, C.c_float)))
    return out


def write_rgb(path: str, channels: np.ndarray):
    """[3, ny, nx] channels to an 8-bit RGB TIFF, values capped at 255 and
    truncated (color_denoising.h:72-77, Save2DTifImage_RGB)."""
    ch = np.ascontiguousarray(channels, np.float32)
    sh = np.array([ch.shape[2], ch.shape[1]], np.int32)
    capi.check(capi.lib().inlm_rgb_tiff_write(_b(path), capi.ptr(ch, C.c_float), capi.ptr(sh, C.c_int32)))


def denoise_rgb_files(inpath: str, params, outpath: Optional[str] = None, device: int = 0, independent_channels=False,
                      stabilize_variance=True, average_channelsigma=True, **options):
    """The reference program's colour mode (main.cpp:254-257) through
    inlm_denoise_rgb_files: every RGB TIFF of inpath to
    outpath/<name>_denoised.tif (and denoise_iterations/<name>_iterNN.tif
    unless cleanup)."""
    from .denoiser import rgb_options
    o = io_options(**options)
    c = rgb_options(independent_channels, stabilize_variance, average_channelsigma)
    capi.check(capi.lib().inlm_denoise_rgb_files(C.byref(params.to_c()), C.byref(c), device, _b(inpath),
                                                 _b(outpath) if outpath else None, C.byref(o)))

