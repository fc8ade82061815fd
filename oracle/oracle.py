"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg may
import this module, and only as the checker.  It wraps two libraries built by
oracle/Makefile:

* ``liborc.so`` -- the restatement of the reference CPU path
  (oracle/inlm_oracle.c), FP32-faithful and FP64;
* ``_ref/libinlm_ref.so`` -- the reference's own ``iternlm_prepare.cpp`` and
  ``iternlm_cpu_kernels.cpp`` compiled from its sources, when present.

All volumes are numpy arrays in C order ``[nz, ny, nx]`` (x fastest), which
is the reference layout ``idx = z*nx*ny + y*nx + x``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORC = None
_REF = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")


def orc():
    global _ORC
    if _ORC is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liborc.so is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        L.orc_search_offsets.argtypes = [_i32p, C.c_int, _i32p, C.c_int]
        L.orc_patch_offsets.argtypes = [_i32p, C.c_int, _i32p, C.c_int]
        L.orc_distweight.argtypes = [_i32p, C.c_int, C.c_float, C.c_int, _f32p, C.c_int]
        L.orc_literal_weights.argtypes = [_f32p, C.c_int, _f32p]
        L.orc_cpu_weight_kind.argtypes = [_i32p, C.c_int]
        L.orc_pad_reflect.argtypes = [_f32p, _i32p, _i32p, _f32p]
        L.orc_pade_f32.argtypes = [C.c_float]
        L.orc_pade_f32.restype = C.c_float
        L.orc_pade_f64.argtypes = [C.c_double]
        L.orc_pade_f64.restype = C.c_double
        L.orc_filter_rgb_f32.argtypes = [_f32p, _f32p, _i32p, _i32p, _f32p, C.c_int, C.c_int, _i64p, C.c_int, _i64p,
                                         _f32p, _f32p]
        L.orc_filter_rgb_f32.restype = None
        L.orc_filter_f32.argtypes = [_f32p, _f32p, _i32p, _i32p, _f32p, C.c_float, C.c_int, _i64p, C.c_int,
                                     _i64p, _f32p, C.c_int, _f32p, C.c_int, C.c_int, C.c_int]
        L.orc_filter_f64.argtypes = [_f32p, _f32p, _i32p, _i32p, _f64p, C.c_double, C.c_int, _i64p, C.c_int,
                                     _i64p, _f32p, C.c_int, _f64p, C.c_int, C.c_int, C.c_int]
        L.orc_weight_mass_f32.argtypes = [_f32p, _i32p, _i32p, _f32p, C.c_float, C.c_int, _i64p, C.c_int, _i64p,
                                          _f32p, C.c_int, _i32p, _f64p]
        L.orc_requant_u16.argtypes = [_f32p, C.c_longlong, _u16p]
        L.orc_draw_samples.argtypes = [_i32p, C.c_longlong, C.c_int, C.c_uint64, _i64p]
        L.orc_draw_samples_mask.argtypes = [_i32p, C.c_longlong, C.c_int, C.c_int, C.c_uint64, _i64p]
        L.orc_patchstd_f32.argtypes = [_f32p, _i32p, _i64p, C.c_longlong, C.c_int, _f32p]
        L.orc_patchstd_f64.argtypes = [_f32p, _i32p, _i64p, C.c_longlong, C.c_int, _f64p]
        L.orc_create2clusters_f32.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float),
                                              C.POINTER(C.c_float)]
        L.orc_create2clusters_f64.argtypes = [_f64p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                              C.POINTER(C.c_double)]
        L.orc_zwindow.argtypes = [_i64p, _f32p, _f64p, C.c_longlong, _i32p, C.c_int, C.c_int, _f64p]
        _ORC = L
    return _ORC


def ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libinlm_ref.so"))


def ref():
    """The reference's own CPU kernels (oracle/_ref), or an error if not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libinlm_ref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libinlm_ref.so is not built")
        L = C.CDLL(path)
        L.ref_tables.argtypes = [_i32p, _i32p, _i32p, C.c_int, C.c_float, C.POINTER(C.c_int), _i64p,
                                 C.POINTER(C.c_int), _i64p, _f32p, C.c_int]
        L.ref_pad_reflective.argtypes = [_f32p, _i32p, _i32p, _f32p]
        L.ref_filter.argtypes = [C.c_int, _f32p, _f32p, _i32p, _i32p, _i32p, C.c_int, C.c_float, _f32p, _f32p,
                                 C.c_int]
        L.ref_filter.restype = C.c_int
        _REF = L
    return _REF


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


# ---------------------------------------------------------------------------
# Parameters
# ---------------------------------------------------------------------------
@dataclass
class Params:
    """The fields of protocol::DenoiseParameters (denoise_parameters.h:9-74)
    that the hot path reads."""
    radius_searchspace: tuple = (10, 10, 10)
    radius_patchspace: tuple = (1, 1, 1)
    nslices: int = 11
    beta: float = 1.0
    z_anisotropy: float = 1.0
    alpha: float = 0.5
    maxiterations: int = 4
    blockwise: bool = True  # cpu.blockwise: selects the CPU dispatch table
    weight_mode: int = 0  # 0 Pade+cutoff (CPU), 1 expf (reference GPU)
    noise_mode: str = "global"
    zwindow: int = 100
    n_samples: int = 100000
    patchsize: int = 15
    stds_from_mean: float = 1.0
    sigma: tuple = (0.0, 0.0)  # manual and semimanual modes
    extra: dict = field(default_factory=dict)

    def padding(self):
        rs, rp = self.radius_searchspace, self.radius_patchspace
        return (rs[0] + rp[0], rs[1] + rp[1], min(self.nslices // 2, rs[2]) + rp[2])


# ---------------------------------------------------------------------------
# Tables
# ---------------------------------------------------------------------------
def search_offsets(rs, nslices=11) -> np.ndarray:
    """(dx,dy,dz) per search offset in the reference order."""
    L = orc()
    cap = (2 * rs[0] + 1) * (2 * rs[1] + 1) * (2 * rs[2] + 1)
    buf = np.zeros(3 * cap, np.int32)
    n = L.orc_search_offsets(_i32(rs), nslices, buf, cap)
    return buf[: 3 * n].reshape(n, 3)


def patch_offsets(rp, nslices=11) -> np.ndarray:
    L = orc()
    cap = (2 * rp[0] + 1) * (2 * rp[1] + 1) * (2 * rp[2] + 1) + 1
    buf = np.zeros(3 * cap, np.int32)
    n = L.orc_patch_offsets(_i32(rp), nslices, buf, cap)
    return buf[: 3 * n].reshape(n, 3)


def distweight(rp, nslices=11, z_aniso=1.0, nz_padded=2) -> np.ndarray:
    L = orc()
    cap = (2 * rp[0] + 1) * (2 * rp[1] + 1) * (2 * rp[2] + 1) + 1
    buf = np.zeros(cap, np.float32)
    n = L.orc_distweight(_i32(rp), nslices, z_aniso, nz_padded, buf, cap)
    return buf[:n].copy()


def literal_weights(w) -> np.ndarray:
    w = np.ascontiguousarray(w, np.float32)
    out = np.zeros_like(w)
    orc().orc_literal_weights(w, len(w), out)
    return out


def cpu_weight_table(p: Params, nz_padded: int) -> np.ndarray:
    """The weight table the reference CPU path applies for these parameters
    (Denoiser/iternlm_cpu.cpp:97-114, :529-548); with z anisotropy the literal
    radii but (1,1,3) take the computed table of the reference GPU class."""
    rp = p.radius_patchspace
    kind = orc().orc_cpu_weight_kind(_i32(rp), int(p.blockwise))
    if kind < 0:
        raise ValueError(f"patch radius {rp} in block mode reaches a kernel with a different tap count")
    wc = distweight(rp, p.nslices, p.z_anisotropy, nz_padded)
    if kind == 2:
        return np.full_like(wc, np.float32(1.0) / np.float32(7.0))
    if kind == 1 and (p.z_anisotropy == 1.0 or tuple(rp) == (1, 1, 3)):
        return literal_weights(distweight(rp, p.nslices, 1.0, nz_padded))
    # anisotropic voxels: the computed table, as the reference GPU class's
    # generic kernel applies it (Denoiser/iternlm_gpu.cu:1470, :1700-1720)
    return wc


def linear(offs: np.ndarray, nxp: int, nyp: int) -> np.ndarray:
    o = offs.astype(np.int64)
    return np.ascontiguousarray(o[:, 2] * nxp * nyp + o[:, 1] * nxp + o[:, 0])


# ---------------------------------------------------------------------------
# Padding / filter
# ---------------------------------------------------------------------------
def pad_reflect(vol: np.ndarray, pad) -> np.ndarray:
    """pad = (px, py, pz) symmetric, or 6-tuple {x0,y0,z0,x1,y1,z1}."""
    vol = np.ascontiguousarray(vol, np.float32)
    nz, ny, nx = vol.shape
    if len(pad) == 3:
        pad = (pad[0], pad[1], pad[2], pad[0], pad[1], pad[2])
    out = np.zeros((nz + pad[2] + pad[5], ny + pad[1] + pad[4], nx + pad[0] + pad[3]), np.float32)
    rc = orc().orc_pad_reflect(vol, _i32([nx, ny, nz]), _i32(pad), out)
    if rc != 0:
        raise ValueError("reflective padding needs at least 2 voxels on every padded axis")
    return out


def filter_iteration(R: np.ndarray, P: np.ndarray, sigma, p: Params, *, precision="f32", weights=None,
                     z_range=None, nthreads=0, offsets=None) -> np.ndarray:
    """One iterNLM filter pass of the restated CPU path.

    R is the original noisy volume, P the previous iterate (R at iteration
    1), sigma a per-slice list.  precision 'f32' is operation-faithful to the
    reference CPU path; 'f64' is the FP64 oracle of the same algorithm.
    `offsets` (indices into the search table) restricts the pass to a subset
    of the search, as the pruned filter visits it.
    """
    R = np.ascontiguousarray(R, np.float32)
    P = np.ascontiguousarray(P, np.float32)
    nz, ny, nx = R.shape
    pad = p.padding()
    Rp, Pp = pad_reflect(R, pad), pad_reflect(P, pad)
    nzp, nyp, nxp = Rp.shape
    S = search_offsets(p.radius_searchspace, p.nslices)
    PT = patch_offsets(p.radius_patchspace, p.nslices)
    soff, poff = linear(S, nxp, nyp), linear(PT, nxp, nyp)
    if offsets is not None:
        soff = np.ascontiguousarray(soff[np.asarray(offsets, np.int64)])
    w = cpu_weight_table(p, nzp) if weights is None else np.ascontiguousarray(weights, np.float32)
    z0, z1 = (0, nz) if z_range is None else z_range
    shape = _i32([nx, ny, nz])
    if precision == "f32":
        out = np.zeros_like(R)
        sg = np.ascontiguousarray(sigma, np.float32)
        orc().orc_filter_f32(Rp, Pp, shape, _i32(pad), sg, p.beta, len(soff), soff, len(poff), poff, w,
                             p.weight_mode, out, z0, z1, nthreads)
    else:
        out = np.zeros(R.shape, np.float64)
        sg = np.ascontiguousarray(sigma, np.float64)
        orc().orc_filter_f64(Rp, Pp, shape, _i32(pad), sg, p.beta, len(soff), soff, len(poff), poff, w,
                             p.weight_mode, out, z0, z1, nthreads)
    return out


def reflect_index(n: int, pad: int) -> np.ndarray:
    """Source index of every padded position along one axis of length n, from
    the oracle's own reflect padding applied to the indices."""
    idx = np.arange(n, dtype=np.float32).reshape(n, 1, 1)
    return pad_reflect(idx, (0, 0, pad))[:, 0, 0].astype(np.int64)


def filter_window(R: np.ndarray, P: np.ndarray, sigma, p: Params, z_range, y_range, *, precision="f32",
                  weights=None, nthreads=0) -> np.ndarray:
    """filter_iteration restricted to the output window z in [z0, z1), y in
    [y0, y1), all x, on a volume too large to pad whole: the window and its
    halo are gathered through the reflect padding's index maps, so every
    output value equals filter_iteration's at that voxel.  Returns the window
    (z1 - z0, y1 - y0, nx)."""
    nz, ny, nx = R.shape
    z0, z1 = z_range
    y0, y1 = y_range
    if not (0 <= z0 < z1 <= nz and 0 <= y0 < y1 <= ny):
        raise ValueError("window outside the volume")
    px, py, pz = p.padding()
    zi = reflect_index(nz, pz)[z0:z1 + 2 * pz]
    yi = reflect_index(ny, py)[y0:y1 + 2 * py]
    xi = reflect_index(nx, px)
    ix = np.ix_(zi, yi, xi)
    Rw = np.ascontiguousarray(np.asarray(R, np.float32)[ix])
    Pw = np.ascontiguousarray(np.asarray(P, np.float32)[ix])
    nxp, nyp = nx + 2 * px, (y1 - y0) + 2 * py
    S = search_offsets(p.radius_searchspace, p.nslices)
    PT = patch_offsets(p.radius_patchspace, p.nslices)
    soff, poff = linear(S, nxp, nyp), linear(PT, nxp, nyp)
    # the table of the whole padded volume (the r=1 rule depends on its depth)
    w = cpu_weight_table(p, nz + 2 * pz) if weights is None else np.ascontiguousarray(weights, np.float32)
    shape = _i32([nx, y1 - y0, z1 - z0])
    dt = np.float32 if precision == "f32" else np.float64
    sg = np.ascontiguousarray(np.asarray(sigma)[z0:z1], dt)
    out = np.zeros((z1 - z0, y1 - y0, nx), dt)
    fn = orc().orc_filter_f32 if precision == "f32" else orc().orc_filter_f64
    fn(Rw, Pw, shape, _i32((px, py, pz)), sg, p.beta, len(soff), soff, len(poff), poff, w, p.weight_mode, out,
       0, z1 - z0, nthreads)
    return out


def filter_rgb(R: np.ndarray, P: np.ndarray, channel_sigma, p: Params, independent=False) -> np.ndarray:
    """One pass of the colour path's three-channel filter
    (iternlm_cpu.cpp:932-1075), operation-faithful FP32.  R, P: [3, ny, nx]
    (or [3, nz, ny, nx]) channels; channel_sigma: three noise levels.  The
    colour path uses the computed patch weights (setup_distweight)."""
    R = np.ascontiguousarray(R, np.float32)
    P = np.ascontiguousarray(P, np.float32)
    if R.ndim == 3:
        R, P = R[:, None], P[:, None]
    _, nz, ny, nx = R.shape
    pad = p.padding()
    Rp = np.ascontiguousarray(np.stack([pad_reflect(R[c], pad) for c in range(3)]))
    Pp = np.ascontiguousarray(np.stack([pad_reflect(P[c], pad) for c in range(3)]))
    _, nzp, nyp, nxp = Rp.shape
    S = search_offsets(p.radius_searchspace, p.nslices)
    PT = patch_offsets(p.radius_patchspace, p.nslices)
    soff, poff = linear(S, nxp, nyp), linear(PT, nxp, nyp)
    w = distweight(p.radius_patchspace, p.nslices, p.z_anisotropy, nzp)
    s = np.asarray(channel_sigma, np.float32)
    mult = np.array([np.float32(-1.0) / (s[c] * s[c] * np.float32(p.beta)) for c in range(3)], np.float32)
    out = np.zeros(R.shape, np.float32)
    orc().orc_filter_rgb_f32(Rp, Pp, _i32([nx, ny, nz]), _i32(pad), mult, int(independent), len(soff), soff,
                             len(poff), poff, w, out)
    return out[:, 0] if out.shape[1] == 1 else out


def weight_mass(P: np.ndarray, sigma, p: Params, tile=(32, 8, 8)) -> np.ndarray:
    """Per-tile weight mass of every search offset ([tiles, n_search],
    float64) for previous iterate P: the labels of the learned proposal."""
    P = np.ascontiguousarray(P, np.float32)
    nz, ny, nx = P.shape
    pad = p.padding()
    Pp = pad_reflect(P, pad)
    nzp, nyp, nxp = Pp.shape
    S = search_offsets(p.radius_searchspace, p.nslices)
    PT = patch_offsets(p.radius_patchspace, p.nslices)
    soff, poff = linear(S, nxp, nyp), linear(PT, nxp, nyp)
    w = cpu_weight_table(p, nzp)
    nt = -(-nx // tile[0]) * -(-ny // tile[1]) * -(-nz // tile[2])
    out = np.zeros((nt, len(soff)), np.float64)
    orc().orc_weight_mass_f32(Pp, _i32([nx, ny, nz]), _i32(pad), np.ascontiguousarray(sigma, np.float32), p.beta,
                              len(soff), soff, len(poff), poff, w, p.weight_mode, _i32(list(tile)), out)
    return out


def ref_filter_iteration(R, P, sigma, p: Params, nthreads=0) -> np.ndarray:
    """One iteration through the reference's own specialised CPU kernel."""
    rp = tuple(p.radius_patchspace)
    kern = int("".join(str(v) for v in rp))
    R = np.ascontiguousarray(R, np.float32)
    nz, ny, nx = R.shape
    pad = p.padding()
    Rp, Pp = pad_reflect(R, pad), pad_reflect(np.ascontiguousarray(P, np.float32), pad)
    out = np.zeros_like(R)
    rc = ref().ref_filter(kern, Rp, Pp, _i32([nx, ny, nz]), _i32(p.radius_searchspace), _i32(rp), p.nslices,
                          p.beta, np.ascontiguousarray(sigma, np.float32), out, nthreads)
    if rc != 0:
        raise ValueError(f"no specialised reference kernel for patch radius {rp}")
    return out


def ref_tables(shape_padded, rs, rp, nslices=11, z_aniso=1.0):
    cap = 16384
    soff = np.zeros(cap, np.int64)
    poff = np.zeros(cap, np.int64)
    w = np.zeros(cap, np.float32)
    nS, nP = C.c_int(), C.c_int()
    ref().ref_tables(_i32(shape_padded), _i32(rs), _i32(rp), nslices, z_aniso, C.byref(nS), soff, C.byref(nP),
                     poff, w, cap)
    return soff[: nS.value].copy(), poff[: nP.value].copy(), w[: nP.value].copy()


def ref_pad_reflect(vol, pad6):
    vol = np.ascontiguousarray(vol, np.float32)
    nz, ny, nx = vol.shape
    out = np.zeros((nz + pad6[2] + pad6[5], ny + pad6[1] + pad6[4], nx + pad6[0] + pad6[3]), np.float32)
    ref().ref_pad_reflective(vol, _i32([nx, ny, nz]), _i32(pad6), out)
    return out


def requant_u16(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(x.shape, np.uint16)
    orc().orc_requant_u16(x.reshape(-1), x.size, out.reshape(-1))
    return out


# ---------------------------------------------------------------------------
# Noise estimator
# ---------------------------------------------------------------------------
def draw_samples(shape_xyz, n, patchsize=15, seed=0, mask_diameter=0) -> np.ndarray:
    """drawsamples (noiselevel.cpp:192-234) on SplitMix64(seed); with
    mask_diameter > 0 the corners stay inside the circular field of view."""
    out = np.zeros(n, np.int64)
    rc = orc().orc_draw_samples_mask(_i32(shape_xyz), n, patchsize, int(mask_diameter), seed, out)
    if rc != 0:
        raise ValueError("volume too small for the noise patch, or no corner inside the mask")
    return out


def patchstd(vol, samples, patchsize=15, precision="f32"):
    vol = np.ascontiguousarray(vol, np.float32)
    nz, ny, nx = vol.shape
    samples = np.ascontiguousarray(samples, np.int64)
    if precision == "f32":
        out = np.zeros(len(samples), np.float32)
        orc().orc_patchstd_f32(vol, _i32([nx, ny, nz]), samples, len(samples), patchsize, out)
    else:
        out = np.zeros(len(samples), np.float64)
        orc().orc_patchstd_f64(vol, _i32([nx, ny, nz]), samples, len(samples), patchsize, out)
    return out


def create2clusters(values, fs=-1, ls=-1, precision="f32"):
    if precision == "f32":
        v = np.ascontiguousarray(values, np.float32)
        m, s = C.c_float(), C.c_float()
        orc().orc_create2clusters_f32(v, len(v), fs, ls, C.byref(m), C.byref(s))
    else:
        v = np.ascontiguousarray(values, np.float64)
        m, s = C.c_double(), C.c_double()
        orc().orc_create2clusters_f64(v, len(v), fs, ls, C.byref(m), C.byref(s))
    return m.value, s.value


def zwindow(samples, values, shape_xyz, windowsize=100, precision="f32"):
    samples = np.ascontiguousarray(samples, np.int64)
    nz = shape_xyz[2]
    out = np.zeros(3 * nz, np.float64)
    vf = np.ascontiguousarray(values, np.float32)
    vd = np.ascontiguousarray(values, np.float64)
    orc().orc_zwindow(samples, vf, vd, len(samples), _i32(shape_xyz), windowsize, int(precision != "f32"), out)
    return out.reshape(nz, 3)


def noiselevel(vol, samples, p: Params, precision="f32", which=0, state=None) -> np.ndarray:
    """get_noiselevel (Noise/noiselevel.cpp:9-95) on a sample list: global
    and z-adaptive give sigma = mean + stds_from_mean * std of the lower
    cluster; manual gives sigma[0] on the first estimate (which=0) and
    sigma[1] later; semimanual (:56-79) gives sigma[0] first and records in
    state["ref"] the first sample whose patch std is closest to it, and
    later that sample's patch std."""
    nz, ny, nx = vol.shape
    dt = np.float32 if precision == "f32" else np.float64
    if p.noise_mode == "manual":
        return np.full(nz, p.sigma[0 if which == 0 else 1], dt)
    ps = patchstd(vol, samples, p.patchsize, precision)
    if p.noise_mode == "semimanual":
        if which == 0:
            best, ref = 1e9, 0
            for i, v in enumerate(ps):
                d = abs(float(v) - float(np.float32(p.sigma[0])))
                if d < best:
                    best, ref = d, i
            state["ref"] = ref
            return np.full(nz, p.sigma[0], dt)
        return np.full(nz, ps[state["ref"]], dt)
    if p.noise_mode == "z-adaptive":
        zw = zwindow(samples, ps, (nx, ny, nz), p.zwindow, precision)
        if precision == "f32":
            return (zw[:, 1].astype(np.float32) + np.float32(p.stds_from_mean) * zw[:, 2].astype(np.float32))
        return zw[:, 1] + p.stds_from_mean * zw[:, 2]
    m, s = create2clusters(ps, precision=precision)
    if precision == "f32":
        return np.full(nz, np.float32(m) + np.float32(p.stds_from_mean) * np.float32(s), np.float32)
    return np.full(nz, m + p.stds_from_mean * s, np.float64)


def blend_sigma(s0, s1, alpha):
    """sigma = alpha*s0 + (1-alpha)*s1 in float (main.cpp:443-444)."""
    a = np.float32(alpha)
    return (a * np.asarray(s0, np.float32) + (np.float32(1.0) - a) * np.asarray(s1, np.float32)).astype(np.float32)


def iternlm(raw: np.ndarray, p: Params, *, sigma0=None, sigma1=None, samples=None, precision="f32",
            nthreads=0):
    """The main.cpp:325-545 schedule: sigma0 drives iteration 1; sigma1 is
    re-estimated on iteration 1's output with the same samples and blended
    with alpha for iterations >= 2.  Fixed sigma lists may be injected."""
    raw = np.ascontiguousarray(raw, np.float32)
    nz, ny, nx = raw.shape
    state = {}
    if sigma0 is None:
        sigma0 = noiselevel(raw, samples, p, precision, which=0, state=state)
    prev = raw
    out = filter_iteration(raw, prev, sigma0, p, precision="f32", nthreads=nthreads)
    if p.maxiterations > 1:
        if sigma1 is None:
            s1 = noiselevel(out, samples, p, precision, which=1, state=state)
            sigma1 = blend_sigma(sigma0, s1, p.alpha)
        for _ in range(2, p.maxiterations + 1):
            out = filter_iteration(raw, out, sigma1, p, precision="f32", nthreads=nthreads)
    return out
