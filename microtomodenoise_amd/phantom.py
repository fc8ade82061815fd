"""Synthetic tomography phantoms for tests and benchmarks.

Deterministic given a seed: geometry (spheres, ellipsoids, streaks) is drawn
with numpy's PCG64 generator and rasterised on the CPU into a label volume;
noise is white Gaussian, correlated by a separable 3D Gaussian blur, added,
then clipped (and rounded for uint16 data).  The white-noise sigma is the
pre-blur value; the blur lowers the per-voxel std.

backend="numpy" draws the noise with numpy (bit-reproducible everywhere);
backend="torch" draws it on the GPU with torch (fast for 10^9-voxel volumes;
reproducible for a given torch version).  Volumes are float32 arrays of shape
(nz, ny, nx).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np


@dataclass
class PhantomSpec:
    shape: tuple  # (nx, ny, nz)
    seed: int = 0
    kind: str = "spheres"  # spheres | ellipsoids
    radii: tuple = (3, 12)  # sphere radii or ellipsoid semi-axes range
    fill: float = 0.4  # target volume fraction of objects (spheres)
    levels: Sequence[float] = (20000.0, 40000.0)  # background, object level(s)
    noise_sigma: float = 1500.0
    noise_sigma_z: Optional[tuple] = None  # (sigma at z=0, sigma at z=nz-1): linear in z
    blur_sigma: float = 0.8
    streaks: int = 0
    streak_len: tuple = (50, 300)
    streak_amp: float = 3000.0
    uint16: bool = True
    max_objects: int = 200000


# Configurations of the benchmark suite (shape given as nx, ny, nz).
CONFIGS = {
    "cfg1": PhantomSpec((128, 128, 64), seed=0, radii=(3, 12), noise_sigma=1500.0, blur_sigma=0.8),
    "cfg2": PhantomSpec((512, 512, 512), seed=1, radii=(3, 30), noise_sigma=2000.0, blur_sigma=1.0, streaks=200),
    "cfg3": PhantomSpec((1024, 1024, 512), seed=2, kind="ellipsoids", radii=(5, 60), levels=(0.2, 0.5, 0.9),
                        noise_sigma_z=(0.02, 0.06), blur_sigma=1.2, uint16=False),
    "headline": PhantomSpec((1024, 1024, 512), seed=3, radii=(3, 30), noise_sigma=2000.0, blur_sigma=1.0,
                            streaks=200, uint16=False),
    "cfg5": PhantomSpec((2048, 2048, 256), seed=4, radii=(3, 30), levels=(10000.0, 25000.0, 40000.0, 55000.0),
                        noise_sigma=2500.0, blur_sigma=1.0, streaks=1000, streak_amp=4000.0),
}


def _spheres(rng, shape, rmin, rmax, fill, max_objects):
    nx, ny, nz = shape
    target = fill * nx * ny * nz
    radii = []
    vol = 0.0
    while vol < target * 1.15 and len(radii) < max_objects:
        r = rng.uniform(rmin, rmax)
        radii.append(r)
        vol += 4.0 / 3.0 * math.pi * r ** 3
    radii.sort(reverse=True)
    cell = 2.0 * rmax
    grid = {}
    placed = []
    filled = 0.0
    for r in radii:
        if filled >= target:
            break
        for _ in range(30):
            c = rng.uniform([0, 0, 0], [nx, ny, nz])
            key = tuple((c // cell).astype(int))
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for (c2, r2) in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                            if (c2[0] - c[0]) ** 2 + (c2[1] - c[1]) ** 2 + (c2[2] - c[2]) ** 2 < (r + r2) ** 2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                grid.setdefault(key, []).append((c, r))
                placed.append((c, r))
                filled += 4.0 / 3.0 * math.pi * r ** 3
                break
    return placed


def _raster_sphere(lab, c, r, value):
    nz, ny, nx = lab.shape
    x0, x1 = max(int(math.floor(c[0] - r)), 0), min(int(math.ceil(c[0] + r)) + 1, nx)
    y0, y1 = max(int(math.floor(c[1] - r)), 0), min(int(math.ceil(c[1] + r)) + 1, ny)
    z0, z1 = max(int(math.floor(c[2] - r)), 0), min(int(math.ceil(c[2] + r)) + 1, nz)
    if x0 >= x1 or y0 >= y1 or z0 >= z1:
        return
    zz, yy, xx = np.ogrid[z0:z1, y0:y1, x0:x1]
    m = (xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2 <= r * r
    lab[z0:z1, y0:y1, x0:x1][m] = value


def _raster_ellipsoid(lab, rng, c, axes, value):
    nz, ny, nx = lab.shape
    # random rotation from a normalised quaternion
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, cc, d = q
    rot = np.array([[a * a + b * b - cc * cc - d * d, 2 * (b * cc - a * d), 2 * (b * d + a * cc)],
                    [2 * (b * cc + a * d), a * a - b * b + cc * cc - d * d, 2 * (cc * d - a * b)],
                    [2 * (b * d - a * cc), 2 * (cc * d + a * b), a * a - b * b - cc * cc + d * d]])
    # inside: x^T M x <= 1, M = rot diag(1/a^2) rot^T (ellipsoid frame p = x rot);
    # the half extent along axis j is sqrt((M^-1)_jj) = sqrt(sum_i rot_ji^2 a_i^2)
    a2 = np.asarray(axes, np.float64) ** 2
    M = rot @ np.diag(1.0 / a2) @ rot.T
    ext = np.sqrt((rot * rot) @ a2)
    x0, x1 = max(int(math.floor(c[0] - ext[0])), 0), min(int(math.ceil(c[0] + ext[0])) + 1, nx)
    y0, y1 = max(int(math.floor(c[1] - ext[1])), 0), min(int(math.ceil(c[1] + ext[1])) + 1, ny)
    z0, z1 = max(int(math.floor(c[2] - ext[2])), 0), min(int(math.ceil(c[2] + ext[2])) + 1, nz)
    if x0 >= x1 or y0 >= y1 or z0 >= z1:
        return
    zz = (np.arange(z0, z1) - c[2])[:, None, None]
    yy = (np.arange(y0, y1) - c[1])[None, :, None]
    xx = (np.arange(x0, x1) - c[0])[None, None, :]
    # q = M00 x^2 + x (2 M01 y + 2 M02 z) + (M11 y^2 + M22 z^2 + 2 M12 y z)
    lin = 2.0 * (M[0, 1] * yy + M[0, 2] * zz)
    const = M[1, 1] * yy * yy + M[2, 2] * zz * zz + 2.0 * M[1, 2] * yy * zz
    m = (M[0, 0] * xx + lin) * xx + const <= 1.0
    lab[z0:z1, y0:y1, x0:x1][m] = value


def _streaks(rng, shape, n, lmin, lmax, amp):
    """(z, y, x, amplitude) index arrays of n one-voxel-thick line segments."""
    nx, ny, nz = shape
    zs, ys, xs, amps = [], [], [], []
    for _ in range(n):
        p0 = rng.uniform([0, 0, 0], [nx, ny, nz])
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        L = rng.uniform(lmin, lmax)
        a = amp if rng.random() < 0.5 else -amp
        t = np.arange(0.0, L, 0.5)
        pts = np.round(p0[None, :] + t[:, None] * d[None, :]).astype(np.int64)
        ok = (pts[:, 0] >= 0) & (pts[:, 0] < nx) & (pts[:, 1] >= 0) & (pts[:, 1] < ny) & (pts[:, 2] >= 0) & \
             (pts[:, 2] < nz)
        pts = np.unique(pts[ok], axis=0)
        xs.append(pts[:, 0]); ys.append(pts[:, 1]); zs.append(pts[:, 2])
        amps.append(np.full(len(pts), a, np.float32))
    if not xs:
        e = np.zeros(0, np.int64)
        return e, e, e, np.zeros(0, np.float32)
    return np.concatenate(zs), np.concatenate(ys), np.concatenate(xs), np.concatenate(amps)


def _gauss_kernel(sigma):
    rad = max(1, int(math.ceil(4.0 * sigma)))
    x = np.arange(-rad, rad + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return (k / k.sum()).astype(np.float32)


def clean_volume(spec: PhantomSpec) -> np.ndarray:
    """The noise-free phantom (float32, (nz, ny, nx)), streaks included."""
    rng = np.random.default_rng(spec.seed)
    nx, ny, nz = spec.shape
    lab = np.zeros((nz, ny, nx), np.uint8)
    nlev = len(spec.levels)
    if spec.kind == "spheres":
        for i, (c, r) in enumerate(_spheres(rng, spec.shape, spec.radii[0], spec.radii[1], spec.fill,
                                            spec.max_objects)):
            _raster_sphere(lab, c, r, 1 + (i % (nlev - 1)) if nlev > 2 else 1)
    else:
        target = spec.fill * nx * ny * nz
        vol = 0.0
        count = 0
        while vol < target and count < spec.max_objects:
            axes = rng.uniform(spec.radii[0], spec.radii[1], size=3)
            c = rng.uniform([0, 0, 0], [nx, ny, nz])
            _raster_ellipsoid(lab, rng, c, axes, 1 + int(rng.integers(0, nlev - 1)))
            vol += 4.0 / 3.0 * math.pi * float(np.prod(axes))
            count += 1
    lut = np.asarray(spec.levels, np.float32)
    out = lut[lab]
    if spec.streaks:
        z, y, x, a = _streaks(rng, spec.shape, spec.streaks, spec.streak_len[0], spec.streak_len[1],
                              spec.streak_amp)
        np.add.at(out, (z, y, x), a)
    return out


def _noise_numpy(spec: PhantomSpec, shape_zyx):
    rng = np.random.default_rng(spec.seed + 0x5EED)
    n = rng.standard_normal(size=shape_zyx, dtype=np.float32)
    if spec.blur_sigma > 0:
        from scipy.ndimage import correlate1d
        k = _gauss_kernel(spec.blur_sigma)
        for ax in range(3):
            n = correlate1d(n, k, axis=ax, mode="reflect")
    return n


def _blur_axis_torch(t, k, axis):
    """Correlate along one axis with symmetric (edge-repeating) boundaries,
    as scipy.ndimage mode="reflect", using shifted adds only."""
    import torch
    r = (len(k) - 1) // 2
    n = t.shape[axis]
    left = t.narrow(axis, 0, r).flip(axis)
    right = t.narrow(axis, n - r, r).flip(axis)
    tp = torch.cat([left, t, right], dim=axis)
    out = torch.zeros_like(t)
    for i in range(2 * r + 1):
        out.add_(tp.narrow(axis, i, n), alpha=float(k[i]))
    return out


def _noise_torch(spec: PhantomSpec, shape_zyx, device):
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(spec.seed + 0x5EED)
    n = torch.randn(shape_zyx, generator=g, device=device, dtype=torch.float32)
    if spec.blur_sigma > 0:
        k = _gauss_kernel(spec.blur_sigma)
        for ax in range(3):
            n = _blur_axis_torch(n, k, ax)
    return n


def make(spec: PhantomSpec, backend: str = "numpy", device: str = "cuda", with_clean: bool = False):
    """Noisy phantom as float32 (nz, ny, nx) -- a numpy array for the numpy
    backend, a device tensor for the torch backend.  with_clean also returns
    the noise-free volume (numpy)."""
    nx, ny, nz = spec.shape
    clean = clean_volume(spec)
    zyx = (nz, ny, nx)
    if backend == "numpy":
        n = _noise_numpy(spec, zyx)
        if spec.noise_sigma_z is not None:
            s = np.linspace(spec.noise_sigma_z[0], spec.noise_sigma_z[1], nz, dtype=np.float32)
            vol = clean + n * s[:, None, None]
        else:
            vol = clean + n * np.float32(spec.noise_sigma)
        if spec.uint16:
            vol = np.clip(np.round(vol), 0, 65535)
        else:
            vol = np.clip(vol, 0, None) if spec.levels[0] >= 1 else vol
        vol = vol.astype(np.float32)
    else:
        import torch
        n = _noise_torch(spec, zyx, device)
        c = torch.from_numpy(clean).to(device)
        if spec.noise_sigma_z is not None:
            s = torch.linspace(spec.noise_sigma_z[0], spec.noise_sigma_z[1], nz, device=device)
            vol = c + n * s.view(-1, 1, 1)
        else:
            vol = c + n * spec.noise_sigma
        if spec.uint16:
            vol = torch.clamp(torch.round(vol), 0, 65535)
        elif spec.levels[0] >= 1:
            vol = torch.clamp(vol, min=0)
        vol = vol.contiguous()
        del n, c
    return (vol, clean) if with_clean else vol


def psnr(x, ref, peak: float) -> float:
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    mse = float(np.mean((x - ref) ** 2))
    return float("inf") if mse == 0 else 10.0 * math.log10(peak * peak / mse)
