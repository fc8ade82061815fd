"""The per-tile scale of the FP8-plane form (microtomodenoise_amd/csrc/
inlm_mfma.hip): a patch pair whose distance has one saturated tap must keep
weight exactly zero, as the reference's Padé weight with its cutoff gives
(iternlm_cpu.h:125-126: a saturated difference puts the distance far beyond
the cutoff).

The kernel stores D * S per tap in FP8 e4m3, S = 256 f^2, clamping
(f (P(y+s) - P(y)))^2 to 1, so a saturated tap contributes 256 times the FP8
code of its weight times the common factor c.  The tile scale f^2 = 4^m is
the largest power of four not above k * wmin * |mm| with |mm| = 1/(sigma^2
beta) * mscale at the tile's largest sigma; k is read from the kernel
source.  Planes hold differences of the standardised volume, and f is a
normal FP16 number, so the scale spans f^2 in [4^-12, 4^8]; below that (a
noise level thousands of times the volume's standard deviation) a tap
saturates only at a difference of 4096 standard deviations.  The emulation
below runs the weight step the kernel runs on the
accumulator (FP16 conversion toward zero, u = 1 + d * msu in FP16, weight
zero for u <= 0) for every slice of tiles with z-varying sigma, the smallest
weight of each patch radius's table and every common factor c the host may
choose, and requires u <= 0.  The factor the kernel used before (a scale
from the middle slice's sigma alone, cutoff in [2, 8] plane units) fails
this at patch radius 3, which the last test shows.
"""
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H = 3.56648


def _kernel_factor():
    src = open(os.path.join(ROOT, "microtomodenoise_amd", "csrc", "inlm_mfma.hip")).read()
    m = re.search(r"a\.wmin \* \(([0-9.]+)f / ([0-9.]+)f / 3\.56648f\)", src)
    assert m, "FP8 tile scale not found in the kernel source"
    return float(m.group(1)) / float(m.group(2)) / H


def e4m3(x):
    """Round positive floats to OCP FP8 e4m3 (bias 7, max 448), nearest even."""
    x = np.asarray(x, np.float64)
    out = np.zeros_like(x)
    pos = x > 0
    e = np.floor(np.log2(x[pos]))
    e = np.maximum(e, -6.0)  # subnormals share the smallest exponent
    q = 2.0 ** (e - 3)  # 3 mantissa bits
    out[pos] = np.minimum(np.round(x[pos] / q) * q, 448.0)
    return out


def pow4_below(x):
    """The kernel's 4^m <= x from the float exponent of x, m = floor(e / 2)."""
    e = int(np.floor(np.log2(np.float32(x))))
    e = min(max(e, -24), 16)
    return 4.0 ** (e >> 1)


def saturated_u(k, wmin, c, mm):
    """u = 1 + d msu for a patch pair with one saturated tap of weight wmin,
    per slice of a tile whose slices have the factors mm = -mscale / (sigma^2
    beta)."""
    mm = np.asarray(mm, np.float64)
    x = np.float32(np.abs(mm).min() * wmin * k)
    f2 = pow4_below(x)
    acc = 256.0 * e4m3(wmin * c)  # the accumulator (FP32)
    dinv = 1.0 / (256.0 * f2 * c)
    acc16 = np.float16(np.float32(acc))  # conversion toward zero rounds no higher
    if float(acc16) > acc:
        acc16 = np.nextafter(acc16, np.float16(0))
    msu = (mm * (1.0 / H) * dinv).astype(np.float32).astype(np.float16)
    return (acc16.astype(np.float64) * msu.astype(np.float64) + 1.0).astype(np.float16).astype(np.float64)


def _wmin(rp):
    p = O.Params(radius_patchspace=(rp, rp, rp), blockwise=False)
    w = O.cpu_weight_table(p, 30)
    return float(w[w > 0].min())


@pytest.mark.parametrize("rp", [1, 2, 3])
def test_saturated_tap_gives_zero_weight(rp):
    k = _kernel_factor()
    wmin = _wmin(rp)
    rng = np.random.default_rng(rp)
    # the scale spans the exponents the kernel keeps (f a normal FP16 number,
    # f^2 in [4^-12, 4^8]); sigma varies by up to 3x along the tile
    for _ in range(300):
        x = 4.0 ** rng.uniform(-12, 8.5)
        mm_min = x / (wmin * k)
        mm = -mm_min * rng.uniform(1.0, 3.0, size=8) ** 2
        mm[rng.integers(8)] = -mm_min
        for c in np.arange(1.0, 2.0, 1.0 / 32):
            u = saturated_u(k, wmin, c, mm)
            assert np.all(u <= 0.0), (rp, x, c, u.max())


def test_previous_middle_slice_scale_fails_at_radius_3():
    # the earlier rule: 4^m <= |mm(middle slice)| * 8 / (3.56648 * 256)
    wmin = _wmin(3)
    sig = np.full(8, 5.0)
    mm_mid = 1.0 / 25.0
    x = np.float32(mm_mid * 8 / (H * 256))
    f2 = pow4_below(x)
    acc = 256.0 * e4m3(wmin * 1.0)
    u = 1.0 - acc * mm_mid / (H * 256 * f2)
    assert u > 0.0  # a saturated outer tap kept a nonzero weight
    assert np.all(saturated_u(_kernel_factor(), wmin, 1.0, -1.0 / sig ** 2) <= 0.0)
