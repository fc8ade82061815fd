"""The packed FP16 weight of the matrix-core forms against the reference's
Padé [3/3] weight with its cutoff (iternlm_cpu.h:125-126, restated by the
oracle's orc_pade_f64).

The kernels evaluate w / c0 = Q(u) u + s(u) in FP16, with u = 1 + a / h
(h = 3.56648) clamped to [0, 1], Q the degree-4 polynomial kPadePolyQ5 of
microtomodenoise_amd/csrc/inlm_mfma.hip and s the step clamp(u * 16384).
The coefficients are read from that source file and the evaluation is
emulated operation by operation (each FP16 FMA rounded once), so a changed
coefficient, a dropped Horner term or a wrong step fails the bounds below.
"""
import os
import re

import numpy as np

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H = 3.56648


def _coefficients():
    src = open(os.path.join(ROOT, "microtomodenoise_amd", "csrc", "inlm_mfma.hip")).read()
    m = re.search(r"kPadePolyQ5\[5\]\s*=\s*\{(.*?)\};", src, re.S)
    assert m, "kPadePolyQ5 not found"
    vals = [float(v) for v in re.findall(r"\(_Float16\)\s*([-0-9.eE]+)f", m.group(1))]
    assert len(vals) == 5
    return [np.float16(v) for v in vals]


def _fma16(x, y, z):
    # one rounding to FP16 of the exact product-sum (exact in FP64 for FP16 inputs)
    return (x.astype(np.float64) * y.astype(np.float64) + z.astype(np.float64)).astype(np.float16)


def _kernel_weight(u):
    """w / c0 for FP16 u in [0, 1], in the kernel's order of operations."""
    q = _coefficients()
    p = np.full_like(u, q[4])
    for c in (q[3], q[2], q[1], q[0]):
        p = _fma16(p, u, np.full_like(u, c))
    s = np.clip((u.astype(np.float64) * 16384.0).astype(np.float16), 0, 1).astype(np.float16)
    return _fma16(p, u, s).astype(np.float64)


def _pade(a):
    L = O.orc()
    return np.array([L.orc_pade_f64(float(x)) for x in a])


def test_packed_weight_tracks_pade_within_documented_bound():
    a = np.linspace(-0.999 * H, 0.0, 4001)
    u = (1.0 + a / H).astype(np.float16)
    w = _kernel_weight(u)
    ref = _pade(a)
    # the scale is free (sums and maximum weight scale alike); it is the
    # weight at the cutoff, about 0.025
    c0 = float(np.dot(w, ref) / np.dot(w, w))
    assert 0.024 < c0 < 0.026, c0
    err = w * c0 - ref
    assert np.abs(err).max() < 2.0e-3, np.abs(err).max()
    assert np.sqrt(np.mean(err ** 2)) < 4.0e-4


def test_packed_weight_cutoff_and_peak():
    # at and beyond the cutoff u clamps to 0 and the weight is exactly zero
    assert _kernel_weight(np.array([0.0], np.float16))[0] == 0.0
    # at a = 0 (identical patches) the Padé weight is 1, i.e. 1 / c0 here
    w1 = _kernel_weight(np.array([1.0], np.float16))[0]
    w_cut = _kernel_weight(np.array([2.0 ** -13], np.float16))[0]
    assert 38.0 < w1 / w_cut < 42.0, w1 / w_cut
    # the weight rises monotonically with u (decreasing distance)
    u = np.linspace(2.0 ** -12, 1.0, 2001).astype(np.float16)
    w = _kernel_weight(np.unique(u))
    assert np.all(np.diff(w) >= -1e-2 * w[1:])
