"""The noise estimator's host half without a GPU: the circular sampling mask
(Noise/noiselevel.cpp:204-224), the semimanual mode (:56-79), the z-window
extrapolation into slices without valid samples (:572-603), and the library's
FP64 clustering (inlm_sigma_from_patchstd) against the oracle.  Each oracle
function is pinned here by brute force or by values derived by hand from the
reference's loop, not by retyping its formula."""
import math

import numpy as np
import pytest

from microtomodenoise_amd import _capi as capi
from oracle import oracle as O


def params(mode, **kw):
    p = capi.default_params()
    p.noise_mode = {"global": capi.NOISE_GLOBAL, "z-adaptive": capi.NOISE_ZADAPTIVE, "manual": capi.NOISE_MANUAL,
                    "semimanual": capi.NOISE_SEMIMANUAL}[mode]
    for k, v in kw.items():
        if k == "sigma":
            p.sigma[0], p.sigma[1] = v
        else:
            setattr(p, k, v)
    return p


def corners(samples, nx, ny):
    r = samples % (nx * ny)
    return r % nx, r // nx


def feasible_corners(nx, ny, ps, diameter):
    """Brute force over the sampling range: corners inside the reference's
    circle, r = diameter/2 - sqrt(2 ps^2), compared as squares in float32."""
    r = np.float32(np.float64(np.float32(diameter) / np.float32(2)) - math.sqrt(2 * ps * ps))
    r2 = np.float32(r * r)
    xs, ys = np.meshgrid(np.arange(nx - ps), np.arange(ny - ps), indexing="xy")
    d2 = ((xs - nx // 2) ** 2 + (ys - ny // 2) ** 2).astype(np.float32)
    return set(zip(xs[d2 <= r2].tolist(), ys[d2 <= r2].tolist()))


@pytest.mark.parametrize("shape,ps,diameter", [((120, 100, 7), 15, 110), ((64, 64, 3), 9, 60), ((90, 70, 4), 5, 50)])
def test_mask_draws_only_and_all_feasible_corners(shape, ps, diameter):
    nx, ny, nz = shape
    lib = capi.draw_samples(shape, 60000, ps, seed=5, mask_diameter=diameter)
    orc = O.draw_samples(shape, 60000, ps, seed=5, mask_diameter=diameter)
    assert np.array_equal(lib, orc)
    assert np.all(np.diff(lib) >= 0) and lib.min() >= 0 and lib.max() < nx * ny * nz
    feas = feasible_corners(nx, ny, ps, diameter)
    got = set(zip(*[c.tolist() for c in corners(lib, nx, ny)]))
    # only corners inside the circle, and (with this many samples) every one
    assert got <= feas and got == feas
    # slices stay uniform: the redraw keeps z
    z = lib // (nx * ny)
    assert abs(z.mean() - (nz - 1) / 2) < 0.05 * nz


def test_mask_covering_the_range_changes_nothing():
    # every corner inside the circle: no redraw, the same positions as without a mask
    shape = (80, 60, 5)
    a = capi.draw_samples(shape, 5000, 15, seed=9)
    b = capi.draw_samples(shape, 5000, 15, seed=9, mask_diameter=400)
    c = O.draw_samples(shape, 5000, 15, seed=9, mask_diameter=400)
    assert np.array_equal(a, b) and np.array_equal(a, c)


def test_mask_negative_radius_squares_like_the_reference():
    # diameter 2: r = 1 - sqrt(450) < 0, whose float square is positive, so
    # the reference samples a disk of radius |r| (about 20.2)
    shape = (100, 100, 2)
    s = capi.draw_samples(shape, 4000, 15, seed=2, mask_diameter=2)
    x, y = corners(s, 100, 100)
    d2 = (x - 50) ** 2 + (y - 50) ** 2
    assert d2.max() <= 20.2135 ** 2 and d2.max() > 15 ** 2


def test_mask_without_feasible_corner_is_rejected():
    # 20x20 slice, patch 15: corners x, y in [0, 5), the centre (10, 10) is
    # 6 away; diameter 44 leaves r = 0.79
    shape = (20, 20, 3)
    with pytest.raises(capi.InlmError):
        capi.draw_samples(shape, 10, 15, seed=1, mask_diameter=44)
    with pytest.raises(ValueError):
        O.draw_samples(shape, 10, 15, seed=1, mask_diameter=44)
    assert not feasible_corners(20, 20, 15, 44)


def test_semimanual_reference_sample_and_later_estimates():
    shape = (40, 30, 6)
    samples = np.array([0, 50, 400, 1300, 2500, 3000], np.int64)
    p = params("semimanual", sigma=(3.0, 0.0))
    ref = np.full(1, -1, np.int64)
    # distances to 3.0: 2, 0.5, 4, 0.5, 0.25, 1 -> sample 4
    s0 = capi.sigma_from_patchstd(p, shape, samples, [5.0, 3.5, 7.0, 2.5, 3.25, 4.0], which=0, semimanual_ref=ref)
    assert ref[0] == 4 and np.all(s0 == np.float32(3.0))
    s1 = capi.sigma_from_patchstd(p, shape, samples, [1.0, 2.0, 3.0, 4.0, 5.5, 6.0], which=1, semimanual_ref=ref)
    assert np.all(s1 == np.float32(5.5))
    # equal distances: the first wins (strict comparison, noiselevel.cpp:63)
    ref2 = np.full(1, -1, np.int64)
    capi.sigma_from_patchstd(p, shape, samples, [5.0, 3.5, 2.5, 9.0, 9.0, 9.0], which=0, semimanual_ref=ref2)
    assert ref2[0] == 1
    # a later estimate needs the first one's reference sample
    with pytest.raises(capi.InlmError):
        capi.sigma_from_patchstd(p, shape, samples, np.ones(6), which=1, semimanual_ref=np.full(1, -1, np.int64))


def test_semimanual_oracle_picks_the_same_sample():
    rng = np.random.default_rng(4)
    vol = rng.normal(100.0, 1.0, (5, 40, 40)).astype(np.float32)
    vol[:, 20:, :] += rng.normal(0.0, 9.0, (5, 20, 40)).astype(np.float32)
    samples = O.draw_samples((40, 40, 5), 400, 9, seed=3)
    p = O.Params(noise_mode="semimanual", patchsize=9, sigma=(4.0, 0.0))
    st = {}
    s0 = O.noiselevel(vol, samples, p, "f64", which=0, state=st)
    ps = O.patchstd(vol, samples, 9, "f64")
    assert st["ref"] == int(np.argmin(np.abs(ps - 4.0))) and np.all(s0 == 4.0)
    cp = params("semimanual", sigma=(4.0, 0.0), patchsize=9)
    ref = np.full(1, -1, np.int64)
    capi.sigma_from_patchstd(cp, (40, 40, 5), samples, ps, which=0, semimanual_ref=ref)
    assert ref[0] == st["ref"]
    later = O.noiselevel(vol * 2, samples, p, "f64", which=1, state=st)
    assert later[0] == pytest.approx(2 * ps[st["ref"]], rel=1e-12)


def zwindow_planted():
    # two samples per slice of an 8-slice volume; slices 0-1 have patch std
    # 4, slices 2-5 are flat (std 0: no valid sample), slices 6-7 have 8
    nx = ny = 20
    z = np.repeat(np.arange(8), 2)
    samples = (z * nx * ny + np.tile([0, 1], 8)).astype(np.int64)
    vals = np.where(z < 2, 4.0, np.where(z >= 6, 8.0, 0.0))
    return (nx, ny, 8), samples, vals


def test_zwindow_extrapolation_by_hand():
    # Window 2: slice pos takes samples of slices (pos-1, pos+1] and the first
    # sample beyond (noiselevel.cpp:541-548).  create2clusters (:366-528)
    # takes min and max over the window with its zeros, starts the lower mean
    # at min + (max-min)/4, excludes zeros from the assignment, and returns
    # that starting mean (std 0) when every valid value lands in the upper
    # cluster.  By hand:
    #   slice 0: 4 4 4 4 0 -> start means 1, 3; the 4s go up -> (1, 0)
    #   slice 1: 4 4 0 0 0 -> (1, 0)
    #   slices 2, 3: only zeros -> (0, 0), empty
    #   slice 4: 0 0 0 0 8 (first sample of slice 6) -> start 2, 6 -> (2, 0)
    #   slice 5: 0 0 8 8 8 -> (2, 0)
    #   slices 6, 7: only 8s -> both starting means 8 -> (8, 0)
    # The fill (:572-603) runs in place in ascending order: slice 2 takes its
    # left neighbour (1); slice 3 the mean of its updated left neighbour and
    # its right one, (1 + 2) / 2.
    want = [1, 1, 1, 1.5, 2, 2, 8, 8]
    shape, samples, vals = zwindow_planted()
    out = O.zwindow(samples, vals, shape, windowsize=2, precision="f64")
    assert np.array_equal(out[:, 1], want)
    assert np.array_equal(out[:, 2], np.zeros(8))
    out32 = O.zwindow(samples, vals, shape, windowsize=2, precision="f32")
    assert np.array_equal(out32[:, 1], want)
    # the library's host clustering (z-adaptive, sigma = mean + 1 * std)
    p = params("z-adaptive", zwindow=2, stds_from_mean=1.0)
    sig = capi.sigma_from_patchstd(p, shape, samples, vals)
    assert np.array_equal(sig, np.float32(want))


def test_zwindow_without_any_valid_sample_stays_zero():
    shape, samples, _ = zwindow_planted()
    out = O.zwindow(samples, np.zeros(16), shape, windowsize=2, precision="f64")
    assert np.all(out[:, 1] == 0)
    sig = capi.sigma_from_patchstd(params("z-adaptive", zwindow=2), shape, samples, np.zeros(16))
    assert np.all(sig == 0)


@pytest.mark.parametrize("mode,window", [("global", 100), ("z-adaptive", 100), ("z-adaptive", 7)])
def test_library_clustering_matches_fp64_oracle(mode, window):
    rng = np.random.default_rng(8)
    shape = (200, 150, 40)
    samples = O.draw_samples(shape, 20000, 15, seed=4)
    z = samples // (200 * 150)
    noise = 30.0 + 0.5 * z  # noise std rising along z, plus a texture cluster
    vals = np.where(rng.random(len(samples)) < 0.7, rng.normal(noise, 2.0), rng.normal(4 * noise, 20.0))
    vals[rng.random(len(samples)) < 0.01] = 0.0  # flat patches
    p = params(mode, zwindow=window, stds_from_mean=1.5)
    got = capi.sigma_from_patchstd(p, shape, samples, vals)
    if mode == "global":
        m, s = O.create2clusters(vals, precision="f64")
        want = np.full(shape[2], m + 1.5 * s)
    else:
        zw = O.zwindow(samples, vals, shape, windowsize=window, precision="f64")
        want = zw[:, 1] + 1.5 * zw[:, 2]
    assert np.allclose(got, want, rtol=2e-7, atol=0)
    # and the estimate tracks the planted noise level
    if mode == "z-adaptive" and window == 7:
        assert abs(got[20] - (30.0 + 0.5 * 20 + 1.5 * 2.0)) < 1.5


def test_manual_mode_first_and_later():
    p = params("manual", sigma=(11.0, 22.0))
    assert np.all(capi.sigma_from_patchstd(p, (20, 20, 3), np.zeros(0, np.int64), np.zeros(0), which=0) == 11.0)
    assert np.all(capi.sigma_from_patchstd(p, (20, 20, 3), np.zeros(0, np.int64), np.zeros(0), which=1) == 22.0)
