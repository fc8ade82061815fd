"""CPU tests that pin the oracle (oracle/inlm_oracle.c).

Each check is chosen so that a plausible mistake in the restatement -- a tap
out of order, a fused multiply-add, a wrong weight literal, an off-by-one
window -- fails it: bit-exact agreement with fixtures written by the
reference's own compiled CPU kernels, closed forms, invariants and brute
force on tiny inputs.
"""
import math

import numpy as np
import pytest

from conftest import filter_fixtures, load_npz
from oracle import oracle as O


# ---------------------------------------------------------------------------
# Tables
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("rs,nslices,count", [((5, 5, 5), 11, 514), ((7, 7, 7), 11, 1326), ((7, 7, 7), 15, 1418),
                                              ((10, 10, 10), 11, 3102)])
def test_search_counts(rs, nslices, count):
    S = O.search_offsets(rs, nslices)
    assert len(S) == count
    # symmetric set, centre excluded
    assert set(map(tuple, S)) == set(map(tuple, -S))
    assert not np.any(np.all(S == 0, axis=1))


@pytest.mark.parametrize("rp,count", [((1, 1, 1), 7), ((2, 2, 2), 33), ((3, 3, 3), 115), ((3, 3, 2), 73),
                                      ((2, 2, 1), 15)])
def test_patch_counts(rp, count):
    P = O.patch_offsets(rp)
    assert len(P) == count
    assert tuple(P[0]) == (0, 0, 0)


def test_tables_match_reference_fixtures():
    z = load_npz("tables.npz")
    for i in range(int(z["n"])):
        shp, rs, rp = z[f"c{i}_shape"], tuple(z[f"c{i}_rs"]), tuple(z[f"c{i}_rp"])
        ns, an = int(z[f"c{i}_nslices"]), float(z[f"c{i}_aniso"])
        S = O.linear(O.search_offsets(rs, ns), shp[0], shp[1])
        P = O.linear(O.patch_offsets(rp, ns), shp[0], shp[1])
        w = O.distweight(rp, ns, an, int(shp[2]))
        assert np.array_equal(S, z[f"c{i}_soff"]), i
        assert np.array_equal(P, z[f"c{i}_poff"]), i
        assert np.array_equal(w.view(np.uint32), z[f"c{i}_w"].view(np.uint32)), i


def test_weight_classes_closed_form():
    w = O.distweight((2, 2, 2))
    P = O.patch_offsets((2, 2, 2))
    d = np.sqrt((P ** 2).sum(1).astype(np.float64))
    expect = 1.0 / (2 * d + 1) ** 2
    expect[0] = 1.0 / 9.0  # centre takes the largest weight
    assert np.allclose(w, expect, rtol=1e-6)
    assert np.all(O.distweight((1, 1, 1)) == np.float32(1) / np.float32(7))


def test_literal_rule_examples():
    lit = O.literal_weights(O.distweight((2, 2, 2)))
    assert set(np.unique(lit)) == {np.float32(0.111111), np.float32(0.0682275), np.float32(0.0501801),
                                   np.float32(0.04)}


# ---------------------------------------------------------------------------
# Padding
# ---------------------------------------------------------------------------
def test_pad_matches_reference_fixture():
    z = load_npz("pad.npz")
    for i in range(int(z["n"])):
        out = O.pad_reflect(z[f"c{i}_in"], tuple(z[f"c{i}_pad"]))
        assert np.array_equal(out, z[f"c{i}_out"])


def test_pad_brute_force():
    rng = np.random.default_rng(1)
    v = rng.standard_normal((3, 4, 5)).astype(np.float32)
    out = O.pad_reflect(v, (7, 6, 5))

    def refl(i, n):  # period 2(n-1) mirror without edge repeat
        p = 2 * (n - 1)
        i = i % p
        return i if i < n else p - i

    nz, ny, nx = v.shape
    for z in range(out.shape[0]):
        for y in range(out.shape[1]):
            for x in range(out.shape[2]):
                assert out[z, y, x] == v[refl(z - 5, nz), refl(y - 6, ny), refl(x - 7, nx)]


# ---------------------------------------------------------------------------
# Filter: reference fixtures (bit-exact) and closed forms
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", filter_fixtures())
def test_filter_matches_reference_fixture(name):
    f = load_npz(name)
    p = O.Params(radius_searchspace=tuple(f["rs"]), radius_patchspace=tuple(f["rp"]), nslices=int(f["nslices"]),
                 blockwise=False)
    out1 = O.filter_iteration(f["raw"], f["raw"], f["sigma1"], p)
    assert np.array_equal(out1.view(np.uint32), f["out1"].view(np.uint32))
    out2 = O.filter_iteration(f["raw"], f["out1"], f["sigma2"], p)
    assert np.array_equal(out2.view(np.uint32), f["out2"].view(np.uint32))


def _small(seed=0, shape=(12, 10, 9)):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape[::-1]) * 100 + 1000).astype(np.float32)


P_SMALL = O.Params(radius_searchspace=(2, 2, 2), radius_patchspace=(1, 1, 1), blockwise=False)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_constant_volume_is_fixed_point(prec):
    v = np.full((8, 9, 10), 1234.0, np.float32)
    out = O.filter_iteration(v, v, np.full(8, 50.0, np.float32), P_SMALL, precision=prec)
    assert np.all(out == 1234.0)


def test_sigma_infinite_gives_ball_mean():
    v = _small()
    out = O.filter_iteration(v, v, np.full(v.shape[0], 1e30), P_SMALL, precision="f64")
    S = O.search_offsets(P_SMALL.radius_searchspace)
    pad = O.pad_reflect(v, P_SMALL.padding()).astype(np.float64)
    px, py, pz = P_SMALL.padding()
    nz, ny, nx = v.shape
    acc = v.astype(np.float64).copy()
    for dx, dy, dz in S:
        acc += pad[pz + dz:pz + dz + nz, py + dy:py + dy + ny, px + dx:px + dx + nx]
    assert np.allclose(out, acc / (len(S) + 1), rtol=1e-13)


def test_sigma_zero_gives_identity():
    v = _small(2)
    out = O.filter_iteration(v, v, np.full(v.shape[0], 1e-30, np.float32), P_SMALL)
    assert np.array_equal(out, v)


def test_shift_equivariance_and_convexity():
    v = _small(3).astype(np.float64)
    sig = np.full(v.shape[0], 80.0)
    a = O.filter_iteration(v.astype(np.float32), v.astype(np.float32), sig, P_SMALL, precision="f64")
    b = O.filter_iteration((v + 500).astype(np.float32), (v + 500).astype(np.float32), sig, P_SMALL,
                           precision="f64")
    assert np.allclose(b - 500, a, atol=1e-9)
    assert np.all(a <= v.max() + 1e-9) and np.all(a >= v.min() - 1e-9)


def test_mirror_equivariance():
    v = _small(4)
    sig = np.full(v.shape[0], 60.0)
    a = O.filter_iteration(v, v, sig, P_SMALL, precision="f64")
    b = O.filter_iteration(np.ascontiguousarray(v[:, :, ::-1]), np.ascontiguousarray(v[:, :, ::-1]), sig, P_SMALL,
                           precision="f64")
    assert np.allclose(b[:, :, ::-1], a, rtol=1e-12)


def test_fp32_restatement_within_fp64_budget():
    f = load_npz("filter_p222_s2_u16.npz")
    p = O.Params(radius_searchspace=tuple(f["rs"]), radius_patchspace=tuple(f["rp"]), blockwise=False)
    a = O.filter_iteration(f["raw"], f["raw"], f["sigma1"], p)
    b = O.filter_iteration(f["raw"], f["raw"], f["sigma1"].astype(np.float64), p, precision="f64")
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
    # FP32 accumulation of ~500 terms and the cutoff discontinuity: a voxel
    # whose exponent sits within rounding of -3.56648 may gain or lose one
    # weight of ~0.025.  Everything else agrees to 1e-5.
    assert np.mean(rel < 1e-5) > 0.999


def test_pade_closed_forms():
    assert O.orc().orc_pade_f32(0.0) == 1.0
    assert O.orc().orc_pade_f32(-1.0) == np.float32(71.0 / 193.0)
    assert O.orc().orc_pade_f32(-2.0) == np.float32(40.0 / 296.0)
    assert O.orc().orc_pade_f32(-3.56648) == 0.0  # cutoff is strict
    below = np.nextafter(np.float32(-3.56648), np.float32(0))
    assert 0.02 < O.orc().orc_pade_f32(float(below)) < 0.03
    assert O.orc().orc_pade_f32(-10.0) == 0.0


def test_requant_rounding():
    x = np.array([0.5, 1.5, 2.5, -0.4, -3.0, 65534.5, 70000.0, 1.49999], np.float32)
    assert list(O.requant_u16(x)) == [1, 2, 3, 0, 0, 65535, 65535, 1]


# ---------------------------------------------------------------------------
# Noise estimator
# ---------------------------------------------------------------------------
def test_patchstd_constant_and_ramp():
    nx, ny, nz = 40, 35, 3
    yy, xx = np.mgrid[0:ny, 0:nx]
    a, b = 3.0, -2.0
    vol = np.stack([a * xx + b * yy + 7.0 * z for z in range(nz)]).astype(np.float32)
    samples = O.draw_samples((nx, ny, nz), 200, 15, seed=3)
    s32 = O.patchstd(vol, samples, 15, "f32")
    s64 = O.patchstd(vol, samples, 15, "f64")
    expect = math.sqrt((a * a + b * b) * (15 * 15 - 1) / 12.0)
    assert np.allclose(s64, expect, rtol=1e-12)
    assert np.allclose(s32, expect, rtol=1e-5)
    const = np.full((nz, ny, nx), 5.0, np.float32)
    assert np.all(O.patchstd(const, samples, 15, "f32") == 0)


def test_draw_samples_distribution():
    s = O.draw_samples((100, 80, 30), 50000, 15, seed=11)
    assert np.all(np.diff(s) >= 0)
    x, y, z = s % 100, (s // 100) % 80, s // 8000
    assert x.max() < 85 and y.max() < 65 and z.max() < 30 and s.min() >= 0
    assert abs(x.mean() - 42) < 1.0 and abs(z.mean() - 14.5) < 0.5


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_two_clusters_bimodal(prec):
    rng = np.random.default_rng(5)
    vals = np.concatenate([rng.normal(10, 1, 3000), rng.normal(50, 3, 2000), np.zeros(100)])
    rng.shuffle(vals)
    m, s = O.create2clusters(vals, precision=prec)
    assert abs(m - 10) < 0.1 and abs(s - 1) < 0.1
    # zeros are excluded: dropping them changes nothing
    m2, s2 = O.create2clusters(vals[vals != 0], precision=prec)
    assert m2 == pytest.approx(m, rel=1e-6) and s2 == pytest.approx(s, rel=1e-5)


def test_two_clusters_f32_vs_f64():
    rng = np.random.default_rng(6)
    vals = np.concatenate([rng.normal(300, 20, 4000), rng.normal(1500, 200, 1000)]).astype(np.float32)
    m32, s32 = O.create2clusters(vals, precision="f32")
    m64, s64 = O.create2clusters(vals.astype(np.float64), precision="f64")
    assert m32 == pytest.approx(m64, rel=1e-4) and s32 == pytest.approx(s64, rel=1e-3)


def test_zwindow_includes_one_sample_past_window():
    # two slices' worth of samples; window 2 -> (pos-1, pos+1] plus one more
    nx = ny = 20
    shape = (nx, ny, 6)
    z = np.array([0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5])
    samples = (z * nx * ny).astype(np.int64)
    vals = np.where(z == 3, 100.0, 1.0 + 0.01 * np.arange(len(z)))
    out = O.zwindow(samples, vals, shape, windowsize=2, precision="f64")
    # slice 1 covers z in (0, 2] = indices 2..5 and the reference also takes
    # index 6 (z=3, value 100): the high cluster must see it.
    assert out.shape == (6, 3)
    m_in, _ = O.create2clusters(vals, 2, 6, precision="f64")
    assert out[1, 1] == pytest.approx(m_in)


@pytest.mark.parametrize("rp,rs", [((2, 2, 2), (3, 3, 2)), ((1, 1, 1), (2, 2, 2)), ((3, 3, 3), (2, 2, 1))])
def test_filter_window_equals_whole_volume(rp, rs):
    # the window form gathers its halo through the reflect maps: at the volume
    # faces (the fold) and inside it must reproduce the whole-volume pass
    rng = np.random.default_rng(11)
    R = rng.normal(100.0, 10.0, (9, 17, 13)).astype(np.float32)
    P = (R + rng.normal(0.0, 2.0, R.shape)).astype(np.float32)
    sig = np.linspace(8.0, 12.0, 9).astype(np.float32)
    p = O.Params(radius_searchspace=rs, radius_patchspace=rp, blockwise=False)
    for prec in ("f32", "f64"):
        whole = O.filter_iteration(R, P, sig, p, precision=prec)
        for zr, yr in [((0, 2), (0, 5)), ((3, 6), (6, 11)), ((7, 9), (12, 17)), ((0, 9), (0, 17))]:
            win = O.filter_window(R, P, sig, p, zr, yr, precision=prec)
            assert np.array_equal(win, whole[zr[0]:zr[1], yr[0]:yr[1]]), (prec, zr, yr)
