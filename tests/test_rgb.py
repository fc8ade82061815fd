"""The colour path (Color/color_denoising.h:19-150): the three-channel filter
of iternlm_cpu.cpp:932-1075, its schedule and the RGB TIFF files.

The oracle's three-channel filter has no golden vectors in the reference's
tests, so the CPU tests pin it by what the mathematics fixes: a constant
image is a fixed point, flat weights give the mean over the search disk
(with the centre in the joint mode, without it in the independent mode),
the cutoff gives the identity, channels 0 and 1 commute, independent
channels do not mix, and with two flat channels the joint filter is the
volume filter at sigma * sqrt(3).  The GPU tests compare the HIP path with
the oracle bit for bit (FP32) and the schedule with a restatement built on
the oracle's filter."""
import os

import numpy as np
import pytest
from PIL import Image

from conftest import ROOT  # noqa: F401
from microtomodenoise_amd import DenoiseParameters, InlmError, _capi, denoise_rgb
from microtomodenoise_amd import io as vio
from oracle import oracle as O

F = np.float32


def _op(rs=2, rp=1):
    # the colour path forces 2D: z radii 0, one slice (color_denoising.h:26-29)
    return O.Params(radius_searchspace=(rs, rs, 0), radius_patchspace=(rp, rp, 0), nslices=1)


def _image(ny, nx, seed, peak=200.0):
    """Smooth colour image with Poisson noise, values 0..255."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:ny, 0:nx].astype(np.float64)
    chans = []
    for c in range(3):
        base = np.full((ny, nx), 20.0 + 15 * c)
        for _ in range(4):
            cy, cx = rng.uniform(0, ny), rng.uniform(0, nx)
            r = rng.uniform(3, 0.4 * min(nx, ny))
            base += rng.uniform(40, peak - 40) * (((y - cy) ** 2 + (x - cx) ** 2) < r * r)
        chans.append(np.clip(rng.poisson(np.clip(base, 0, 255)), 0, 255))
    return np.stack(chans).astype(F)


def _bits(a):
    return np.ascontiguousarray(a, F).view(np.uint32)


# ---------------------------------------------------------------------------
# Oracle: closed forms and invariants
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("independent", [False, True])
def test_oracle_constant_image_is_a_fixed_point(independent):
    img = np.stack([np.full((9, 11), v, F) for v in (3.5, 70.25, 201.0)])
    out = O.filter_rgb(img, img, [2.0, 3.0, 4.0], _op(), independent)
    assert np.allclose(out, img, rtol=2e-7, atol=0)


@pytest.mark.parametrize("independent", [False, True])
def test_oracle_flat_weights_give_the_disk_mean(independent):
    img = _image(13, 17, 3)
    op = _op(2, 1)
    out = O.filter_rgb(img, img, [1e18] * 3, op, independent)  # every weight Pade(0) = 1
    pad = op.padding()
    S = O.search_offsets(op.radius_searchspace, op.nslices)
    ny, nx = img.shape[1:]
    for c in range(3):
        Rp = O.pad_reflect(img[c][None], pad)[0].astype(np.float64)
        acc = np.zeros((ny, nx))
        for dx, dy, _ in S:
            acc += Rp[pad[1] + dy: pad[1] + dy + ny, pad[0] + dx: pad[0] + dx + nx]
        n = len(S)
        if not independent:  # the joint mode adds the centre once, at the largest weight
            acc += img[c]
            n += 1
        assert np.allclose(out[c], acc / n, rtol=1e-5), c


@pytest.mark.parametrize("independent", [False, True])
def test_oracle_cutoff_gives_the_identity(independent):
    img = _image(12, 10, 4)
    img += np.random.default_rng(5).uniform(0.01, 0.02, img.shape).astype(F)  # no two patches equal
    out = O.filter_rgb(img, img, [1e-3] * 3, _op(), independent)  # every weight below the cutoff
    assert np.array_equal(_bits(out), _bits(img))


@pytest.mark.parametrize("independent", [False, True])
def test_oracle_channels_0_and_1_commute(independent):
    img = _image(14, 12, 6)
    prev = O.filter_rgb(img, img, [9.0, 11.0, 13.0], _op(), independent)
    out = O.filter_rgb(img, prev, [6.0, 8.0, 10.0], _op(), independent)
    sw = [1, 0, 2]
    out2 = O.filter_rgb(img[sw], prev[sw], [8.0, 6.0, 10.0], _op(), independent)
    assert np.array_equal(_bits(out2), _bits(out[sw]))


def test_oracle_independent_channels_do_not_mix():
    img = _image(14, 12, 7)
    other = img.copy()
    other[1:] = _image(14, 12, 8)[1:]
    a = O.filter_rgb(img, img, [9.0, 9.0, 9.0], _op(), True)
    b = O.filter_rgb(other, other, [9.0, 3.0, 30.0], _op(), True)
    assert np.array_equal(_bits(a[0]), _bits(b[0]))
    assert not np.array_equal(a[1], b[1])


@pytest.mark.parametrize("rs,rp", [(2, 1), (3, 2)])
def test_oracle_joint_with_two_flat_channels_is_the_volume_filter(rs, rp):
    # d1 = d2 = 0, so the joint distance is d0 / 3: the volume filter's
    # weight at sigma * sqrt(3) (same centre rule, computed patch weights)
    img = _image(15, 13, 9)
    img[1] = 40.0
    img[2] = 90.0
    op = _op(rs, rp)
    s = 12.0
    out = O.filter_rgb(img, img, [s, 5.0, 7.0], op, False)
    nzp = 1
    w = O.distweight(op.radius_patchspace, op.nslices, op.z_anisotropy, nzp)
    vol = O.filter_iteration(img[0][None], img[0][None], [s * np.sqrt(3.0)], op, weights=w)[0]
    assert np.allclose(out[0], vol, rtol=2e-5)
    assert np.allclose(out[1], 40.0, rtol=1e-6) and np.allclose(out[2], 90.0, rtol=1e-6)


# ---------------------------------------------------------------------------
# Host side: files and arguments (no GPU)
# ---------------------------------------------------------------------------
def test_rgb_tiff_written_and_read_like_pil(tmp_path):
    img = np.random.default_rng(10).uniform(-20, 300, (3, 7, 9)).astype(F)
    vio.write_rgb(str(tmp_path / "a.tif"), img)
    im = Image.open(tmp_path / "a.tif")
    assert im.mode == "RGB" and im.size == (9, 7)
    want = np.where(img > 0, np.trunc(np.minimum(F(255), img)), 0).astype(np.uint8)  # min(255, x), truncated
    assert np.array_equal(np.asarray(im), np.moveaxis(want, 0, 2))
    assert np.array_equal(vio.read_rgb(str(tmp_path / "a.tif")), want.astype(F))


@pytest.mark.parametrize("mode", ["RGB", "RGBA"])
def test_rgb_tiff_from_pil(tmp_path, mode):
    rng = np.random.default_rng(11)
    pix = rng.integers(0, 256, (6, 5, len(mode)), dtype=np.uint8)
    Image.fromarray(pix, mode).save(tmp_path / "b.tif")
    got = vio.read_rgb(str(tmp_path / "b.tif"))
    assert got.shape == (3, 6, 5)
    assert np.array_equal(got, np.moveaxis(pix[..., :3], 2, 0).astype(F))


def test_grey_tiff_is_not_rgb(tmp_path):
    Image.fromarray(np.zeros((4, 4), np.uint16)).save(tmp_path / "g.tif")
    with pytest.raises(InlmError):
        vio.read_rgb(str(tmp_path / "g.tif"))


def test_rgb_arguments_checked_on_the_host():
    p = DenoiseParameters(radius_searchspace=[2, 2, 2], radius_patchspace=[1, 1, 1], maxiterations=1)
    with pytest.raises(ValueError):
        denoise_rgb(np.zeros((2, 8, 8), F), p)
    for shape in ((3, 1, 30), (3, 30, 1)):
        with pytest.raises(InlmError) as e:
            denoise_rgb(np.zeros(shape, F), p)
        assert e.value.status == _capi.INLM_BAD_ARG
    with pytest.raises(InlmError) as e:  # the 15-pixel noise patch does not fit
        denoise_rgb(np.zeros((3, 12, 40), F), p)
    assert e.value.status == _capi.INLM_BAD_ARG
    p.noiselevel.mode = "manual"
    p.radius_searchspace = [30, 30, 0]
    with pytest.raises(InlmError) as e:  # the search window is wider than the image
        denoise_rgb(np.zeros((3, 20, 20), F), p)
    assert e.value.status == _capi.INLM_BAD_ARG


# ---------------------------------------------------------------------------
# The schedule restated on the oracle's filter
# ---------------------------------------------------------------------------
def _anscombe(v):
    return F(2) * np.sqrt(np.asarray(v, F) + F(0.375))  # colorspace.h:99-101


def _inverse_anscombe(v):
    h = np.asarray(v, F) / F(2)
    return h * h - F(0.375)  # colorspace.h:118-120


def _to_rgb8(v):
    return np.where(v > 0, np.trunc(np.minimum(F(255), v)), 0).astype(np.uint8)


def _channel_sigma(s, average):
    s = np.asarray(s, F).copy()
    if average:
        s[:] = ((s[0] + s[1]) + s[2]) / F(3)
    return s


def restate(img, op, sigmas, maxiter, alpha, continuous, stabilize, average, independent, write_iterations):
    """color_denoising.h:19-150: sigmas(which, image) gives the three channel
    levels of an estimate (which 0 first, 1 later); sigma0 takes the blend."""
    x = _anscombe(img) if stabilize else np.asarray(img, F)
    s0 = _channel_sigma(sigmas(0, x), average)
    P, log, written = x, [], []
    a = F(alpha)
    for it in range(1, maxiter + 1):
        out = O.filter_rgb(x, P, s0, op, independent)
        log.append(s0.copy())
        if write_iterations:
            y = _inverse_anscombe(out) if stabilize else out
            written.append(_to_rgb8(y))
            out = _anscombe(y) if stabilize else y
        if (maxiter > 1) if it == 1 else continuous:
            s1 = _channel_sigma(sigmas(1, out), average)
            s0 = a * s0 + (F(1) - a) * s1
        P = out
    final = _inverse_anscombe(P) if stabilize else P
    return final, np.array(log), written


def _manual(s_first, s_later):
    return lambda which, _img: [s_first if which == 0 else s_later] * 3


def _rgb_params(rs, rp, iters, mode="manual", sigma=(1.0, 1.0), alpha=0.5, continuous=False):
    p = DenoiseParameters(radius_searchspace=[rs, rs, 5], radius_patchspace=[rp, rp, 2], maxiterations=iters,
                          alpha=alpha)
    p.noiselevel.mode = mode
    p.noiselevel.sigma = list(sigma)
    p.noiselevel.continuous_estimate = continuous
    return p


# ---------------------------------------------------------------------------
# GPU: the HIP path against the oracle
# ---------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("rs,rp,shape", [(2, 1, (23, 37)), (4, 1, (40, 48)), (3, 2, (17, 64)), (5, 3, (33, 30))])
@pytest.mark.parametrize("independent", [False, True])
def test_filter_bit_exact_vs_oracle(gpu_available, rs, rp, shape, independent):
    img = _image(*shape, seed=20 + rs + rp)
    p = _rgb_params(rs, rp, 1, sigma=(14.0, 9.0))
    print(f"[gpu] rgb filter rs={rs} rp={rp} shape={shape} independent={independent}", flush=True)
    out, log = denoise_rgb(img, p, independent_channels=independent, stabilize_variance=False,
                           average_channelsigma=False)
    ref = O.filter_rgb(img, img, [14.0] * 3, _op(rs, rp), independent)
    assert np.array_equal(log, np.full((1, 3), 14.0, F))
    assert np.array_equal(_bits(out), _bits(ref))


@pytest.mark.gpu
@pytest.mark.parametrize("continuous", [False, True])
@pytest.mark.parametrize("independent", [False, True])
def test_schedule_manual_sigma_bit_exact(gpu_available, continuous, independent):
    img = _image(36, 44, 31)
    p = _rgb_params(3, 1, 4, sigma=(1.3, 0.7), alpha=0.3, continuous=continuous)
    print(f"[gpu] rgb schedule continuous={continuous} independent={independent}", flush=True)
    out, log = denoise_rgb(img, p, independent_channels=independent)
    ref, rlog, _ = restate(img, _op(3, 1), _manual(1.3, 0.7), 4, 0.3, continuous, True, True, independent, False)
    assert np.array_equal(_bits(log), _bits(rlog))
    assert np.array_equal(_bits(out), _bits(ref))
    if continuous:  # sigma0 takes every blend
        assert len(set(map(float, log[:, 0]))) == 4


def _estimate_oracle(chans, samples, patchsize=15, k=1.0):
    """noiselevel.cpp:96-124 per channel, FP64 statistics."""
    s = []
    for c in range(3):
        sd = O.patchstd(np.ascontiguousarray(chans[c][None], F), samples, patchsize, "f64")
        m, st = O.create2clusters(sd, precision="f64")
        s.append(F(m + np.float64(F(k)) * st))
    return np.array(s, F)


@pytest.mark.gpu
@pytest.mark.parametrize("average", [True, False])
def test_schedule_estimated_sigma(gpu_available, average):
    img = _image(48, 56, 41)
    ny, nx = img.shape[1:]
    p = _rgb_params(3, 1, 3, mode="global")
    print(f"[gpu] rgb estimated sigma average={average}", flush=True)
    out, log = denoise_rgb(img, p, average_channelsigma=average)
    samples = _capi.draw_samples([nx, ny, 1], 2048, 15, p.sample_seed)  # 2048 samples in colour mode
    est0 = _channel_sigma(_estimate_oracle(_anscombe(img), samples), average)
    assert np.allclose(log[0], est0, rtol=1e-6)
    # the filter chain with the logged levels, and the blend of the second
    levels = iter(log)
    x = _anscombe(img)
    P = x
    for it in range(3):
        P = O.filter_rgb(x, P, next(levels), _op(3, 1), False)
        if it == 0:
            s1 = _channel_sigma(_estimate_oracle(P, samples), average)
            assert np.allclose(log[1], F(0.5) * log[0] + F(0.5) * s1, rtol=1e-6)
    assert np.array_equal(log[1], log[2])  # no continuous estimate: the blend holds
    assert np.array_equal(_bits(out), _bits(_inverse_anscombe(P)))


@pytest.mark.gpu
@pytest.mark.parametrize("cleanup", [False, True])
def test_file_program(gpu_available, tmp_path, cleanup):
    imgs = {"leaf": _image(30, 41, 51).astype(np.uint8), "bark": _image(26, 33, 52).astype(np.uint8)}
    inp = tmp_path / "photos"
    inp.mkdir()
    for name, im in imgs.items():
        Image.fromarray(np.moveaxis(im, 0, 2), "RGB").save(inp / f"{name}.tif")
    out = tmp_path / "den"
    p = _rgb_params(3, 1, 3, sigma=(1.1, 0.6))
    print(f"[gpu] rgb file program cleanup={cleanup}", flush=True)
    vio.denoise_rgb_files(str(inp), p, str(out), cleanup=cleanup)
    for name, im in imgs.items():
        final, _, written = restate(im.astype(F), _op(3, 1), _manual(1.1, 0.6), 3, 0.5, False, True, True, False,
                                    not cleanup)
        got = np.asarray(Image.open(out / f"{name}_denoised.tif"))
        assert np.array_equal(got, np.moveaxis(_to_rgb8(final), 0, 2)), name
        for it in range(1, 4):
            f = out / "denoise_iterations" / f"{name}_iter{it:02d}.tif"
            assert os.path.exists(f) != cleanup
            if not cleanup:
                assert np.array_equal(np.asarray(Image.open(f)), np.moveaxis(written[it - 1], 0, 2)), (name, it)


@pytest.mark.gpu
def test_command_line_rgb(gpu_available, tmp_path):
    import subprocess
    from microtomodenoise_amd import build as B
    if not os.path.exists(B.CLI_BIN):
        pytest.skip("command-line program not built (python -m microtomodenoise_amd.build)")
    im = _image(28, 35, 61).astype(np.uint8)
    inp = tmp_path / "photos"
    inp.mkdir()
    Image.fromarray(np.moveaxis(im, 0, 2), "RGB").save(inp / "x.tif")
    out = tmp_path / "den"
    print("[gpu] inlm --rgb", flush=True)
    r = subprocess.run([B.CLI_BIN, str(inp), "-o", str(out), "--rgb", "--independent", "-search", "3", "-patch", "1",
                        "-iter", "2", "-noise", "manual", "-s0", "1.1", "-s1", "0.6"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    final, _, written = restate(im.astype(F), _op(3, 1), _manual(1.1, 0.6), 2, 0.5, False, True, True, True, True)
    assert np.array_equal(np.asarray(Image.open(out / "x_denoised.tif")), np.moveaxis(_to_rgb8(final), 0, 2))
    assert np.array_equal(np.asarray(Image.open(out / "denoise_iterations" / "x_iter02.tif")),
                          np.moveaxis(written[1], 0, 2))
