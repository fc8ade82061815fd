"""The out-of-core slab schedule (inlm_denoise_slabs) against the
whole-volume schedule: the reference's block drivers
(Run_GaussianNoise_GPUBlocks, iternlm_gpu.cu:1729-2115; iternlm_cpu.cpp:403-577)
compute the same filter slab by slab, so the FP32 form must agree bit for
bit for every slab size, including slabs thinner than the filter's z reach
and a ragged last slab, with sigma indexed by global z."""
import numpy as np
import pytest

from conftest import ROOT  # noqa: F401
from microtomodenoise_amd import DenoiseParameters, InlmError, denoise, denoise_slabs
from microtomodenoise_amd import _capi
from microtomodenoise_amd import phantom as ph


def _params(mode="global", precision="fp32", continuous=False, iters=3):
    p = DenoiseParameters(radius_searchspace=[3, 3, 2], radius_patchspace=[1, 1, 1], maxiterations=iters,
                          precision=precision)
    p.noiselevel.mode = mode
    p.noiselevel.n_samples = 20000
    p.noiselevel.zwindow = 5
    p.noiselevel.continuous_estimate = continuous
    return p


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.gpu
@pytest.mark.parametrize("mode,continuous", [("global", False), ("z-adaptive", True)])
def test_slabs_equal_whole_volume_fp32(gpu_available, mode, continuous):
    spec = ph.PhantomSpec((40, 32, 14), seed=95, radii=(2, 8))
    raw = ph.make(spec)
    print(f"[gpu] whole volume {mode}", flush=True)
    ref, rlog = denoise(raw, _params(mode, continuous=continuous, iters=4))
    for slab in (1, 2, 4, 5, 13):  # z reach here is min(11//2, 2) + 1 = 3 slices
        print(f"[gpu] slabs of {slab} slices, {mode}", flush=True)
        got, glog = denoise_slabs(raw, _params(mode, continuous=continuous, iters=4), slab)
        assert np.array_equal(_bits(got), _bits(ref)), slab
        assert np.array_equal(_bits(glog), _bits(rlog)), slab


@pytest.mark.gpu
def test_slabs_fp16_within_budget(gpu_available):
    spec = ph.PhantomSpec((40, 32, 14), seed=96, radii=(2, 8), noise_sigma=1500.0)
    raw = ph.make(spec)
    print("[gpu] fp16 whole volume and slabs of 4", flush=True)
    ref, _ = denoise(raw, _params(precision="fp16", iters=2))
    got, _ = denoise_slabs(raw, _params(precision="fp16", iters=2), 4)
    # each slab is normalised by its own moments: the 16-bit rounding
    # differs, not the filter
    err = np.abs(got.astype(np.float64) - ref) / spec.noise_sigma
    assert err.max() < 0.08 and np.sqrt(np.mean(err ** 2)) < 0.006, (err.max(), np.sqrt(np.mean(err ** 2)))


def test_slab_arguments_checked_on_the_host():
    raw = np.zeros((4, 20, 20), np.float32)
    for slab in (0, -3):
        with pytest.raises(InlmError) as e:
            denoise_slabs(raw, _params(), slab)
        assert e.value.status == _capi.INLM_BAD_ARG


@pytest.mark.gpu
@pytest.mark.parametrize("slab", [1, 3, 6])
def test_slabs_equal_fp32_oracle(gpu_available, slab):
    """The slab schedule against the oracle's schedule on the sigma lists the
    slabs used (z-adaptive, so sigma varies along z and is indexed by global
    z), bit for bit."""
    from oracle import oracle as O
    spec = ph.PhantomSpec((36, 30, 12), seed=97, radii=(2, 8))
    raw = ph.make(spec)
    print(f"[gpu] slabs of {slab} vs the oracle", flush=True)
    got, glog = denoise_slabs(raw, _params("z-adaptive", iters=3), slab)
    assert np.unique(glog[0]).size > 1  # sigma really varies along z
    p = O.Params(radius_searchspace=(3, 3, 2), radius_patchspace=(1, 1, 1), maxiterations=3, blockwise=True)
    want = O.iternlm(raw, p, sigma0=glog[0], sigma1=glog[1])
    assert np.array_equal(_bits(got), _bits(want))


@pytest.mark.gpu
@pytest.mark.parametrize("rp,rs", [((1, 1, 0), (3, 3, 0)), ((1, 1, 0), (3, 3, 2)), ((2, 2, 1), (2, 2, 1))])
def test_iterate_slabs_thin_slabs_fp32_oracle(gpu_available, rp, rs):
    """One slab iteration per call (inlm_iterate_slabs) with one-slice slabs:
    the region keeps at least two slices, so a 2D patch keeps the weight table
    of the whole volume (at patch radius 1 the reference's CPU kernel uses 1/7
    for every tap of a volume with more than one slice)."""
    from microtomodenoise_amd.denoiser import iterate_slabs
    from oracle import oracle as O
    spec = ph.PhantomSpec((30, 26, 7), seed=98, radii=(2, 6))
    raw = ph.make(spec)
    nz = raw.shape[0]
    sig1 = np.linspace(900, 1500, nz).astype(np.float32)
    sig2 = np.linspace(700, 1100, nz).astype(np.float32)
    prm = DenoiseParameters(radius_searchspace=list(rs), radius_patchspace=list(rp), maxiterations=2)
    prm.cpu.blockwise = True
    p = O.Params(radius_searchspace=rs, radius_patchspace=rp, blockwise=True)
    w1 = O.filter_iteration(raw, raw, sig1, p)
    w2 = O.filter_iteration(raw, w1, sig2, p)
    print(f"[gpu] iterate_slabs rp={rp} rs={rs}", flush=True)
    o1 = iterate_slabs(raw, raw, sig1, prm, 1, 1)
    assert np.array_equal(_bits(o1), _bits(w1))
    o2 = iterate_slabs(raw, o1, sig2, prm, 2, 1)
    assert np.array_equal(_bits(o2), _bits(w2))
