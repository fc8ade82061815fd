"""The noise estimator's modes end to end on the GPU (inlm_denoise): the
semimanual mode (Noise/noiselevel.cpp:56-79) and the circular sampling mask
(:204-224), against the oracle's FP64 estimator on the same sample list and
the oracle's FP32 filter on the sigma lists the library used."""
import numpy as np
import pytest

from microtomodenoise_amd import DenoiseParameters, IterativeNLM_GPU
from microtomodenoise_amd import _capi
from microtomodenoise_amd import phantom as ph
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def run_schedule(raw, p):
    nz, ny, nx = raw.shape
    g = IterativeNLM_GPU()
    assert g.configure_device([nx, ny, nz], p) == nz
    out, log = g.denoise(raw)
    g.free_device()
    return out, log


@pytest.mark.parametrize("mode,mask", [("semimanual", 0), ("z-adaptive", 70), ("global", 60), ("semimanual", 64)])
def test_noise_modes_schedule_vs_oracle(gpu_available, mode, mask):
    spec = ph.PhantomSpec((80, 72, 12), seed=61, radii=(3, 10), noise_sigma=1800.0, blur_sigma=0.8)
    raw = ph.make(spec)
    nz, ny, nx = raw.shape
    p = DenoiseParameters(radius_searchspace=[3, 3, 2], radius_patchspace=[1, 1, 1], maxiterations=3)
    p.noiselevel.mode = mode
    p.noiselevel.n_samples = 30000
    p.noiselevel.zwindow = 5
    p.noiselevel.sigma = [900.0, 0.0]
    p.noiselevel.circular_mask_diameter = mask
    print(f"[gpu] noise mode {mode} mask {mask}", flush=True)
    out, log = run_schedule(raw, p)

    samples = O.draw_samples((nx, ny, nz), 30000, 15, seed=0, mask_diameter=mask)
    assert np.array_equal(samples, _capi.draw_samples((nx, ny, nz), 30000, 15, 0, mask))
    op = O.Params(radius_searchspace=(3, 3, 2), radius_patchspace=(1, 1, 1), maxiterations=3, noise_mode=mode,
                  zwindow=5, sigma=(900.0, 0.0), blockwise=True)
    st = {}
    s0 = O.noiselevel(raw, samples, op, "f64", which=0, state=st).astype(np.float32)
    assert np.allclose(log[0], s0, rtol=2e-7, atol=0)
    out1 = O.filter_iteration(raw, raw, log[0], op)
    s1 = O.noiselevel(out1, samples, op, "f64", which=1, state=st)
    want1 = O.blend_sigma(log[0], s1.astype(np.float32), op.alpha)
    assert np.allclose(log[1], want1, rtol=2e-7, atol=0)
    if mode == "semimanual":
        assert np.all(log[0] == np.float32(900.0)) and np.unique(log[1]).size == 1
    want = O.iternlm(raw, op, sigma0=log[0], sigma1=log[1])
    assert np.array_equal(bits(out), bits(want))
