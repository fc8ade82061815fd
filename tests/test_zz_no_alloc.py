"""inlm_patchstd_noalloc (no device allocation, a fixed device array as
scratch) against inlm_patchstd (the handle's allocated buffers): the same
kernel, so the same doubles, bit for bit, on two calls in a row, the second
with fewer samples and another volume; and the limit is enforced on the host."""
import numpy as np
import pytest

from conftest import ROOT  # noqa: F401
from microtomodenoise_amd import DenoiseParameters, IterativeNLM_GPU, _capi
from microtomodenoise_amd import phantom as ph

pytestmark = pytest.mark.gpu


def handle_patchstd(vol, samples, patchsize):
    """inlm_patchstd on a handle holding the volume (the facade's patchstd)."""
    nz, ny, nx = vol.shape
    p = DenoiseParameters(radius_searchspace=[2, 2, 2], radius_patchspace=[1, 1, 1])
    p.noiselevel.patchsize = patchsize
    g = IterativeNLM_GPU()
    assert g.configure_device([nx, ny, nz], p) == nz
    try:
        g.load_raw(vol)
        return g.patchstd(samples, 0)
    finally:
        g.free_device()


def test_noalloc_patchstd_equals_handle_patchstd(gpu_available):
    import torch
    cases = [((96, 80, 24), 60000, 15, 31), ((50, 40, 9), 700, 9, 32)]  # second call: fewer samples
    for (nx, ny, nz), n, ps, seed in cases:
        vol = ph.make(ph.PhantomSpec((nx, ny, nz), seed=seed, radii=(2, 9), noise_sigma=1500.0, blur_sigma=0.8))
        samples = _capi.draw_samples((nx, ny, nz), n, ps, seed=seed)
        print(f"[gpu] no-alloc patchstd {nx}x{ny}x{nz}, {n} samples, patch {ps}", flush=True)
        want = handle_patchstd(vol, samples, ps)
        dev = torch.from_numpy(vol).cuda()
        torch.cuda.synchronize()
        got = _capi.patchstd_noalloc(dev.data_ptr(), (nx, ny, nz), ps, samples)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
        assert np.all(got > 0)


def test_noalloc_rejects_more_samples_than_its_array(gpu_available):
    import torch
    dev = torch.zeros((4, 32, 32), dtype=torch.float32, device="cuda")
    samples = np.zeros(_capi.NOALLOC_MAX_SAMPLES + 1, np.int64)
    with pytest.raises(_capi.InlmError) as e:
        _capi.patchstd_noalloc(dev.data_ptr(), (32, 32, 4), 15, samples)
    assert e.value.status == _capi.INLM_BAD_ARG and "at most" in str(e.value)
