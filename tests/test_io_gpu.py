"""The reference program's whole on-disk run (main.cpp:231-563) through
inlm_denoise_files on the GPU: outputs, noise files and log in the
reference's layout, results identical to the in-memory schedule, resume
from finished iterations and cleanup."""
import os

import numpy as np
import pytest
from PIL import Image

from conftest import ROOT  # noqa: F401
from microtomodenoise_amd import DenoiseParameters, IterativeNLM_GPU
from microtomodenoise_amd import io as vio
from microtomodenoise_amd import phantom as ph

pytestmark = pytest.mark.gpu


def _params():
    p = DenoiseParameters(radius_searchspace=[3, 3, 2], radius_patchspace=[1, 1, 1], maxiterations=3)
    p.noiselevel.mode = "global"
    p.noiselevel.n_samples = 20000
    p.cpu.blockwise = False
    return p


def _in_memory(raw, p):
    nz, ny, nx = raw.shape
    g = IterativeNLM_GPU()
    assert g.configure_device([nx, ny, nz], p) == nz
    out, log = g.denoise(raw)
    g.free_device()
    return out, log


def test_denoise_files_layout_and_results(gpu_available, tmp_path):
    spec = ph.PhantomSpec((48, 40, 12), seed=91, radii=(2, 8))
    raw = ph.make(spec)  # uint16-valued floats
    inp = tmp_path / "data" / "raw"
    vio.write_stack(str(inp), "slice", raw, as16bit=True)
    out = tmp_path / "data" / "result"
    p = _params()
    print("[gpu] denoise_files", flush=True)
    vio.denoise_files(str(inp), p, str(out))
    for it in (1, 2, 3):
        d = out / f"denoise-iteration{it:02d}"
        assert sorted(os.listdir(d)) == [f"denoised{z:04d}.tif" for z in range(raw.shape[0])]
    assert (out / "noise_vs_slice00.csv").exists() and (out / "noise_vs_slice01.csv").exists()
    assert "search space: (3,3,2)" in (tmp_path / "data" / "logfile.txt").read_text()
    got = vio.read_stack(str(out / "denoise-iteration03"))
    ref, log = _in_memory(raw, _params())
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    s0 = vio.read_noise_csv(str(out / "noise_vs_slice00.csv"), raw.shape[0])
    assert np.allclose(s0, log[0], rtol=1e-5)
    im = Image.open(out / "denoise-iteration03" / "denoised0000.tif")
    assert im.mode == "F"

    # resume: the last iteration is gone; iterations 1-2 and the noise files
    # are taken from disk (the CSVs hold 6 significant digits, as the
    # reference writes them), so iteration 3 is one filter pass on the stored
    # iteration-2 result with the sigma blended from the CSV values
    for f in os.listdir(out / "denoise-iteration03"):
        os.remove(out / "denoise-iteration03" / f)
    os.rmdir(out / "denoise-iteration03")
    print("[gpu] denoise_files resume", flush=True)
    vio.denoise_files(str(inp), _params(), str(out), resume=True)
    again = vio.read_stack(str(out / "denoise-iteration03"))
    nz, ny, nx = raw.shape
    s1 = vio.read_noise_csv(str(out / "noise_vs_slice01.csv"), nz)
    a = np.float32(0.5)
    sig = (a * s0 + (np.float32(1) - a) * s1).astype(np.float32)
    g = IterativeNLM_GPU()
    assert g.configure_device([nx, ny, nz], _params()) == nz
    g.load_raw(raw)
    g.set_result(2, vio.read_stack(str(out / "denoise-iteration02")))
    g.set_sigma(sig)
    g.Run_GaussianNoise(3)
    want = g.get_result()
    g.free_device()
    assert np.array_equal(again.view(np.uint32), want.view(np.uint32))
    assert np.abs(again - ref).max() <= 1e-3 * np.abs(ref).max()


def test_denoise_files_cleanup_16bit_and_substack(gpu_available, tmp_path):
    spec = ph.PhantomSpec((40, 32, 14), seed=92, radii=(2, 8))
    raw = ph.make(spec)
    inp = tmp_path / "raw"
    vio.write_stack(str(inp), "s", raw, as16bit=True)
    out = tmp_path / "res"
    print("[gpu] denoise_files cleanup", flush=True)
    vio.denoise_files(str(inp), _params(), str(out), cleanup=True, save_16bit=True, firstslice=3, lastslice=10)
    names = sorted(os.listdir(out))
    assert names == [f"denoised{z:04d}.tif" for z in range(3, 11)]  # no iteration folders, no noise files
    got = vio.read_stack(str(out))
    ref, _ = _in_memory(raw[3:11], _params())
    want = np.clip(np.sign(ref) * np.floor(np.abs(ref.astype(np.float64)) + 0.5), 0, 65535)
    assert np.array_equal(got, want.astype(np.float32))
