"""The command-line program (microtomodenoise_amd/inlm): the reference's flags
(main.cpp:41-219) parsed into the library's parameters, checked without a GPU
through --print-params, and one end-to-end run on the GPU against the
in-memory schedule."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT  # noqa: F401
from microtomodenoise_amd import DenoiseParameters, IterativeNLM_GPU
from microtomodenoise_amd import build as B
from microtomodenoise_amd import io as vio
from microtomodenoise_amd import phantom as ph

CLI = B.CLI_BIN


def run(*args, timeout=60):
    if not os.path.exists(CLI):
        pytest.skip("command-line program not built (python -m microtomodenoise_amd.build)")
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)


def params(*args):
    r = run(*args, "--print-params")
    assert r.returncode == 0, r.stderr
    return dict(line.split("=", 1) for line in r.stdout.strip().splitlines())


def test_defaults_are_the_reference_defaults():
    # denoise_parameters.h: search 10, patch 1, 4 iterations, alpha 0.5,
    # nslices 11, z-adaptive noise over 1e5 samples of 15x15, blockwise CPU
    # dispatch; io: no resume, full stack, 32-bit output
    d = params("in/")
    assert d["input"] == "in/" and d["output"] == ""
    assert d["search"] == "10,10,10" and d["patch"] == "1,1,1"
    assert d["maxiterations"] == "4" and d["alpha"] == "0.5" and d["nslices"] == "11"
    assert d["noise"] == "z-adaptive" and d["n_samples"] == "100000" and d["patchsize"] == "15"
    assert d["stds_from_mean"] == "1" and d["continuous"] == "0" and d["cpu_blockwise"] == "1"
    assert d["resume"] == "0" and d["firstslice"] == "-1" and d["lastslice"] == "-1"
    assert d["save_16bit"] == "0" and d["cleanup"] == "0" and d["precision"] == "fp32" and d["weight"] == "pade"


@pytest.mark.parametrize("args,want", [
    (("-i", "a", "-o", "b"), {"input": "a", "output": "b"}),
    (("-input", "a", "-output", "b"), {"input": "a", "output": "b"}),
    (("x", "-r", "true"), {"resume": "1"}),
    (("x", "-resume", "True", "-r", "false"), {"resume": "0"}),
    (("x", "-r", "maybe"), {"resume": "0"}),  # neither true nor false: ignored, as in the reference
    (("x", "--resume"), {"resume": "1"}),
    (("x", "-fs", "3", "-lastslice", "10"), {"firstslice": "3", "lastslice": "10"}),
    (("x", "--16bit", "--cleanup"), {"save_16bit": "1", "cleanup": "1"}),
    (("x", "-aniso", "2.5"), {"z_anisotropy": "2.5"}),
    (("x", "-iterations", "7", "-nslices", "5", "-a", "0.25"), {"maxiterations": "7", "nslices": "5", "alpha": "0.25"}),
    (("x", "-search", "7", "-search2", "3"), {"search": "7,7,3"}),
    (("x", "-search0", "4", "-search1", "5"), {"search": "4,5,10"}),
    (("x", "-patch", "2", "-patch0", "3"), {"patch": "3,2,2"}),
    (("x", "-noise", "global", "-s", "120"), {"noise": "global", "sigma": "120,120"}),
    (("x", "-noisemode", "manual", "-s0", "10", "-sigma1", "20"), {"noise": "manual", "sigma": "10,20"}),
    (("x", "-nsamples", "5000", "-noisepatch", "9", "-noiseshift", "2"),
     {"n_samples": "5000", "patchsize": "9", "stds_from_mean": "2"}),
    (("x", "--continuous"), {"continuous": "1"}),
    (("x", "-noise", "semimanual", "-s", "300"), {"noise": "semimanual", "sigma": "300,300"}),
    (("x", "-noise", "auto"), {"noise": "global"}),
    (("x", "-mask", "900"), {"mask_diameter": "900"}),
    (("x", "-mask_diameter", "0"), {"mask_diameter": "0"}),
    (("x", "--noblocks"), {"cpu_blockwise": "0"}),
    (("x", "--noblocks", "--blocks"), {"cpu_blockwise": "1"}),
    (("x", "-gpu0", "2"), {"device": "2"}),
    (("x", "-precision", "fp16", "-weight", "exp", "-seed", "9"), {"precision": "fp16", "weight": "exp", "seed": "9"}),
    (("x", "--verbose", "-threads", "256", "-cpus", "8", "--unroll"), {"verbose": "1"}),
    (("x", "--rgb", "--independent", "--nopoisson", "--noaverage"),
     {"rgb": "1", "independent_channels": "1", "stabilize_variance": "0", "average_channelsigma": "0"}),
    (("x", "--color"), {"rgb": "1", "independent_channels": "0", "stabilize_variance": "1", "average_channelsigma": "1"}),
])
def test_flags(args, want):
    d = params(*args)
    for k, v in want.items():
        assert d[k] == v, (k, d[k], v)


def test_preserve_detail_is_a_preset_that_later_flags_override():
    # main.cpp:195-202: continuous estimate, alpha 0, 4 stds, 6 iterations
    d = params("x", "--preserve_detail")
    assert (d["continuous"], d["alpha"], d["stds_from_mean"], d["maxiterations"]) == ("1", "0", "4", "6")
    d = params("x", "--preserve_detail", "-iter", "3", "-a", "0.5")
    assert (d["maxiterations"], d["alpha"]) == ("3", "0.5")
    d = params("x", "-iter", "3", "--preserve_detail")
    assert d["maxiterations"] == "6"


def test_first_positional_argument_is_the_input():
    d = params("first", "second")
    assert d["input"] == "first"
    d = params("-i", "given", "other")
    assert d["input"] == "given"


@pytest.mark.parametrize("args,msg", [
    (("x", "-search"), "missing value"),
    (("x", "-noise", "median"), "noise estimation mode"),
    (("x", "-precision", "fp8"), "precision"),
    ((), "no input"),
])
def test_rejected_arguments(args, msg):
    r = run(*args)
    assert r.returncode == 2 and msg in r.stderr, (r.returncode, r.stderr)


def test_help():
    r = run("-h")
    assert r.returncode == 0 and "usage: inlm" in r.stdout


@pytest.mark.gpu
def test_cli_run_matches_in_memory_schedule(gpu_available, tmp_path):
    spec = ph.PhantomSpec((36, 28, 10), seed=93, radii=(2, 7))
    raw = ph.make(spec)
    inp = tmp_path / "scan" / "raw"
    vio.write_stack(str(inp), "slice", raw, as16bit=True)
    out = tmp_path / "scan" / "den"
    print("[gpu] inlm command line", flush=True)
    r = run(str(inp), "-o", str(out), "-search", "3", "-patch", "1", "-iter", "2", "-noise", "global",
            "-nsamples", "20000", "--noblocks", timeout=120)
    assert r.returncode == 0, r.stderr
    got = vio.read_stack(str(out / "denoise-iteration02"))
    p = DenoiseParameters(radius_searchspace=[3, 3, 3], radius_patchspace=[1, 1, 1], maxiterations=2)
    p.noiselevel.mode = "global"
    p.noiselevel.n_samples = 20000
    p.cpu.blockwise = False
    nz, ny, nx = raw.shape
    g = IterativeNLM_GPU()
    assert g.configure_device([nx, ny, nz], p) == nz
    ref, _ = g.denoise(raw)
    g.free_device()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert "search space: (3,3,3)" in (tmp_path / "scan" / "logfile.txt").read_text()
