"""The C++ drop-in (include/iternlm_gpu.h) driven through every GPU branch of
the reference's main.cpp:311-545 by tests/cpp/facade_main.cpp: the
whole-volume Run_GaussianNoise, the host-RAM Run_GaussianNoise_GPUBlocks
(forced by a device memory limit below the volume's footprint, so that
configure_device returns fewer slices than the stack has), the file-list
Run_GaussianNoise_GPUBlocks (gpu.blockwise_host), and the resume path.  The
written iterations are checked bit for bit against the FP32 oracle, which is
pinned against the reference's own compiled CPU kernels."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_npz
from microtomodenoise_amd import _capi
from microtomodenoise_amd import io as vio
from oracle import oracle as O

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "cpp", "facade_main")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def setup_case(tmp_path, name="filter_p222_s3_f32.npz"):
    f = load_npz(name)
    raw = f["raw"].astype(np.float32)
    raw.tofile(tmp_path / "raw.bin")
    f["sigma1"].astype(np.float32).tofile(tmp_path / "sigma1.bin")
    f["sigma2"].astype(np.float32).tofile(tmp_path / "sigma2.bin")
    vio.write_stack(str(tmp_path / "raw"), "raw", raw)
    return f, raw


def run(tmp_path, f, raw, blockwise_host=0, memory_limit=0, resume=0):
    assert os.path.exists(EXE), "tests/cpp/facade_main is not built (run __graft_entry__.build())"
    nz, ny, nx = raw.shape
    args = [EXE, str(tmp_path), str(nx), str(ny), str(nz), str(int(f["rs"][0])), str(int(f["rp"][0])), "0",
            str(blockwise_host), str(memory_limit), str(resume)]
    print("[gpu] facade", " ".join(args[2:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split("\n")
    npar = int(next(x for x in lines if x.startswith("n_parallelslices")).split()[1])
    resumed = [int(x.split()[3]) for x in lines if x.startswith("iteration ") and "resumed" in x]
    return npar, resumed


def oracle_two_iterations(f, raw):
    p = O.Params(radius_searchspace=tuple(int(v) for v in f["rs"]), radius_patchspace=tuple(int(v) for v in f["rp"]),
                 blockwise=False)
    o1 = O.filter_iteration(raw, raw, f["sigma1"], p)
    o2 = O.filter_iteration(raw, o1, f["sigma2"], p)
    return o1, o2


def written(tmp_path, it):
    return vio.read_stack(str(tmp_path / "out" / f"denoise-iteration{it:02d}"))


def footprint_limit(raw, f, slices):
    """A device memory limit that holds slabs of exactly `slices` output
    slices for this case (from the library's own capacity query)."""
    nz, ny, nx = raw.shape
    p = _capi.default_params()
    for i in range(3):
        p.radius_searchspace[i] = int(f["rs"][i])
        p.radius_patchspace[i] = int(f["rp"][i])
    p.cpu_blockwise = 0
    lo, hi = 1, 1 << 34
    while hi - lo > 1:  # smallest limit with capacity >= slices
        mid = (lo + hi) // 2
        try:
            n = _capi.capacity_slices(p, (nx, ny, nz), 0, mid)
        except _capi.InlmError:
            n = 0
        if n >= slices:
            hi = mid
        else:
            lo = mid
    assert _capi.capacity_slices(p, (nx, ny, nz), 0, hi) == slices
    return hi


def test_whole_volume_path(gpu_available, tmp_path):
    f, raw = setup_case(tmp_path)
    npar, resumed = run(tmp_path, f, raw)
    assert npar == raw.shape[0] and resumed == [0, 0]
    o1, o2 = oracle_two_iterations(f, raw)
    assert np.array_equal(bits(written(tmp_path, 1)), bits(o1))
    assert np.array_equal(bits(written(tmp_path, 2)), bits(o2))
    assert np.array_equal(bits(o1), bits(f["out1"])) and np.array_equal(bits(o2), bits(f["out2"]))


@pytest.mark.parametrize("slices", [1, 3, 5])
def test_host_blocks_path(gpu_available, tmp_path, slices):
    f, raw = setup_case(tmp_path)
    limit = footprint_limit(raw, f, slices)
    npar, _ = run(tmp_path, f, raw, memory_limit=limit)
    assert npar == slices < raw.shape[0]
    o1, o2 = oracle_two_iterations(f, raw)
    assert np.array_equal(bits(written(tmp_path, 1)), bits(o1))
    assert np.array_equal(bits(written(tmp_path, 2)), bits(o2))


@pytest.mark.parametrize("slices", [2, 16])
def test_file_blocks_path(gpu_available, tmp_path, slices):
    f, raw = setup_case(tmp_path)
    limit = footprint_limit(raw, f, slices) if slices < raw.shape[0] else 0
    npar, _ = run(tmp_path, f, raw, blockwise_host=1, memory_limit=limit)
    assert npar == min(slices, raw.shape[0])
    o1, o2 = oracle_two_iterations(f, raw)
    assert np.array_equal(bits(written(tmp_path, 1)), bits(o1))
    assert np.array_equal(bits(written(tmp_path, 2)), bits(o2))


@pytest.mark.parametrize("blocks", [False, True])
def test_resume_reads_finished_iteration(gpu_available, tmp_path, blocks):
    f, raw = setup_case(tmp_path)
    limit = footprint_limit(raw, f, 3) if blocks else 0
    run(tmp_path, f, raw, memory_limit=limit)
    o1, o2 = oracle_two_iterations(f, raw)
    # iteration 2 lost; iteration 1 is replaced on disk by a marked copy so
    # that reading it back (not recomputing it) is visible in iteration 2
    d2 = tmp_path / "out" / "denoise-iteration02"
    for p in d2.iterdir():
        p.unlink()
    marked = o1 + np.float32(7.0)
    vio.write_stack(str(tmp_path / "out" / "denoise-iteration01"), "denoised", marked)
    _, resumed = run(tmp_path, f, raw, memory_limit=limit, resume=1)
    assert resumed == [1, 0]
    assert np.array_equal(bits(written(tmp_path, 1)), bits(marked))
    p = O.Params(radius_searchspace=tuple(int(v) for v in f["rs"]), radius_patchspace=tuple(int(v) for v in f["rp"]),
                 blockwise=False)
    want2 = O.filter_iteration(raw, marked, f["sigma2"], p)
    assert np.array_equal(bits(written(tmp_path, 2)), bits(want2))
