"""GPU parity tests: the HIP path through the C ABI against the reference's
golden fixtures (bit-exact) and the oracle (bit-exact for the FP32 form,
stated tolerances otherwise)."""
import ctypes as C

import numpy as np
import pytest

from conftest import filter_fixtures, load_npz
from microtomodenoise_amd import DenoiseParameters, IterativeNLM_GPU, _capi
from microtomodenoise_amd import phantom as ph
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def make_gpu(shape_xyz, rs, rp, *, blockwise=False, weight_fn="pade", kernel=0, noise="global", nslices=11,
             n_samples=100000, zwindow=100, **kw):
    p = DenoiseParameters(radius_searchspace=[int(v) for v in rs], radius_patchspace=[int(v) for v in rp],
                          nslices=nslices, weight_fn=weight_fn, kernel_select=kernel)
    p.cpu.blockwise = blockwise
    p.noiselevel.mode = noise
    p.noiselevel.n_samples = n_samples
    p.noiselevel.zwindow = zwindow
    for k, v in kw.items():
        setattr(p, k, v)
    g = IterativeNLM_GPU()
    assert g.configure_device([int(v) for v in shape_xyz], p) == shape_xyz[2]
    return g, p


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("name", filter_fixtures())
def test_golden_fixture_bit_exact(gpu_available, name):
    f = load_npz(name)
    raw = f["raw"]
    nz, ny, nx = raw.shape
    g, _ = make_gpu((nx, ny, nz), f["rs"], f["rp"])
    print(f"[gpu] fixture {name} shape {raw.shape}", flush=True)
    g.set_sigma(f["sigma1"])
    g.Run_GaussianNoise(1, raw)
    out1 = g.get_result()
    g.set_sigma(f["sigma2"])
    g.Run_GaussianNoise(2)
    out2 = g.get_result()
    g.free_device()
    assert np.array_equal(bits(out1), bits(f["out1"])), f"iteration 1 differs in {np.sum(out1 != f['out1'])} voxels"
    assert np.array_equal(bits(out2), bits(f["out2"])), f"iteration 2 differs in {np.sum(out2 != f['out2'])} voxels"


@pytest.mark.parametrize("rp", [(1, 1, 1), (2, 2, 2), (3, 3, 3)])
def test_tiled_equals_pervoxel_on_ragged_shape(gpu_available, rp):
    spec = ph.PhantomSpec((37, 29, 19), seed=21, radii=(2, 7), streaks=4, streak_len=(5, 20))
    raw = ph.make(spec)
    sig = np.linspace(900, 1600, 19).astype(np.float32)
    outs = []
    for kernel in (0, 1):
        print(f"[gpu] ragged rp={rp} kernel={kernel}", flush=True)
        g, _ = make_gpu((37, 29, 19), (3, 3, 3), rp, kernel=kernel)
        g.set_sigma(sig)
        g.Run_GaussianNoise(1, raw)
        outs.append(g.get_result())
        g.free_device()
    assert np.array_equal(bits(outs[0]), bits(outs[1]))
    p = O.Params(radius_searchspace=(3, 3, 3), radius_patchspace=rp, blockwise=False)
    ref = O.filter_iteration(raw, raw, sig, p)
    assert np.array_equal(bits(outs[0]), bits(ref))


def test_config1_parity_vs_oracle(gpu_available):
    """Config 1 (128x128x64 uint16 phantom, patch r=1, search r=5, 4
    iterations, fixed sigma lists) against the FP32 restatement: every voxel
    bit-identical, hence within 1e-4 relative and 1 LSB after uint16
    re-quantisation."""
    raw = ph.make(ph.CONFIGS["cfg1"])
    nz, ny, nx = raw.shape
    s0 = np.full(nz, 1500.0, np.float32)
    s1 = np.full(nz, 1100.0, np.float32)
    g, _ = make_gpu((nx, ny, nz), (5, 5, 5), (1, 1, 1), blockwise=True)
    p = O.Params(radius_searchspace=(5, 5, 5), radius_patchspace=(1, 1, 1), blockwise=True)
    prev = raw
    for it in range(1, 5):
        sig = s0 if it == 1 else s1
        print(f"[gpu] config1 iteration {it}", flush=True)
        g.set_sigma(sig)
        g.Run_GaussianNoise(it, raw if it == 1 else None)
        got = g.get_result()
        ref = O.filter_iteration(raw, prev, sig, p)
        rel = np.max(np.abs(got.astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-30))
        lsb = np.max(np.abs(O.requant_u16(got).astype(np.int32) - O.requant_u16(ref).astype(np.int32)))
        assert rel <= 1e-4 and lsb <= 1
        assert np.array_equal(bits(got), bits(ref)), f"iteration {it}: {np.sum(got != ref)} voxels differ"
        prev = ref
    g.free_device()


def test_exp_weight_mode_close_to_libm(gpu_available):
    """exp mode (the reference GPU's weight): parity unpinned against CUDA's
    expf.  Checked against the oracle with libm expf: the two expf differ by
    an ulp here and there, which moves a voxel by at most 1e-4 relative (the
    plain path's parity bound) and leaves 99% of voxels within 4 ulp."""
    f = load_npz("filter_p222_s2_u16.npz")
    raw = f["raw"]
    nz, ny, nx = raw.shape
    g, _ = make_gpu((nx, ny, nz), f["rs"], f["rp"], weight_fn="exp")
    g.set_sigma(f["sigma1"])
    g.Run_GaussianNoise(1, raw)
    got = g.get_result()
    g.free_device()
    p = O.Params(radius_searchspace=tuple(f["rs"]), radius_patchspace=tuple(f["rp"]), blockwise=False,
                 weight_mode=1)
    ref = O.filter_iteration(raw, raw, f["sigma1"], p)
    rel = np.abs(got.astype(np.float64) - ref) / np.abs(ref)
    assert rel.max() < 1e-4
    assert np.mean(rel < 4 * 2.0 ** -23) > 0.99


def test_patchstd_and_noise_estimate_vs_oracle(gpu_available):
    spec = ph.PhantomSpec((96, 80, 40), seed=31, radii=(3, 9), noise_sigma=1500.0, blur_sigma=0.8)
    raw = ph.make(spec)
    nz, ny, nx = raw.shape
    samples = _capi.draw_samples([nx, ny, nz], 20000, 15, 7)
    for mode in ("global", "z-adaptive"):
        print(f"[gpu] noise {mode}", flush=True)
        g, params = make_gpu((nx, ny, nz), (2, 2, 2), (1, 1, 1), noise=mode, zwindow=10)
        g.load_raw(raw)
        sd = g.patchstd(samples, 0)
        ref = O.patchstd(raw, samples, 15, "f64")
        assert np.allclose(sd, ref, rtol=1e-12, atol=1e-9)
        est = g.estimate_noise(samples, 0)
        g.free_device()
        op = O.Params(noise_mode=mode, zwindow=10)
        oref = O.noiselevel(raw, samples, op, precision="f64")
        assert np.allclose(est, oref.astype(np.float32), rtol=1e-6)


def test_full_schedule_vs_oracle(gpu_available):
    """main.cpp:325-545 schedule (global noise, 3 iterations): the GPU's FP64
    noise estimate and blended sigma feed the FP32 filter; the oracle runs the
    same schedule with the same samples."""
    spec = ph.PhantomSpec((64, 48, 24), seed=41, radii=(3, 8), noise_sigma=1500.0, blur_sigma=0.8)
    raw = ph.make(spec)
    nz, ny, nx = raw.shape
    print("[gpu] full schedule", flush=True)
    g, params = make_gpu((nx, ny, nz), (3, 3, 3), (1, 1, 1), blockwise=True, maxiterations=3, n_samples=5000)
    out, log = g.denoise(raw)
    g.free_device()
    samples = _capi.draw_samples([nx, ny, nz], params.noiselevel.n_samples, 15, params.sample_seed)
    op = O.Params(radius_searchspace=(3, 3, 3), radius_patchspace=(1, 1, 1), blockwise=True, maxiterations=3,
                  noise_mode="global")
    s0 = O.noiselevel(raw, samples, op, precision="f64").astype(np.float32)
    assert np.allclose(log[0], s0, rtol=1e-6)
    ref = O.iternlm(raw, op, sigma0=log[0], sigma1=log[1])
    assert np.array_equal(bits(out), bits(ref))


def test_division_selftest_exhaustive(gpu_available):
    """The parity kernel's shortened division equals IEEE division for the
    Pade weight at every float exponent in (-3.56648, 0]."""
    mism, checked, ex = C.c_uint64(), C.c_uint64(), C.c_uint32()
    _capi.check(_capi.lib().inlm_selftest_division(0, C.byref(mism), C.byref(checked), C.byref(ex)))
    assert checked.value > 1_000_000_000
    assert mism.value == 0, f"{mism.value} mismatches, e.g. bits {ex.value:#x}"


# Error budgets of the matrix-core forms against the FP64 oracle of the same
# algorithm, in units of the phantom's white-noise sigma (one iteration, same
# sigma list).  FP32_FAST differs by FP16 D planes and fused arithmetic; FP16
# and BF16 also store the volumes in 16 bits (BF16 keeps 8 significant bits,
# so its budget is the loosest); FP16_FP8 keeps FP16 storage and rounds the
# squared differences and the patch weights to FP8 (4 significant bits).
# The largest single errors come from patch pairs at the weight's cutoff,
# where the reference's Pade weight jumps by 0.025 of the maximum: rounding
# moves a distance across it and switches one term of the sum, which moves
# the output by at most 0.025 |R - out| / sum(w) <= 0.025 * contrast (about
# 0.33 noise sigma for these phantoms, 20000 between levels at sigma 1500).
# Hence the fraction of voxels beyond 0.1 sigma next to max and RMS.
FAST_BUDGET = {"fp32_fast": (0.35, 2e-4), "fp16": (0.35, 0.002), "bf16": (0.8, 0.008), "fp16_fp8": (0.8, 0.01)}
# largest fraction of voxels further than 0.1 noise sigma from the oracle
FAST_FRAC = {"fp32_fast": 0.001, "fp16": 0.001, "bf16": 0.001, "fp16_fp8": 0.001}


@pytest.mark.parametrize("form", ["fp32_fast", "fp16", "bf16", "fp16_fp8"])
@pytest.mark.parametrize("rp,rs,nx", [((1, 1, 1), (3, 3, 3), 40), ((2, 2, 2), (3, 3, 3), 40),
                                      ((3, 3, 3), (2, 2, 2), 40), ((2, 2, 2), (1, 1, 1), 40),
                                      ((2, 2, 2), (3, 3, 3), 41), ((1, 1, 1), (2, 2, 2), 37)])
def test_fast_forms_vs_fp64_oracle(gpu_available, form, rp, rs, nx):
    # nx odd: the padded row pitch of the 16-bit forms gains an unused column
    spec = ph.PhantomSpec((nx, 30, 20), seed=51, radii=(2, 7), noise_sigma=1500.0, blur_sigma=0.8, streaks=3,
                          streak_len=(5, 15))
    raw = ph.make(spec)
    nz, ny, nx = raw.shape
    sig = np.full(nz, 700.0, np.float32)
    print(f"[gpu] fast form {form} rp={rp} rs={rs} nx={nx}", flush=True)
    g, _ = make_gpu((nx, ny, nz), rs, rp, precision=form)
    g.set_sigma(sig)
    g.Run_GaussianNoise(1, raw)
    got = g.get_result().astype(np.float64)
    g.free_device()
    p = O.Params(radius_searchspace=rs, radius_patchspace=rp, blockwise=False)
    ref = O.filter_iteration(raw, raw, sig.astype(np.float64), p, precision="f64")
    err = np.abs(got - ref) / spec.noise_sigma
    mx, rms, frac = float(err.max()), float(np.sqrt(np.mean(err ** 2))), float(np.mean(err > 0.1))
    print(f"[gpu]   max {mx:.3e} rms {rms:.3e} (noise-sigma units), {frac:.4f} of voxels beyond 0.1", flush=True)
    bmax, brms = FAST_BUDGET[form]
    assert mx <= bmax and rms <= brms and frac <= FAST_FRAC[form], (mx, rms, frac)


@pytest.mark.parametrize("rp", [(2, 2, 2), (3, 3, 3)])
@pytest.mark.parametrize("form", ["fp32_fast", "fp16", "fp16_fp8"])
def test_fast_forms_anisotropic_patch(gpu_available, form, rp):
    # z anisotropy makes the weights depend on (dx^2 + dy^2, dz^2) rather
    # than on the distance alone: the per-pattern (not per-class) operands
    spec = ph.PhantomSpec((40, 30, 20), seed=52, radii=(2, 7), noise_sigma=1500.0, blur_sigma=0.8)
    raw = ph.make(spec)
    nz, ny, nx = raw.shape
    rs = (3, 3, 2)
    sig = np.full(nz, 700.0, np.float32)
    nzp = nz + 2 * (min(11 // 2, rs[2]) + rp[2])
    iso = O.cpu_weight_table(O.Params(radius_searchspace=rs, radius_patchspace=rp, blockwise=False), nzp)
    ani = O.cpu_weight_table(O.Params(radius_searchspace=rs, radius_patchspace=rp, blockwise=False,
                                      z_anisotropy=2.0), nzp)
    assert not np.allclose(iso, ani, rtol=1e-3)  # the weights really depend on it
    print(f"[gpu] anisotropic patch, fast form {form}", flush=True)
    g, _ = make_gpu((nx, ny, nz), rs, rp, precision=form, z_anisotropy=2.0)
    g.set_sigma(sig)
    g.Run_GaussianNoise(1, raw)
    got = g.get_result().astype(np.float64)
    g.free_device()
    p = O.Params(radius_searchspace=rs, radius_patchspace=rp, blockwise=False, z_anisotropy=2.0)
    ref = O.filter_iteration(raw, raw, sig.astype(np.float64), p, precision="f64")
    err = np.abs(got - ref) / spec.noise_sigma
    mx, rms, frac = float(err.max()), float(np.sqrt(np.mean(err ** 2))), float(np.mean(err > 0.1))
    print(f"[gpu]   max {mx:.3e} rms {rms:.3e} (noise-sigma units), {frac:.4f} of voxels beyond 0.1", flush=True)
    bmax, brms = FAST_BUDGET[form]
    assert mx <= bmax and rms <= brms and frac <= FAST_FRAC[form], (mx, rms, frac)


@pytest.mark.parametrize("rp", [(2, 2, 2), (3, 3, 3)])
@pytest.mark.parametrize("form", ["fp16", "bf16", "fp16_fp8"])
def test_fast_forms_zvarying_sigma(gpu_available, form, rp):
    # sigma rising threefold along z: the per-tile plane scales take the
    # tile's largest sigma, so slices with small sigma keep their cutoff
    # (a saturated tap still gives weight zero there)
    spec = ph.PhantomSpec((44, 30, 24), seed=53, radii=(2, 7), noise_sigma=1500.0, blur_sigma=0.8, streaks=3,
                          streak_len=(5, 15))
    raw = ph.make(spec)
    nz, ny, nx = raw.shape
    rs = (3, 3, 2)
    sig = np.linspace(400.0, 1200.0, nz).astype(np.float32)
    print(f"[gpu] z-varying sigma, fast form {form} rp={rp}", flush=True)
    g, _ = make_gpu((nx, ny, nz), rs, rp, precision=form)
    g.set_sigma(sig)
    g.Run_GaussianNoise(1, raw)
    got = g.get_result().astype(np.float64)
    g.free_device()
    p = O.Params(radius_searchspace=rs, radius_patchspace=rp, blockwise=False)
    ref = O.filter_iteration(raw, raw, sig.astype(np.float64), p, precision="f64")
    err = np.abs(got - ref) / spec.noise_sigma
    mx, rms, frac = float(err.max()), float(np.sqrt(np.mean(err ** 2))), float(np.mean(err > 0.1))
    print(f"[gpu]   max {mx:.3e} rms {rms:.3e} (noise-sigma units), {frac:.4f} of voxels beyond 0.1", flush=True)
    bmax, brms = FAST_BUDGET[form]
    assert mx <= bmax and rms <= brms and frac <= FAST_FRAC[form], (mx, rms, frac)


@pytest.mark.parametrize("form", ["fp16", "bf16"])
@pytest.mark.parametrize("shape,rs", [((40, 30, 20), (3, 3, 3)), ((41, 35, 13), (4, 4, 2)), ((16, 16, 8), (2, 2, 2)),
                                      ((67, 18, 9), (7, 7, 5)), ((70, 33, 17), (7, 3, 2)), ((24, 20, 10), (1, 1, 1))])
def test_block_kernel_equals_row_kernel(gpu_available, form, shape, rs):
    """At patch radius 2 the 16-bit forms run the 4x4-block kernel (seven
    MFMAs per half over 32x16x8 tiles); the row kernel (kernel_select 2,
    eleven MFMAs over 32x8x8 tiles) stages the same FP16 planes with the same
    per-tile scale, so the two differ only in the order of FP32 sums: ragged
    shapes, both storage types.  The search shapes give the FP16 form's run
    moves runs of every length from 1 (radius 1: no moves) to 8 (radius 7),
    and an anisotropic search."""
    spec = ph.PhantomSpec(shape, seed=53, radii=(2, 6), noise_sigma=1500.0, blur_sigma=0.8, streaks=2,
                          streak_len=(5, 12))
    raw = ph.make(spec)
    nz, ny, nx = raw.shape
    sig = np.linspace(600.0, 900.0, nz).astype(np.float32)
    outs = {}
    for ks in (0, 2):
        print(f"[gpu] {form} shape {shape} rs {rs} kernel_select {ks}", flush=True)
        g, _ = make_gpu((nx, ny, nz), rs, (2, 2, 2), precision=form, kernel=ks)
        g.set_sigma(sig)
        g.Run_GaussianNoise(1, raw)
        g.set_sigma(sig * np.float32(0.9))
        g.Run_GaussianNoise(2)
        outs[ks] = g.get_result().astype(np.float64)
        g.free_device()
    err = np.abs(outs[0] - outs[2]) / spec.noise_sigma
    mx, rms = float(err.max()), float(np.sqrt(np.mean(err ** 2)))
    print(f"[gpu]   block vs row: max {mx:.3e} rms {rms:.3e} (noise-sigma units)", flush=True)
    assert mx < 0.02 and rms < 2e-4, (mx, rms)


@pytest.mark.parametrize("form", ["fp16", "bf16", "fp16_fp8"])
def test_fast_forms_repeat_bit_for_bit(gpu_available, form):
    # the 16-bit forms normalise by the volume's FP64 moments; their
    # reduction runs in a fixed order, so two handles on the same volume
    # give the same result bit for bit
    spec = ph.PhantomSpec((200, 96, 40), seed=57, radii=(2, 9), noise_sigma=1500.0, blur_sigma=0.8)
    raw = ph.make(spec)
    nz, ny, nx = raw.shape
    sig = np.full(nz, 900.0, np.float32)
    print(f"[gpu] repeatability of {form}", flush=True)
    outs = []
    for _ in range(2):
        g, _ = make_gpu((nx, ny, nz), (3, 3, 3), (2, 2, 2), precision=form)
        g.set_sigma(sig)
        g.Run_GaussianNoise(1, raw)
        outs.append(g.get_result())
        g.free_device()
    assert np.array_equal(bits(outs[0]), bits(outs[1]))
