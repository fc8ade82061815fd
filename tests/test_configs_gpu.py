"""The benchmark suite's configurations at their full workloads on the GPU.

Each configuration runs its phantom recipe (microtomodenoise_amd.phantom.
CONFIGS), its noise mode, alpha and iteration count through the schedule of
main.cpp:325-545 (sigma0 on the original, sigma1 re-estimated after
iteration 1 and blended with alpha), and checks

* the noise estimate against the FP64 oracle on the same samples;
* every FP32 iteration bit for bit against the oracle on windows of the
  volume (a run of slices times a row, all x, at both volume corners and
  inside), with the previous GPU iterate as the oracle's input: each pass is
  then checked exactly, and so is the chain;
* full-volume properties the filter fixes whatever the size: every value is
  finite and lies within the [min, max] of the original over the voxel's
  search box (the output is a convex combination of original values inside
  the search ball), and mirroring the volume in x mirrors the output (up to
  the order of FP32 sums: 1e-4 relative per voxel);
* the 16-bit forms' max and RMS error against the FP64 oracle on windows,
  after iteration 1 and after the last one (from the form's own previous
  iterate), in units of the phantom's noise level, with the fraction of
  voxels beyond 0.1 of it, and their PSNR against the clean phantom next to
  the plain FP32 path's after the configured iterations.

The learned-proposal configuration trains on 64 phantoms of 256^3 (seeds
100-163) and is evaluated on the held-out 1024x1024x512 phantom (seed 3).
Windows keep the oracle's share of each test to seconds."""
import dataclasses
import time

import numpy as np
import pytest

from conftest import ROOT  # noqa: F401
from microtomodenoise_amd import DenoiseParameters, IterativeNLM_GPU
from microtomodenoise_amd import _capi
from microtomodenoise_amd import phantom as ph
from oracle import oracle as O

pytestmark = pytest.mark.gpu

T0 = time.perf_counter()


def say(msg):
    print(f"[gpu {time.perf_counter() - T0:7.1f}s] {msg}", flush=True)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def make_phantom(name):
    import torch
    spec = ph.CONFIGS[name]
    vol, clean = ph.make(spec, backend="torch", device="cuda", with_clean=True)
    torch.cuda.synchronize()
    say(f"{name} phantom {spec.shape} ready")
    return spec, vol, clean


def params_for(rs, rp, form, noise, alpha):
    p = DenoiseParameters(radius_searchspace=[rs] * 3, radius_patchspace=[rp] * 3, precision=form, alpha=alpha)
    p.cpu.blockwise = False
    p.noiselevel.mode = noise
    return p


def oparams(rs, rp, noise="global"):
    return O.Params(radius_searchspace=(rs,) * 3, radius_patchspace=(rp,) * 3, blockwise=False, noise_mode=noise)


def run_schedule(vol, shape, rs, rp, form, noise, alpha, iters, sigmas=None, host_iters=(), label=""):
    """The schedule on a resident volume.  sigmas=None estimates sigma0 and
    sigma1 (the product's estimator); otherwise (sigma0, sigma1) are used as
    given.  Returns the sigma lists, the samples, host copies of the iterates
    in host_iters and the final result on the device."""
    import torch
    nx, ny, nz = shape
    p = params_for(rs, rp, form, noise, alpha)
    g = IterativeNLM_GPU()
    assert g.configure_device([nx, ny, nz], p) == nz
    out = torch.empty_like(vol)
    res = {"host": {}, "samples": None}
    try:
        g.load_raw_device(vol.data_ptr())
        if sigmas is None:
            samples = _capi.draw_samples([nx, ny, nz], p.noiselevel.n_samples, p.noiselevel.patchsize, 0)
            res["samples"] = samples
            s0 = g.estimate_noise(samples, 0)
            sig1 = None
        else:
            s0, sig1 = sigmas
        for it in range(1, iters + 1):
            g.set_sigma(s0 if it == 1 else sig1)
            g.Run_GaussianNoise(it)
            if it == 1 and sigmas is None:
                s1 = g.estimate_noise(samples, 1)
                res["s1"] = s1
                a = np.float32(alpha)
                sig1 = (a * s0 + (np.float32(1) - a) * s1).astype(np.float32)
            if it in host_iters:
                g.get_result_device(out.data_ptr())
                res["host"][it] = out.cpu().numpy()
            say(f"{label} {form} iteration {it} filter {g.last_filter_ms():.0f} ms")
        g.get_result_device(out.data_ptr())
    finally:
        g.free_device()
    res.update(s0=s0, sig1=sig1, out=out)
    return res


def windows(shape, nz_win, rows=1):
    nx, ny, nz = shape
    zm, ym = nz // 2 - nz_win // 2, ny // 2
    return [((0, nz_win), (0, rows)), ((zm, zm + nz_win), (ym, ym + rows)), ((nz - nz_win, nz), (ny - rows, ny))]


def check_fp32_windows(raw, iterates, sig0, sig1, rs, rp, shape, nz_win, label):
    """iterates[k] = host iterate k (k >= 1); iterate 0 is the original."""
    p = oparams(rs, rp)
    for it in sorted(iterates):
        prev = raw if it == 1 else iterates[it - 1]
        sig = sig0 if it == 1 else sig1
        for zr, yr in windows(shape, nz_win):
            want = O.filter_window(raw, prev, sig, p, zr, yr)
            got = iterates[it][zr[0]:zr[1], yr[0]:yr[1]]
            assert np.array_equal(bits(got), bits(want)), (label, it, zr, yr)
        say(f"{label} FP32 iteration {it}: windows bit-exact against the oracle")


def check_noise(raw, samples, s0, noise, rs, rp, label, which=0):
    p = oparams(rs, rp, noise)
    want = O.noiselevel(raw, samples, p, "f64", which=which)
    assert np.allclose(s0, want, rtol=1e-5, atol=0), (label, np.abs(s0 / want - 1).max())
    say(f"{label} sigma{which} {float(s0[0]):.5g}..{float(s0[-1]):.5g} matches the FP64 oracle "
        f"(max rel {float(np.abs(s0 / want - 1).max()):.2e})")


def search_box_bounds(vol, rs, nslices=11):
    """[min, max] of the original over each voxel's search box (reflect
    padded): the search ball lies inside the box, so these bound the filter."""
    import torch
    import torch.nn.functional as F
    rz = min(nslices // 2, rs)
    v = vol[None, None]
    vp = F.pad(v, (rs, rs, rs, rs, rz, rz), mode="reflect")
    k = (2 * rz + 1, 2 * rs + 1, 2 * rs + 1)
    hi = F.max_pool3d(vp, k, stride=1)[0, 0]
    lo = -F.max_pool3d(-vp, k, stride=1)[0, 0]
    del vp
    torch.cuda.synchronize()
    return lo, hi


def check_properties_fp32(vol, out, rs, label):
    import torch
    assert bool(torch.isfinite(out).all()), label
    lo, hi = search_box_bounds(vol, rs)
    tol = 1e-6 * torch.maximum(lo.abs(), hi.abs())
    bad = int(((out < lo - tol) | (out > hi + tol)).sum())
    del lo, hi, tol
    assert bad == 0, (label, bad)
    say(f"{label} FP32 output finite and inside the search-box range of the original")


def check_mirror(vol, shape, rs, rp, sig0, out1, label):
    """Mirroring in x reorders the FP32 sums over offsets and taps, so
    distances change in their last bits.  The reference's weight jumps at its
    cutoff (the Pade form is about 0.025 there, iternlm_cpu.h:125-126), so a
    last-bit change can switch one term on or off: rare voxels differ by such
    a term, the rest by rounding."""
    import torch
    flipped = torch.flip(vol, dims=[2]).contiguous()
    r = run_schedule(flipped, shape, rs, rp, "fp32", "global", 0.5, 1, sigmas=(sig0, sig0), label=label + " mirrored")
    back = torch.flip(r["out"], dims=[2])
    # relative per voxel, values near zero measured against 1% of the range
    floor = 0.01 * float(vol.abs().max())
    relt = (back - out1).abs() / out1.abs().clamp_min(floor)
    rel = relt.max().item()
    beyond = int((relt > 1e-5).sum())
    same = float((back == out1).double().mean())
    del flipped, back, r, relt
    say(f"{label} mirror in x: max rel difference {rel:.2e}, {beyond} voxels beyond 1e-5, "
        f"{same:.4f} of voxels bit-equal")
    assert rel <= 1e-2 and beyond <= 1e-4 * out1.numel()


def error_stats(got, want, noise_z):
    e = np.abs(got.astype(np.float64) - want) / noise_z[:, None, None]
    return float(e.max()), float(np.sqrt(np.mean(e ** 2))), float(np.mean(e > 0.1))


def noise_levels(spec):
    nz = spec.shape[2]
    if spec.noise_sigma_z is not None:
        return np.linspace(spec.noise_sigma_z[0], spec.noise_sigma_z[1], nz)
    return np.full(nz, spec.noise_sigma)


def psnr_of(out, clean_t, peak):
    import torch
    mse = float(torch.mean((out.double() - clean_t.double()) ** 2))
    return 10.0 * np.log10(peak * peak / mse)


# 16-bit forms against the FP64 oracle, per pass, in noise-level units
# (max, RMS, fraction of voxels beyond 0.1 noise level)
# (the budgets of the small-volume form tests; the largest single errors are
# patch pairs at the weight's cutoff, where the reference's weight jumps by
# 0.025 of its maximum: one switched term moves a voxel by at most 0.025
# times the local contrast)
BUDGET16 = {"fp16": (0.35, 0.002, 0.001), "bf16": (0.8, 0.008, 0.001)}


def check_16bit(form, vol, raw, spec, rs, rp, sigs, iters, clean_t, peak, plain_psnr, label, nz_win=16):
    import torch
    r = run_schedule(vol, spec.shape, rs, rp, form, "global", 0.5, iters, sigmas=sigs, host_iters=(1, iters - 1),
                     label=label)
    assert bool(torch.isfinite(r["out"]).all())
    noise_z = noise_levels(spec)
    p = oparams(rs, rp)
    bmax, brms, bfrac = BUDGET16[form]
    stats = []
    for it, prev in ((1, raw), (iters, r["host"][iters - 1])):
        got_vol = r["host"][1] if it == 1 else r["out"].cpu().numpy()
        sig = sigs[0] if it == 1 else sigs[1]
        for zr, yr in windows(spec.shape, nz_win)[1:2]:
            want = O.filter_window(raw, prev, sig, p, zr, yr, precision="f64")
            mx, rms, frac = error_stats(got_vol[zr[0]:zr[1], yr[0]:yr[1]], want, noise_z[zr[0]:zr[1]])
            stats.append((it, mx, rms, frac))
            say(f"{label} {form} pass {it} vs FP64 oracle: max {mx:.3e} rms {rms:.3e} (noise-level units), "
                f"{frac:.4f} of voxels beyond 0.1")
    ps = psnr_of(r["out"], clean_t, peak)
    say(f"{label} {form} PSNR after {iters} iterations {ps:.4f} dB (plain FP32 {plain_psnr:.4f}, "
        f"loss {plain_psnr - ps:.4f} dB)")
    for it, mx, rms, frac in stats:
        assert mx <= bmax and rms <= brms and frac <= bfrac, (label, form, it, mx, rms, frac)
    return plain_psnr - ps


def full_config(name, rs, rp, noise, alpha, iters, forms, nz_win, fp16_budget_db=0.05):
    import torch
    spec, vol, clean = make_phantom(name)
    shape = spec.shape
    raw = vol.cpu().numpy()
    plain = run_schedule(vol, shape, rs, rp, "fp32", noise, alpha, iters, host_iters=tuple(range(1, iters + 1)),
                         label=name)
    check_noise(raw, plain["samples"], plain["s0"], noise, rs, rp, name)
    check_noise(plain["host"][1], plain["samples"], plain["s1"], noise, rs, rp, name, which=1)
    check_fp32_windows(raw, plain["host"], plain["s0"], plain["sig1"], rs, rp, shape, nz_win, name)
    for it in (1, iters):
        check_properties_fp32(vol, torch.from_numpy(plain["host"][it]).cuda(), rs, f"{name} iteration {it}")
    check_mirror(vol, shape, rs, rp, plain["s0"], torch.from_numpy(plain["host"][1]).cuda(), name)
    clean_t = torch.from_numpy(clean).cuda()
    peak = float(clean.max() - clean.min()) if max(spec.levels) <= 1.0 else 65535.0
    plain_psnr = psnr_of(plain["out"], clean_t, peak)
    sigs = (plain["s0"], plain["sig1"])
    del plain
    losses = {}
    for f in forms:
        losses[f] = check_16bit(f, vol, raw, spec, rs, rp, sigs, iters, clean_t, peak, plain_psnr, name)
    if "fp16" in losses:
        assert losses["fp16"] <= fp16_budget_db
    return losses


@pytest.mark.timeout(300)
def test_cfg2_512_cube_p2_s5(gpu_available):
    """512^3 uint16 phantom with streaks, patch r=2, search r=5, 4 iterations,
    alpha 0.5, global noise; FP32 and the FP16 form."""
    full_config("cfg2", 5, 2, "global", 0.5, 4, ("fp16",), nz_win=16)


@pytest.mark.timeout(400)
def test_headline_p2_s7(gpu_available):
    """The benchmark's workload: 1024x1024x512 float32 phantom (spheres, 200
    streaks), patch r=2, search r=7, 4 iterations, alpha 0.5, global noise;
    FP32, FP16 and BF16."""
    full_config("headline", 7, 2, "global", 0.5, 4, ("fp16", "bf16"), nz_win=16)


@pytest.mark.timeout(400)
def test_cfg3_precision_forms_p3_s7_zadaptive(gpu_available):
    """1024x1024x512 three-phase float32 phantom, noise rising along z, patch
    r=3, search r=7, 4 iterations, alpha 0.5, z-adaptive noise (window 100);
    FP32, FP16 and BF16."""
    full_config("cfg3", 7, 3, "z-adaptive", 0.5, 4, ("fp16", "bf16"), nz_win=16)


@pytest.mark.timeout(400)
def test_cfg5_large_volume_p2_s7_six_iterations(gpu_available):
    """2048x2048x256 uint16 phantom with four grey levels and 1000 streaks,
    patch r=2, search r=7, 6 iterations, alpha 0.6, z-adaptive noise; FP32,
    BF16 (the configuration's form) and FP16."""
    full_config("cfg5", 7, 2, "z-adaptive", 0.6, 6, ("bf16", "fp16"), nz_win=16)


@pytest.mark.timeout(400)
def test_cfg4_learned_proposal_training_and_held_out(gpu_available):
    """The proposal network trained on 64 phantoms of 256^3 built like the
    512^3 configuration (seeds 100-163), labelled by the FP32 filter's weight
    mass (patch r=2, search r=7), and evaluated on the held-out 1024x1024x512
    phantom (seed 3) at K=128: captured weight mass and PSNR of the pruned
    FP16 filter against the K offsets nearest the centre, and the 0.95
    fallback with a 2% audit.  With every tile falling back the pruned run
    must equal the full search bit for bit."""
    import torch

    from microtomodenoise_amd import proposal as PR
    rs, rp, K = 7, 2, 128
    params = params_for(rs, rp, "fp32", "global", 0.5)
    base = ph.CONFIGS["cfg2"]
    data = []
    ent = []
    for i in range(64):
        spec = dataclasses.replace(base, shape=(256, 256, 256), seed=100 + i, streaks=25)
        vol = ph.make(spec, backend="torch", device="cuda")
        mass, _ = PR.make_labels(vol, params, iteration=1)
        pm = mass.double() / mass.double().sum(dim=1, keepdim=True)
        ent.append(float(-(pm * torch.log(pm.clamp_min(1e-300))).sum(dim=1).mean()))
        data.append((vol, mass))
        if i % 16 == 15:
            say(f"cfg4 training volume {i + 1}/64 labelled")
    ns = data[0][1].shape[1]
    net = PR.OffsetProposal(ns, seed=1)
    hist = PR.train(net, data, steps=256, lr=3e-3)
    floor = float(np.mean(ent))
    say(f"cfg4 trained: loss {hist[0]:.4f} -> {np.mean(hist[-16:]):.4f}, label entropy floor {floor:.4f}, "
        f"uniform {np.log(ns):.4f}")
    assert np.mean(hist[-16:]) < hist[0]
    del data

    spec, vol, clean = make_phantom("headline")
    nx, ny, nz = spec.shape
    mass, sigma = PR.make_labels(vol, params, iteration=1)
    mass = mass.cpu().numpy().astype(np.float64)
    probs = net.predict(vol.data_ptr(), [nx, ny, nz])
    S = O.search_offsets((rs,) * 3)
    nt = mass.shape[0]
    learned = PR.lists_from_scores(probs, K)
    nearest = PR.same_lists(nt, PR.nearest_offsets(S, K))
    cap_l, cap_n = PR.captured_mass(mass, *learned), PR.captured_mass(mass, *nearest)
    say(f"cfg4 held-out captured mass at K={K}: learned {cap_l:.4f}, nearest {cap_n:.4f}")
    del mass, probs
    clean_t = torch.from_numpy(clean).cuda()

    def one_pass(lists=None, proposal=False):
        g = IterativeNLM_GPU()
        p = params_for(rs, rp, "fp16", "global", 0.5)
        assert g.configure_device([nx, ny, nz], p) == nz
        out = torch.empty_like(vol)
        try:
            g.load_raw_device(vol.data_ptr())
            g.set_sigma(sigma)
            if lists is not None:
                g.set_offset_lists(lists[0], lists[1], K)
            if proposal:
                g.set_proposal(net, K, threshold=0.95, audit_frac=0.02, seed=7)
            g.Run_GaussianNoise(1)
            g.get_result_device(out.data_ptr())
            st = g.proposal_stats() if proposal else None
        finally:
            g.free_device()
        return out, st

    full, _ = one_pass()
    out_l, _ = one_pass(learned)
    out_n, _ = one_pass(nearest)
    out_f, st = one_pass(proposal=True)
    ps = {k: psnr_of(v, clean_t, 65535.0) for k, v in (("full", full), ("learned", out_l), ("nearest", out_n),
                                                         ("fallback0.95", out_f))}
    say("cfg4 held-out PSNR after one FP16 pass: " + ", ".join(f"{k} {v:.3f}" for k, v in ps.items()))
    fb = st["fallback_tiles"] / max(st["tiles"], 1)
    say(f"cfg4 fallback 0.95: {fb:.4f} of tiles fall back, {st['audited_tiles'] / max(st['tiles'], 1):.4f} audited, "
        f"mean predicted mass {st['mean_predicted_mass']:.4f}")
    net.close()
    assert cap_l > cap_n
    assert ps["learned"] > ps["nearest"] and ps["learned"] < ps["full"]
    if st["fallback_tiles"] >= st["tiles"]:
        assert bool(torch.equal(out_f, full))
