"""GPU tests of the learned offset proposal and the pruned matrix-core path,
through the C ABI: weight-mass labels against the oracle, the network's loss,
probabilities and gradients against the CPU oracle (oracle/proposal_oracle.py,
itself pinned against torch autograd and finite differences), the pruned
filter against the FP64 oracle restricted to each tile's offsets, the
fallback / audit rules, and the gain of learned lists over the nearest-offset
lists that need no network."""
import numpy as np
import pytest
import torch

from conftest import ROOT  # noqa: F401
from microtomodenoise_amd import DenoiseParameters, IterativeNLM_GPU
from microtomodenoise_amd import phantom as ph
from microtomodenoise_amd import proposal as PR
from oracle import oracle as O
from oracle import proposal_oracle as PO

pytestmark = pytest.mark.gpu


def make_gpu(shape_xyz, rs, rp, precision="fp32"):
    p = DenoiseParameters(radius_searchspace=list(rs), radius_patchspace=list(rp), precision=precision)
    p.cpu.blockwise = False
    p.noiselevel.mode = "global"
    g = IterativeNLM_GPU()
    assert g.configure_device(list(shape_xyz), p) == shape_xyz[2]
    return g, p


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_weight_mass_matches_oracle_and_keeps_the_result(gpu_available):
    spec = ph.PhantomSpec((70, 20, 19), seed=61, radii=(2, 7), streaks=3, streak_len=(5, 15))
    raw = ph.make(spec)
    nz, ny, nx = raw.shape
    rs, rp = (3, 3, 3), (2, 2, 2)
    sig = np.linspace(900, 1400, nz).astype(np.float32)
    print(f"[gpu] weight mass {raw.shape} rs={rs} rp={rp}", flush=True)
    g, _ = make_gpu((nx, ny, nz), rs, rp)
    g.load_raw(raw)
    g.set_sigma(sig)
    m1 = g.weight_mass(1)
    out1 = g.get_result()
    m2 = g.weight_mass(2)
    g.free_device()
    h, _ = make_gpu((nx, ny, nz), rs, rp)
    h.set_sigma(sig)
    h.Run_GaussianNoise(1, raw)
    ref1 = h.get_result()
    h.free_device()
    assert np.array_equal(bits(out1), bits(ref1)), "the labelled pass must give the plain pass's result"
    p = O.Params(radius_searchspace=rs, radius_patchspace=rp, blockwise=False)
    for got, P in ((m1, raw), (m2, out1)):
        want = O.weight_mass(P, sig, p)
        assert got.shape == want.shape
        rel = np.abs(got - want) / np.maximum(want, 1e-30)
        print(f"[gpu]   max rel {rel.max():.3e}", flush=True)
        assert rel.max() <= 2e-5


def _net_case(seed=0):
    ns = 37
    shape = (16, 24, 40)  # z, y, x: partial x tile (40 = 32 + 8), 3 y tiles, 2 z tiles
    rng = np.random.default_rng(seed)
    vol = (rng.normal(size=shape) * 300 + 2000).astype(np.float32)
    vol[:, 10:14, :] += 900.0  # some structure
    _, nt = PO.tile_index(shape)
    mass = rng.uniform(0.0, 4.0, size=(nt, ns)).astype(np.float32)
    mass[:, :5] += 20.0
    return ns, shape, vol, mass


def test_proposal_gradients_match_oracle(gpu_available):
    ns, shape, vol, mass = _net_case(1)
    nz, ny, nx = shape
    net = PR.OffsetProposal(ns, seed=11)
    params = net.get_params()
    assert params.size == PO.num_params(ns)
    vd = torch.from_numpy(vol).cuda()
    md = torch.from_numpy(mass).cuda()
    print("[gpu] proposal gradient check", flush=True)
    loss = net.train_step(vd.data_ptr(), [nx, ny, nz], md.data_ptr(), 0.0)
    grads = net.grads()
    probs = net.predict(vd.data_ptr(), [nx, ny, nz])
    net.close()
    lo, go, qo = PO.forward_backward(vol, params, ns, mass, bf16=True)
    le, ge, _ = PO.forward_backward(vol, params, ns, mass, bf16=False)
    print(f"[gpu]   loss gpu {loss:.7f} oracle(bf16) {lo:.7f} exact {le:.7f}", flush=True)
    assert abs(loss - lo) <= 2e-4 * abs(lo)
    assert np.abs(probs - qo).max() <= 2e-3 * qo.max()
    for name, sl in PO.block_slices(ns):
        e_b = np.linalg.norm(grads[sl] - go[sl]) / np.linalg.norm(go[sl])
        e_x = np.linalg.norm(grads[sl] - ge[sl]) / np.linalg.norm(ge[sl])
        print(f"[gpu]   {name}: rel err vs bf16 oracle {e_b:.2e}, vs exact {e_x:.2e}", flush=True)
        assert e_b <= 2e-2, name
        assert e_x <= 1e-1, name


def test_adam_update(gpu_available):
    ns, shape, vol, mass = _net_case(2)
    nz, ny, nx = shape
    net = PR.OffsetProposal(ns, seed=12)
    p0 = net.get_params()
    vd = torch.from_numpy(vol).cuda()
    md = torch.from_numpy(mass).cuda()
    lr = 1e-3
    l0 = net.train_step(vd.data_ptr(), [nx, ny, nz], md.data_ptr(), lr)
    g = net.grads()
    p1 = net.get_params()
    # first Adam step with bias correction: p -= lr * g / (|g| + eps)
    want = p0.astype(np.float64) - lr * g / (np.abs(g) + 1e-8)
    assert np.allclose(p1, want, rtol=1e-5, atol=1e-6)
    losses = [l0] + [net.train_step(vd.data_ptr(), [nx, ny, nz], md.data_ptr(), 3e-3) for _ in range(60)]
    net.close()
    print(f"[gpu] loss {losses[0]:.5f} -> {losses[-1]:.5f}", flush=True)
    assert losses[-1] < losses[0]


def test_training_repeats_bit_for_bit(gpu_available):
    """The weight gradients are summed in a fixed order (per-block partials,
    then one reduction), so two networks of one seed trained on the same
    data end with identical parameters, gradients and losses."""
    ns, shape, vol, mass = _net_case(3)
    nz, ny, nx = shape
    vd = torch.from_numpy(vol).cuda()
    md = torch.from_numpy(mass).cuda()
    runs = []
    for rep in range(2):
        print(f"[gpu] training run {rep}", flush=True)
        net = PR.OffsetProposal(ns, seed=13)
        losses = [net.train_step(vd.data_ptr(), [nx, ny, nz], md.data_ptr(), 3e-3) for _ in range(20)]
        runs.append((np.array(losses, np.float32), net.grads(), net.get_params()))
        net.close()
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(bits(a), bits(b))


# max, rms error against the FP64 oracle in noise-sigma units.  FP32_FAST
# keeps the volumes in FP32 (FP16 difference planes only) and is held to the
# full-search budget; with 40 of 114 offsets a voxel's output rests on fewer
# weights, so one FP16 rounding that moves a distance across the weight
# cutoff moves the output further than in the full search: the FP16 form's
# maximum is looser, its rms is not.
PRUNED_BUDGET = {"fp32_fast": (0.05, 0.004), "fp16": (0.5, 0.006), "fp16_fp8": (1.0, 0.02)}


@pytest.mark.parametrize("form", ["fp32_fast", "fp16", "fp16_fp8"])
def test_pruned_lists_match_oracle_per_tile(gpu_available, form):
    spec = ph.PhantomSpec((64, 24, 16), seed=71, radii=(2, 7), noise_sigma=1500.0, blur_sigma=0.8, streaks=3,
                          streak_len=(5, 15))
    raw = ph.make(spec)
    nz, ny, nx = raw.shape
    rs, rp = (3, 3, 3), (2, 2, 2)
    sig = np.full(nz, 700.0, np.float32)
    S = O.search_offsets(rs)
    ns = len(S)
    K = 40
    near = PR.nearest_offsets(S, K)
    rng = np.random.default_rng(5)
    rand = np.sort(rng.choice(ns, size=K, replace=False)).astype(np.int32)
    g, _ = make_gpu((nx, ny, nz), rs, rp, precision=form)
    nt = g.tile_count()
    assert nt == 2 * 3 * 2
    lists = np.zeros((nt, K), np.int32)
    counts = np.full(nt, K, np.int32)
    kind = np.array([t % 3 for t in range(nt)])  # 0 nearest, 1 random, 2 whole search
    for t in range(nt):
        lists[t] = near if kind[t] == 0 else rand
        if kind[t] == 2:
            counts[t] = ns
    print(f"[gpu] pruned per-tile lists {form} {raw.shape} K={K} of {ns}", flush=True)
    g.set_offset_lists(lists, counts, K)
    g.set_sigma(sig)
    g.Run_GaussianNoise(1, raw)
    got = g.get_result().astype(np.float64)
    g.free_device()
    p = O.Params(radius_searchspace=rs, radius_patchspace=rp, blockwise=False)
    refs = [O.filter_iteration(raw, raw, sig.astype(np.float64), p, precision="f64", offsets=o)
            for o in (near, rand, None)]
    tix, _ = PO.tile_index(raw.shape)
    ref = np.choose(kind[tix], refs)
    err = np.abs(got - ref) / spec.noise_sigma
    mx, rms = float(err.max()), float(np.sqrt(np.mean(err ** 2)))
    print(f"[gpu]   max {mx:.3e} rms {rms:.3e}", flush=True)
    bmax, brms = PRUNED_BUDGET[form]
    assert mx <= bmax and rms <= brms
    # the lists matter: filtering with the other list moves the result by far
    # more than the form's rounding does
    sep = float(np.sqrt(np.mean(((refs[0] - refs[1]) / spec.noise_sigma) ** 2)))
    print(f"[gpu]   rms difference between the two lists' results {sep:.3e}", flush=True)
    assert sep > 10 * rms


@pytest.mark.parametrize("form", ["fp16", "fp16_fp8"])
def test_proposal_fallback_audit_and_topk(gpu_available, form):
    spec = ph.PhantomSpec((64, 32, 24), seed=81, radii=(2, 7), streaks=3, streak_len=(5, 15))
    raw = ph.make(spec)
    nz, ny, nx = raw.shape
    rs, rp = (3, 3, 3), (2, 2, 2)
    sig = np.full(nz, 800.0, np.float32)
    ns = len(O.search_offsets(rs))
    K = 24
    net = PR.OffsetProposal(ns, seed=3)
    rd = torch.from_numpy(raw).cuda()
    probs = net.predict(rd.data_ptr(), [nx, ny, nz])
    print(f"[gpu] proposal modes {form} {raw.shape} K={K} of {ns}", flush=True)

    def run(mode):
        g, _ = make_gpu((nx, ny, nz), rs, rp, precision=form)
        g.load_raw(raw)
        g.set_sigma(sig)
        if mode == "full":
            pass
        elif mode == "lists":
            lists, counts = PR.lists_from_scores(probs, K)
            g.set_offset_lists(lists, counts, K)
        else:
            thr, audit = mode
            g.set_proposal(net, K, threshold=thr, audit_frac=audit, seed=9)
        g.Run_GaussianNoise(1)
        out = g.get_result()
        st = g.proposal_stats()
        g.free_device()
        return out, st

    full, _ = run("full")
    fb, st_fb = run((1.0, 0.0))
    assert st_fb["fallback_tiles"] == st_fb["tiles"] and st_fb["mean_offsets_visited"] == ns
    assert np.array_equal(bits(fb), bits(full)), "all tiles falling back must equal the whole search"
    au, st_au = run((0.0, 1.0))
    assert st_au["audited_tiles"] == st_au["tiles"]
    assert np.array_equal(bits(au), bits(full)), "all tiles audited must equal the whole search"
    pr, st_pr = run((0.0, 0.0))
    assert st_pr["fallback_tiles"] == 0 and st_pr["mean_offsets_visited"] == K
    want_mass = np.sort(probs, axis=1)[:, ::-1][:, :K].sum(axis=1).mean()
    assert abs(st_pr["mean_predicted_mass"] - want_mass) <= 1e-4
    li, _ = run("lists")
    assert np.array_equal(bits(pr), bits(li)), "device top-K must match host top-K of the same prediction"
    assert not np.array_equal(bits(pr), bits(full))
    # a partial audit draws about that fraction of tiles
    _, st_part = run((0.0, 0.5))
    assert 0.2 * st_part["tiles"] <= st_part["audited_tiles"] <= 0.8 * st_part["tiles"]
    net.close()


def test_learned_lists_capture_more_mass_than_nearest(gpu_available):
    """The learned proposal against the lists that need no network (the K
    offsets nearest the centre, and the K offsets of highest mean training
    mass), on a held-out phantom: captured true weight mass at equal K
    (strictly more), and PSNR of the pruned FP16 filter against the clean
    phantom (strictly higher than with either list that needs no network)."""
    rs, rp = (7, 7, 7), (2, 2, 2)
    K = 128
    shape = (64, 64, 32)
    params = DenoiseParameters(radius_searchspace=list(rs), radius_patchspace=list(rp))
    params.cpu.blockwise = False
    params.noiselevel.mode = "global"
    data = []
    for seed in (100, 101, 102, 103):
        spec = ph.PhantomSpec(shape, seed=seed, radii=(3, 14), noise_sigma=2000.0, blur_sigma=1.0, streaks=20,
                              streak_len=(10, 40))
        vol = torch.from_numpy(ph.make(spec)).cuda()
        mass, _ = PR.make_labels(vol, params, iteration=1)
        data.append((vol, mass))
    ns = data[0][1].shape[1]
    print(f"[gpu] learned vs nearest: {len(data)} training volumes {shape}, {ns} offsets, K={K}", flush=True)
    net = PR.OffsetProposal(ns, seed=1)
    hist = PR.train(net, data, steps=400, lr=3e-3)
    print(f"[gpu]   loss {hist[0]:.4f} -> {np.mean(hist[-20:]):.4f}", flush=True)
    mean_mass = sum(m.sum(dim=0) for _, m in data).cpu().numpy()

    spec = ph.PhantomSpec(shape, seed=200, radii=(3, 14), noise_sigma=2000.0, blur_sigma=1.0, streaks=20,
                          streak_len=(10, 40))
    raw, clean = ph.make(spec, with_clean=True)
    vol = torch.from_numpy(raw).cuda()
    mass, sigma = PR.make_labels(vol, params, iteration=1)
    mass = mass.cpu().numpy().astype(np.float64)
    nx, ny, nz = shape
    probs = net.predict(vol.data_ptr(), [nx, ny, nz])
    S = O.search_offsets(rs)
    nt = mass.shape[0]
    learned = PR.lists_from_scores(probs, K)
    nearest = PR.same_lists(nt, PR.nearest_offsets(S, K))
    popular = PR.same_lists(nt, np.sort(np.argsort(-mean_mass, kind="stable")[:K]))
    oracle_best = PR.lists_from_scores(mass, K)
    cap = {name: PR.captured_mass(mass, *l) for name, l in
           (("learned", learned), ("nearest", nearest), ("popular", popular), ("best", oracle_best))}
    print("[gpu]   captured mass " + ", ".join(f"{k} {v:.4f}" for k, v in cap.items()), flush=True)

    def psnr_of(lists):
        g, _ = make_gpu(shape, rs, rp, precision="fp16")
        g.load_raw(raw)
        g.set_sigma(sigma)
        if lists is not None:
            g.set_offset_lists(lists[0], lists[1], K)
        g.Run_GaussianNoise(1)
        out = g.get_result().astype(np.float64)
        g.free_device()
        return 10 * np.log10(65535.0 ** 2 / np.mean((out - clean) ** 2))

    ps = {name: psnr_of(l) for name, l in (("full", None), ("learned", learned), ("nearest", nearest),
                                             ("popular", popular))}
    print("[gpu]   PSNR " + ", ".join(f"{k} {v:.3f}" for k, v in ps.items()), flush=True)
    net.close()
    assert cap["learned"] > cap["nearest"] and cap["learned"] > cap["popular"]
    # what the network gains over the lists that need none: more of the true
    # weight mass kept, and a pruned filter closer to the clean phantom
    assert ps["learned"] > ps["nearest"] and ps["learned"] > ps["popular"] and ps["learned"] < ps["full"]
