"""Pins of the learned-proposal oracle (oracle/proposal_oracle.py) and of the
weight-mass labels (oracle weight_mass), on the CPU.

The network oracle is checked against torch's own conv3d / autograd in FP64
(a library routine for every convolution, ReLU, pooling and softmax
cross-entropy step), against finite differences, and its BF16 rounding
against torch's bfloat16 conversion.  The weight-mass oracle is checked
against closed forms (a constant previous iterate gives every weight
pade(0) = 1) and brute force on single tiles.
"""
import numpy as np
import pytest
import torch

from conftest import ROOT  # noqa: F401
from oracle import oracle as O
from oracle import proposal_oracle as PO


def _rand_params(ns, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    n = PO.num_params(ns)
    p = np.zeros(n, np.float32)
    for name, sl in PO.block_slices(ns):
        fan = {"w1": 27, "wh": 16}.get(name, 27 * 16)
        p[sl] = (rng.uniform(-1, 1, sl.stop - sl.start) * np.sqrt(6.0 / fan) * scale).astype(np.float32)
        if name.startswith("b"):
            p[sl] = (rng.uniform(-0.1, 0.1, sl.stop - sl.start)).astype(np.float32)
    return p


def _torch_loss(vol, params, ns, mass):
    """The same network written with torch FP64 autograd."""
    w1, b1, ws, bs, wh, bh = PO.unpack(params, ns)
    x = torch.from_numpy(PO.normalise(vol).astype(np.float64))[None, None]
    tp = [torch.tensor(a, dtype=torch.float64, requires_grad=True)
          for a in [w1, b1] + [t for l in range(3) for t in (ws[l], bs[l])] + [wh, bh]]

    def wconv(w, cin):  # [27, cin, cout] -> [cout, cin, 3, 3, 3] (tap = (dz+1)*9+(dy+1)*3+(dx+1))
        return w.reshape(3, 3, 3, cin, 16).permute(4, 3, 0, 1, 2)

    a = torch.relu(torch.nn.functional.conv3d(x, wconv(tp[0], 1), tp[1], padding=1))
    for l in range(3):
        a = torch.relu(torch.nn.functional.conv3d(a, wconv(tp[2 + 2 * l], 16), tp[3 + 2 * l], padding=1))
    tix, nt = PO.tile_index(vol.shape)
    tix = torch.from_numpy(tix.ravel())
    a4 = a[0].reshape(16, -1).T
    feat = torch.zeros(nt, 16, dtype=torch.float64).index_add(0, tix, a4)
    cnt = torch.bincount(tix, minlength=nt).double()
    feat = feat / cnt[:, None]
    logits = tp[9] + feat @ tp[8]
    logq = torch.log_softmax(logits, dim=1)
    m = torch.from_numpy(np.asarray(mass, np.float64))
    p = m / m.sum(dim=1, keepdim=True)
    loss = -(p * logq).sum() / nt
    loss.backward()
    grads = torch.cat([t.grad.reshape(-1) for t in tp]).numpy()
    return float(loss.detach()), grads, torch.exp(logq).detach().numpy()


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(size=10000).astype(np.float32) * 10 ** rng.uniform(-3, 3, 10000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38, -2.5], np.float32)])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(PO.bf16_round(x), ref)


def test_conv_and_its_gradients_match_torch():
    rng = np.random.default_rng(2)
    shape = (5, 6, 7)
    x = rng.normal(size=shape + (16,))
    w = rng.normal(size=(27, 16, 16))
    b = rng.normal(size=16)
    xt = torch.tensor(x.transpose(3, 0, 1, 2)[None], requires_grad=True)
    wt = torch.tensor(w.reshape(3, 3, 3, 16, 16).transpose(4, 3, 0, 1, 2).copy(), requires_grad=True)
    yt = torch.nn.functional.conv3d(xt, wt, torch.tensor(b), padding=1)
    y = PO.conv(x, w, b)
    assert np.allclose(y, yt[0].detach().numpy().transpose(1, 2, 3, 0), rtol=1e-12, atol=1e-10)
    g = rng.normal(size=shape + (16,))
    yt.backward(torch.tensor(g.transpose(3, 0, 1, 2)[None]))
    dx = PO.conv_dgrad(g, w)
    assert np.allclose(dx, xt.grad[0].numpy().transpose(1, 2, 3, 0), rtol=1e-12, atol=1e-10)
    dw, db = PO.conv_wgrad(x, g)
    dwt = wt.grad.numpy().transpose(2, 3, 4, 1, 0).reshape(27, 16, 16)
    assert np.allclose(dw, dwt, rtol=1e-12, atol=1e-9)
    assert np.allclose(db, g.reshape(-1, 16).sum(0))


@pytest.mark.parametrize("shape", [(8, 9, 10), (9, 17, 40)])
def test_network_loss_and_gradients_match_torch_autograd(shape):
    ns = 11
    rng = np.random.default_rng(3)
    vol = (rng.normal(size=shape) * 100 + 1000).astype(np.float32)
    params = _rand_params(ns, 4)
    _, nt = PO.tile_index(shape)
    mass = rng.uniform(0.1, 5.0, size=(nt, ns))
    loss, grads, q = PO.forward_backward(vol, params, ns, mass, bf16=False)
    lt, gt, qt = _torch_loss(vol, params, ns, mass)
    assert abs(loss - lt) <= 1e-12 * abs(lt)
    assert np.allclose(q, qt, rtol=1e-12, atol=1e-15)
    for name, sl in PO.block_slices(ns):
        err = np.abs(grads[sl] - gt[sl]).max()
        assert err <= 1e-10 * max(np.abs(gt[sl]).max(), 1e-30), name


def test_network_gradient_finite_differences():
    ns = 7
    shape = (8, 8, 9)
    rng = np.random.default_rng(5)
    vol = rng.normal(size=shape).astype(np.float32)
    params = _rand_params(ns, 6).astype(np.float64)
    _, nt = PO.tile_index(shape)
    mass = rng.uniform(0.1, 5.0, size=(nt, ns))
    _, grads, _ = PO.forward_backward(vol, params, ns, mass, bf16=False)
    h = 1e-6
    for name, sl in PO.block_slices(ns):
        for i in rng.choice(np.arange(sl.start, sl.stop), size=4, replace=False):
            pp, pm = params.copy(), params.copy()
            pp[i] += h
            pm[i] -= h
            lp = PO.forward_backward(vol, pp, ns, mass, bf16=False)[0]
            lm = PO.forward_backward(vol, pm, ns, mass, bf16=False)[0]
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grads[i]) <= 1e-6 + 1e-4 * abs(grads[i]), (name, i, fd, grads[i])


def test_bf16_forward_is_close_to_exact():
    ns = 9
    shape = (8, 16, 33)
    rng = np.random.default_rng(7)
    vol = rng.normal(size=shape).astype(np.float32)
    params = _rand_params(ns, 8)
    _, nt = PO.tile_index(shape)
    mass = rng.uniform(0.1, 5.0, size=(nt, ns))
    l0, g0, _ = PO.forward_backward(vol, params, ns, mass, bf16=False)
    l1, g1, _ = PO.forward_backward(vol, params, ns, mass, bf16=True)
    assert abs(l1 - l0) < 1e-2 * abs(l0)
    assert np.linalg.norm(g1 - g0) < 0.05 * np.linalg.norm(g0)


# ---------------------------------------------------------------------------
# weight-mass labels
# ---------------------------------------------------------------------------

def test_weight_mass_constant_volume_is_voxel_count():
    p = O.Params(radius_searchspace=(3, 3, 3), radius_patchspace=(1, 1, 1))
    shape = (24, 13, 70)  # partial tiles in x (70 = 2*32 + 6) and y (13 = 8 + 5)
    P = np.full(shape, 7.0, np.float32)
    m = O.weight_mass(P, np.full(shape[0], 0.5, np.float32), p)
    tix, nt = PO.tile_index(shape)
    cnt = np.bincount(tix.ravel(), minlength=nt)
    assert m.shape == (nt, len(O.search_offsets((3, 3, 3))))
    assert np.array_equal(m, np.repeat(cnt[:, None].astype(np.float64), m.shape[1], axis=1))


def test_weight_mass_tile_numbering_and_brute_force():
    p = O.Params(radius_searchspace=(2, 2, 2), radius_patchspace=(1, 1, 1))
    shape = (24, 16, 64)
    rng = np.random.default_rng(9)
    P = np.full(shape, 3.0, np.float32)
    P[:8] += rng.normal(size=(8,) + shape[1:]).astype(np.float32)  # only the first row of z tiles is noisy
    sig = np.full(shape[0], 0.8, np.float32)
    m = O.weight_mass(P, sig, p)
    tix, nt = PO.tile_index(shape)
    cnt = np.bincount(tix.ravel(), minlength=nt)
    # tiles of z >= 16 are beyond reach (search 2 + patch 1) of the noise
    far = np.unique(tix[16:])
    assert np.array_equal(m[far], np.repeat(cnt[far, None].astype(np.float64), m.shape[1], axis=1))
    # brute force, FP64 weights, for one noisy tile and a few offsets
    S = O.search_offsets((2, 2, 2))
    PT = O.patch_offsets((1, 1, 1))
    w = O.cpu_weight_table(p, shape[0] + 2 * p.padding()[2])
    Pp = O.pad_reflect(P, p.padding())
    px, py, pz = p.padding()
    t = 1  # x tile 1 of the first y/z row
    for k in (0, 5, len(S) - 1):
        dx, dy, dz = S[k]
        tot = 0.0
        for z in range(0, 8):
            mm = -1.0 / (float(sig[z]) ** 2)
            for y in range(0, 8):
                for x in range(32, 64):
                    d = 0.0
                    for (ax, ay, az), wp in zip(PT, w):
                        a = Pp[z + pz + dz + az, y + py + dy + ay, x + px + dx + ax]
                        b = Pp[z + pz + az, y + py + ay, x + px + ax]
                        d += (float(a) - float(b)) ** 2 * float(wp)
                    tot += O.orc().orc_pade_f64(d * mm)
        assert abs(m[t, k] - tot) <= 1e-4 * tot
