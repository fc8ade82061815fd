"""CPU oracle of the learned offset proposal (test infrastructure only).

Restates, in numpy FP64, the network and loss that csrc/inlm_learn.hip runs
on the GPU, with its analytic backward pass:

  x   = (vol - mean) / std                      (FP64 moments, FP32 input)
  a1  = relu(conv1(x))        1 -> 16 channels, 3x3x3, zero padding
  a2..a4 = relu(conv(a_prev)) 16 -> 16 channels, 3x3x3, zero padding
  f_t = mean of a4 over tile t (32x8x8, partial at the edges)
  l_t = bh + f_t . Wh,  q_t = softmax(l_t)
  L   = -(1/T) sum_t sum_s p_ts log q_ts,  p_t = mass_t / sum(mass_t)

With bf16=True the activations, the gradients at the convolution outputs and
the 16-channel weights are rounded to bfloat16 (round to nearest even) where
the GPU stores them in BF16, so the GPU result can be checked tightly; with
bf16=False it is the exact FP64 function, which the finite-difference test
pins.  The reference has no learned part; nothing here cites it.

Parameter vector layout: w1[27][16] b1[16]; three times w[27][16 in][16 out]
b[16]; Wh[16][ns] bh[ns].  Tap t = (dz+1)*9 + (dy+1)*3 + (dx+1).
"""
from __future__ import annotations

import numpy as np

C = 16
NTAP = 27
TILE = (32, 8, 8)  # x, y, z


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 values rounded to bfloat16 (round to nearest, ties to even)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    u = ((u + r) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32)


def unpack(params: np.ndarray, ns: int):
    p = np.asarray(params, np.float64)
    o = 0
    w1 = p[o:o + NTAP * C].reshape(NTAP, C); o += NTAP * C
    b1 = p[o:o + C]; o += C
    ws, bs = [], []
    for _ in range(3):
        ws.append(p[o:o + NTAP * C * C].reshape(NTAP, C, C)); o += NTAP * C * C
        bs.append(p[o:o + C]); o += C
    wh = p[o:o + C * ns].reshape(C, ns); o += C * ns
    bh = p[o:o + ns]; o += ns
    assert o == p.size, (o, p.size)
    return w1, b1, ws, bs, wh, bh


def num_params(ns: int) -> int:
    return NTAP * C + C + 3 * (NTAP * C * C + C) + C * ns + ns


def taps():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(NTAP)]  # (dz, dy, dx)


def _shift(xp: np.ndarray, dz: int, dy: int, dx: int, shape):
    """View of the zero-padded array xp (one-voxel halo) shifted by (dz,dy,dx)."""
    nz, ny, nx = shape
    return xp[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]


def _pad(a: np.ndarray) -> np.ndarray:
    pw = [(1, 1), (1, 1), (1, 1)] + [(0, 0)] * (a.ndim - 3)
    return np.pad(a, pw)


def conv(x: np.ndarray, w: np.ndarray, b) -> np.ndarray:
    """x [nz,ny,nx,cin] (or [nz,ny,nx] for cin=1), w [27,cin,cout]."""
    shape = x.shape[:3]
    xp = _pad(x)
    out = np.zeros(shape + (w.shape[-1],), np.float64)
    for t, (dz, dy, dx) in enumerate(taps()):
        xs = _shift(xp, dz, dy, dx, shape)
        if x.ndim == 3:
            out += xs[..., None] * w[t].reshape(-1)
        else:
            out += xs @ w[t]
    if b is not None:
        out += b
    return out


def conv_wgrad(x: np.ndarray, g: np.ndarray):
    """dW[t][cin][cout] = sum_v x[v + off(t)][cin] g[v][cout]."""
    shape = g.shape[:3]
    xp = _pad(x if x.ndim == 4 else x[..., None])
    cin = xp.shape[-1]
    dw = np.zeros((NTAP, cin, g.shape[-1]), np.float64)
    gf = g.reshape(-1, g.shape[-1])
    for t, (dz, dy, dx) in enumerate(taps()):
        xs = _shift(xp, dz, dy, dx, shape).reshape(-1, cin)
        dw[t] = xs.T @ gf
    return dw, gf.sum(axis=0)


def conv_dgrad(g: np.ndarray, w: np.ndarray) -> np.ndarray:
    """dX[v][cin] = sum_t sum_cout g[v - off(t)][cout] w[t][cin][cout]."""
    shape = g.shape[:3]
    gp = _pad(g)
    out = np.zeros(shape + (w.shape[1],), np.float64)
    for t, (dz, dy, dx) in enumerate(taps()):
        out += _shift(gp, -dz, -dy, -dx, shape) @ w[t].T
    return out


def tile_index(shape_zyx):
    nz, ny, nx = shape_zyx
    ntx, nty, ntz = -(-nx // TILE[0]), -(-ny // TILE[1]), -(-nz // TILE[2])
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    t = ((z // TILE[2]) * nty + (y // TILE[1])) * ntx + (x // TILE[0])
    return t, ntx * nty * ntz


def normalise(vol: np.ndarray) -> np.ndarray:
    v = np.asarray(vol, np.float32)
    m = float(np.sum(v, dtype=np.float64)) / v.size
    var = max(float(np.sum(v.astype(np.float64) ** 2)) / v.size - m * m, 0.0)
    shift = np.float32(m)
    scale = np.float32(1.0 / np.sqrt(var)) if var > 0 else np.float32(1.0)
    return ((v - shift) * scale).astype(np.float32)


def forward_backward(vol: np.ndarray, params: np.ndarray, ns: int, mass=None, bf16: bool = True):
    """Returns (loss, grads, probs); loss/grads are None without mass.
    vol is z-y-x float32; mass [tiles, ns]."""
    def rnd(a):  # BF16 storage of the GPU (FP32 value rounded), or exact
        return bf16_round(a.astype(np.float32)).astype(np.float64) if bf16 else a

    w1, b1, ws, bs, wh, bh = unpack(params, ns)
    wq = [rnd(w) for w in ws]
    x = normalise(vol).astype(np.float64)
    acts = [rnd(np.maximum(conv(x, w1[:, None, :], b1), 0))]
    for l in range(3):
        acts.append(rnd(np.maximum(conv(acts[-1], wq[l], bs[l]), 0)))
    tix, nt = tile_index(vol.shape)
    cnt = np.bincount(tix.ravel(), minlength=nt).astype(np.float64)
    a4 = acts[-1].reshape(-1, C)
    feat = np.zeros((nt, C))
    np.add.at(feat, tix.ravel(), a4)
    feat /= cnt[:, None]
    logits = bh[None, :] + feat @ wh
    mx = logits.max(axis=1, keepdims=True)
    lse = mx + np.log(np.exp(logits - mx).sum(axis=1, keepdims=True))
    logq = logits - lse
    q = np.exp(logq)
    if mass is None:
        return None, None, q
    m = np.asarray(mass, np.float64)
    tm = m.sum(axis=1, keepdims=True)
    p = np.where(tm > 0, m / np.where(tm > 0, tm, 1), 1.0 / ns)
    loss = float(-(p * logq).sum() / nt)
    dl = (q - p) / nt
    g_wh = feat.T @ dl
    g_bh = dl.sum(axis=0)
    dfeat = dl @ wh.T
    # gradient at the pre-activation of the last convolution
    g = rnd((dfeat / cnt[:, None])[tix.ravel()].reshape(acts[-1].shape) * (acts[-1] > 0))
    g_ws, g_bs = [None] * 3, [None] * 3
    for l in (2, 1, 0):
        g_ws[l], g_bs[l] = conv_wgrad(acts[l], g)
        g = rnd(conv_dgrad(g, wq[l]) * (acts[l] > 0))
    g_w1, g_b1 = conv_wgrad(x, g)
    grads = np.concatenate([g_w1.reshape(-1), g_b1] +
                           [a for l in range(3) for a in (g_ws[l].reshape(-1), g_bs[l])] +
                           [g_wh.reshape(-1), g_bh])
    return loss, grads, q


def block_slices(ns: int):
    """Name and slice of each parameter block in the flat vector."""
    out, o = [], 0
    for name, n in [("w1", NTAP * C), ("b1", C)] + [(f"{k}{l + 2}", s) for l in range(3)
                                                      for k, s in (("w", NTAP * C * C), ("b", C))] + \
            [("wh", C * ns), ("bh", ns)]:
        out.append((name, slice(o, o + n)))
        o += n
    return out
