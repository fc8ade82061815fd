"""Learned offset proposal for the pruned matrix-core filter.

A small 3D CNN (``inlm_proposal_*`` in include/inlm_capi.h, kernels in
csrc/inlm_learn.hip) predicts, for every 32x8x8 filter tile, how the filter's
weight mass spreads over the search offsets.  The filter then visits only the
K most probable offsets of a tile, falls back to the whole search where those
K carry less than a threshold of the predicted mass, and audits a random
fraction of tiles with the whole search each iteration.

Labels come from the FP32 parity filter itself (``inlm_weight_mass``): for a
tile and an offset, the sum over the tile's voxels of the weight the offset
received.  Training is softmax cross-entropy against the normalised mass,
with Adam, entirely in the library's own HIP kernels.  Device buffers are
plain device pointers; torch tensors serve only as device memory here.

The reference has no counterpart: its GPU filter always visits every offset
(iternlm_gpu.cu:80-174).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _capi as capi


class OffsetProposal:
    """Handle of one proposal network (``inlm_proposal``)."""

    def __init__(self, n_search: int, device: int = 0, seed: int = 0):
        self.n_search = int(n_search)
        self.handle = C.c_void_p()
        capi.check_proposal(capi.lib().inlm_proposal_create(self.n_search, device, seed, C.byref(self.handle)))

    def close(self):
        if self.handle:
            capi.lib().inlm_proposal_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        capi.check_proposal(st, self.handle)

    @property
    def num_params(self) -> int:
        return int(capi.lib().inlm_proposal_num_params(self.handle))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        self._check(capi.lib().inlm_proposal_params(self.handle, capi.ptr(out, C.c_float), 0))
        return out

    def set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {p.size}")
        self._check(capi.lib().inlm_proposal_params(self.handle, capi.ptr(p, C.c_float), 1))

    def grads(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        self._check(capi.lib().inlm_proposal_grads(self.handle, capi.ptr(out, C.c_float)))
        return out

    def train_step(self, vol_ptr: int, shape_xyz: Sequence[int], mass_ptr: int, lr: float) -> float:
        """Forward + backward on a device volume and its device labels; Adam
        update when lr > 0.  Returns the loss before the update."""
        sh = np.asarray(shape_xyz, np.int32)
        loss = C.c_float()
        self._check(capi.lib().inlm_proposal_train_step(self.handle, C.c_void_p(vol_ptr), capi.ptr(sh, C.c_int32),
                                                        C.c_void_p(mass_ptr), float(lr), C.byref(loss)))
        return float(loss.value)

    def predict(self, vol_ptr: int, shape_xyz: Sequence[int]) -> np.ndarray:
        sh = np.asarray(shape_xyz, np.int32)
        nx, ny, nz = (int(v) for v in shape_xyz)
        nt = ((nx + 31) // 32) * ((ny + 7) // 8) * ((nz + 7) // 8)
        out = np.empty((nt, self.n_search), np.float32)
        self._check(capi.lib().inlm_proposal_predict(self.handle, C.c_void_p(vol_ptr), capi.ptr(sh, C.c_int32),
                                                     out.ctypes.data_as(C.c_void_p), 0))
        return out


# ---------------------------------------------------------------------------
# Offset lists without a network, for comparison
# ---------------------------------------------------------------------------

def nearest_offsets(search_dxyz: np.ndarray, K: int) -> np.ndarray:
    """Indices of the K search offsets closest to the centre (Euclidean, ties
    in search order), ascending."""
    d2 = (np.asarray(search_dxyz, np.int64) ** 2).sum(axis=1)
    order = np.argsort(d2, kind="stable")[:K]
    return np.sort(order).astype(np.int32)


def lists_from_scores(scores: np.ndarray, K: int):
    """Per-tile top-K lists (ascending search index) from a [tiles, n_search]
    score array; counts are all K."""
    idx = np.argsort(-scores, axis=1, kind="stable")[:, :K]
    lists = np.sort(idx, axis=1).astype(np.int32)
    counts = np.full(scores.shape[0], K, np.int32)
    return lists, counts


def same_lists(ntiles: int, idx: np.ndarray):
    """One offset list shared by every tile."""
    K = len(idx)
    return np.tile(np.asarray(idx, np.int32), (ntiles, 1)), np.full(ntiles, K, np.int32)


def captured_mass(mass: np.ndarray, lists: np.ndarray, counts: np.ndarray) -> float:
    """Fraction of the total weight mass that the lists keep (tiles whose
    count is the whole search keep all of theirs)."""
    ns = mass.shape[1]
    kept = 0.0
    for t in range(mass.shape[0]):
        if counts[t] >= ns:
            kept += float(mass[t].sum())
        else:
            kept += float(mass[t, lists[t, : counts[t]]].sum())
    return kept / float(mass.sum())


# ---------------------------------------------------------------------------
# Training
# ---------------------------------------------------------------------------

def make_labels(vol, params, iteration: int = 1, sigma: Optional[np.ndarray] = None, device: int = 0):
    """Weight-mass labels of a device volume (torch tensor, float32, z-y-x)
    from the FP32 parity filter.  Iteration 1 runs on the original; iteration
    2 first runs iteration 1 with the same sigma.  sigma defaults to the
    library's noise estimate of the original.  Returns (mass tensor on the
    device [tiles, n_search], sigma)."""
    import torch

    from .denoiser import IterativeNLM_GPU

    nz, ny, nx = vol.shape
    g = IterativeNLM_GPU()
    if g.configure_device([nx, ny, nz], params) < 0:
        raise RuntimeError("configure_device failed")
    try:
        g.load_raw_device(vol.data_ptr())
        if sigma is None:
            samples = capi.draw_samples([nx, ny, nz], params.noiselevel.n_samples, params.noiselevel.patchsize, 0)
            sigma = g.estimate_noise(samples, 0)
        g.set_sigma(sigma)
        ns, _ = g.table_sizes()
        mass = torch.empty((g.tile_count(), ns), dtype=torch.float32, device=vol.device)
        for it in range(1, iteration):
            g.Run_GaussianNoise(it)
        g.weight_mass(iteration, mass.data_ptr())
    finally:
        g.free_device()
    return mass, sigma


def train(proposal: OffsetProposal, data, steps: int, lr: float = 3e-3, log=None, log_every: int = 50):
    """Adam over `steps` steps, cycling through `data` = [(volume tensor,
    mass tensor)].  Returns the loss history."""
    hist = []
    for s in range(steps):
        vol, mass = data[s % len(data)]
        nz, ny, nx = vol.shape
        loss = proposal.train_step(vol.data_ptr(), [nx, ny, nz], mass.data_ptr(), lr)
        hist.append(loss)
        if log is not None and (s % log_every == 0 or s == steps - 1):
            log(f"[proposal] step {s} loss {loss:.5f}")
    return hist
