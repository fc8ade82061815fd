"""Writes the golden fixtures under tests/golden/ from the reference's own CPU
code, compiled from its sources by oracle/Makefile into oracle/_ref.

Run in a container that has the reference tree:
    make -C oracle && python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  tables.npz      search/patch linear offsets and computed weights from
                  setup_searchspace/setup_patchspace/setup_distweight
                  (iternlm_prepare.cpp:102-263) for several radii;
  pad.npz         pad_reflective (iternlm_prepare.cpp:6-48) on small volumes;
  filter_*.npz    two iterations of the reference's specialised CPU kernels
                  (iternlm_cpu_kernels.cpp) on small synthetic volumes with a
                  fixed per-slice sigma list: inputs, sigma lists, outputs.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from microtomodenoise_amd import phantom as ph  # noqa: E402

TABLE_CASES = [
    # (shape_padded, rs, rp, nslices, z_aniso)
    ((40, 36, 30), (5, 5, 5), (1, 1, 1), 11, 1.0),
    ((50, 44, 30), (7, 7, 7), (2, 2, 2), 11, 1.0),
    ((50, 44, 36), (7, 7, 7), (2, 2, 2), 15, 1.0),
    ((60, 60, 40), (10, 10, 10), (1, 1, 1), 11, 1.0),
    ((40, 40, 30), (2, 3, 4), (3, 3, 3), 11, 1.0),
    ((40, 40, 30), (3, 3, 3), (3, 3, 2), 11, 1.0),
    ((40, 40, 30), (3, 3, 3), (2, 2, 1), 11, 1.0),
    ((40, 40, 30), (3, 3, 3), (1, 1, 2), 11, 1.0),
    ((40, 40, 30), (3, 3, 3), (2, 2, 2), 11, 2.5),
    ((40, 40, 30), (3, 3, 3), (4, 4, 4), 11, 1.0),
]

# (name, shape xyz, rs, rp, uint16-valued?, sigma base)
FILTER_CASES = [
    ("p111_s3_u16", (24, 20, 16), (3, 3, 3), (1, 1, 1), True, 1200.0),
    ("p111_s5_u16", (32, 32, 24), (5, 5, 5), (1, 1, 1), True, 1200.0),
    ("p222_s2_u16", (24, 20, 16), (2, 2, 2), (2, 2, 2), True, 1500.0),
    ("p222_s3_f32", (24, 20, 16), (3, 3, 3), (2, 2, 2), False, 0.08),
    ("p333_s2_u16", (24, 20, 16), (2, 2, 2), (3, 3, 3), True, 2200.0),
    ("p112_s2_u16", (24, 20, 16), (2, 2, 2), (1, 1, 2), True, 1300.0),
    ("p113_s2_f32", (24, 20, 16), (2, 2, 2), (1, 1, 3), False, 0.07),
    ("p221_s3_u16", (24, 20, 16), (3, 3, 2), (2, 2, 1), True, 1400.0),
    ("p331_s2_u16", (24, 20, 16), (2, 2, 2), (3, 3, 1), True, 1800.0),
    ("p332_s2_f32", (24, 20, 16), (2, 2, 2), (3, 3, 2), False, 0.1),
]


def volume(shape, seed, u16):
    if u16:
        spec = ph.PhantomSpec(shape, seed=seed, radii=(2, 6), noise_sigma=1500.0, blur_sigma=0.8, streaks=3,
                              streak_len=(5, 15))
    else:
        spec = ph.PhantomSpec(shape, seed=seed, kind="ellipsoids", radii=(2, 8), levels=(0.2, 0.5, 0.9),
                              noise_sigma=0.05, blur_sigma=1.0, uint16=False)
    return ph.make(spec)


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref is not built (needs the reference tree)")
    tabs = {}
    for i, (shp, rs, rp, ns, an) in enumerate(TABLE_CASES):
        soff, poff, w = O.ref_tables(shp, rs, rp, ns, an)
        tabs[f"c{i}_shape"] = np.asarray(shp, np.int32)
        tabs[f"c{i}_rs"] = np.asarray(rs, np.int32)
        tabs[f"c{i}_rp"] = np.asarray(rp, np.int32)
        tabs[f"c{i}_nslices"] = np.int32(ns)
        tabs[f"c{i}_aniso"] = np.float32(an)
        tabs[f"c{i}_soff"], tabs[f"c{i}_poff"], tabs[f"c{i}_w"] = soff, poff, w
    np.savez_compressed(os.path.join(HERE, "tables.npz"), n=np.int32(len(TABLE_CASES)), **tabs)

    rng = np.random.default_rng(7)
    pads = {}
    for i, (shape, pad6) in enumerate([((5, 4, 3), (3, 2, 2, 4, 3, 2)), ((7, 6, 5), (6, 5, 4, 6, 5, 4)),
                                       ((2, 3, 2), (5, 5, 5, 5, 5, 5))]):
        v = rng.standard_normal((shape[2], shape[1], shape[0])).astype(np.float32)
        pads[f"c{i}_in"] = v
        pads[f"c{i}_pad"] = np.asarray(pad6, np.int32)
        pads[f"c{i}_out"] = O.ref_pad_reflect(v, pad6)
    np.savez_compressed(os.path.join(HERE, "pad.npz"), n=np.int32(3), **pads)

    for k, (name, shape, rs, rp, u16, sbase) in enumerate(FILTER_CASES):
        raw = volume(shape, 100 + k, u16)
        nz = shape[2]
        sig1 = (sbase * (1.0 + 0.3 * np.sin(np.arange(nz) * 0.7))).astype(np.float32)
        sig2 = (0.8 * sig1).astype(np.float32)
        p = O.Params(radius_searchspace=rs, radius_patchspace=rp, nslices=11, blockwise=False)
        out1 = O.ref_filter_iteration(raw, raw, sig1, p)
        out2 = O.ref_filter_iteration(raw, out1, sig2, p)
        np.savez_compressed(os.path.join(HERE, f"filter_{name}.npz"), raw=raw, sigma1=sig1, sigma2=sig2,
                            out1=out1, out2=out2, rs=np.asarray(rs, np.int32), rp=np.asarray(rp, np.int32),
                            nslices=np.int32(11))
        zero = int(np.sum(out1 == raw))
        print(f"{name}: shape {shape} zero-weight voxels {zero}/{raw.size}")


if __name__ == "__main__":
    main()
