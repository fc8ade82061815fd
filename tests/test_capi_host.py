"""CPU tests of the shipped library: it loads, exports every symbol the
header declares, rejects bad arguments, reports NO_DEVICE without a GPU, and
its host-side logic (tables, weight tables, sample positions) agrees with the
oracle and the reference fixtures.  No compute call runs here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, load_npz
from microtomodenoise_amd import _capi
from oracle import oracle as O


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "inlm_capi.h")).read()
    return sorted(set(re.findall(r"INLM_API\s+[\w\s\*]+?\b(inlm_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    syms = header_symbols()
    assert len(syms) >= 15
    assert sorted(_capi.EXPORTS) == syms
    L = _capi.lib()
    for s in syms:
        assert hasattr(L, s), s


def test_default_params_are_reference_defaults():
    p = _capi.default_params()
    assert p.alpha == 0.5 and p.maxiterations == 4 and p.nslices == 11
    assert list(p.radius_searchspace) == [10, 10, 10] and list(p.radius_patchspace) == [1, 1, 1]
    assert p.beta == 1.0 and p.z_anisotropy == 1.0
    assert p.noise_mode == _capi.NOISE_ZADAPTIVE and p.zwindow == 100 and p.n_samples == 100000
    assert p.patchsize == 15 and p.stds_from_mean == 1.0 and p.cpu_blockwise == 1


def _create(p, shape):
    h = C.c_void_p()
    sh = np.asarray(shape, np.int32)
    st = _capi.lib().inlm_create(C.byref(p), _capi.ptr(sh, C.c_int32), 0, C.byref(h))
    if st == 0:
        _capi.lib().inlm_destroy(h)
    return st


def test_bad_arguments_rejected_before_device():
    L = _capi.lib()
    p = _capi.default_params()
    assert _create(p, [0, 10, 10]) == _capi.INLM_BAD_ARG
    p2 = _capi.default_params()
    p2.radius_patchspace[:] = [2, 2, 3]
    p2.cpu_blockwise = 1
    assert _create(p2, [32, 32, 32]) == _capi.INLM_BAD_ARG
    assert b"undefined" in L.inlm_last_error(None)
    p3 = _capi.default_params()
    p3.radius_searchspace[:] = [3, 3, 3]
    assert _create(p3, [1, 32, 32]) == _capi.INLM_BAD_ARG  # reflect padding of a 1-voxel axis


def test_no_device_status_and_fallback_signal():
    # asked at run time, not at collection: a collection-time call would
    # start the library's HIP runtime before torch's in a GPU session
    if _capi.lib().inlm_device_count() > 0:
        pytest.skip("a GPU is visible")
    p = _capi.default_params()
    p.radius_searchspace[:] = [2, 2, 2]
    assert _create(p, [16, 16, 16]) == _capi.INLM_NO_DEVICE
    from microtomodenoise_amd import DenoiseParameters, IterativeNLM_GPU
    params = DenoiseParameters()
    assert IterativeNLM_GPU().configure_device([16, 16, 16], params) == -1
    assert params.gpu.n_gpus == 0  # as the reference, which then takes the CPU path


def test_product_tables_match_reference_fixtures():
    z = load_npz("tables.npz")
    for i in range(int(z["n"])):
        shp, rs, rp = z[f"c{i}_shape"], z[f"c{i}_rs"], z[f"c{i}_rp"]
        p = _capi.default_params()
        p.radius_searchspace[:] = [int(v) for v in rs]
        p.radius_patchspace[:] = [int(v) for v in rp]
        p.nslices = int(z[f"c{i}_nslices"])
        p.z_anisotropy = float(z[f"c{i}_aniso"])
        p.cpu_blockwise = 1
        # unpadded shape whose padded shape is the fixture's
        pz = min(p.nslices // 2, int(rs[2])) + int(rp[2])
        unp = [int(shp[0]) - 2 * (int(rs[0]) + int(rp[0])), int(shp[1]) - 2 * (int(rs[1]) + int(rp[1])),
               int(shp[2]) - 2 * pz]
        if tuple(rp) == (2, 2, 3):
            continue
        S, PT, w = _capi.tables(p, unp)
        assert np.array_equal(O.linear(S, shp[0], shp[1]), z[f"c{i}_soff"]), i
        assert np.array_equal(O.linear(PT, shp[0], shp[1]), z[f"c{i}_poff"]), i
        kind = O.orc().orc_cpu_weight_kind(np.asarray(rp, np.int32), 1)
        # the fixture holds the reference's setup_distweight table: the
        # product applies it for computed radii and, with z anisotropy, for
        # the literal radii too (the reference GPU class's generic kernel)
        if kind == 0 or (kind == 1 and p.z_anisotropy != 1.0):
            assert np.array_equal(w.view(np.uint32), z[f"c{i}_w"].view(np.uint32)), i


def test_anisotropy_reaches_the_literal_radii():
    # rp (2,2,2) is a literal-table radius of the CPU kernels; z anisotropy
    # must still change its weights (only z taps lose weight), while (1,1,3)
    # keeps its literal table as the reference GPU class's patch113 kernel does
    def table(rp, aniso):
        p = _capi.default_params()
        p.radius_searchspace[:] = [3, 3, 3]
        p.radius_patchspace[:] = list(rp)
        p.z_anisotropy = aniso
        return _capi.tables(p, [30, 30, 30])
    _, PT, w1 = table((2, 2, 2), 1.0)
    _, _, w2 = table((2, 2, 2), 2.0)
    dz = PT[:, 2] != 0
    assert np.all(w2[dz] < w1[dz]) and np.allclose(w2[~dz], w1[~dz], rtol=1e-6)
    op = O.Params(radius_searchspace=(3, 3, 3), radius_patchspace=(2, 2, 2), z_anisotropy=2.0)
    assert np.array_equal(w2, O.cpu_weight_table(op, 40))
    _, _, a = table((1, 1, 3), 1.0)
    _, _, b = table((1, 1, 3), 2.0)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("rp,blockwise", [((1, 1, 1), 1), ((2, 2, 2), 1), ((3, 3, 3), 0), ((3, 3, 3), 1),
                                          ((1, 1, 2), 0), ((2, 2, 1), 1), ((3, 3, 2), 0), ((3, 3, 2), 1),
                                          ((2, 3, 4), 1)])
def test_product_weight_table_matches_oracle_policy(rp, blockwise):
    p = _capi.default_params()
    p.radius_searchspace[:] = [3, 3, 3]
    p.radius_patchspace[:] = list(rp)
    p.cpu_blockwise = blockwise
    _, _, w = _capi.tables(p, [40, 40, 40])
    op = O.Params(radius_searchspace=(3, 3, 3), radius_patchspace=rp, blockwise=bool(blockwise))
    nzp = 40 + 2 * op.padding()[2]
    assert np.array_equal(w.view(np.uint32), O.cpu_weight_table(op, nzp).view(np.uint32))


@pytest.mark.parametrize("seed", [0, 1, 12345])
def test_product_samples_equal_oracle_samples(seed):
    a = _capi.draw_samples([100, 90, 20], 20000, 15, seed)
    b = O.draw_samples((100, 90, 20), 20000, 15, seed)
    assert np.array_equal(a, b)


def test_product_cluster_matches_oracle_fp64():
    # exercised through the GPU path on the box; here through the oracle's
    # FP64 twin to pin the shared algorithm on a planted bimodal set
    rng = np.random.default_rng(9)
    vals = np.concatenate([rng.normal(100, 5, 5000), rng.normal(400, 40, 3000)])
    m, s = O.create2clusters(vals, precision="f64")
    assert abs(m - 100) < 1 and abs(s - 5) < 0.5
