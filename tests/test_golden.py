"""Golden vectors (tests/golden/*.json, written by tests/golden/make_fixtures.py).

CPU: the oracle still reproduces the stored |M|^2 (reference constants and exact
constants), the stored jamps, and the hand-restated reference mockup for gg -> tt~.
GPU: the fused HIP kernels against the stored vectors (FP64 1e-12, FP32 stated bars).
"""

import json
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PROCS = ["gg_ttx", "gg_ttxg", "gg_ttxgg", "gg_ttxggg"]


def load(proc):
    with open(os.path.join(HERE, proc + ".json")) as f:
        g = json.load(f)
    g["all_ps"] = np.array(g["all_ps"])
    g["jamp"] = np.array(g["jamp_re"]) + 1j * np.array(g["jamp_im"])
    return g


def _params(g, key):
    p = g[key]
    rp = np.array([p["mdl_MT"], p["mdl_WT"]])
    coup = {k: complex(*v) for k, v in p["couplings"].items()}
    return rp, coup


@pytest.mark.parametrize("proc", PROCS)
def test_oracle_reproduces_golden(proc):
    g = load(proc)
    ps = g["all_ps"]
    rp, coup = _params(g, "params_ref_compat")
    me = O.smatrix(proc, ps, rp, coup, ref_compat=True)
    np.testing.assert_allclose(me, g["smatrix_ref_compat"], rtol=1e-14, atol=0)
    rp_e, coup_e = _params(g, "params_exact")
    me_e = O.smatrix(proc, ps, rp_e, coup_e, ref_compat=False)
    np.testing.assert_allclose(me_e, g["smatrix_exact"], rtol=1e-14, atol=0)
    m, j = O.matrix_hel(proc, ps, np.array(g["helicity"], dtype=np.float64), rp, coup, ref_compat=True,
                        want_jamp=True)
    np.testing.assert_allclose(m, g["matrix_hel_ref_compat"], rtol=1e-14, atol=0)
    np.testing.assert_allclose(j, g["jamp"], rtol=1e-13, atol=1e-13 * np.abs(g["jamp"]).max())
    # the two constant sets differ by the float32-rounded sqrt(1/2) per gluon polarisation
    # (and the 2-ulp gs literal): (SQH32/SQH)^(2 n_g) - 1
    ng = len(ps[0]) - 2
    sqh32 = float(np.float32(np.sqrt(0.5)))
    expect = (sqh32 / np.sqrt(0.5)) ** (2 * ng)
    gs_ratio = (1.2177157847767195 / (2 * np.sqrt(np.pi * 0.118))) ** (2 * ng)
    np.testing.assert_allclose(np.array(g["smatrix_ref_compat"]) / np.array(g["smatrix_exact"]),
                               expect * gs_ratio, rtol=2e-14)


def test_mockup_matches_golden():
    """The reference's committed gg -> tt~ module, restated by hand, on the stored points."""
    g = load("gg_ttx")
    np.testing.assert_allclose(g["mockup_ref_compat"], g["smatrix_ref_compat"], rtol=1e-14, atol=0)
    me, _ = O.smatrix_gg_ttx_mockup(g["all_ps"], ref_compat=True)
    np.testing.assert_array_equal(me, np.array(g["mockup_ref_compat"]))


GPU_PROCS = ["gg_ttx", "gg_ttxg", "gg_ttxgg", "gg_ttxggg"]


@pytest.mark.gpu
@pytest.mark.parametrize("proc", GPU_PROCS)
@pytest.mark.parametrize("precision", [64, 32])
def test_gpu_smatrix_golden(dev, proc, precision):
    import torch
    from madflow_amd.matrix import load_matrix
    print("case golden", proc, precision, flush=True)
    g = load(proc)
    m, _ = load_matrix(proc, precision=precision, ref_compat=True)
    rp, coup = _params(g, "params_ref_compat")
    vals = [rp[0], rp[1]] + [coup[c] for c in m.params[2:]]
    got = m.smatrix(torch.from_numpy(g["all_ps"]).to(dev), *vals).cpu().numpy()
    rel = np.abs(got / np.array(g["smatrix_ref_compat"]) - 1)
    print(proc, precision, "golden max rel", rel.max(), flush=True)
    assert rel.max() < (1e-12 if precision == 64 else 1e-4)
    if precision == 64:
        mh, jh = m._matrix(torch.from_numpy(g["all_ps"]).to(dev), g["helicity"], tuple(vals), return_jamp=True)
        np.testing.assert_allclose(mh.cpu().numpy(), g["matrix_hel_ref_compat"], rtol=1e-12, atol=0)
        jh = jh.cpu().numpy()
        scale = np.abs(g["jamp"]).max(axis=1, keepdims=True)
        assert np.max(np.abs(jh - g["jamp"]) / scale) < 1e-11
