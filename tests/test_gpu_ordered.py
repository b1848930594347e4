"""Colour-ordered kernels (Berends-Giele per colour ordering + MFMA colour contraction)
against the CPU oracle's Feynman-diagram evaluation.

g g > t t~ g g runs through the colour-ordered kernels by default (the drop-in
Matrix_1_gg_ttxgg), so the same physics is checked against the oracle and against the
fused Feynman-diagram kernel (kernels="fused"); gg_ttxggg (120 colour flows) is only
evaluated this way.  Bars: FP64 1e-12 relative per
event, jamps 1e-11 of the event's largest jamp; FP32 2e-6 median / 5e-5 worst event.
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.helpers import phase_space_points
from madflow_amd.matrix import load_matrix

CASES = [("gg_ttxgg", "gg_ttxgg", 128), ("gg_ttxggg", "gg_ttxggg", 32)]


def _vals(m):
    rp, coup = O.default_params(ref_gs=True)
    return rp, coup, [rp[0], rp[1]] + [coup[c] for c in m.params[2:]]


@pytest.mark.gpu
@pytest.mark.parametrize("name,oracle_name,nev", CASES)
def test_ordered_fp64_parity(dev, name, oracle_name, nev):
    print("case ordered fp64", name, flush=True)
    m, _ = load_matrix(name, precision=64, ref_compat=True)
    ps, _ = phase_space_points(oracle_name, nev, seed=21)
    rp, coup, vals = _vals(m)
    ref = O.smatrix(oracle_name, ps, rp, coup, ref_compat=True)
    got = m.smatrix(torch.from_numpy(ps).to(dev), *vals).cpu().numpy()
    rel = np.abs(got / ref - 1)
    print(name, "fp64 max rel", rel.max(), flush=True)
    assert rel.max() < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name,oracle_name,nev", CASES)
def test_ordered_fp32_tolerance(dev, name, oracle_name, nev):
    print("case ordered fp32", name, flush=True)
    m, _ = load_matrix(name, precision=32, ref_compat=True)
    ps, _ = phase_space_points(oracle_name, nev, seed=22)
    rp, coup, vals = _vals(m)
    ref = O.smatrix(oracle_name, ps, rp, coup, ref_compat=True)
    got = m.smatrix(torch.from_numpy(ps).to(dev), *vals).cpu().numpy()
    rel = np.abs(got / ref - 1)
    print(name, "fp32 median rel", np.median(rel), "max rel", rel.max(), flush=True)
    assert np.median(rel) < 2e-6
    assert rel.max() < 5e-5


@pytest.mark.gpu
@pytest.mark.parametrize("name,oracle_name,nev", CASES)
def test_ordered_jamps_per_helicity(dev, name, oracle_name, nev):
    """Each colour-ordered amplitude equals the oracle's jamp of that ordering (the sum of
    its Feynman diagrams with their colour coefficients)."""
    print("case ordered jamps", name, flush=True)
    m, _ = load_matrix(name, precision=64, ref_compat=True)
    ps, _ = phase_space_points(oracle_name, 16, seed=23)
    rp, coup, vals = _vals(m)
    tps = torch.from_numpy(ps).to(dev)
    for row in (0, len(m.helicities) // 2 + 3, len(m.helicities) - 1):
        hel = m.helicities[row]
        ref, jref = O.matrix_hel(oracle_name, ps, hel, rp, coup, ref_compat=True, want_jamp=True)
        got, jgot = m._matrix(tps, hel, tuple(vals), return_jamp=True)
        assert np.max(np.abs(got.cpu().numpy() / ref - 1)) < 1e-12
        scale = np.max(np.abs(jref), axis=1, keepdims=True)
        err = np.max(np.abs(jgot.cpu().numpy() - jref) / scale)
        print(name, "row", row, "jamp err", err, flush=True)
        assert err < 1e-11


@pytest.mark.gpu
def test_ordered_matches_fused_kernel(dev):
    """Same events through both g g > t t~ g g kernels, FP32: they agree to FP32 rounding."""
    print("case ordered vs fused", flush=True)
    ps, _ = phase_space_points("gg_ttxgg", 4096, seed=24)
    tps = torch.from_numpy(ps).to(dev)
    mo, _ = load_matrix("gg_ttxgg", precision=32, ref_compat=True)
    mf, _ = load_matrix("gg_ttxgg", precision=32, ref_compat=True, kernels="fused")
    assert (mo.registry_name, mf.registry_name) == ("gg_ttxgg", "gg_ttxgg_fused")
    _, _, vals = _vals(mf)
    a = mo.smatrix(tps, *vals).cpu().numpy()
    b = mf.smatrix(tps, *vals).cpu().numpy()
    rel = np.abs(a / b - 1)
    print("ordered vs fused fp32 median", np.median(rel), "max", rel.max(), flush=True)
    assert np.median(rel) < 2e-6 and rel.max() < 5e-5


@pytest.mark.gpu
@pytest.mark.parametrize("precision,rows", [(32, 40), (64, 40), (32, 200), (32, 400)])
def test_ordered_small_workspace_splits(dev, precision, rows):
    """A workspace of a few tiles forces event sub-chunks and helicity groups (with the
    colour kernel accumulating across groups): the result is unchanged.  40, 200 and 400
    event-rows for 100 events give groups of 1, 2 and 4 helicity rows, i.e. part of, one
    half of, and all four top / anti-top rows that share one gluon helicity set."""
    print("case ordered workspace split", precision, rows, flush=True)
    m, _ = load_matrix("gg_ttxggg", precision=precision, ref_compat=True)
    ps, _ = phase_space_points("gg_ttxggg", 100, seed=25)
    tps = torch.from_numpy(ps).to(dev)
    _, _, vals = _vals(m)
    full = m.smatrix(tps, *vals).cpu().numpy()
    m2, _ = load_matrix("gg_ttxggg", precision=precision, ref_compat=True)
    row = 120 * 2 * (4 if precision == 32 else 8)
    m2.workspace_limit = rows * row
    small = m2.smatrix(tps, *vals).cpu().numpy()
    np.testing.assert_allclose(small, full, rtol=1e-6 if precision == 32 else 1e-13, atol=0)
