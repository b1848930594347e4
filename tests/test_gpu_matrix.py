"""GPU parity of the fused smatrix kernels against the CPU oracle.

Bars: FP64 build 1e-12 relative per event (reference constants, ref_compat), FP32 build
2e-6 relative on the median and 5e-5 on the worst event (FP32 amplitudes and colour sum).
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.helpers import phase_space_points
from madflow_amd.matrix import load_matrix

PROCS = ["gg_ttx", "gg_ttxg", "gg_ttxgg"]
NEV = {"gg_ttx": 512, "gg_ttxg": 256, "gg_ttxgg": 64}


def _fused(name, **kw):
    """The process's fused Feynman-diagram kernel (g g > t t~ g g defaults to the
    colour-ordered kernels, tested in test_gpu_ordered.py)."""
    return load_matrix(name, kernels="fused", **kw)


def _params(matrix, ref_gs=True):
    rp, coup = O.default_params(ref_gs=ref_gs)
    vals = [rp[0], rp[1]] + [coup[c] for c in matrix.params[2:]]
    return rp, coup, vals


@pytest.mark.gpu
@pytest.mark.parametrize("name", PROCS)
def test_smatrix_fp64_parity(dev, name):
    print("case smatrix fp64", name, flush=True)
    m, _ = _fused(name, precision=64, ref_compat=True)
    ps, _ = phase_space_points(name, NEV[name], seed=11)
    rp, coup, vals = _params(m)
    ref = O.smatrix(name, ps, rp, coup, ref_compat=True)
    got = m.smatrix(torch.from_numpy(ps).to(dev), *vals).cpu().numpy()
    rel = np.abs(got / ref - 1)
    print(name, "fp64 max rel", rel.max(), flush=True)
    assert rel.max() < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", PROCS)
def test_smatrix_fp32_tolerance(dev, name):
    print("case smatrix fp32", name, flush=True)
    m, _ = _fused(name, precision=32, ref_compat=True)
    ps, _ = phase_space_points(name, NEV[name], seed=12)
    rp, coup, vals = _params(m)
    ref = O.smatrix(name, ps, rp, coup, ref_compat=True)
    got = m.smatrix(torch.from_numpy(ps).to(dev), *vals).cpu().numpy()
    rel = np.abs(got / ref - 1)
    print(name, "fp32 median rel", np.median(rel), "max rel", rel.max(), flush=True)
    assert np.median(rel) < 2e-6
    assert rel.max() < 5e-5


@pytest.mark.gpu
@pytest.mark.parametrize("name", PROCS)
def test_matrix_jamp_per_helicity(dev, name):
    """matrix(hel) and the jamps of single helicity rows equal the oracle's."""
    print("case matrix/jamp", name, flush=True)
    m, _ = _fused(name, precision=64, ref_compat=True)
    ps, _ = phase_space_points(name, 32, seed=13)
    rp, coup, vals = _params(m)
    tps = torch.from_numpy(ps).to(dev)
    for row in (0, len(m.helicities) // 3, len(m.helicities) - 1):
        hel = m.helicities[row]
        ref, jref = O.matrix_hel(name, ps, hel, rp, coup, ref_compat=True, want_jamp=True)
        got, jgot = m._matrix(tps, hel, tuple(vals), return_jamp=True)
        got = got.cpu().numpy()
        jgot = jgot.cpu().numpy()
        assert np.max(np.abs(got / ref - 1)) < 1e-12
        scale = np.max(np.abs(jref), axis=1, keepdims=True)
        # jamps carry cancellations between diagrams: bar on the largest jamp of the event
        assert np.max(np.abs(jgot - jref) / scale) < 1e-11


@pytest.mark.gpu
@pytest.mark.parametrize("name", PROCS)
def test_gauge_invariance_gpu(dev, name):
    """Replacing a gluon polarisation by its momentum (nhel = 4) kills every jamp.

    The Ward identity needs a zero top width (a Breit-Wigner width breaks it at
    O(m Gamma / (p^2 - m^2)) ~ 1e-3), and helicity rows whose amplitudes vanish
    identically are skipped by using a row that is non-zero for every process."""
    print("case brst", name, flush=True)
    m, _ = _fused(name, precision=64, ref_compat=False)
    ps, _ = phase_space_points(name, 16, seed=14)
    rp, coup, vals = _params(m, ref_gs=False)
    vals[1] = 0.0
    tps = torch.from_numpy(ps).to(dev)
    hel = list(m.helicities[5])
    _, jphys = m._matrix(tps, hel, tuple(vals), return_jamp=True)
    for leg in (0, 1):
        h = list(hel)
        h[leg] = 4
        _, jbrst = m._matrix(tps, h, tuple(vals), return_jamp=True)
        ratio = (jbrst.abs().max(dim=1).values / jphys.abs().max(dim=1).values).cpu().numpy()
        print(name, "leg", leg, "BRST ratio", ratio.max(), flush=True)
        assert ratio.max() < 1e-11


@pytest.mark.gpu
def test_running_couplings_per_event(dev):
    """Per-event couplings (running alpha_s) give the oracle's per-event values."""
    name = "gg_ttxg"
    m, mod = _fused(name, precision=64, ref_compat=True)
    model = mod.get_model_param()
    ps, _ = phase_space_points(name, 64, seed=15)
    alphas = np.linspace(0.1, 0.13, 64)
    params = model.evaluate(torch.tensor(alphas))
    got = m.smatrix(torch.from_numpy(ps).to(dev), *params).cpu().numpy()
    ref = np.zeros(64)
    for i, a in enumerate(alphas):
        rp, coup = O.default_params(alpha_s=a)
        ref[i] = O.smatrix(name, ps[i:i + 1], rp, coup, ref_compat=True)[0]
    assert np.max(np.abs(got / ref - 1)) < 1e-12


@pytest.mark.gpu
def test_empty_batch(dev):
    m, _ = load_matrix("gg_ttx")
    rp, coup, vals = _params(m)
    out = m.smatrix(torch.zeros((0, 4, 4), dtype=torch.float64, device=dev), *vals)
    assert out.shape == (0,)
