"""Per-event error of the FP32 build of every process against the FP64 oracle, on larger
samples than the parity tests: random RAMBO points, and the subset that passes the
benchmark cuts (pT(g) > 30 GeV, dR(g,g) > 0.4).  Prints one line per process with the
median, 99th percentile and maximum relative error.  Bars: median 2e-6, 99th percentile
1e-5, worst event 5e-5 (measured: worst 7e-7 to 1.9e-5 depending on the process)."""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.helpers import phase_space_points
from madflow_amd.matrix import load_matrix

# (process, kernel family): the default kernels of every process, and the fused
# g g > t t~ g g kernel beside its colour-ordered default
CASES = [("gg_ttx", None, 8192), ("gg_ttxg", None, 8192), ("gg_ttxgg", None, 8192), ("gg_ttxgg", "fused", 8192),
         ("gg_ttxggg", None, 1024), ("uux_ttx", None, 8192), ("uux_ttxg", None, 8192), ("uux_ttxgg", None, 8192),
         ("ug_ttxu", None, 8192), ("uxg_ttxux", None, 8192)]


def _passes_cuts(ps, pt_min=30.0, dr_min=0.4):
    g = ps[:, 4:]                                   # outgoing gluons
    pt = np.hypot(g[..., 1], g[..., 2])
    ok = np.all(pt > pt_min, axis=1)
    pabs = np.linalg.norm(g[..., 1:], axis=-1)
    eta = np.arctanh(np.clip(g[..., 3] / pabs, -1 + 1e-15, 1 - 1e-15))
    phi = np.arctan2(g[..., 2], g[..., 1])
    for a in range(g.shape[1]):
        for b in range(a + 1, g.shape[1]):
            dphi = np.abs(phi[:, a] - phi[:, b])
            dphi = np.minimum(dphi, 2 * np.pi - dphi)
            ok &= np.hypot(eta[:, a] - eta[:, b], dphi) > dr_min
    return ok


@pytest.mark.gpu
@pytest.mark.parametrize("name,kernels,nev", CASES)
def test_fp32_error_against_oracle(dev, name, kernels, nev):
    print("case fp32 error", name, kernels, flush=True)
    oname = name
    m, _ = load_matrix(name, precision=32, ref_compat=True, kernels=kernels)
    name = m.registry_name
    ps, _ = phase_space_points(oname, nev, seed=31)
    rp, coup = O.default_params(ref_gs=True)
    vals = [rp[0], rp[1]] + [coup[c] for c in m.params[2:]]
    ref = O.smatrix(oname, ps, rp, coup, ref_compat=True, nthreads=16)
    got = m.smatrix(torch.from_numpy(ps).to(dev), *vals).cpu().numpy()
    rel = np.abs(got / ref - 1)
    cut = _passes_cuts(ps) if ps.shape[1] > 4 else np.ones(len(ps), dtype=bool)
    line = "fp32 error %s: %d points median %.2e p99 %.2e max %.2e" % (
        name, len(rel), np.median(rel), np.quantile(rel, 0.99), rel.max())
    if ps.shape[1] > 4:
        rc = rel[cut]
        line += "; %d post-cut median %.2e p99 %.2e max %.2e" % (len(rc), np.median(rc), np.quantile(rc, 0.99),
                                                                 rc.max())
    print(line, flush=True)
    assert np.median(rel) < 2e-6
    assert np.quantile(rel, 0.99) < 1e-5
    assert rel.max() < 5e-5
