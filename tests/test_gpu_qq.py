"""GPU parity of the generated u u~ > t t~ (+ g, + g g) and u g > t t~ u kernels (two fermion lines, rational
jamp coefficients, rows that vanish by chirality skipped in the kernel) against the CPU
oracle, which sums every row.

Bars: FP64 1e-12 relative per event (reference constants); jamps 1e-11 of the event's
largest jamp; FP32 2e-6 median and 5e-5 worst event."""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.helpers import phase_space_points
from madflow_amd.matrix import load_matrix

PROCS = ["uux_ttx", "uux_ttxg", "uux_ttxgg", "ug_ttxu", "uxg_ttxux"]


def _vals(m, ref_gs=True):
    rp, coup = O.default_params(ref_gs=ref_gs)
    return rp, coup, [rp[0], rp[1]] + [coup[c] for c in m.params[2:]]


def _cancellation(name, ps, rp, coup):
    """Per event, how much the Feynman-diagram sums cancel: sum over live rows of
    sum_ij |CF_ij| (sum_d |c_id A_d|)(sum_d |c_jd A_d|) over |M|^2 (1 when nothing
    cancels).  The amplitudes A_d come from the oracle with one jamp per amplitude."""
    import copy
    from fractions import Fraction
    d = O.load_process(name)
    na = d["ngraphs"]
    m = copy.deepcopy(d)
    m["jamp"] = [[[a, "1", 0]] for a in range(na)]
    m["ncolor"], m["color_num"], m["color_denom"] = na, np.eye(na, dtype=int).tolist(), [1] * na
    key = name + "_amplitudes"
    O._cache[key] = O.EncodedProcess(m)
    try:
        C = np.zeros((len(d["jamp"]), na), complex)
        for c, terms in enumerate(d["jamp"]):
            for a, re, im in terms:
                C[c, a] += float(Fraction(re)) + 1j * float(Fraction(im))
        cf = np.array(d["color_num"], float) / np.array(d["color_denom"], float)[:, None]
        num = np.zeros(len(ps))
        den = np.zeros(len(ps))
        for h in d["helicities"]:
            if (h[0] != h[4]) if name in ("ug_ttxu", "uxg_ttxux") else (h[0] == h[1]):    # vanish by chirality
                continue
            _, A = O.matrix_hel(key, ps, h, rp, coup, ref_compat=True, want_jamp=True)
            A = np.asarray(A)
            J, Jabs = A @ C.T, np.abs(A) @ np.abs(C).T
            num += np.einsum("ei,ij,ej->e", Jabs, np.abs(cf), Jabs)
            den += np.einsum("ei,ij,ej->e", np.conj(J), cf, J).real
    finally:
        del O._cache[key]
    return num / den


@pytest.mark.gpu
@pytest.mark.parametrize("name", PROCS)
def test_qq_fp64_parity(dev, name):
    """1e-12 per event where the diagram sums cancel less than 100-fold, 1e-14 kappa beyond
    (kappa the cancellation factor): rounding at 1e-16 per operation is amplified by
    kappa whatever the order of the operations, and the kernel's order is not the
    oracle's.  Among the 512 uncut u u~ > t t~ g g points the largest kappa is 335."""
    print("case qq fp64", name, flush=True)
    m, _ = load_matrix(name, precision=64, ref_compat=True)
    ps, _ = phase_space_points(name, 512, seed=51)
    rp, coup, vals = _vals(m)
    ref = O.smatrix(name, ps, rp, coup, ref_compat=True)
    got = m.smatrix(torch.from_numpy(ps).to(dev), *vals).cpu().numpy()
    rel = np.abs(got / ref - 1)
    kappa = _cancellation(name, ps, rp, coup)
    w = int(np.argmax(rel))
    print(name, "fp64 max rel %.3g at event %d (cancellation %.1f); events with cancellation > 100: %d, their max "
          "rel / cancellation %.3g" % (rel.max(), w, kappa[w], int(np.sum(kappa > 100)),
                                       float(np.max(rel / kappa, initial=0.0, where=kappa > 100))), flush=True)
    assert np.all(rel < np.maximum(1e-12, 1e-14 * kappa))


@pytest.mark.gpu
@pytest.mark.parametrize("name", PROCS)
def test_qq_fp32_tolerance(dev, name):
    print("case qq fp32", name, flush=True)
    m, _ = load_matrix(name, precision=32, ref_compat=True)
    ps, _ = phase_space_points(name, 4096, seed=52)
    rp, coup, vals = _vals(m)
    ref = O.smatrix(name, ps, rp, coup, ref_compat=True)
    got = m.smatrix(torch.from_numpy(ps).to(dev), *vals).cpu().numpy()
    rel = np.abs(got / ref - 1)
    print(name, "fp32 median rel", np.median(rel), "max rel", rel.max(), flush=True)
    assert np.median(rel) < 2e-6
    assert rel.max() < 5e-5


@pytest.mark.gpu
@pytest.mark.parametrize("name", PROCS)
def test_qq_rows_and_jamps(dev, name):
    """matrix(hel) and the jamps of every helicity row, live or vanishing, equal the
    oracle's (rows are evaluated one by one here, none skipped)."""
    print("case qq rows", name, flush=True)
    m, _ = load_matrix(name, precision=64, ref_compat=True)
    ps, _ = phase_space_points(name, 32, seed=53)
    rp, coup, vals = _vals(m)
    tps = torch.from_numpy(ps).to(dev)
    for hel in m.helicities:
        ref, jref = O.matrix_hel(name, ps, hel, rp, coup, ref_compat=True, want_jamp=True)
        got, jgot = m._matrix(tps, hel, tuple(vals), return_jamp=True)
        got, jgot = got.cpu().numpy(), jgot.cpu().numpy()
        if (hel[0] != hel[4]) if name in ("ug_ttxu", "uxg_ttxux") else (hel[0] == hel[1]):    # vanish by chirality
            assert np.all(got == 0.0) and np.all(ref == 0.0)
            continue
        assert np.max(np.abs(got / ref - 1)) < 1e-12
        scale = np.max(np.abs(jref), axis=1, keepdims=True)
        assert np.max(np.abs(jgot - jref) / scale) < 1e-11


@pytest.mark.gpu
def test_qq_gauge_invariance_gpu(dev):
    """u u~ > t t~ g with the gluon's polarisation replaced by its momentum (nhel = 4):
    every jamp vanishes (zero top width)."""
    print("case qq brst", flush=True)
    m, _ = load_matrix("uux_ttxg", precision=64, ref_compat=False)
    ps, _ = phase_space_points("uux_ttxg", 16, seed=54)
    rp, coup, vals = _vals(m, ref_gs=False)
    vals[1] = 0.0
    tps = torch.from_numpy(ps).to(dev)
    hel = list(m.helicities[0])
    assert hel[0] != hel[1]
    _, jphys = m._matrix(tps, hel, tuple(vals), return_jamp=True)
    hel[4] = 4
    _, jbrst = m._matrix(tps, hel, tuple(vals), return_jamp=True)
    scale = jphys.abs().max().item()
    assert scale > 0
    assert jbrst.abs().max().item() < 1e-11 * scale


@pytest.mark.gpu
def test_uux_ttx_cross_section(dev):
    """u u~ > t t~ at sqrt(s) = 1 TeV, alpha_s = 0.118, m_t = 173, Gamma_t = 0: the closed
    form sigma = 8 pi alpha_s^2 / (27 s) beta (1 + rho / 2) = 5.0185 pb; VEGAS with 1e6
    points x 5 iterations must agree within 3 sigma."""
    import math
    from madflow_amd.vegas import MEIntegrand, VegasFlow, GEV2PB
    print("case sigma uux_ttx", flush=True)
    ig = MEIntegrand("uux_ttx", precision=64, mapping="rambo", seed=2, sqrts=1000.0, ref_compat=False, width=0.0)
    vf = VegasFlow(ig.ndim, 1000000)
    vf.compile(ig)
    vf.run_integration(2)
    vf.freeze_grid()
    res, err = vf.run_integration(5)
    s, m, a = 1e6, 173.0, 0.118
    rho = 4 * m * m / s
    closed = 8 * math.pi * a * a / (27 * s) * math.sqrt(1 - rho) * (1 + rho / 2) * GEV2PB
    print("sigma", res, "+-", err, "closed", closed, flush=True)
    assert abs(res - closed) < 3 * err
    assert err / res < 2e-3


def _toy_quark_pdf(x, q2, pid):
    """x f(x) of a toy valence-like quark (pid > 0) and a softer antiquark (pid < 0),
    flavour-independent; Q^2 is ignored."""
    import torch as _t
    if pid > 0:
        return 2.0 * _t.sqrt(x) * (1 - x) ** 3
    return 0.5 * (1 - x) ** 7


@pytest.mark.gpu
def test_hadronic_uux_ttx_cross_section(dev):
    """Hadronic mode with a quark-initiated process: the luminosity sums the four flavour
    pairs (u u~, c c~, d d~, s s~) and their mirrors (madflow_exec.py:141-155,440-454),
    so sigma = 8 int dx1 dx2 q(x1) qbar(x2) sigma_hat(x1 x2 s) for the toy densities
    (q(x) qbar(y) + qbar(x) q(y) summed over 4 pairs, the two orders equal after x1 <-> x2
    in the symmetric integral); VEGAS must agree with the direct 2D integral within 3 sigma."""
    import math
    from scipy import integrate
    from madflow_amd.vegas import MEIntegrand, VegasFlow, GEV2PB
    print("case hadronic sigma uux_ttx", flush=True)
    ig = MEIntegrand("uux_ttx", precision=64, mapping="hadronic", seed=3, sqrts=13000.0, ref_compat=False,
                     width=0.0, pdf=_toy_quark_pdf, scale=173.0)
    assert ig.flavours == ([2, 4, 1, 3, -2, -4, -1, -3], [-2, -4, -1, -3, 2, 4, 1, 3])
    vf = VegasFlow(ig.ndim, 1000000)
    vf.compile(ig)
    vf.run_integration(3)
    vf.freeze_grid()
    res, err = vf.run_integration(5)
    m, a, s = 173.0, 0.118, 13000.0 ** 2

    def sig_hat(sh):
        rho = 4 * m * m / sh
        return 8 * math.pi * a * a / (27 * sh) * math.sqrt(max(1 - rho, 0.0)) * (1 + rho / 2) * GEV2PB

    t = 4 * m * m / s
    q = lambda x: 2.0 * math.sqrt(x) * (1 - x) ** 3 / x
    qb = lambda x: 0.5 * (1 - x) ** 7 / x
    direct, _ = integrate.dblquad(
        lambda u2, u1: (q(math.exp(u1)) * qb(math.exp(u2)) + qb(math.exp(u1)) * q(math.exp(u2)))
        * math.exp(u1 + u2) * sig_hat(math.exp(u1 + u2) * s),
        math.log(t), 0.0, lambda u1: math.log(t) - u1, lambda u1: 0.0, epsabs=0, epsrel=1e-7)
    direct *= 4
    print("hadronic sigma", res, "+-", err, "direct", direct, flush=True)
    assert abs(res - direct) < 3 * err + 1e-6 * direct
    assert err / res < 1e-2
