"""Pins the oracle's VEGAS pieces (vegasflow is not in the reference tree and cannot be
imported: parity with it is unpinned; the published algorithm is checked instead):
map jacobians, refinement invariants, and adaptive integration of closed-form integrals.
RAMBO on diet is checked against the closed-form massless volume and against RAMBO by
equal integrals."""

import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import minkowski


def test_map_uniform_grid_is_identity():
    edges = np.tile(np.linspace(0, 1, 51), (3, 1))
    y = np.random.default_rng(0).random((100, 3))
    x, jac, bins = O.vegas_map(edges, y)
    np.testing.assert_allclose(x, y, rtol=1e-14)
    np.testing.assert_allclose(jac, 1.0, rtol=1e-13)
    assert np.array_equal(bins, np.minimum((y * 50).astype(int), 49))


def test_map_jacobian_integrates_to_one():
    rng = np.random.default_rng(1)
    g = np.sort(rng.random((2, 49)), axis=1)
    edges = np.concatenate([np.zeros((2, 1)), g, np.ones((2, 1))], axis=1)
    # the map is piecewise linear: E_y[jac * 1] = volume of the x cube = 1 exactly per bin
    y = (np.arange(50) + 0.5)[:, None] / 50 * np.ones((1, 2))
    _, jac, _ = O.vegas_map(edges, y)
    per_bin = (edges[:, 1:] - edges[:, :-1]) * 50
    np.testing.assert_allclose(jac, per_bin[0] * per_bin[1], rtol=1e-13)


def test_refine_flat_histogram_keeps_grid():
    edges = np.tile(np.linspace(0, 1, 51), (2, 1))
    out = O.vegas_refine(edges, np.ones((2, 50)))
    np.testing.assert_allclose(out, edges, atol=1e-14)


def test_refine_monotone_and_concentrates():
    edges = np.tile(np.linspace(0, 1, 51), (1, 1))
    hist = np.exp(-((np.arange(50) + 0.5) / 50 - 0.3) ** 2 / 0.001)[None]
    out = O.vegas_refine(edges, hist)
    assert np.all(np.diff(out[0]) > 0) and out[0, 0] == 0 and out[0, -1] == 1
    w = np.diff(out[0])
    assert w[np.argmin(np.abs((out[0, :-1] + w / 2) - 0.3))] < 0.02 < w.max()


def test_refine_hand_worked_smoothing():
    """Lepage's refinement worked by hand on 5 uniform bins, histogram (0, 0, 1, 0, 0).

    Smoothing (ends: mean of 2, inside: mean of 3) gives (0, 1/3, 1/3, 1/3, 0), so
    r = (0, 1/3, 1/3, 1/3, 0): bins 1..3 get the same damped importance m and the end bins
    none (r = 0).  Each new bin holds 3m/5: the new edges fall 3/5, 6/5, 9/5 and 12/5 of
    an old bin (width 0.2) after x = 0.2, i.e. at 0.32, 0.44, 0.56, 0.68, whatever the
    damping exponent.  Without the smoothing, r = (0, 0, 1, 0, 0) has no bin with
    0 < r < 1, every importance is zero and the grid would stay uniform."""
    edges = np.linspace(0.0, 1.0, 6)[None]
    for alpha in (1.5, 0.7):
        out = O.vegas_refine(edges, np.array([[0.0, 0.0, 1.0, 0.0, 0.0]]), alpha=alpha)
        np.testing.assert_allclose(out[0], [0.0, 0.32, 0.44, 0.56, 0.68, 1.0], rtol=0, atol=1e-15)


def test_refine_hand_worked_damping():
    """Lepage's refinement worked by hand on 3 uniform bins, histogram (0, 0, 6), alpha 1.5.

    Smoothed (0, 2, 3), r = (0, 0.4, 0.6).  Damped importances m = ((1 - r) / ln(1/r))^1.5:
    m0 = 0, m1 = (0.6 / ln 2.5)^1.5 = 0.529879..., m2 = (0.4 / ln(5/3))^1.5 = 0.692917...
    Each new bin holds A = (m1 + m2) / 3.  Edge 1 lies in old bin 1 at the fraction A / m1;
    edge 2 in old bin 2 at the fraction (2 A - m1) / m2:
        x1 = 1/3 + (1/3) (1 + m2/m1) / 3          = 0.5897430785...
        x2 = 2/3 + (1/3) (2 (m1 + m2) / 3 - m1) / m2 = 0.8039212697...
    Without the damping (m = r) the edges would be 0.6111 and 0.8148; without the smoothing
    r = (0, 0, 1) damps to zero everywhere and the grid would not move."""
    m1 = (0.6 / math.log(2.5)) ** 1.5
    m2 = (0.4 / math.log(5.0 / 3.0)) ** 1.5
    x1 = 1 / 3 + (1 + m2 / m1) / 9
    x2 = 2 / 3 + (2 * (m1 + m2) / 3 - m1) / m2 / 3
    assert abs(x1 - 0.5897430785377977) < 1e-15 and abs(x2 - 0.803921269685988) < 1e-15
    out = O.vegas_refine(np.array([[0.0, 1 / 3, 2 / 3, 1.0]]), np.array([[0.0, 0.0, 6.0]]), alpha=1.5)
    np.testing.assert_allclose(out[0], [0.0, x1, x2, 1.0], rtol=0, atol=1e-15)


def _vegas(f, ndim, n, iters, seed=3):
    edges = np.tile(np.linspace(0, 1, 51), (ndim, 1))
    res = []
    for it in range(iters):
        y = O.philox_uniforms(n, ndim, seed, it)
        x, jac, bins = O.vegas_map(edges, y)
        v = f(x) * jac
        res.append((v.mean(), v.var(ddof=1) / n))
        hist = np.stack([np.bincount(bins[:, d], weights=v * v, minlength=50) for d in range(ndim)])
        edges = O.vegas_refine(edges, hist)
    w = np.array([1 / r[1] for r in res[2:]])
    i = np.sum([r[0] / r[1] for r in res[2:]]) / w.sum()
    return i, math.sqrt(1 / w.sum()), res


def test_gaussian_integral_and_variance_reduction():
    """Narrow separable Gaussian: closed-form integral; adaptation cuts the variance."""
    s = 0.05
    norm = (math.erf(0.5 / (s * math.sqrt(2))) * s * math.sqrt(2 * math.pi)) ** 4

    def f(x):
        return np.exp(-np.sum((x - 0.5) ** 2, axis=1) / (2 * s * s))

    i, e, res = _vegas(f, 4, 50000, 8)
    assert abs(i - norm) < 4 * e
    assert res[-1][1] < 0.05 * res[0][1]


@pytest.mark.parametrize("n", [3, 4, 5])
def test_diet_volume_and_equivalence(n):
    """RAMBO on diet: flat weight equal to the massless volume, momentum conservation and
    massless legs; the integral of a smooth observable agrees with RAMBO's."""
    sqrts = 1000.0
    cfg = dict(n_out=n, ndim=3 * n - 4, nbins=50, mapping=1, sqrts=sqrts, masses=[0.0] * n, gev2pb=1.0,
               pt_min=0.0, abs_eta_max=0.0, dr_min=0.0, cut_first_leg=n + 2, seed=4)
    edges = np.tile(np.linspace(0, 1, 51), (cfg["ndim"], 1))
    ps, wps, flag = O.generate(cfg, edges, 0, 20000, 0)
    assert flag.all()
    vol = sqrts ** (2 * n - 4) / 2 / math.factorial(n - 1) / math.factorial(n - 2) / (4 * math.pi) ** (2 * n - 3)
    np.testing.assert_allclose(wps * 2 * sqrts ** 2, vol, rtol=1e-12)
    out = ps[:, 2:]
    np.testing.assert_allclose(out.sum(axis=1), np.broadcast_to([sqrts, 0, 0, 0], (20000, 4)), atol=1e-9 * sqrts)
    np.testing.assert_allclose(minkowski(out, out), 0.0, atol=1e-7 * sqrts ** 2)
    # observable: mean of (p1.p2)^2 / s^2 over phase space, RAMBO vs diet
    obs_d = (minkowski(out[:, 0], out[:, 1]) / sqrts ** 2) ** 2
    cfg_r = dict(cfg, ndim=4 * n, mapping=0)
    ps_r, _, _ = O.generate(cfg_r, np.tile(np.linspace(0, 1, 51), (4 * n, 1)), 0, 20000, 0)
    o = ps_r[:, 2:]
    obs_r = (minkowski(o[:, 0], o[:, 1]) / sqrts ** 2) ** 2
    err = math.hypot(obs_d.std() / math.sqrt(len(obs_d)), obs_r.std() / math.sqrt(len(obs_r)))
    assert abs(obs_d.mean() - obs_r.mean()) < 4 * err


def test_massive_diet_on_shell():
    cfg = dict(n_out=4, ndim=8, nbins=50, mapping=1, sqrts=3000.0, masses=[173.0, 173.0, 0.0, 0.0], gev2pb=1.0,
               pt_min=30.0, abs_eta_max=0.0, dr_min=0.4, cut_first_leg=4, seed=2)
    ps, wps, flag = O.generate(cfg, np.tile(np.linspace(0, 1, 51), (8, 1)), 0, 5000, 0)
    out = ps[:, 2:]
    np.testing.assert_allclose(minkowski(out, out), np.broadcast_to([173.0 ** 2, 173.0 ** 2, 0, 0], (5000, 4)),
                               atol=1e-6 * 3000.0 ** 2)
    pt = np.hypot(out[:, 2:, 1], out[:, 2:, 2])
    assert np.all(pt[flag == 1] > 30.0)
    assert 0.2 < flag.mean() < 1.0


def test_unweight_oracle_statistics():
    """Hit-or-miss keep mask: acceptance min(1, w/wmax) per bin of w/wmax, zero weights never kept."""
    rng = np.random.default_rng(8)
    n = 400_000
    w = rng.random(n) * 2.0
    w[:1000] = 0.0
    keep = O.unweight(w, 1.0, seed=3, iteration=2)
    assert keep[:1000].sum() == 0
    r = np.minimum(1.0, w[1000:])
    k = keep[1000:].astype(float)
    for lo in np.arange(0, 1.0, 0.1):
        sel = (r >= lo) & (r < lo + 0.1)
        p = r[sel].mean()
        assert abs(k[sel].mean() - p) < 5 * np.sqrt(p * (1 - p) / sel.sum()) + 1e-12
    assert k[r >= 1.0].all()
    # the stream is the Philox stream (ev, 0xffffffff, iteration) of the seed: shifting ev0 shifts it
    a = O.unweight(w[10:20], 1.0, seed=3, iteration=2, ev0=10)
    assert np.array_equal(a, keep[10:20])
