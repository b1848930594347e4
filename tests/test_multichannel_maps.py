"""CPU pins of the multi-channel phase-space maps (numpy restatement of the device maps,
tests/multichannel_reference.py), independent of the GPU:

* every channel's map inverts: momenta -> cube point -> the same momenta, and the weight
  dPhi/dx computed from the momenta equals the forward weight;
* momentum conservation and on-shell legs;
* every channel estimates the same phase-space volume of the cut region (the volume is
  fixed by the mathematics, so a wrong jacobian factor in any map shows up here), and so
  does the multi-channel mixture with its weights 1 / sum_j alpha_j g_j;
* a planted mistake (a missing 1/(2 pi) in the s-channel's jacobian) moves that channel's
  volume by far more than its error.
"""

import numpy as np
import pytest

from tests import multichannel_reference as R
from tests.test_gpu_multichannel import _cuts, _random_grids

SQRTS, MASSES = 3000.0, [173.0, 173.0, 0.0, 0.0]


@pytest.fixture(scope="module")
def chs():
    return R.Channels(4, MASSES, SQRTS, pt_min=30.0, dr_min=0.4)


def test_round_trip_and_weights(chs):
    rng = np.random.default_rng(1)
    for c in range(len(chs.chan)):
        x = rng.random((4000, chs.ndim))
        p, w = chs.forward(c, x)
        xi, wi, ok = chs.inverse(c, p)
        assert ok.all()
        assert np.abs(xi - x).max() < 1e-9, c
        np.testing.assert_allclose(wi, w, rtol=1e-9)
        np.testing.assert_allclose(p.sum(axis=1), np.tile([SQRTS, 0, 0, 0], (len(p), 1)), atol=1e-9 * SQRTS)
        m2 = p[..., 0] ** 2 - np.sum(p[..., 1:] ** 2, axis=-1)
        np.testing.assert_allclose(m2, np.tile(np.square(MASSES), (len(p), 1)), atol=1e-6 * SQRTS ** 2 * 1e-3)


def _volume(chs, c, n, seed, jac_fix=1.0):
    rng = np.random.default_rng(seed)
    p, w = chs.forward(c, rng.random((n, chs.ndim)))
    f = w * jac_fix * _cuts(p)
    return f.mean(), f.std() / np.sqrt(n)


def test_channels_share_the_volume(chs):
    ref, eref = _volume(chs, 0, 200000, 2)
    for c in range(1, len(chs.chan)):
        v, e = _volume(chs, c, 200000, 2 + c)
        assert abs(v - ref) < 4 * np.hypot(e, eref), (c, v, ref)


def test_mixture_weight_is_unbiased(chs):
    rng = np.random.default_rng(9)
    grids = _random_grids(len(chs.chan), chs.ndim, 20, rng)
    alpha = np.array([0.3, 0.3, 0.2, 0.2])
    ref, eref = _volume(chs, 0, 200000, 2)
    tot, var = 0.0, 0.0
    for c in range(len(chs.chan)):
        n = 60000
        p, w, r = R.mc_weight(chs, grids, alpha, c, rng.random((n, chs.ndim)))
        np.testing.assert_allclose(r @ alpha, 1.0, rtol=1e-12)
        f = w * _cuts(p)
        tot += alpha[c] * f.mean()
        var += alpha[c] ** 2 * f.var() / n
    assert abs(tot - ref) < 4 * np.hypot(np.sqrt(var), eref), (tot, ref)


def test_volume_check_catches_a_wrong_jacobian(chs):
    ref, eref = _volume(chs, 0, 200000, 2)
    v, e = _volume(chs, 3, 200000, 5, jac_fix=2 * np.pi)
    assert abs(v - ref) > 20 * np.hypot(e, eref)
