"""Oracle phase space: RAMBO against the reference's own closed-form checks
(python_package/madflow/tests/test_ps.py), on-shell / conservation invariants, the
reference's batch-global Newton quirk, and Philox4x32-10 known-answer vectors."""

import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import minkowski


def massless_volume(n, e):
    """test_ps.py:9-16: 1/2/(4 pi)^(2n-3) E^(2n-4) / (n-1)!/(n-2)!"""
    return pow(e, 2 * n - 4) / 2.0 / math.factorial(n - 1) / math.factorial(n - 2) / pow(4 * np.pi, 2 * n - 3)


@pytest.mark.parametrize("n", range(2, 8))
def test_rambo_massless_volume(n):
    """test_ps.py:19-39 (rtol 1e-6 there; exact constants here give 1e-13)."""
    sqrts = 7e3
    x = np.random.default_rng(n).random((3, 4 * n))
    p, w = O.rambo(x, n, sqrts, None)
    assert p.shape == (3, n, 4)
    np.testing.assert_allclose(w, massless_volume(n, sqrts), rtol=1e-13)
    # reference constants (float32 pi) still pass the reference's own tolerance
    _, w32 = O.rambo(x, n, sqrts, None, ref_compat=True)
    np.testing.assert_allclose(w32, massless_volume(n, sqrts), rtol=1e-6)


def test_rambo_variable_sqrts():
    """test_ps.py:36-39: per-event sqrt(s)."""
    rng = np.random.default_rng(5)
    s = rng.random(13) * 7e3
    x = rng.random((13, 28))
    _, w = O.rambo(x, 7, s, None)
    np.testing.assert_allclose(w, [massless_volume(7, e) for e in s], rtol=1e-12)


@pytest.mark.parametrize("masses", [[50.0, 125.0], [173.0, 173.0], [173.0, 173.0, 0.0], [173.0, 173.0, 0.0, 0.0],
                                    [173.0, 173.0, 0.0, 0.0, 0.0], [10.0, 20.0, 30.0]])
def test_rambo_massive_invariants(masses):
    """test_ps.py:68-78: p^2 = m^2 (there atol=rtol=1e-4); plus sum p = (sqrts, 0)."""
    n = len(masses)
    sqrts = 3000.0
    x = np.random.default_rng(n).random((500, 4 * n))
    p, w = O.rambo(x, n, sqrts, masses)
    m2 = minkowski(p, p)
    np.testing.assert_allclose(m2, np.broadcast_to(np.array(masses) ** 2, m2.shape), atol=1e-6 * sqrts ** 2)
    tot = p.sum(axis=1)
    np.testing.assert_allclose(tot[:, 0], sqrts, rtol=1e-13)
    assert np.max(np.abs(tot[:, 1:])) < 1e-9 * sqrts
    assert np.all(w > 0)


def test_massive_two_body_weight_closed_form():
    """Two-body phase space volume: beta/(8 pi) for equal masses (flat in cos theta)."""
    sqrts, m = 1000.0, 173.0
    x = np.random.default_rng(1).random((50, 8))
    _, w = O.rambo(x, 2, sqrts, [m, m])
    beta = math.sqrt(1 - 4 * m * m / sqrts ** 2)
    np.testing.assert_allclose(w, beta / (8 * np.pi), rtol=1e-12)


def test_reference_newton_quirk_only_with_ref_compat():
    """The reference stops Newton for the whole batch once one event converged
    (phasespace.py:90-103); in ref_compat mode the oracle reproduces that (events are
    left slightly off-shell), in the exact mode every event converges."""
    masses = [173.0, 173.0, 0.0, 0.0]
    x = np.random.default_rng(9).random((2000, 16))
    p_ref, _ = O.rambo(x, 4, 3000.0, masses, ref_compat=True)
    p_ex, _ = O.rambo(x, 4, 3000.0, masses, ref_compat=False)
    dev_ref = np.abs(p_ref.sum(axis=1)[:, 0] - 3000.0).max()
    dev_ex = np.abs(p_ex.sum(axis=1)[:, 0] - 3000.0).max()
    assert dev_ex < 1e-9
    assert dev_ref > 1e3 * dev_ex


def test_philox_known_answers():
    """Random123 known-answer vectors for philox4x32-10."""
    assert list(O.philox_raw([0, 0, 0, 0], [0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(O.philox_raw([0xffffffff] * 4, [0xffffffff] * 2)) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6,
                                                                      0x6d5451fd]
    assert list(O.philox_raw([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0])) == [
        0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_philox_uniforms_range_and_moments():
    u = O.philox_uniforms(200000, 3, 1, 0)
    assert u.min() > 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 3e-3
    assert abs(u.var() - 1 / 12) < 3e-3


def test_cut_variables():
    p = np.array([10.0, 3.0, 4.0, 5.0])
    pabs = math.sqrt(50)
    assert abs(O.eta(p) - 0.5 * math.log((pabs + 5) / (pabs - 5))) < 1e-14
    q = np.array([10.0, -3.0, 4.0, -5.0])
    dphi = abs(math.atan2(4, 3) - math.atan2(4, -3))
    assert abs(O.delta_r(p, q) - math.hypot(2 * O.eta(p), dphi)) < 1e-12


def _hadronic_cfg(lab_frame, shat_min=346.0 ** 2, n_out=2, masses=(173.0, 173.0)):
    return dict(n_out=n_out, ndim=4 * n_out + 2, nbins=50, mapping=2, sqrts=13000.0,
                masses=list(masses), gev2pb=389379365.6, pt_min=0.0, abs_eta_max=0.0, dr_min=0.0,
                cut_first_leg=n_out + 2, seed=5, shat_min=shat_min, lab_frame=lab_frame)


def test_hadronic_tau_jacobian_closed_form():
    """ramboflow's (x1, x2) sampling (phasespace.py:213-232): the weight (1 - tau_min)
    (-ln tau) times a function of (x1, x2) averages to its integral over x1 x2 > tau_min.
    For massless 2-body RAMBO the phase-space weight is the constant 1/(8 pi), so
    wps 2 shat 8 pi / gev2pb isolates the (x1, x2) weight; its mean must be the area
    1 - tau_min + tau_min ln tau_min, and its x1-weighted mean the first moment."""
    cfg = _hadronic_cfg(0, shat_min=0.1 * 13000.0 ** 2, masses=(0.0, 0.0))
    n = 400000
    edges = np.tile(np.linspace(0, 1, 51), (cfg["ndim"], 1))
    ps, wps, flag, x12 = O.generate(cfg, edges, 0, n, 1)
    s = cfg["sqrts"] ** 2
    shat = x12[:, 0] * x12[:, 1] * s
    wt = wps * 2 * shat * 8 * math.pi / cfg["gev2pb"]
    t = cfg["shat_min"] / s
    assert np.all(x12[:, 0] * x12[:, 1] >= t * (1 - 1e-12)) and np.all(x12 <= 1)
    area = 1 - t + t * math.log(t)
    assert abs(wt.mean() - area) < 4 * wt.std() / math.sqrt(n)
    first = 0.5 * (1 - t) ** 2                       # = int_t^1 x1 (1 - t / x1) dx1
    m1 = (wt * x12[:, 0]).mean()
    assert abs(m1 - first) < 4 * (wt * x12[:, 0]).std() / math.sqrt(n)


@pytest.mark.parametrize("n_out,masses", [(2, (173.0, 173.0)), (3, (173.0, 173.0, 0.0))])
def test_hadronic_lab_frame(n_out, masses):
    """_boost_to_lab (phasespace.py:322-357): the lab-frame partons carry x1 sqrt(s)/2 and
    x2 sqrt(s)/2 along +-z, momentum is conserved, invariant masses do not change, and the
    partonic-frame output is the same event before the boost."""
    edges = np.tile(np.linspace(0, 1, 51), (4 * n_out + 2, 1))
    com = O.generate(_hadronic_cfg(0, n_out=n_out, masses=masses), edges, 0, 2000, 3)
    lab = O.generate(_hadronic_cfg(1, n_out=n_out, masses=masses), edges, 0, 2000, 3)
    np.testing.assert_array_equal(com[3], lab[3])
    np.testing.assert_array_equal(com[1], lab[1])
    x1, x2 = lab[3][:, 0], lab[3][:, 1]
    e = 6500.0
    np.testing.assert_allclose(lab[0][:, 0, 0], x1 * e, rtol=1e-12)
    np.testing.assert_allclose(lab[0][:, 0, 3], x1 * e, rtol=1e-12)
    np.testing.assert_allclose(lab[0][:, 1, 0], x2 * e, rtol=1e-12)
    np.testing.assert_allclose(lab[0][:, 1, 3], -x2 * e, rtol=1e-12)
    p_in = lab[0][:, 0] + lab[0][:, 1]
    np.testing.assert_allclose(lab[0][:, 2:].sum(axis=1), p_in, atol=1e-8 * 13000)
    for i in range(n_out + 2):
        np.testing.assert_allclose(minkowski(lab[0][:, i], lab[0][:, i]), minkowski(com[0][:, i], com[0][:, i]),
                                   atol=1e-7 * 13000 ** 2 / 1e4)
    # partonic frame: symmetric partons of energy sqrt(shat)/2
    np.testing.assert_allclose(com[0][:, 0, 0], 0.5 * np.sqrt(x1 * x2) * 13000.0, rtol=1e-12)
