"""Pins the oracle's g g > t t~ smatrix: against the committed reference module (restated by
hand: python_package/madflow/tests/mockup_debug_me.py:419-541) and against the textbook
LO matrix element (Ellis-Stirling-Webber) and its integrated cross section."""

import math

import numpy as np

from oracle import oracle as O
from tests.helpers import minkowski, phase_space_points


def test_generated_program_equals_reference_module():
    """The generated g g > t t~ program is the mockup's call sequence: bitwise equal
    |M|^2 and jamps (reference constants, reference gs literal)."""
    ps, _ = phase_space_points("gg_ttx", 200, seed=3)
    ref, jref = O.smatrix_gg_ttx_mockup(ps, ref_compat=True)
    rp, coup = O.default_params(ref_gs=True)
    gen = O.smatrix("gg_ttx", ps, rp, coup, ref_compat=True)
    assert np.array_equal(gen, ref)
    _, jgen = O.matrix_hel("gg_ttx", ps, [-1, -1, -1, 1], rp, coup, ref_compat=True, want_jamp=True)
    assert np.array_equal(jgen, jref)


def esw(ps, alpha_s=0.118, m=173.0):
    """sum-averaged |M|^2 for g g -> t t~ (Gamma_t = 0):
    g^4 (1/(6 t1 t2) - 3/8)(t1^2 + t2^2 + rho - rho^2/(4 t1 t2))."""
    g2 = 4 * math.pi * alpha_s
    s = minkowski(ps[:, 0] + ps[:, 1], ps[:, 0] + ps[:, 1])
    t1 = 2 * minkowski(ps[:, 0], ps[:, 2]) / s
    t2 = 2 * minkowski(ps[:, 1], ps[:, 2]) / s
    rho = 4 * m * m / s
    return g2 ** 2 * (1 / (6 * t1 * t2) - 3 / 8) * (t1 ** 2 + t2 ** 2 + rho - rho ** 2 / (4 * t1 * t2))


def test_textbook_matrix_element():
    ps, _ = phase_space_points("gg_ttx", 500, seed=4)
    rp, coup = O.default_params(wt=0.0)
    me = O.smatrix("gg_ttx", ps, rp, coup, ref_compat=False)
    np.testing.assert_allclose(me, esw(ps), rtol=1e-11)


def test_textbook_cross_section():
    """sigma(gg->tt~) at sqrt(s)=1 TeV: pi alpha_s^2/(3s)[(1+rho+rho^2/16) ln((1+b)/(1-b)) - b(7/4+31 rho/16)]
    = 11.3683 pb; integrate the oracle |M|^2 over the exact 2-body phase space (Gauss-Legendre in cos theta)."""
    sqrts, m, a_s = 1000.0, 173.0, 0.118
    s = sqrts ** 2
    rho = 4 * m * m / s
    b = math.sqrt(1 - rho)
    closed = math.pi * a_s ** 2 / (3 * s) * ((1 + rho + rho * rho / 16) * math.log((1 + b) / (1 - b))
                                             - b * (7 / 4 + 31 * rho / 16))
    xg, wg = np.polynomial.legendre.leggauss(64)
    pabs = b * sqrts / 2
    ps = np.zeros((64, 4, 4))
    ps[:, 0] = [sqrts / 2, 0, 0, sqrts / 2]
    ps[:, 1] = [sqrts / 2, 0, 0, -sqrts / 2]
    st = np.sqrt(1 - xg ** 2)
    ps[:, 2, 0] = sqrts / 2
    ps[:, 2, 1] = pabs * st
    ps[:, 2, 3] = pabs * xg
    ps[:, 3, 0] = sqrts / 2
    ps[:, 3, 1] = -pabs * st
    ps[:, 3, 3] = -pabs * xg
    rp, coup = O.default_params(wt=0.0)
    me = O.smatrix("gg_ttx", ps, rp, coup)
    # dPhi2 = b/(8 pi) dcos/2 ; sigma = 1/(2s) int |M|^2 dPhi2
    sigma = np.sum(wg * me) * b / (16 * math.pi) / (2 * s)
    assert abs(sigma / closed - 1) < 1e-10
    assert abs(closed * 389379365.6 - 11.3683) < 1e-4


def test_flop_counts_are_stable():
    """The algorithmic flop counts used for the roofline are reproducible."""
    f = O.flops_per_me("gg_ttx")
    assert f["total"] == O.flops_per_me("gg_ttx")["total"]
    assert f["total"] > 1e4
