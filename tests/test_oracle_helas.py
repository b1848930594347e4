"""Pins the oracle's HELAS external wavefunctions (restated from
python_package/madflow/wavefunctions_flow.py) with closed-form properties:

* Dirac equation:  ubar(p)(pslash - m) = 0 (oxxxxx, nsf=+1),  (pslash + m) v(p) = 0 (ixxxxx, nsf=-1),
  (pslash - m) u(p) = 0 (ixxxxx, nsf=+1),  vbar(p)(pslash + m) = 0 (oxxxxx, nsf=-1)
* completeness: sum_l u ubar = pslash + m, sum_l v vbar = pslash - m
* polarisation vectors: eps.p = 0, eps.eps* = -1, transverse completeness, BRST eps = p/p0
* every branch of the reference (p at rest, p along +-z, pT = 0) stays finite and valid
"""

import numpy as np
import pytest

from oracle import oracle as O
from tests.bruteforce import METRIC, mdot, slash

M = 173.0


def momenta(mass, kind):
    rng = np.random.default_rng(17)
    out = []
    for _ in range(6):
        v = rng.normal(size=3) * 300
        out.append(v)
    out += [np.array([0, 0, 250.0]), np.array([0, 0, -250.0]), np.array([1e-3, 0, 400.0])]
    if mass > 0:
        out.append(np.zeros(3))
    ps = []
    for v in out:
        e = np.sqrt(v @ v + mass * mass)
        ps.append(np.concatenate([[e], v]))
    return np.array(ps)


def spinor(kind, p, mass, hel, nsf):
    return O.wavefunction(kind, p[None], mass, hel, nsf)[0][2:6]


@pytest.mark.parametrize("mass", [M, 0.0])
def test_dirac_equation(mass):
    for p in momenta(mass, "f"):
        ps = slash(p.astype(complex))
        one = np.eye(4)
        for hel in (-1, 1):
            ubar = spinor("o", p, mass, hel, 1)
            v = spinor("i", p, mass, hel, -1)
            u = spinor("i", p, mass, hel, 1)
            vbar = spinor("o", p, mass, hel, -1)
            scale = p[0] + mass
            assert np.abs(ubar @ (ps - mass * one)).max() < 1e-12 * scale * np.sqrt(scale)
            assert np.abs((ps + mass * one) @ v).max() < 1e-12 * scale * np.sqrt(scale)
            assert np.abs((ps - mass * one) @ u).max() < 1e-12 * scale * np.sqrt(scale)
            assert np.abs(vbar @ (ps + mass * one)).max() < 1e-12 * scale * np.sqrt(scale)


@pytest.mark.parametrize("mass", [M, 0.0])
def test_spin_sums(mass):
    """sum over helicities of u ubar = pslash + m and v vbar = pslash - m (HELAS sign conventions)."""
    for p in momenta(mass, "f")[:7]:
        ps = slash(p.astype(complex))
        su = sum(np.outer(spinor("i", p, mass, h, 1), spinor("o", p, mass, h, 1)) for h in (-1, 1))
        sv = sum(np.outer(spinor("i", p, mass, h, -1), spinor("o", p, mass, h, -1)) for h in (-1, 1))
        np.testing.assert_allclose(su, ps + mass * np.eye(4), atol=1e-10 * p[0])
        np.testing.assert_allclose(sv, ps - mass * np.eye(4), atol=1e-10 * p[0])


@pytest.mark.parametrize("nsv", [-1, 1])
def test_gluon_polarisations(nsv):
    for p in momenta(0.0, "v"):
        epss = []
        for hel in (-1, 1):
            w = O.wavefunction("v", p[None], 0.0, hel, nsv)[0]
            eps = w[2:6]
            # slots carry nsv * p
            np.testing.assert_allclose([w[0].real, w[1].real, w[1].imag, w[0].imag], nsv * p, atol=1e-12)
            assert abs(mdot(eps, p)) < 1e-12 * p[0]
            assert abs(mdot(eps, eps.conj()) + 1) < 1e-12
            assert abs(eps[0]) < 1e-15
            epss.append(eps)
        # transverse completeness: sum_l eps^i eps^j* = delta^ij - p^i p^j / |p|^2
        pv = p[1:] / np.linalg.norm(p[1:])
        s = sum(np.outer(e[1:], e[1:].conj()) for e in epss)
        np.testing.assert_allclose(s, np.eye(3) - np.outer(pv, pv), atol=1e-12)
        # the two helicities are orthogonal
        assert abs(mdot(epss[0], epss[1].conj())) < 1e-12
        # BRST: nhel = 4 gives p / p0
        w4 = O.wavefunction("v", p[None], 0.0, 4, nsv)[0]
        np.testing.assert_allclose(w4[2:6], p / p[0], atol=1e-15)


def test_massive_vector_polarisations():
    """Massive vectors (not used by the QCD processes) keep eps.p = 0, |eps|^2 = -1 away from rest."""
    mv = 91.0
    for p in momenta(mv, "v")[:9]:
        for hel in (-1, 0, 1):
            eps = O.wavefunction("v", p[None], mv, hel, 1)[0][2:6]
            assert abs(mdot(eps, p)) < 1e-10 * p[0]
            assert abs(mdot(eps, eps.conj()) + 1) < 1e-10


def test_sqh_reference_rounding():
    """ref_compat uses float_me(sqrt(0.5)) = 0.70710676908493042 (wavefunctions_flow.py:10)."""
    p = np.array([[100.0, 30.0, 40.0, 50.0]])
    a = O.wavefunction("v", p, 0.0, 1, 1, ref_compat=False)[0]
    b = O.wavefunction("v", p, 0.0, 1, 1, ref_compat=True)[0]
    ratio = np.abs(b[3:6]) / np.abs(a[3:6])
    np.testing.assert_allclose(ratio, 0.70710676908493042 / np.sqrt(0.5), rtol=1e-15)


@pytest.mark.parametrize("nss", [1, -1])
def test_scalar_wavefunction_layout(nss):
    """sxxxxx: the momentum slots follow the same HELAS convention as the vector and fermion
    wavefunctions of the same particle direction (slots [0], [1] = sign * (E + i pz, px + i py)),
    and the scalar component is 1."""
    p = momenta(125.0, "s")
    s = O.wavefunction("s", p, 125.0, 0, nss)
    v = O.wavefunction("v", p, 125.0, 1, nss)
    assert s.shape == (len(p), 3)
    np.testing.assert_array_equal(s[:, :2], v[:, :2])
    np.testing.assert_array_equal(s[:, 2], np.ones(len(p)))
    # the slots give back the momentum: P = (Re s0, Re s1, Im s1, Im s0) * sign
    P = nss * np.stack([s[:, 0].real, s[:, 1].real, s[:, 1].imag, s[:, 0].imag], axis=1)
    np.testing.assert_array_equal(P, p)
