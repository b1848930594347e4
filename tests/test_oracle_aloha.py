"""Pins the oracle's ALOHA routines against explicit Dirac-matrix / metric contractions.

Conventions fixed here (slots [0],[1] carry the outgoing momentum s of a line):
  FFV1_0(F1,F2,V)  = C (-i) F2 . Vslash . F1      (F2 = row spinor, F1 = column spinor)
  FFV1_1(F2,V)     = i C F2 . Vslash . (sslash + M) / (s^2 - M^2 + i M W)
  FFV1_2(F1,V)     = i C (-sslash + M) . Vslash . F1 / (s^2 - M^2 + i M W)
  FFV1P0_3(F1,F2)  = -i C (F2 gamma^mu F1) / (s^2 - M^2 + i M W)
  VVV1P0_1, VVVV*P0_1 = -i C L^mu / (s^2 - M^2 + i M W)
  VVV1_0, VVVV*_0  = -i C L
with the SM UFO Lorentz structures L.  The reference's own FFV1_0/FFV1_1/FFV1_2/VVV1P0_1
(mockup_debug_me.py:34-397) must equal these; the routines the reference does not ship
are also checked through diagram identities (same amplitude computed two ways).
"""

import numpy as np
import pytest

from oracle import oracle as O
from tests.bruteforce import G, METRIC, mdot, slash

rng = np.random.default_rng(2024)
C = 0.7 + 0.3j
M, W = 173.0, 1.5


def rw(n=3):
    return rng.normal(size=(n, 6)) + 1j * rng.normal(size=(n, 6))


def slot(w):
    return np.array([w[0].real, w[1].real, w[1].imag, w[0].imag])


def den(s, m=M, w=W):
    return mdot(s, s) - m * m + 1j * m * w


def vvv1(e1, e2, e3, p1, p2, p3):
    """UFO VVV1 with all three vectors contracted."""
    return (mdot(e1, e2) * (mdot(p1, e3) - mdot(p2, e3)) + mdot(e1, e3) * (mdot(p3, e2) - mdot(p1, e2))
            + mdot(e2, e3) * (mdot(p2, e1) - mdot(p3, e1)))


def test_ffv1_0():
    F1, F2, V = rw(), rw(), rw()
    got = O.aloha("FFV1_0", [F1, F2, V], C)
    for e in range(3):
        ref = C * (-1j) * (F2[e, 2:] @ slash(V[e, 2:]) @ F1[e, 2:])
        assert abs(got[e] - ref) < 1e-12 * abs(ref)


def test_ffv1_1_and_2():
    F, V = rw(), rw()
    o1 = O.aloha("FFV1_1", [F, V], C, M, W)
    o2 = O.aloha("FFV1_2", [F, V], C, M, W)
    for e in range(3):
        s = slot(F[e]) + slot(V[e])
        r1 = 1j * C * (F[e, 2:] @ slash(V[e, 2:]) @ (slash(s) + M * np.eye(4))) / den(s)
        r2 = 1j * C * ((-slash(s) + M * np.eye(4)) @ slash(V[e, 2:]) @ F[e, 2:]) / den(s)
        assert np.allclose(o1[e, 2:], r1, rtol=1e-12, atol=1e-14 * np.abs(r1).max())
        assert np.allclose(o2[e, 2:], r2, rtol=1e-12, atol=1e-14 * np.abs(r2).max())
        assert np.allclose(o1[e, :2], F[e, :2] + V[e, :2])
        assert np.allclose(o2[e, :2], F[e, :2] + V[e, :2])


def test_ffv1p0_3():
    F1, F2 = rw(), rw()
    out = O.aloha("FFV1P0_3", [F1, F2], C, 0.0, 0.0)
    for e in range(3):
        s = slot(F1[e]) + slot(F2[e])
        J = np.array([F2[e, 2:] @ G[m] @ F1[e, 2:] for m in range(4)])
        ref = -1j * C * J / mdot(s, s)
        assert np.allclose(out[e, 2:], ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_vvv1p0_1_and_vvv1_0():
    V2, V3, V1 = rw(), rw(), rw()
    out = O.aloha("VVV1P0_1", [V2, V3], C, 0.0, 0.0)
    amp = O.aloha("VVV1_0", [V1, V2, V3], C)
    for e in range(3):
        p2, p3 = slot(V2[e]), slot(V3[e])
        p1 = -(p2 + p3)
        e2, e3 = V2[e, 2:], V3[e, 2:]
        # L^mu with leg 1 open: differentiate the contracted structure
        L = np.array([vvv1(np.eye(4)[m] * METRIC[m, m], e2, e3, p1, p2, p3) for m in range(4)])
        ref = -1j * C * L / mdot(p1, p1)
        assert np.allclose(out[e, 2:], ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())
        ra = -1j * C * vvv1(V1[e, 2:], e2, e3, slot(V1[e]), p2, p3)
        assert abs(amp[e] - ra) < 1e-11 * abs(ra)


@pytest.mark.parametrize("which,pairs", [(1, ((0, 3), (1, 2), (0, 2), (1, 3))),
                                         (3, ((0, 3), (1, 2), (0, 1), (2, 3))),
                                         (4, ((0, 2), (1, 3), (0, 1), (2, 3)))])
def test_vvvv(which, pairs):
    """VVVV1 = g14 g23 - g13 g24, VVVV3 = g14 g23 - g12 g34, VVVV4 = g13 g24 - g12 g34."""
    V1, V2, V3, V4 = rw(), rw(), rw(), rw()
    amp = O.aloha("VVVV%d_0" % which, [V1, V2, V3, V4], C)
    cur = O.aloha("VVVV%dP0_1" % which, [V2, V3, V4], C, 0.0, 0.0)
    for e in range(3):
        ws = [V1[e, 2:], V2[e, 2:], V3[e, 2:], V4[e, 2:]]
        (a, b), (c, d), (f, g), (h, k) = pairs
        L = mdot(ws[a], ws[b]) * mdot(ws[c], ws[d]) - mdot(ws[f], ws[g]) * mdot(ws[h], ws[k])
        assert abs(amp[e] - (-1j) * C * L) < 1e-11 * abs(L)
        # off-shell: replace leg 1 by a unit vector e_mu with lowered index
        s1 = -(slot(V2[e]) + slot(V3[e]) + slot(V4[e]))
        Lmu = []
        for m in range(4):
            ws[0] = np.eye(4)[m] * METRIC[m, m]
            Lmu.append(mdot(ws[a], ws[b]) * mdot(ws[c], ws[d]) - mdot(ws[f], ws[g]) * mdot(ws[h], ws[k]))
        ref = -1j * C * np.array(Lmu) / mdot(s1, s1)
        assert np.allclose(cur[e, 2:], ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())


def test_diagram_identities():
    """The same tree amplitude closed at different vertices must agree:
    FFV1_0(F1,F2,VVV1P0_1(A,B)) = VVV1_0(FFV1P0_3(F1,F2),A,B) and the same with VVVV."""
    F1, F2, A, B, Cv = rw(), rw(), rw(), rw(), rw()
    F2[:, :2] = -(F1[:, :2] + A[:, :2] + B[:, :2])      # momentum conservation
    x = O.aloha("FFV1_0", [F1, F2, O.aloha("VVV1P0_1", [A, B], -1.2, 0.0, 0.0)], 1.2j)
    y = O.aloha("VVV1_0", [O.aloha("FFV1P0_3", [F1, F2], 1.2j, 0.0, 0.0), A, B], -1.2)
    assert np.allclose(x, y, rtol=1e-11)
    F2[:, :2] = -(F1[:, :2] + A[:, :2] + B[:, :2] + Cv[:, :2])
    for k in (1, 3, 4):
        x = O.aloha("FFV1_0", [F1, F2, O.aloha("VVVV%dP0_1" % k, [A, B, Cv], 1.44j, 0.0, 0.0)], 1.2j)
        y = O.aloha("VVVV%d_0" % k, [O.aloha("FFV1P0_3", [F1, F2], 1.2j, 0.0, 0.0), A, B, Cv], 1.44j)
        assert np.allclose(x, y, rtol=1e-11)
