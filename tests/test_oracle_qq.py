"""Pins the oracle's q q~ > t t~ (+ g) processes (two fermion lines), for which the
reference ships no vectors:

* u u~ > t t~ against the textbook spin- and colour-averaged |M|^2 (Ellis, Stirling,
  Webber, QCD and Collider Physics, table 10.1):
  g^4 (4/9) (tau1^2 + tau2^2 + rho/2), tau_i = 2 p_i.p_t / s, rho = 4 m_t^2 / s (zero width);
* every helicity row of u u~ > t t~ and u u~ > t t~ g against the explicit-colour
  brute force (tests/bruteforce.py: each Feynman diagram written out with Dirac and
  Gell-Mann matrices, no colour basis, no jamps);
* gauge (BRST) invariance of the radiated gluon, Lorentz invariance, the helicity-sum
  denominator 4 x 9 x k!, and the rows that vanish by chirality on the massless line.
"""

import itertools
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import bruteforce as BF
from tests.helpers import phase_space_points, minkowski
from tests.test_oracle_multijet import _boost


def test_uux_ttx_closed_form():
    ps, _ = phase_space_points("uux_ttx", 64, seed=41)
    rp, coup = O.default_params(wt=0.0)
    me = O.smatrix("uux_ttx", ps, rp, coup)
    s = minkowski(ps[:, 0] + ps[:, 1], ps[:, 0] + ps[:, 1])
    t1 = 2 * minkowski(ps[:, 0], ps[:, 2]) / s
    t2 = 2 * minkowski(ps[:, 1], ps[:, 2]) / s
    rho = 4 * 173.0 ** 2 / s
    gs = 2 * np.sqrt(np.pi * 0.118)
    np.testing.assert_allclose(me, gs ** 4 * 4 / 9 * (t1 ** 2 + t2 ** 2 + rho / 2), rtol=1e-13)


@pytest.mark.parametrize("name", ["uux_ttx", "uux_ttxg"])
def test_qq_bruteforce_every_row(name):
    ps, _ = phase_space_points(name, 2, seed=42)
    d = O.load_process(name)
    rp, coup = O.default_params()
    for e in range(2):
        for hel in d["helicities"]:
            bf = BF.qq_matrix_hel(ps[e], hel)
            orc = O.matrix_hel(name, ps[e:e + 1], hel, rp, coup)[0]
            if bf == 0.0:
                assert orc == 0.0, (hel, orc)
            else:
                assert abs(orc / bf - 1) < 1e-12, (hel, orc, bf)


def test_qq_vanishing_rows():
    """Exactly the rows whose incoming quark and antiquark carry the same helicity label
    vanish (chirality on the massless line); the generated kernels skip those rows."""
    from madflow_amd.procgen.emit import zero_helicity_pairs
    for name in ("uux_ttx", "uux_ttxg"):
        d = O.load_process(name)
        assert zero_helicity_pairs(d) == [(0, 1)]
        ps, _ = phase_space_points(name, 4, seed=43)
        rp, coup = O.default_params()
        for hel in d["helicities"]:
            v = O.matrix_hel(name, ps, hel, rp, coup)
            if hel[0] == hel[1]:
                assert np.all(v == 0.0), hel
            else:
                assert np.all(v > 0.0), hel
    assert zero_helicity_pairs(O.load_process("gg_ttxg")) == []


def test_qq_brst_gauge_invariance():
    ps, _ = phase_space_points("uux_ttxg", 3, seed=44)
    d = O.load_process("uux_ttxg")
    rp, coup = O.default_params(wt=0.0)
    for row in (0, 9, 22):
        hel = list(d["helicities"][row])
        if hel[0] == hel[1]:
            continue
        _, jp = O.matrix_hel("uux_ttxg", ps, hel, rp, coup, want_jamp=True)
        h = list(hel)
        h[4] = 4
        _, jb = O.matrix_hel("uux_ttxg", ps, h, rp, coup, want_jamp=True)
        assert np.abs(jp).max() > 0
        assert np.abs(jb).max() < 1e-11 * np.abs(jp).max(), row


@pytest.mark.parametrize("name", ["uux_ttx", "uux_ttxg"])
def test_qq_lorentz_invariance(name):
    ps, _ = phase_space_points(name, 4, seed=45)
    rp, coup = O.default_params(wt=0.0)
    a = O.smatrix(name, ps, rp, coup)
    b = O.smatrix(name, _boost(ps, [0.2, 0.1, -0.3]), rp, coup)
    np.testing.assert_allclose(a, b, rtol=1e-10)


@pytest.mark.parametrize("name,k", [("uux_ttx", 0), ("uux_ttxg", 1)])
def test_qq_helicity_sum(name, k):
    """smatrix = sum over all 2^nexternal rows / (4 spins x 9 colours x k!)."""
    ps, _ = phase_space_points(name, 3, seed=46)
    d = O.load_process(name)
    n = d["nexternal"]
    rows = [list(h) for h in itertools.product([-1, 1], repeat=n)]
    assert sorted(map(tuple, d["helicities"])) == sorted(map(tuple, rows))
    assert d["denominator"] == 4 * 9 * math.factorial(k)
    rp, coup = O.default_params()
    tot = sum(O.matrix_hel(name, ps, h, rp, coup) for h in rows) / d["denominator"]
    np.testing.assert_allclose(O.smatrix(name, ps, rp, coup), tot, rtol=1e-13)


def test_qq_bruteforce_catches_a_dropped_diagram():
    """The brute force is sensitive to each diagram: removing the three-gluon diagram's
    amplitude from the oracle's jamps changes |M|^2 by far more than the 1e-12 bar."""
    import copy
    d = copy.deepcopy(O.load_process("uux_ttxg"))
    ggg = [op["out"] for op in d["ops"] if op["kind"] == "amp"
           and any(o2["out"] == op["args"][2] and o2["routine"] == "VVV1P0_1" for o2 in d["ops"])]
    assert len(ggg) == 1
    d["jamp"] = [[t for t in terms if t[0] != ggg[0]] for terms in d["jamp"]]
    enc = O.EncodedProcess(d)
    O._cache["uux_ttxg_mutant"] = enc
    try:
        ps, _ = phase_space_points("uux_ttxg", 1, seed=47)
        rp, coup = O.default_params()
        hel = d["helicities"][0]
        bf = BF.qq_matrix_hel(ps[0], hel)
        mut = O.matrix_hel("uux_ttxg_mutant", ps, hel, rp, coup)[0]
        assert abs(mut / bf - 1) > 1e-3
    finally:
        del O._cache["uux_ttxg_mutant"]
