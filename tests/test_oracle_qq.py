"""Pins the oracle's q q~ > t t~ (+ g, + g g) processes (two fermion lines), for which the
reference ships no vectors:

* u u~ > t t~ against the textbook spin- and colour-averaged |M|^2 (Ellis, Stirling,
  Webber, QCD and Collider Physics, table 10.1):
  g^4 (4/9) (tau1^2 + tau2^2 + rho/2), tau_i = 2 p_i.p_t / s, rho = 4 m_t^2 / s (zero width);
* every helicity row of u u~ > t t~ and u u~ > t t~ g against the explicit-colour
  brute force (tests/bruteforce.py: each Feynman diagram written out with Dirac and
  Gell-Mann matrices, no colour basis, no jamps);
* gauge (BRST) invariance of every radiated gluon, Lorentz invariance, the helicity-sum
  denominator 4 x 9 x k!, and the rows that vanish by chirality on the massless line;
* u u~ > t t~ g g: Bose symmetry of the two gluons, and its soft-gluon limit against the
  eikonal factor applied to the explicit-colour brute force of u u~ > t t~ g (colour
  operators on every leg, colour conservation checked), which fixes its normalisation,
  colour factors and the diagrams that emit the soft gluon from external legs; BRST fixes
  the rest;
* u g > t t~ u and u~ g > t t~ u~ against their brute force row by row, and against each
  other by charge conjugation (the top and antitop momenta exchanged).
"""

import itertools
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import bruteforce as BF
from tests.helpers import phase_space_points, minkowski
from tests.test_oracle_multijet import _boost


QG = ("ug_ttxu", "uxg_ttxux")


def _vanishes(name, hel):
    """A helicity row that vanishes by chirality: the (anti)quark that scatters flips its
    label, or the annihilating quark and antiquark share theirs."""
    return (hel[0] != hel[4]) if name in QG else (hel[0] == hel[1])


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


@pytest.mark.parametrize("name", ["uux_ttx", "uux_ttxg", "ug_ttxu", "uxg_ttxux"])
def test_qq_bruteforce_every_row(name):
    """u g > t t~ u (the crossing: incoming gluon, outgoing quark) and u~ g > t t~ u~
    against their own brute force, written out with the crossed spinors."""
    ps, _ = phase_space_points(name, 2, seed=42)
    d = O.load_process(name)
    rp, coup = O.default_params()
    brute = {"ug_ttxu": BF.qg_matrix_hel, "uxg_ttxux": lambda p, h: BF.qg_matrix_hel(p, h, anti=True)}.get(
        name, BF.qq_matrix_hel)
    for e in range(2):
        for hel in d["helicities"]:
            bf = brute(ps[e], hel)
            orc = O.matrix_hel(name, ps[e:e + 1], hel, rp, coup)[0]
            if bf == 0.0:
                assert orc == 0.0, (hel, orc)
            else:
                assert abs(orc / bf - 1) < 1e-12, (hel, orc, bf)


def test_qq_vanishing_rows():
    """Exactly the rows whose incoming quark and antiquark carry the same helicity label
    vanish (chirality on the massless line); the generated kernels skip those rows."""
    from madflow_amd.procgen.emit import zero_helicity_pairs
    for name in ("uux_ttx", "uux_ttxg", "uux_ttxgg"):
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


@pytest.mark.parametrize("name", QG)
def test_qg_vanishing_rows(name):
    """u g > t t~ u and u~ g > t t~ u~: the (anti)quark scatters, so the rows in which the
    incoming and outgoing one carry different helicity labels vanish, and only those."""
    from madflow_amd.procgen.emit import same_helicity_pairs, zero_helicity_pairs
    d = O.load_process(name)
    assert same_helicity_pairs(d) == [(0, 4)] and zero_helicity_pairs(d) == []
    ps, _ = phase_space_points(name, 4, seed=43)
    rp, coup = O.default_params()
    for hel in d["helicities"]:
        v = O.matrix_hel(name, ps, hel, rp, coup)
        if hel[0] != hel[4]:
            assert np.all(v == 0.0), hel
        else:
            assert np.all(v > 0.0), hel


@pytest.mark.parametrize("name,gluon", [("uux_ttxg", 4), ("uux_ttxgg", 4), ("uux_ttxgg", 5), ("ug_ttxu", 1),
                                        ("uxg_ttxux", 1)])
def test_qq_brst_gauge_invariance(name, gluon):
    """Replacing a gluon's polarisation by its momentum (nhel = 4) makes every jamp vanish:
    a diagram missing from the gauge-invariant set, or one with a wrong relative sign,
    breaks this."""
    ps, _ = phase_space_points(name, 3, seed=44)
    d = O.load_process(name)
    rp, coup = O.default_params(wt=0.0)
    for row in (0, 9, 22, 37, 50):
        if row >= len(d["helicities"]):
            continue
        hel = list(d["helicities"][row])
        if _vanishes(name, hel):
            continue                                  # a row that vanishes by chirality
        _, jp = O.matrix_hel(name, ps, hel, rp, coup, want_jamp=True)
        h = list(hel)
        h[gluon] = 4
        _, jb = O.matrix_hel(name, ps, h, rp, coup, want_jamp=True)
        assert np.abs(jp).max() > 0
        assert np.abs(jb).max() < 1e-11 * np.abs(jp).max(), row


@pytest.mark.parametrize("name", ["uux_ttx", "uux_ttxg", "uux_ttxgg", "ug_ttxu", "uxg_ttxux"])
def test_qq_lorentz_invariance(name):
    ps, _ = phase_space_points(name, 4, seed=45)
    rp, coup = O.default_params(wt=0.0)
    a = O.smatrix(name, ps, rp, coup)
    b = O.smatrix(name, _boost(ps, [0.2, 0.1, -0.3]), rp, coup)
    np.testing.assert_allclose(a, b, rtol=1e-10)


@pytest.mark.parametrize("name,k", [("uux_ttx", 0), ("uux_ttxg", 1), ("uux_ttxgg", 2), ("ug_ttxu", 0),
                                    ("uxg_ttxux", 0)])
def test_qq_helicity_sum(name, k):
    """smatrix = sum over all 2^nexternal rows / (4 spins x colours x k!), the colours
    3 x 3 for q q~ and 3 x 8 for q g."""
    ps, _ = phase_space_points(name, 3, seed=46)
    d = O.load_process(name)
    n = d["nexternal"]
    rows = [list(h) for h in itertools.product([-1, 1], repeat=n)]
    assert sorted(map(tuple, d["helicities"])) == sorted(map(tuple, rows))
    assert d["denominator"] == 4 * (24 if name in QG else 9) * math.factorial(k)
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


def test_uux_ttxgg_bose_symmetry():
    ps, _ = phase_space_points("uux_ttxgg", 4, seed=48)
    rp, coup = O.default_params()
    sw = ps.copy()
    sw[:, [4, 5]] = ps[:, [5, 4]]
    np.testing.assert_allclose(O.smatrix("uux_ttxgg", ps, rp, coup), O.smatrix("uux_ttxgg", sw, rp, coup), rtol=1e-12)


def _colour_ops(M):
    """T_i^b M of the five legs of the brute-force u u~ > t t~ g tensor M[qb, q, t, tb, a]:
    the antiquark's index transforms with T, the quark's with -T^T, the top's with T, the
    anti-top's with -T^T, the gluon's with (T^b)_{ac} = -i f^{bac}."""
    T, F = BF.T, BF.F
    return [np.einsum("bij,j...->bi...", T, M),
            -np.einsum("bji,xj...->bxi...", T, M),
            np.einsum("bij,xyj...->bxyi...", T, M),
            -np.einsum("bji,xyzj...->bxyzi...", T, M),
            np.einsum("bac,xyzwc->bxyzwa", -1j * F, M)]


def _soft_ratio(lam, seed=5, sqrts=3000.0, name="uux_ttxgg"):
    """|M|^2 of u u~ > t t~ g g with gluon 5 of energy lam sqrt(s), summed over every
    helicity and colour, over its eikonal limit
        -g^2 sum_ij (p_i.p_j) / ((p_i.k)(p_j.k)) <M| T_i.T_j |M>
    with M the explicit-colour brute force of u u~ > t t~ g at the other momenta (the
    remaining legs recoil: the t t~ g system is a RAMBO point at the reduced mass, boosted
    to P - k).  Returns the ratio and the colour-conservation residual |sum_i T_i M| / |M|."""
    rng = np.random.default_rng(seed)
    E = sqrts / 2
    p0, p1 = np.array([E, 0, 0, E]), np.array([E, 0, 0, -E])
    nv = rng.normal(size=3)
    k = lam * sqrts * np.concatenate([[1.0], nv / np.linalg.norm(nv)])
    Q = p0 + p1 - k
    ps, _ = phase_space_points("uux_ttxg", 1, seed=seed + 100, sqrts=np.sqrt(minkowski(Q, Q)))
    fin = _boost(ps[0, 2:], Q[1:] / Q[0])
    p6 = np.vstack([p0, p1, fin, k])
    p5 = np.vstack([p0, p1, fin])
    rp, coup = O.default_params(wt=0.0)
    full = sum(O.matrix_hel(name, p6[None], h, rp, coup)[0] for h in O.load_process("uux_ttxgg")["helicities"])
    gs = 2 * np.sqrt(np.pi * 0.118)
    legs = [p5[1], p5[0], p5[2], p5[3], p5[4]]          # momenta of the axes qb, q, t, tb, a
    eik, cons = 0.0, 0.0
    for h in itertools.product([-1, 1], repeat=5):
        M = BF.qq_amplitude_tensor(p5, list(h), 173.0, 0.0, gs, np.sqrt(0.5))
        if not np.any(M):
            continue
        ops = _colour_ops(M)
        cons = max(cons, np.abs(sum(ops)).max() / np.abs(M).max())
        for i in range(5):
            for j in range(5):
                cij = np.sum(np.conj(ops[i]) * ops[j]).real
                eik -= gs ** 2 * minkowski(legs[i], legs[j]) / (minkowski(legs[i], k) * minkowski(legs[j], k)) * cij
    return full / eik, cons


def test_uux_ttxgg_soft_limit():
    """The ratio tends to 1 linearly in the soft gluon's energy (measured: 1 + 2.9e-2 at
    lam = 1e-3, 1 + 2.9e-4 at 1e-5)."""
    r3, c3 = _soft_ratio(1e-3)
    r5, c5 = _soft_ratio(1e-5)
    assert c3 < 1e-13 and c5 < 1e-13
    assert abs(r5 - 1) < 1e-3
    assert 30 < (r3 - 1) / (r5 - 1) < 300


def test_soft_limit_catches_a_dropped_emission():
    """Dropping the amplitudes in which gluon 5 leaves the top line directly (from the
    oracle's jamps) moves the soft ratio far from 1."""
    import copy
    d = copy.deepcopy(O.load_process("uux_ttxgg"))
    wf = {op["out"] for op in d["ops"] if op["kind"] == "wf" and sorted(op.get("legs", [])) == [2, 5]}
    amps = {op["out"] for op in d["ops"] if op["kind"] == "amp" and wf & set(op["args"])}
    assert amps
    d["jamp"] = [[t for t in terms if t[0] not in amps] for terms in d["jamp"]]
    O._cache["uux_ttxgg_mutant"] = O.EncodedProcess(d)
    try:
        r, _ = _soft_ratio(1e-5, name="uux_ttxgg_mutant")
        assert abs(r - 1) > 0.05
    finally:
        del O._cache["uux_ttxgg_mutant"]


def test_qg_bruteforce_catches_a_dropped_diagram():
    """The crossed brute force is sensitive to each diagram: removing the three-gluon
    (t-channel) diagram from the oracle's u g > t t~ u jamps changes |M|^2 by far more
    than the 1e-12 bar."""
    import copy
    d = copy.deepcopy(O.load_process("ug_ttxu"))
    ggg = [op["out"] for op in d["ops"] if op["kind"] == "amp"
           and any(o2["out"] == op["args"][2] and o2["routine"] == "VVV1P0_1" for o2 in d["ops"])]
    assert len(ggg) == 1
    d["jamp"] = [[t for t in terms if t[0] != ggg[0]] for terms in d["jamp"]]
    O._cache["ug_ttxu_mutant"] = O.EncodedProcess(d)
    try:
        ps, _ = phase_space_points("ug_ttxu", 1, seed=47)
        rp, coup = O.default_params()
        hel = [h for h in d["helicities"] if h[0] == h[4]][0]
        bf = BF.qg_matrix_hel(ps[0], hel)
        mut = O.matrix_hel("ug_ttxu_mutant", ps, hel, rp, coup)[0]
        assert abs(mut / bf - 1) > 1e-3
    finally:
        del O._cache["ug_ttxu_mutant"]


def test_qg_charge_conjugation():
    """C invariance of QCD: |M|^2 of u~ g > t t~ u~ at (p0, p1, p2, p3, p4) equals |M|^2 of
    u g > t t~ u with the top and antitop momenta exchanged, row by row with the same
    helicity labels (the antitop's label moving with its momentum), and after the sum.
    The two programs are generated with different spinors and fermion arrows, so a wrong
    spinor kind, propagator momentum or colour assignment in either breaks this."""
    ps, _ = phase_space_points("uxg_ttxux", 4, seed=48)
    swapped = np.ascontiguousarray(ps[:, [0, 1, 3, 2, 4]])
    rp, coup = O.default_params()
    d = O.load_process("uxg_ttxux")
    for hel in d["helicities"]:
        a = O.matrix_hel("uxg_ttxux", ps, hel, rp, coup)
        b = O.matrix_hel("ug_ttxu", swapped, [hel[0], hel[1], hel[3], hel[2], hel[4]], rp, coup)
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=0.0)
    np.testing.assert_allclose(O.smatrix("uxg_ttxux", ps, rp, coup), O.smatrix("ug_ttxu", swapped, rp, coup),
                               rtol=1e-13)
    # the unswapped momenta are a different configuration: the identity is not trivial
    assert np.max(np.abs(O.smatrix("uxg_ttxux", ps, rp, coup) / O.smatrix("ug_ttxu", ps, rp, coup) - 1)) > 1e-3
