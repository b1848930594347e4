"""Shared input generators for the tests (phase-space points from the oracle's RAMBO)."""

import numpy as np

from oracle import oracle as O

SQRTS = {"gg_ttx": 1000.0, "gg_ttxg": 2000.0, "gg_ttxgg": 3000.0, "gg_ttxggg": 5000.0,
         "uux_ttx": 1000.0, "uux_ttxg": 2000.0, "uux_ttxgg": 3000.0, "ug_ttxu": 2000.0,
         "uxg_ttxux": 2000.0}
NOUT = {"gg_ttx": 2, "gg_ttxg": 3, "gg_ttxgg": 4, "gg_ttxggg": 5, "uux_ttx": 2, "uux_ttxg": 3, "uux_ttxgg": 4,
        "ug_ttxu": 3, "uxg_ttxux": 3}
MT = 173.0


def phase_space_points(name, nev, seed=7, sqrts=None):
    """(nev, nexternal, 4) momenta: incoming partons (gluons or u, u~) along +-z,
    outgoing t, t~, gluons from the oracle's massive RAMBO (per-event Newton)."""
    sqrts = sqrts or SQRTS[name]
    n = NOUT[name]
    rng = np.random.default_rng(seed)
    x = rng.random((nev, 4 * n))
    masses = [MT, MT] + [0.0] * (n - 2)
    p_out, wt = O.rambo(x, n, sqrts, masses)
    pin = np.zeros((nev, 2, 4))
    pin[:, 0] = [sqrts / 2, 0, 0, sqrts / 2]
    pin[:, 1] = [sqrts / 2, 0, 0, -sqrts / 2]
    return np.ascontiguousarray(np.concatenate([pin, p_out], axis=1)), wt


def minkowski(a, b):
    return a[..., 0] * b[..., 0] - np.sum(a[..., 1:] * b[..., 1:], axis=-1)
