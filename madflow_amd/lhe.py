"""Les Houches event files of unweighted events.

madflow writes the events of an integration through MadGraph's ``lhe_parser``
(lhe_writer.py:94-356): weighted events go to ``weighted_events.lhe.gz`` and MadGraph's
hit-or-miss unweighting writes ``unweighted_events.lhe.gz``, where every kept event
carries the cross section as its weight (lhe_writer.py:341-350).  Here the unweighting
runs on the device (``MEIntegrand.unweighted_events``) and this module writes the kept
events directly, gzip-compressed, in the Les Houches Accord v3 layout
(<init> block, then one <event> block per event: NUP IDPRUP XWGTUP SCALUP AQEDUP AQCDUP,
then IDUP ISTUP MOTHUP1 MOTHUP2 ICOLUP1 ICOLUP2 PUP1..5 VTIMUP SPINUP per particle).

As in the reference's writer (lhe_writer.py:151-200), mothers and colour lines are 0
and the scale and couplings are written as given (0 by default); the spin entry is 9
(unpolarised, the LHA code for "no helicity information").  The incoming partons are
written as gluons (the reference's writer labels them as protons), tops as 6 / -6.
"""

import gzip
import math

import numpy as np

# PDG codes of the external legs of each process (the u-quark member of the quark flavours a
# process serves): g g > t t~ g, u u~ > t t~, u g > t t~ u ...
PDG = {
    "gg_ttx": [21, 21, 6, -6],
    "gg_ttxg": [21, 21, 6, -6, 21],
    "gg_ttxgg": [21, 21, 6, -6, 21, 21],
    "gg_ttxggg": [21, 21, 6, -6, 21, 21, 21],
    "uux_ttx": [2, -2, 6, -6],
    "uux_ttxg": [2, -2, 6, -6, 21],
    "uux_ttxgg": [2, -2, 6, -6, 21, 21],
    "ug_ttxu": [2, 21, 6, -6, 2],
    "uxg_ttxux": [-2, 21, 6, -6, -2],
}


def _mass(p):
    m2 = p[0] * p[0] - p[1] * p[1] - p[2] * p[2] - p[3] * p[3]
    return math.sqrt(m2) if m2 > 0 else 0.0


def write_lhe(path, process, momenta, sigma_pb, err_pb, sqrts, scale=0.0, aqed=0.0, aqcd=0.0, process_id=1,
              weights=None, wmax=None):
    """Write unweighted events to ``path`` (gzip if it ends in .gz).

    momenta: (nev, nexternal, 4) array (E, px, py, pz) in GeV, legs in the process
    order (two incoming first).  Every event gets the weight sigma_pb (unweighted events,
    IDWTUP = 3), except that with ``weights`` (the kept events' weights w, as returned by
    MEIntegrand.unweighted_events) and ``wmax`` an event with w > wmax, which hit-or-miss
    keeps with probability 1 < w / wmax, carries sigma_pb * w / wmax so the sample stays
    unbiased; the file then declares weighted events (IDWTUP = 4) and its header gives the
    number of over-weight events.  Returns the number of events written."""
    pdg = PDG[process]
    ps = np.asarray(momenta, dtype=np.float64)
    if ps.ndim != 3 or ps.shape[1] != len(pdg) or ps.shape[2] != 4:
        raise ValueError("momenta must have shape (nev, %d, 4)" % len(pdg))
    xw = np.full(len(ps), float(sigma_pb))
    if weights is not None:
        if wmax is None or not wmax > 0:
            raise ValueError("weights need the positive maximum weight wmax they were unweighted against")
        w = np.asarray(weights, dtype=np.float64).reshape(-1)
        if w.shape != (len(ps),):
            raise ValueError("weights must have one entry per event")
        xw = float(sigma_pb) * np.maximum(1.0, w / float(wmax))
    over = int(np.count_nonzero(xw != float(sigma_pb)))
    idwtup = 4 if over else 3
    opener = gzip.open if str(path).endswith(".gz") else open
    ebeam = 0.5 * float(sqrts)
    with opener(path, "wt") as f:
        f.write('<LesHouchesEvents version="3.0">\n')
        f.write("<header>\n<madflow_amd process=\"%s\" events=\"%d\" overweight_events=\"%d\"/>\n</header>\n"
                % (process, len(ps), over))
        f.write("<init>\n")
        f.write(" %d %d %.10e %.10e 0 0 0 0 %d 1\n" % (pdg[0], pdg[1], ebeam, ebeam, idwtup))
        f.write(" %.10e %.10e %.10e %d\n" % (sigma_pb, err_pb, sigma_pb, process_id))
        f.write("</init>\n")
        for ev, wt in zip(ps, xw):
            f.write("<event>\n")
            f.write(" %d %d %+.10e %.10e %.10e %.10e\n" % (len(pdg), process_id, wt, scale, aqed, aqcd))
            for leg, p in enumerate(ev):
                status = -1 if leg < 2 else 1
                f.write(" %8d %2d 0 0 0 0 %+.12e %+.12e %+.12e %.12e %.12e 0.0 9.0\n" % (
                    pdg[leg], status, p[1], p[2], p[3], p[0], _mass(p)))
            f.write("</event>\n")
        f.write("</LesHouchesEvents>\n")
    return len(ps)


def read_lhe(path):
    """Read back (process_line, events) from a file written by write_lhe: events is a list
    of (event_line, [(pid, status, px, py, pz, E, m), ...])."""
    opener = gzip.open if str(path).endswith(".gz") else open
    init, events, cur, block = None, [], None, None
    with opener(path, "rt") as f:
        for line in f:
            s = line.strip()
            if s == "<init>":
                block, init = "init", []
            elif s == "</init>":
                block = None
            elif s == "<event>":
                block, cur = "event", None
            elif s == "</event>":
                events.append(cur)
                block = None
            elif block == "init":
                init.append([float(x) for x in s.split()])
            elif block == "event":
                v = s.split()
                if cur is None:
                    cur = ([float(x) for x in v], [])
                else:
                    cur[1].append((int(v[0]), int(v[1])) + tuple(float(x) for x in v[6:11]))
    return init, events
