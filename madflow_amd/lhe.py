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


def _write(path, process, ps, xw, sigma_pb, err_pb, sqrts, idwtup, header, scale, aqed, aqcd, process_id):
    """The Les Houches file itself: one event per row of ps with weight xw."""
    pdg = PDG[process]
    opener = gzip.open if str(path).endswith(".gz") else open
    ebeam = 0.5 * float(sqrts)
    with opener(path, "wt") as f:
        f.write('<LesHouchesEvents version="3.0">\n')
        f.write("<header>\n%s\n</header>\n" % header)
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


def _momenta(process, momenta):
    pdg = PDG[process]
    ps = np.asarray(momenta, dtype=np.float64)
    if ps.ndim != 3 or ps.shape[1] != len(pdg) or ps.shape[2] != 4:
        raise ValueError("momenta must have shape (nev, %d, 4)" % len(pdg))
    return ps


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
    ps = _momenta(process, momenta)
    xw = np.full(len(ps), float(sigma_pb))
    if weights is not None:
        if wmax is None or not wmax > 0:
            raise ValueError("weights need the positive maximum weight wmax they were unweighted against")
        w = np.asarray(weights, dtype=np.float64).reshape(-1)
        if w.shape != (len(ps),):
            raise ValueError("weights must have one entry per event")
        xw = float(sigma_pb) * np.maximum(1.0, w / float(wmax))
    over = int(np.count_nonzero(xw != float(sigma_pb)))
    _write(path, process, ps, xw, float(sigma_pb), float(err_pb), sqrts, 4 if over else 3,
           '<madflow_amd process="%s" events="%d" overweight_events="%d"/>' % (process, len(ps), over),
           scale, aqed, aqcd, process_id)
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


def wmax_for_target(w, event_target=0):
    """The maximum weight that hit-or-miss unweighting of the positive weights w is run
    against: the largest weight, or for 0 < event_target < len(w) the w_max at which the
    expected number of kept events, sum min(1, w / w_max), equals event_target.  That sum
    falls as w_max grows; above the largest weight it is sum(w) / w_max."""
    pos = np.asarray(w, dtype=np.float64)
    wmax = float(pos.max())
    if 0 < event_target <= pos.sum() / wmax:
        return float(pos.sum()) / event_target
    if 0 < event_target < len(pos):
        lo, hi = float(pos.min()), wmax
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if np.minimum(1.0, pos / mid).sum() > event_target:
                lo = mid
            else:
                hi = mid
        return hi
    return wmax


class LheWriter:
    """madflow.lhe_writer.LheWriter (lhe_writer.py:94-356): a context manager around an
    integration.  The integrand hands each batch of weighted events to lhe_parser(all_ps,
    res); store_result((sigma, err)) records the result.  When the block ends, the events
    go to <folder>/Events/<run>/weighted_events.lhe.gz, each with its weight.  Unless
    no_unweight, they are then unweighted on the device (mfx_unweight, hit-or-miss) into
    unweighted_events.lhe.gz, every kept event carrying the cross section, as the
    reference's do_unweighting writes them (lhe_writer.py:319-356).  The reference hands
    the weights to MadGraph's lhe_parser, which is not available here.  Here the
    maximum weight is the largest weight, or with 0 < event_target < (events of positive
    weight) the weight at which the expected number of kept events, sum min(1, w / wmax), is
    event_target; events above it are kept with weight sigma x w / wmax (write_lhe).
    ``process`` (a key of PDG) and ``sqrts`` are needed for the particle codes and the beam
    energies; the reference writes protons and a t t~ pair for every process
    (lhe_writer.py:171-184)."""

    def __init__(self, folder, run="run_01", no_unweight=False, event_target=0, process=None, sqrts=None, seed=1):
        from pathlib import Path
        if process not in PDG or sqrts is None:
            raise ValueError("LheWriter needs the process (one of %s) and sqrt(s)" % sorted(PDG))
        self.folder, self.run = Path(folder), run
        self.no_unweight, self.event_target = bool(no_unweight), int(event_target)
        self.process, self.sqrts, self.seed = process, float(sqrts), int(seed)
        lhe_folder = self.folder / "Events" / run
        lhe_folder.mkdir(parents=True, exist_ok=True)
        self.lhe_path = lhe_folder / "weighted_events.lhe.gz"
        self.unweighted_path = lhe_folder / "unweighted_events.lhe.gz"
        self._ps, self._w = [], []
        self.cross = self.err = None
        self.stats = None

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_value, exc_traceback):
        ps, w = self._events()
        sigma = self.cross if self.cross is not None else float(w.sum() / max(1, len(w)))
        err = self.err if self.err is not None else 0.0
        _write(self.lhe_path, self.process, ps, w, sigma, err, self.sqrts, 4,
               '<madflow_amd process="%s" weighted_events="%d"/>' % (self.process, len(ps)), 0.0, 0.0, 0.0, 1)
        if not self.no_unweight and exc_type is None:
            self.stats = self.do_unweighting(self.event_target)
        return False

    def lhe_parser(self, all_ps, res):
        """Store a batch of weighted events: momenta (nev, nexternal, 4) and weights (nev,),
        device tensors or arrays.  Returns 0, as the reference's does."""
        ps = all_ps.detach().cpu().numpy() if hasattr(all_ps, "detach") else np.asarray(all_ps)
        w = res.detach().cpu().numpy() if hasattr(res, "detach") else np.asarray(res)
        self._ps.append(_momenta(self.process, ps))
        self._w.append(np.asarray(w, dtype=np.float64).reshape(-1))
        if self._w[-1].shape != (len(self._ps[-1]),):
            raise ValueError("lhe_parser: one weight per event")
        return 0.0

    def store_result(self, result):
        """Record (cross section, statistical error) for the files' headers and weights."""
        self.cross, self.err = float(result[0]), float(result[1])

    def dump_result(self, filename):
        """Write the cross section and its error, one per line (lhe_writer.py:272-281)."""
        np.savetxt(str(filename), np.array([self.cross, self.err], dtype=np.float64))

    def _events(self):
        if not self._ps:
            return np.zeros((0, len(PDG[self.process]), 4)), np.zeros(0)
        return np.concatenate(self._ps), np.concatenate(self._w)

    def do_unweighting(self, event_target=0):
        """Unweight the stored events into unweighted_events.lhe.gz; returns (kept, weighted)."""
        import torch
        from .vegas import unweight
        ps, w = self._events()
        if self.cross is None:
            raise ValueError("store_result((sigma, err)) before unweighting")
        pos = w[w > 0]
        if not len(pos):
            write_lhe(self.unweighted_path, self.process, ps[:0], self.cross, self.err or 0.0, self.sqrts)
            return 0, len(w)
        wmax = wmax_for_target(pos, event_target)
        wt = torch.from_numpy(w).to("cuda")
        idx, _ = unweight(wt, wmax, self.seed)
        k = idx.long().cpu().numpy()
        write_lhe(self.unweighted_path, self.process, ps[k], self.cross, self.err or 0.0, self.sqrts,
                  weights=w[k], wmax=wmax)
        return len(k), len(w)
