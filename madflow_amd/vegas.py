"""Device VEGAS integration of the matrix elements (the integration loop madflow runs
through vegasflow: scripts/madflow_exec.py:474-527, utilities.py:56-90).

``MEIntegrand`` fuses, per iteration, Philox uniforms -> VEGAS map -> RAMBO -> cuts ->
compaction (``mfx_ps_generate``), the smatrix kernel on the passing events, and the
compensated accumulation of sum f, sum f^2 and the per-dimension histograms
(``mfx_vegas_accumulate``); the grid is refined on the device (``mfx_vegas_refine``).

``VegasFlow`` mirrors the vegasflow interface used by the reference
(``VegasFlow(n_dim, n_events)``, ``compile(integrand)``, ``run_integration(n_iter)``,
``freeze_grid()``): iterations are combined with 1/sigma^2 weights.
"""

import ctypes
import math
import warnings
import time

import numpy as np
import torch

from . import _lib
from .matrix import load_matrix

GEV2PB = 389379365.6
BINS_MAX = 50
ALPHA = 1.5


MAX_CUTS = 16
CUT_VARIABLES = {"pt": 0, "mt": 1, "mt2": 2}


class Cut(ctypes.Structure):
    _fields_ = [("variable", ctypes.c_int), ("particle", ctypes.c_int), ("use_min", ctypes.c_int),
                ("use_max", ctypes.c_int), ("min_val", ctypes.c_double), ("max_val", ctypes.c_double)]


def add_cut(p, nexternal, variable, particle=None, min_val=None, max_val=None):
    """Append register_cut's cut (phasespace.py:424-478) to the PsParams p of a process
    with nexternal legs: min_val < variable(particle) < max_val, variable "pt", "mt" or
    "mt2", particle an outgoing leg (2 .. nexternal - 1).  The reference's errors for an
    unknown variable or a particle past the last leg, and its warning (the cut ignored)
    when neither bound is given."""
    if variable not in CUT_VARIABLES:
        raise ValueError(f"{variable} is not implemented")
    if particle is None:
        raise ValueError("register_cut needs the particle the cut applies to")
    if particle >= nexternal:
        raise ValueError(f"Cannot apply cuts to particle {particle}, python idx starts at 0!")
    if particle < 2:
        raise ValueError("cuts apply to the outgoing particles, 2 .. %d" % (nexternal - 1))
    if min_val is None and max_val is None:
        warnings.warn(f"Cut for {variable} has no min or max val, ignoring")
        return
    if p.ncuts >= MAX_CUTS:
        raise ValueError("at most %d cuts" % MAX_CUTS)
    k = p.cuts[p.ncuts]
    k.variable, k.particle = CUT_VARIABLES[variable], int(particle)
    k.use_min, k.min_val = int(min_val is not None), float(min_val if min_val is not None else 0.0)
    k.use_max, k.max_val = int(max_val is not None), float(max_val if max_val is not None else 0.0)
    p.ncuts += 1


class PsParams(ctypes.Structure):
    _fields_ = [
        ("n_out", ctypes.c_int), ("ndim", ctypes.c_int), ("nbins", ctypes.c_int), ("mapping", ctypes.c_int),
        ("sqrts", ctypes.c_double), ("masses", ctypes.c_double * 8), ("gev2pb", ctypes.c_double),
        ("pt_min", ctypes.c_double), ("abs_eta_max", ctypes.c_double), ("dr_min", ctypes.c_double),
        ("cut_first_leg", ctypes.c_int), ("seed", ctypes.c_uint32),
        ("shat_min", ctypes.c_double), ("lab_frame", ctypes.c_int), ("reserved", ctypes.c_int),
        ("ncuts", ctypes.c_int), ("reserved2", ctypes.c_int), ("cuts", Cut * MAX_CUTS),
    ]


def _sig(lib):
    vp, i64, u64, u32, i, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER(PsParams)
    lib.mfx_integrand_workspace_size.argtypes = [i64, i, i]
    lib.mfx_integrand_workspace_size.restype = i64
    lib.mfx_ps_generate.argtypes = [P, vp, u64, i64, u32, vp, vp, vp, vp, vp, vp, vp, vp, vp, i64,
                                    ctypes.POINTER(i64), vp]
    lib.mfx_ps_generate.restype = i
    lib.mfx_ps_generate_xrand.argtypes = [P, vp, i64, vp, vp, vp, vp, vp, vp, vp, i64, ctypes.POINTER(i64), vp]
    lib.mfx_ps_generate_xrand.restype = i
    lib.mfx_vegas_accumulate.argtypes = [P, u64, i64, u32, vp, vp, vp, vp, vp, vp, vp, i64, vp, vp]
    lib.mfx_vegas_accumulate.restype = i
    lib.mfx_vegas_refine.argtypes = [vp, vp, i, i, d, vp]
    lib.mfx_vegas_refine.restype = i


PROCESS_SETUP = {
    # name: n_out, masses, default sqrt(s)
    "gg_ttx": (2, [173.0, 173.0]),
    "gg_ttxg": (3, [173.0, 173.0, 0.0]),
    "gg_ttxgg": (4, [173.0, 173.0, 0.0, 0.0]),
    "gg_ttxggg": (5, [173.0, 173.0, 0.0, 0.0, 0.0]),
    "uux_ttx": (2, [173.0, 173.0]),
    "uux_ttxg": (3, [173.0, 173.0, 0.0]),
    "uux_ttxgg": (4, [173.0, 173.0, 0.0, 0.0]),
    "ug_ttxu": (3, [173.0, 173.0, 0.0]),
    "uxg_ttxux": (3, [173.0, 173.0, 0.0]),
}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def initial_flavours(matrix, explicit=False):
    """(beam-1 flavours, beam-2 flavours) of a drop-in class's initial states, the mirror
    pairs appended when mirror_initial_states is set (madflow_exec.py:141-155); None for
    processes with two incoming gluons only, unless ``explicit``."""
    states = [tuple(s) for s in matrix.initial_states]
    if all(s == (21, 21) for s in states) and not explicit:
        return None
    f1, f2 = [a for a, _ in states], [b for _, b in states]
    if matrix.mirror_initial_states:
        f1, f2 = f1 + [b for _, b in states], f2 + [a for a, _ in states]
    return f1, f2


def unweight(w, wmax, seed, iteration=0, ev0=0):
    """Hit-or-miss unweighting on the device (mfx_unweight): event e of the weights w
    (device f64) is kept with probability min(1, w_e / wmax), from its own Philox uniform
    (key seed, counter ev0 + e, iteration).  Returns (indices of the kept events in order,
    keep mask)."""
    lib = _lib.load()
    _sig(lib)
    w = _lib.require_device_tensor(w, torch.float64, "w")
    n = w.shape[0]
    keep = torch.empty(n, dtype=torch.uint8, device=w.device)
    idx = torch.empty(n, dtype=torch.int32, device=w.device)
    ws = torch.empty(lib.mfx_unweight_workspace_size(n), dtype=torch.uint8, device=w.device)
    nk = ctypes.c_int64(0)
    _lib.check(lib.mfx_unweight(_ptr(w), n, float(wmax), int(seed) & 0xFFFFFFFF, iteration, ev0, _ptr(keep), _ptr(idx),
                                _ptr(ws), ws.numel(), ctypes.byref(nk), _lib.stream_ptr(w.device)))
    return idx[:nk.value], keep


class MEIntegrand:
    """sigma = int dPhi |M|^2 / (2 s) * GeV^-2->pb for one process at fixed sqrt(s)
    (mapping "rambo" or "diet"), or the hadronic
    sigma = int dx1 dx2 (x1 f(x1) x2 f(x2) / (x1 x2)) dPhi |M|^2 / (2 shat) * GeV^-2->pb
    at the hadronic sqrt(s) (mapping "hadronic", ramboflow: phasespace.py:255-317; the
    integrand of madflow_exec.py:422-470).

    Hadronic mode: ``pdf(x, q2)`` returns x f_g(x, Q^2) for device tensors (the gluon of
    both beams; pdfflow's xfxQ2 in the reference) for processes with two incoming gluons;
    for quark-initiated processes ``pdf(x, q2, pid)`` returns x f(x, Q^2) of the PDG flavour
    pid, and the luminosity sums the process's flavour pairs, mirrors included (the
    reference's luminosity, madflow_exec.py:141-155,440-454); ``scale`` None is the reference's dynamic
    scale Q = sum of the outgoing transverse masses / 2 (madflow_exec.py:427-430), a number
    a fixed Q; ``alphas(q2)`` (optional) gives alpha_s(Q^2) per event, which then sets the
    couplings of every event, else alpha_s stays frozen.  ``lab_frame`` boosts the
    momenta to the lab frame; None (the default) does so in hadronic mode, as the
    reference's integration script does (``PhaseSpaceGenerator(..., com_output=False)``,
    madflow_exec.py:389, so its LHE momenta are lab-frame), while the phase-space class
    alone defaults to the partonic frame (phasespace.py:377).  The cuts act in the partonic
    frame either way.  ``kernels`` selects the matrix element's kernel family (the
    drop-in class's default when None).

    ``process`` may also be a list of processes with the same outgoing masses: the
    integrand is then the sum over them of luminosity x |M|^2 at each phase-space point,
    as the reference's integrand sums its matrices (madflow_exec.py:453-465); in hadronic
    mode the pdf then takes (x, q2, pid) for every one of them, gluons as pid 21.  The
    first process sets the leg order of the momenta (the cuts and event output)."""

    def __init__(self, process, sqrts, precision=64, mapping="rambo", pt_min=0.0, abs_eta_max=0.0, dr_min=0.0,
                 seed=1, nbins=BINS_MAX, alpha_s=0.118, device="cuda", events_limit=1 << 22, ref_compat=True,
                 width=None, pdf=None, scale=None, alphas=None, lab_frame=None, shat_min=None, kernels=None):
        procs = [process] if isinstance(process, str) else list(process)
        if not procs:
            raise ValueError("MEIntegrand needs a process")
        n_out, masses = PROCESS_SETUP[procs[0]]
        for q in procs[1:]:
            if PROCESS_SETUP[q] != (n_out, masses):
                raise ValueError("the processes of one integrand share its phase space: %s and %s have different "
                                 "outgoing masses" % (procs[0], q))
        self.process = process
        self.device = torch.device(device)
        self.hadronic = mapping == "hadronic"
        if self.hadronic and pdf is None:
            raise ValueError("the hadronic mapping needs a pdf(x, q2) -> x f(x, Q^2)")
        self.pdf, self.scale, self.alphas = pdf, scale, alphas
        self.width = width
        # one entry per process: (matrix, model, parameters, initial flavours)
        self.subprocesses = []
        for q in procs:
            matrix, mod = load_matrix(q, precision=precision, ref_compat=ref_compat, kernels=kernels)
            model = mod.get_model_param()
            if alphas is None:
                model.freeze_alpha_s(alpha_s)
                params = list(model.evaluate())
            else:
                params = list(model.evaluate(torch.tensor([float(alpha_s)], dtype=torch.float64)))
            if width is not None:
                params[1] = float(width)
            self.subprocesses.append((matrix, model, params, initial_flavours(matrix, explicit=len(procs) > 1)))
        self.matrix, self.model, self.params, self.flavours = self.subprocesses[0]
        p = PsParams()
        p.n_out = n_out
        p.mapping = {"rambo": 0, "diet": 1, "hadronic": 2}[mapping]
        p.ndim = (4 * n_out, 3 * n_out - 4, 4 * n_out + 2)[p.mapping]
        p.shat_min = float(sum(masses) ** 2 if shat_min is None else shat_min)
        p.lab_frame = int(bool(self.hadronic if lab_frame is None else lab_frame))
        p.nbins = nbins
        p.sqrts = float(sqrts)
        for k, m in enumerate(masses):
            p.masses[k] = m
        p.gev2pb = GEV2PB
        p.pt_min = pt_min
        p.abs_eta_max = abs_eta_max
        p.dr_min = dr_min
        p.cut_first_leg = 4 if (pt_min > 0 or abs_eta_max > 0 or dr_min > 0) else n_out + 2
        p.seed = seed
        self.p = p
        self.ndim = p.ndim
        self.nbins = nbins
        self.nexternal = n_out + 2
        self.lib = _lib.load()
        _sig(self.lib)
        self.grid = torch.linspace(0, 1, nbins + 1, dtype=torch.float64, device=self.device).repeat(self.ndim, 1)
        self.grid = self.grid.contiguous()
        self.events_limit = int(events_limit)
        self._bufs = None
        self.me_timer = None      # list of (start event, end event, n) around smatrix launches
        self.last_chunk_npass = 0

    def register_cut(self, variable, particle=None, min_val=None, max_val=None):
        """PhaseSpaceGenerator.register_cut (phasespace.py:424-478): keep only the events
        with min_val < variable(particle) < max_val (either bound may be None), variable
        "pt", "mt" or "mt2" (m^2 + pt^2) of the outgoing external leg ``particle`` (2 ..
        n_out + 1, in the process's leg order), in the partonic c.m. frame; the cuts are
        ANDed with each other and with the pt_min / abs_eta_max / dr_min cuts.  As in the
        reference, an unknown variable or a particle past the last leg raises ValueError,
        and a cut with neither bound is ignored with a warning.  A cut needs its particle
        (the reference's particle=None form has no well-defined event mask) and an
        outgoing one (the incoming partons have pt = 0)."""
        add_cut(self.p, self.nexternal, variable, particle, min_val, max_val)

    def clear_cuts(self):
        """Remove every cut registered with register_cut (phasespace.py:396-402)."""
        self.p.ncuts = 0

    def _alloc(self, n):
        if self._bufs is not None and self._bufs["n"] >= n:
            return self._bufs
        d = self.device
        ws = self.lib.mfx_integrand_workspace_size(n, self.ndim, self.nbins)
        self._bufs = dict(
            n=n,
            all_ps=torch.empty((n, self.nexternal, 4), dtype=torch.float64, device=d),
            dense=torch.empty((n, self.nexternal, 4), dtype=torch.float64, device=d),
            wps=torch.empty(n, dtype=torch.float64, device=d),
            flag=torch.empty(n, dtype=torch.uint8, device=d),
            slot=torch.empty(n, dtype=torch.int32, device=d),
            ws=torch.empty(ws, dtype=torch.uint8, device=d),
            result=torch.empty(6 + self.ndim * self.nbins, dtype=torch.float64, device=d),
            f=torch.empty(n, dtype=torch.float64, device=d),
            x12=torch.empty((n, 2), dtype=torch.float64, device=d),
        )
        return self._bufs

    def chunk(self, ev0, n, iteration, keep_f=False, y_in=None, w_in=None):
        """One chunk of events; returns (result vector on host, npass, f tensor or None).

        y_in (n, ndim) float32 / w_in (n,) float64: points and extra weights of an external
        proposal (the flow) instead of the Philox stream."""
        b = self._alloc(n)
        s = _lib.stream_ptr(self.device)
        npass = ctypes.c_int64(0)
        _lib.check(self.lib.mfx_ps_generate(
            ctypes.byref(self.p), _ptr(self.grid), ev0, n, iteration, _ptr(y_in), _ptr(w_in), _ptr(b["all_ps"]),
            _ptr(b["wps"]), _ptr(b["flag"]), _ptr(b["slot"]), _ptr(b["dense"]), _ptr(b["x12"]), _ptr(b["ws"]),
            b["ws"].numel(), ctypes.byref(npass), s))
        npass = npass.value
        self.last_chunk_npass = npass
        if len(self.subprocesses) > 1:
            return self._chunk_sum(b, ev0, n, iteration, keep_f, y_in, npass, s)
        params, f_in = self.params, None
        if self.hadronic and npass:
            params, lumi = self._hadronic_factors(b, n, npass)
        if self.me_timer is not None and npass:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            me = self.matrix.smatrix(b["dense"][:npass], *params)
            e1.record()
            self.me_timer.append((e0, e1, npass))
        else:
            # with no passing event every slot is -1 and nothing reads the matrix elements, but
            # the library still needs a device pointer, which an empty tensor does not have
            me = self.matrix.smatrix(b["dense"][:npass], *params) if npass else b["wps"][:1]
        if self.hadronic:
            # f = luminosity x |M|^2 x weight (madflow_exec.py:455-459), scattered to full length
            f_in = torch.zeros(n, dtype=torch.float64, device=self.device)
            if npass:
                ok = b["flag"][:n] == 1
                f_in[ok] = me * lumi * b["wps"][:n][ok]
        _lib.check(self.lib.mfx_vegas_accumulate(
            ctypes.byref(self.p), ev0, n, iteration, _ptr(y_in), _ptr(b["slot"]), _ptr(me), _ptr(b["wps"]),
            _ptr(f_in), _ptr(b["f"]) if keep_f else None, _ptr(b["ws"]), b["ws"].numel(), _ptr(b["result"]), s))
        return b["result"].cpu().numpy().copy(), npass, (b["f"][:n] if keep_f else None)

    def unweight(self, w, wmax, iteration, ev0=0):
        """Hit-or-miss unweighting of the weights w (device f64) against wmax on the
        device: returns (indices of the kept events in order, keep mask)."""
        return unweight(w, wmax, self.p.seed, iteration, ev0)

    def unweighted_events(self, n_events, iteration, wmax, flow=None, mix=0.2, seed=1234):
        """Generate n_events points on the frozen grid (from the flow's defensive mixture
        if ``flow`` is given, as in proposal_iteration), weight them, and keep each with
        probability min(1, w / wmax) on the device.  Returns (momenta of the kept events,
        (K, nexternal, 4) host array, their weights, the number of over-weight events).
        Points, weights and keep decisions are reproducible: they depend only on the
        seeds, the iteration and the event index."""
        from .phasespace_gpu import philox_uniforms
        kept, wts, over = [], [], 0
        ev0 = 0
        while ev0 < n_events:
            n = min(self.events_limit, n_events - ev0)
            if flow is not None:
                u = philox_uniforms(n, self.ndim, seed, iteration, ev0=ev0, dtype=torch.float32)
                ys, winv = flow.sample_mixture(u, u, mix)
                _, _, f = self.chunk(ev0, n, iteration, keep_f=True, y_in=ys.contiguous(), w_in=winv.contiguous())
            else:
                _, _, f = self.chunk(ev0, n, iteration, keep_f=True)
            idx, _ = self.unweight(f, wmax, iteration, ev0)
            il = idx.long()
            kept.append(self._bufs["all_ps"][:n].index_select(0, il).cpu().numpy())
            fw = f.index_select(0, il)
            wts.append(fw.cpu().numpy())
            over += int((fw > wmax).sum())
            ev0 += n
        return np.concatenate(kept), np.concatenate(wts), over

    def iteration(self, n_events, iteration, train=True, keep_f=False):
        """One VEGAS iteration of n_events; returns a dict with I, var, max weight, counts."""
        tot = None
        npass = 0
        fs = []
        ev0 = 0
        while ev0 < n_events:
            n = min(self.events_limit, n_events - ev0)
            r, k, f = self.chunk(ev0, n, iteration, keep_f)
            tot = r if tot is None else _merge(tot, r)
            npass += k
            if keep_f:
                fs.append(f.clone())
            ev0 += n
        N = float(n_events)
        mean = tot[0] / N
        var = max((tot[1] / N - mean * mean) / (N - 1.0), 0.0)
        if train:
            hist = torch.from_numpy(tot[6:].copy()).to(self.device)
            _lib.check(self.lib.mfx_vegas_refine(_ptr(self.grid), _ptr(hist), self.ndim, self.nbins, ALPHA,
                                                 _lib.stream_ptr(self.device)))
        out = dict(I=mean, var=var, sigma=math.sqrt(var), wmax=tot[2], nonzero=tot[3], npass=npass, n=n_events,
                   sum_f=tot[0])
        if keep_f:
            out["f"] = torch.cat(fs)
        return out

    def _chunk_sum(self, b, ev0, n, iteration, keep_f, y_in, npass, s):
        """The rest of chunk() for several processes: f = sum over them of luminosity x
        |M|^2 (1 at fixed energy) at each passing point, times its weight."""
        me = torch.zeros(max(npass, 1), dtype=torch.float64, device=self.device)   # a pointer also for npass 0
        if npass:
            if self.me_timer is not None:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
            for sub in self.subprocesses:
                if self.hadronic:
                    params, lumi = self._hadronic_factors(b, n, npass, sub)
                    me += sub[0].smatrix(b["dense"][:npass], *params) * lumi
                else:
                    me += sub[0].smatrix(b["dense"][:npass], *sub[2])
            if self.me_timer is not None:
                e1.record()
                self.me_timer.append((e0, e1, npass))
        _lib.check(self.lib.mfx_vegas_accumulate(
            ctypes.byref(self.p), ev0, n, iteration, _ptr(y_in), _ptr(b["slot"]), _ptr(me), _ptr(b["wps"]), None,
            _ptr(b["f"]) if keep_f else None, _ptr(b["ws"]), b["ws"].numel(), _ptr(b["result"]), s))
        return b["result"].cpu().numpy().copy(), npass, (b["f"][:n] if keep_f else None)

    def _hadronic_factors(self, b, n, npass, sub=None):
        """Couplings and parton luminosity of the passing events of a chunk: the scale Q
        (fixed, or the dynamic sum of outgoing transverse masses / 2), alpha_s(Q^2) if a
        running alpha_s is given, and x1 f(x1) x2 f(x2) / (x1 x2) at Q^2, for the process
        ``sub`` (an entry of subprocesses; the first by default)."""
        _, model, params0, flavours = sub if sub is not None else self.subprocesses[0]
        d = b["dense"][:npass]
        x = b["x12"][:n][b["flag"][:n] == 1]
        if self.scale is None:
            out = d[:, 2:, :]
            mt = torch.sqrt(torch.clamp(out[..., 0] ** 2 - out[..., 3] ** 2, min=0.0)).sum(dim=1)
            q2 = (0.5 * mt) ** 2
        else:
            q2 = torch.full((npass,), float(self.scale) ** 2, dtype=torch.float64, device=self.device)
        params = params0
        if self.alphas is not None:
            params = list(model.evaluate(self.alphas(q2).to(torch.float64).reshape(-1)))
            if self.width is not None:
                params[1] = float(self.width)
        x1, x2 = x[:, 0], x[:, 1]
        if flavours is None:
            lumi = self.pdf(x1, q2) * self.pdf(x2, q2) / (x1 * x2)
        else:
            # sum over the initial-state flavour pairs of the process, mirror pairs included
            # (madflow_exec.py:141-155,440-454; pdfflow's xfxQ2(pid, x, q2) per beam)
            f1, f2 = flavours
            cache = {}

            def xf(pid, xv, side):
                if (pid, side) not in cache:
                    cache[(pid, side)] = self.pdf(xv, q2, pid)
                return cache[(pid, side)]

            lumi = sum(xf(a, x1, 1) * xf(b, x2, 2) for a, b in zip(f1, f2)) / (x1 * x2)
        return params, lumi

    def proposal_iteration(self, flow, n_events, iteration, mix=0.2, seed=1234):
        """One iteration of n_events points drawn from the defensive mixture
        q' = (1 - mix) q + mix of a trained flow q on this integrand's (frozen) VEGAS cube,
        each weighted by f(y) / q'(y); the estimate is unbiased for any q, and mix = 1 is
        plain VEGAS on the frozen grid.  Uniforms: Philox stream (seed, iteration).  The grid
        is not refined.  Returns the same dict as iteration()."""
        from .phasespace_gpu import philox_uniforms
        if flow.D != self.ndim:
            raise ValueError("proposal_iteration: the flow has %d dimensions, the cube %d" % (flow.D, self.ndim))
        tot = None
        npass = 0
        ev0 = 0
        while ev0 < n_events:
            n = min(self.events_limit, n_events - ev0)
            u = philox_uniforms(n, self.ndim, seed, iteration, ev0=ev0, dtype=torch.float32)
            ys, winv = flow.sample_mixture(u, u, mix)
            r, k, _ = self.chunk(ev0, n, iteration, y_in=ys.contiguous(), w_in=winv.contiguous())
            tot = r if tot is None else _merge(tot, r)
            npass += k
            ev0 += n
        N = float(n_events)
        mean = tot[0] / N
        var = max((tot[1] / N - mean * mean) / (N - 1.0), 0.0)
        return dict(I=mean, var=var, sigma=math.sqrt(var), wmax=tot[2], nonzero=tot[3], npass=npass, n=n_events,
                    sum_f=tot[0])


def _merge(a, b):
    c = a.copy()
    c[0] = math.fsum([a[0], b[0]])
    c[1] = math.fsum([a[1], b[1]])
    c[2] = max(a[2], b[2])
    c[3] = a[3] + b[3]
    c[6:] = a[6:] + b[6:]
    return c


class VegasFlow:
    """vegasflow-style driver: VegasFlow(n_dim, n_events).compile(integrand);
    run_integration(n_iter) -> (result, error); freeze_grid()."""

    def __init__(self, n_dim=None, n_events=1000000, events_limit=None):
        self.n_dim = n_dim
        self.n_events = int(n_events)
        self.events_limit = events_limit
        self.integrand = None
        self.frozen = False
        self.iteration_index = 0
        self.history = []

    def compile(self, integrand):
        if not isinstance(integrand, MEIntegrand):
            raise TypeError("compile() takes an MEIntegrand (device-fused integrand)")
        if self.n_dim is not None and self.n_dim != integrand.ndim:
            raise ValueError("n_dim %d does not match the integrand's %d" % (self.n_dim, integrand.ndim))
        if self.events_limit:
            integrand.events_limit = int(self.events_limit)
        self.integrand = integrand

    def freeze_grid(self):
        self.frozen = True

    def run_integration(self, n_iter, log=None):
        res = []
        for _ in range(n_iter):
            t0 = time.perf_counter()
            r = self.integrand.iteration(self.n_events, self.iteration_index, train=not self.frozen)
            r["seconds"] = time.perf_counter() - t0
            self.iteration_index += 1
            res.append(r)
            self.history.append(r)
            if log:
                log("iteration %d: %.6g +- %.3g  (%.2fs, %d/%d pass)" % (
                    self.iteration_index, r["I"], r["sigma"], r["seconds"], r["npass"], r["n"]))
        return combine(res)


def combine(results):
    """1/sigma^2-weighted combination of iterations -> (result, error)."""
    w = [1.0 / r["var"] for r in results if r["var"] > 0]
    if not w:
        return results[-1]["I"], 0.0
    num = math.fsum(r["I"] / r["var"] for r in results if r["var"] > 0)
    den = math.fsum(w)
    return num / den, math.sqrt(1.0 / den)
