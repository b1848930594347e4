"""Multi-channel phase space with one VEGAS grid per channel (no reference counterpart:
``doc/source/usage.rst:84-99`` names RAMBO + VEGAS as madflow's only phase space).

The channels of g g > t t~ + gluons (and u u~ > t t~ + gluons) at fixed sqrt(s):

* flat: RAMBO on diet over every outgoing leg (the plain ``mapping="diet"`` sampling);
* one t-channel per outgoing gluon: its pT ~ 1/pT^2 from the pT cut to the kinematic
  limit, rapidity and azimuth flat, the other legs by diet (initial-state emission);
* one s-channel per gluon pair: m^2 ~ 1/m^2 from the smallest mass the pT and Delta R
  cuts allow, isotropic decay (final-state g -> g g splitting).

Each iteration draws n_c = alpha_c N points from channel c (stratified), each with weight
f / g, g = sum_j alpha_j g_j (Kleiss and Pittau, Comput. Phys. Commun. 83 (1994) 141); the
kernels evaluate every channel's density at every point through the inverse maps.  Grid c
is refined from the histogram of its own points (Ohl, Comput. Phys. Commun. 120 (1999)
13), and the weights follow alpha_c <- alpha_c sqrt(W_c), W_c = <(f/g)^2 g_c / g>, with a
floor alpha_min so that no channel dies.  The kernels are ``mfx_mc_generate`` and
``mfx_mc_moments``; smatrix and the accumulation are the plain integrand's.
"""

import ctypes
import math

import numpy as np
import torch

from . import _lib
from .vegas import MEIntegrand, _merge, _ptr, ALPHA, PROCESS_SETUP


class McChannel(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("a", ctypes.c_int), ("b", ctypes.c_int), ("reserved", ctypes.c_int),
                ("lo", ctypes.c_double), ("hi", ctypes.c_double)]


def _mc_sig(lib):
    vp, i64, u64, u32, i, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_double
    from .vegas import PsParams
    lib.mfx_mc_generate.argtypes = [ctypes.POINTER(PsParams), vp, i, vp, d, i, vp, u64, i64, u32, vp, vp, vp, vp, vp,
                                    vp, vp, vp, i64, ctypes.POINTER(i64), vp]
    lib.mfx_mc_generate.restype = i
    lib.mfx_mc_generate_points.argtypes = [ctypes.POINTER(PsParams), vp, i, vp, d, i, vp, u64, i64, u32, vp, vp, vp, vp,
                                           vp, vp, vp, vp, vp, vp, i64, ctypes.POINTER(i64), vp]
    lib.mfx_mc_generate_points.restype = i
    lib.mfx_mc_moments.argtypes = [i64, i, vp, vp, vp, vp, vp, i64, vp, vp]
    lib.mfx_mc_moments.restype = i
    lib.mfx_mc_generate_learned.argtypes = [ctypes.POINTER(PsParams), vp, i, vp, d, i, vp, u64, i64, u32, vp, vp, vp,
                                            vp, vp, vp, vp, vp, vp, vp, i64, ctypes.POINTER(i64), vp]
    lib.mfx_mc_generate_learned.restype = i
    lib.mfx_mc_combine.argtypes = [i64, i, d, vp, vp, vp, vp, vp]
    lib.mfx_mc_combine.restype = i


def default_channels(masses, sqrts, pt_min=0.0, dr_min=0.0):
    """[(type, a, b, lo, hi)] for outgoing masses: flat, a t-channel per massless leg, an
    s-channel per pair of massless legs.  The lower ends follow the cuts (pT > pt_min;
    m_ab^2 >= (pT dR)^2 roughly, with 10% margin), else 1e-3 sqrt(s)."""
    masses = [float(m) for m in masses]
    s = float(sqrts) ** 2
    gl = [i for i, m in enumerate(masses) if m == 0.0]
    out = [(0, 0, 0, 0.0, 0.0)]
    emax = (s - sum(masses) ** 2) / (2.0 * sqrts)
    pt_lo = pt_min if pt_min > 0 else 1e-3 * sqrts
    for a in gl:
        out.append((1, a, 0, pt_lo, emax))
    m_lo = 0.9 * pt_min * dr_min if pt_min > 0 and dr_min > 0 else 1e-3 * sqrts
    for k, a in enumerate(gl):
        for b in gl[k + 1:]:
            others = sum(masses) - masses[a] - masses[b]
            out.append((2, a, b, m_lo * m_lo, (sqrts - others) ** 2))
    return out


class MultiChannelIntegrand(MEIntegrand):
    """MEIntegrand over the multi-channel phase space; the keywords are MEIntegrand's.
    mapping "diet" (the default): fixed sqrt(s), the channels on the diet cube of
    3 n_out - 4 dimensions.  mapping "hadronic": the hadronic sqrt(s), with two more
    dimensions in front that give x1, x2 as in MEIntegrand's hadronic mapping, the
    channels at sqrt(x1 x2 s) (their ranges capped by the kinematic limit there), and the
    luminosity, scale and running alpha_s of MEIntegrand's hadronic mode.  ``channels``
    overrides default_channels (built at the given sqrt(s)); ``adapt_alpha`` switches the
    channel-weight adaptation; ``alpha_min`` is the weight floor."""

    def __init__(self, process, sqrts, channels=None, adapt_alpha=True, alpha_min=0.02, **kw):
        if not isinstance(process, str):
            raise NotImplementedError("the multi-channel phase space takes one process")
        mapping = kw.get("mapping", "diet")
        if mapping not in ("diet", "hadronic"):
            raise ValueError("the multi-channel phase space runs on the diet cube: mapping 'diet' (fixed sqrt(s)) "
                             "or 'hadronic'")
        kw["mapping"] = mapping
        super().__init__(process, sqrts, **kw)
        _mc_sig(self.lib)
        n_out, masses = PROCESS_SETUP[process]
        if self.hadronic:
            # x1, x2, then the diet cube (the library's mapping 3)
            self.p.mapping = 3
            self.p.ndim = 3 * n_out - 2
            self.ndim = self.p.ndim
            self.grid = torch.linspace(0, 1, self.nbins + 1, dtype=torch.float64,
                                       device=self.device).repeat(self.ndim, 1).contiguous()
        pt_min, dr_min = kw.get("pt_min", 0.0), kw.get("dr_min", 0.0)
        self.channels = list(channels) if channels is not None else default_channels(masses, sqrts, pt_min, dr_min)
        self.nch = len(self.channels)
        if not 1 <= self.nch <= 8:
            raise ValueError("between 1 and 8 channels")
        arr = (McChannel * self.nch)()
        for c, (t, a, b, lo, hi) in enumerate(self.channels):
            arr[c].type, arr[c].a, arr[c].b, arr[c].lo, arr[c].hi = t, a, b, lo, hi
        self._chan_arr = arr
        self.alpha = np.full(self.nch, 1.0 / self.nch)
        self.adapt_alpha = adapt_alpha
        self.alpha_min = float(alpha_min)
        self.grids = self.grid.unsqueeze(0).repeat(self.nch, 1, 1).contiguous()
        self._ratios = None
        self.history_alpha = []
        self.flow = None          # learned channel (attach_flow)
        self.alpha_learned = 0.0
        self.channel_flows = {}   # channel -> flow on that channel's cube (attach_channel_flow)
        self.channel_flow_mix = 0.1

    def _alpha_all(self):
        """Weights of every channel of the mixture: the analytic ones, then the learned one."""
        if self.flow is None:
            return self.alpha.copy()
        return np.concatenate([(1.0 - self.alpha_learned) * self.alpha, [self.alpha_learned]])

    def _counts(self, alpha, n_events):
        """Points per channel: alpha_c N, rounded, at least 1000, summing to N."""
        k = len(alpha)
        n = np.maximum(np.floor(alpha * n_events).astype(np.int64), min(1000, n_events // k))
        n[np.argmax(alpha)] += n_events - int(n.sum())
        return n

    # ---------------------------------------------------------------- learned channel
    def cube0_chunk(self, ev0, n, iteration, keep_f=False, y_in=None, w_in=None):
        """MEIntegrand.chunk on channel 0's cube alone (its frozen grid, the flat map, no
        mixture weight): the integrand values that train a learned density on that cube."""
        if self.channels[0][0] != 0:
            raise ValueError("the learned channel needs channel 0 to be the flat one")
        saved = self.grid
        self.grid = self.grids[0].contiguous()
        try:
            return MEIntegrand.chunk(self, ev0, n, iteration, keep_f=keep_f, y_in=y_in, w_in=w_in)
        finally:
            self.grid = saved

    def attach_flow(self, flow, alpha_learned=0.5):
        """Add a learned density on channel 0's cube as one more channel of the mixture,
        with weight alpha_learned (the analytic channels share the rest in their current
        proportions).  Grids are not refined while it is attached."""
        if self.channels[0][0] != 0:
            raise ValueError("the learned channel needs channel 0 to be the flat one")
        if self.hadronic:
            raise ValueError("the learned channel runs at fixed sqrt(s)")
        if self.channel_flows:
            raise ValueError("channel flows are attached")
        if self.nch + 1 > 8:
            raise ValueError("at most 7 analytic channels beside the learned one")
        if flow.D != self.ndim:
            raise ValueError("the flow has %d dimensions, the cube %d" % (flow.D, self.ndim))
        if not 0.0 < alpha_learned < 1.0:
            raise ValueError("alpha_learned must lie in (0, 1)")
        self.flow = flow
        self.alpha_learned = float(alpha_learned)

    def attach_channel_flow(self, c, flow, mix=0.1):
        """Draw channel c's points from ``flow`` on its cube (before its frozen grid) in a
        defensive mixture q' = (1 - mix) q + mix with the uniform density; each weight is
        multiplied by 1 / q'(y), so the stratified estimate stays unbiased while the
        channel weights alpha_j g_j / sum alpha g keep their meaning.  While a channel flow
        is attached the channel weights stay as they are and so does that channel's grid
        (the flow was trained on it); training iterations still refine the other grids."""
        if not 0 <= c < self.nch:
            raise ValueError("channel %d out of range" % c)
        if self.flow is not None:
            raise ValueError("a learned channel is attached")
        if flow.D != self.ndim:
            raise ValueError("the flow has %d dimensions, the cube %d" % (flow.D, self.ndim))
        if not 0.0 < mix < 1.0:
            raise ValueError("mix must lie in (0, 1)")
        self.channel_flows[c] = flow
        self.channel_flow_mix = float(mix)

    def _chunk_learned(self, c, ev0, n, iteration, scale, keep_f):
        """One chunk of channel c of the mixture with the learned channel (c == nch)."""
        from .phasespace_gpu import philox_uniforms
        b = self._alloc(n)
        ncol = self.nch + 1
        if self._ratios is None or self._ratios.numel() < n * ncol:
            self._ratios = torch.empty((n, ncol), dtype=torch.float64, device=self.device)
        if getattr(self, "_y0", None) is None or self._y0.shape[0] < n:
            self._y0 = torch.empty((n, self.ndim), dtype=torch.float32, device=self.device)
            self._gd = torch.empty((n, 2), dtype=torch.float64, device=self.device)
        s = _lib.stream_ptr(self.device)
        y_in = None
        if c == self.nch:
            u = philox_uniforms(n, self.ndim, self.p.seed, iteration, ev0=ev0, dtype=torch.float32)
            y_in, _ = self.flow.sample(u, logq=False)
            y_in = y_in.contiguous()
        alpha = np.ascontiguousarray((1.0 - self.alpha_learned) * self.alpha, dtype=np.float64)
        npass = ctypes.c_int64(0)
        _lib.check(self.lib.mfx_mc_generate_learned(
            ctypes.byref(self.p), ctypes.cast(self._chan_arr, ctypes.c_void_p), self.nch,
            alpha.ctypes.data_as(ctypes.c_void_p), float(scale), 0 if c == self.nch else c, _ptr(self.grids), ev0, n,
            iteration, _ptr(y_in), _ptr(b["all_ps"]), _ptr(b["wps"]), _ptr(b["flag"]), _ptr(b["slot"]),
            _ptr(b["dense"]), _ptr(self._ratios), _ptr(self._y0), _ptr(self._gd), _ptr(b["ws"]), b["ws"].numel(),
            ctypes.byref(npass), s))
        lq = self.flow.log_prob(self._y0[:n].contiguous())
        _lib.check(self.lib.mfx_mc_combine(n, self.nch, self.alpha_learned, _ptr(lq), _ptr(self._gd), _ptr(b["wps"]),
                                           _ptr(self._ratios), s))
        npass = npass.value
        self.last_chunk_npass = npass
        me = self._smatrix_timed(b, npass)
        _lib.check(self.lib.mfx_vegas_accumulate(
            ctypes.byref(self.p), ev0, n, iteration, _ptr(y_in), _ptr(b["slot"]), _ptr(me), _ptr(b["wps"]), None,
            _ptr(b["f"]) if keep_f else None, _ptr(b["ws"]), b["ws"].numel(), _ptr(b["result"]), s))
        mom = torch.empty(ncol, dtype=torch.float64, device=self.device)
        _lib.check(self.lib.mfx_mc_moments(n, ncol, _ptr(b["slot"]), _ptr(me), _ptr(b["wps"]), _ptr(self._ratios),
                                           _ptr(b["ws"]), b["ws"].numel(), _ptr(mom), s))
        return (b["result"].cpu().numpy().copy(), npass, mom.cpu().numpy(),
                (b["f"][:n].clone() if keep_f else None))

    def _smatrix_timed(self, b, npass, params=None):
        params = self.params if params is None else params
        if self.me_timer is not None and npass:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            me = self.matrix.smatrix(b["dense"][:npass], *params)
            e1.record()
            self.me_timer.append((e0, e1, npass))
            return me
        # no passing event: every slot is -1 and the matrix elements are not read, but the
        # library needs a device pointer, which an empty tensor does not have
        return self.matrix.smatrix(b["dense"][:npass], *params) if npass else b["wps"][:1]

    def chunk_channel(self, c, ev0, n, iteration, scale, keep_f=False, y_in=None, w_in=None):
        """One chunk of channel c: (accumulated result vector, npass, moments (nch,), f).
        y_in / w_in: the points on channel c's cube and 1 / their density there (a proposal
        other than the uniform one); by default the Philox uniforms."""
        b = self._alloc(n)
        if self._ratios is None or self._ratios.shape[0] < n:
            self._ratios = torch.empty((n, self.nch), dtype=torch.float64, device=self.device)
        s = _lib.stream_ptr(self.device)
        npass = ctypes.c_int64(0)
        alpha = np.ascontiguousarray(self.alpha, dtype=np.float64)
        if y_in is None:
            _lib.check(self.lib.mfx_mc_generate(
                ctypes.byref(self.p), ctypes.cast(self._chan_arr, ctypes.c_void_p), self.nch,
                alpha.ctypes.data_as(ctypes.c_void_p), float(scale), c, _ptr(self.grids), ev0, n, iteration,
                _ptr(b["all_ps"]), _ptr(b["wps"]), _ptr(b["flag"]), _ptr(b["slot"]), _ptr(b["dense"]),
                _ptr(self._ratios), _ptr(b["x12"]), _ptr(b["ws"]), b["ws"].numel(), ctypes.byref(npass), s))
        else:
            y_in = _lib.require_device_tensor(y_in, torch.float32, "y_in")
            w_in = None if w_in is None else _lib.require_device_tensor(w_in, torch.float64, "w_in")
            _lib.check(self.lib.mfx_mc_generate_points(
                ctypes.byref(self.p), ctypes.cast(self._chan_arr, ctypes.c_void_p), self.nch,
                alpha.ctypes.data_as(ctypes.c_void_p), float(scale), c, _ptr(self.grids), ev0, n, iteration,
                _ptr(y_in), _ptr(w_in), _ptr(b["all_ps"]), _ptr(b["wps"]), _ptr(b["flag"]), _ptr(b["slot"]),
                _ptr(b["dense"]), _ptr(self._ratios), _ptr(b["x12"]), _ptr(b["ws"]), b["ws"].numel(),
                ctypes.byref(npass), s))
        npass = npass.value
        self.last_chunk_npass = npass
        if self.hadronic and npass:
            # couplings at each event's scale, and the parton luminosity folded into the
            # matrix elements of the passing events (compact order)
            params, lumi = self._hadronic_factors(b, n, npass)
            me = self._smatrix_timed(b, npass, params) * lumi
        else:
            me = self._smatrix_timed(b, npass)
        _lib.check(self.lib.mfx_vegas_accumulate(
            ctypes.byref(self.p), ev0, n, iteration, _ptr(y_in), _ptr(b["slot"]), _ptr(me), _ptr(b["wps"]), None,
            _ptr(b["f"]) if keep_f else None, _ptr(b["ws"]), b["ws"].numel(), _ptr(b["result"]), s))
        mom = torch.empty(self.nch, dtype=torch.float64, device=self.device)
        _lib.check(self.lib.mfx_mc_moments(n, self.nch, _ptr(b["slot"]), _ptr(me), _ptr(b["wps"]),
                                           _ptr(self._ratios), _ptr(b["ws"]), b["ws"].numel(), _ptr(mom), s))
        return (b["result"].cpu().numpy().copy(), npass, mom.cpu().numpy(),
                (b["f"][:n].clone() if keep_f else None))

    @staticmethod
    def channel_offset(c):
        """Philox counter offset of channel c's points (disjoint streams per channel)."""
        return c << 40

    def _chunk_any(self, c, ev0, n, iteration, scale, keep_f):
        """One chunk of channel c of whatever the mixture holds: an analytic channel (drawn
        uniformly on its cube, or from its attached flow), or the learned channel."""
        if self.flow is not None:
            return self._chunk_learned(c, ev0, n, iteration, scale, keep_f)
        if c in self.channel_flows:
            # the channel's own flow on its cube (defensive mixture with the uniform)
            from .phasespace_gpu import philox_uniforms
            u = philox_uniforms(n, self.ndim, self.p.seed, iteration, ev0=ev0, dtype=torch.float32)
            ys, winv = self.channel_flows[c].sample_mixture(u, u, self.channel_flow_mix)
            return self.chunk_channel(c, ev0, n, iteration, scale, keep_f, y_in=ys, w_in=winv)
        return self.chunk_channel(c, ev0, n, iteration, scale, keep_f)

    def unweighted_events(self, n_events, iteration, wmax):
        """Unweighted events of the multi-channel phase space: n_events points shared among
        the channels in the proportions of the current channel weights (as one iteration
        would draw them), each with its stratified weight (the weights whose mean over all
        points is the estimate of sigma), kept with probability min(1, w / wmax) on the
        device.  Returns (momenta of the kept events, (K, nexternal, 4) host array, their
        weights, the number of kept events with w > wmax), as MEIntegrand.unweighted_events.
        Points, weights and keep decisions depend only on the seed, the iteration and the
        event index."""
        alpha_used = self._alpha_all()
        counts = self._counts(alpha_used, int(n_events))
        N = float(n_events)
        kept, wts, over = [], [], 0
        for c in range(len(alpha_used)):
            nc = int(counts[c])
            scale = alpha_used[c] * N / nc
            ev = 0
            while ev < nc:
                n = min(self.events_limit, nc - ev)
                ev0 = self.channel_offset(c) + ev
                _, _, _, f = self._chunk_any(c, ev0, n, iteration, scale, True)
                idx, _ = self.unweight(f, wmax, iteration, ev0)
                il = idx.long()
                kept.append(self._bufs["all_ps"][:n].index_select(0, il).cpu().numpy())
                fw = f.index_select(0, il)
                wts.append(fw.cpu().numpy())
                over += int((fw > wmax).sum())
                ev += n
        return np.concatenate(kept), np.concatenate(wts), over

    def chunk(self, *args, **kw):
        """Not available here: a chunk belongs to one channel (chunk_channel); the flat
        channel's cube alone is cube0_chunk."""
        raise NotImplementedError("MultiChannelIntegrand: use chunk_channel(c, ...) or cube0_chunk")

    def proposal_iteration(self, *args, **kw):
        """Not available here: a learned density joins the multi-channel phase space as a
        channel (attach_flow) or as one channel's proposal (attach_channel_flow)."""
        raise NotImplementedError("MultiChannelIntegrand: use attach_flow or attach_channel_flow, then iteration()")

    def iteration(self, n_events, iteration, train=True, keep_f=False):
        """One iteration of n_events points over the channels; the same dict as
        MEIntegrand.iteration plus the channel weights used (``alpha``)."""
        learned = self.flow is not None
        alpha_used = self._alpha_all()
        counts = self._counts(alpha_used, int(n_events))
        N = float(n_events)
        K = len(alpha_used)
        I_parts, var_parts, wmax, nonzero, npass_tot = [], [], 0.0, 0.0, 0
        moments = np.zeros(K)
        fs = []
        for c in range(K):
            nc = int(counts[c])
            scale = alpha_used[c] * N / nc
            tot = None
            ev = 0
            while ev < nc:
                n = min(self.events_limit, nc - ev)
                r, k, mom, f = self._chunk_any(c, self.channel_offset(c) + ev, n, iteration, scale, keep_f)
                tot = r if tot is None else _merge(tot, r)
                moments += mom / N
                npass_tot += k
                if keep_f:
                    fs.append(f)
                ev += n
            # stratum c: the mean of its f estimates alpha_c * N / n_c * int g_c f / g
            mean_c = tot[0] / nc
            var_c = max((tot[1] / nc - mean_c * mean_c) / (nc - 1.0), 0.0)
            I_parts.append(tot[0] / N)
            var_parts.append(var_c * (nc / N) ** 2)
            wmax = max(wmax, tot[2])
            nonzero += tot[3]
            if train and not learned and c not in self.channel_flows:
                hist = torch.from_numpy(tot[6:].copy()).to(self.device)
                _lib.check(self.lib.mfx_vegas_refine(_ptr(self.grids[c]), _ptr(hist), self.ndim, self.nbins, ALPHA,
                                                     _lib.stream_ptr(self.device)))
        I = math.fsum(I_parts)
        var = math.fsum(var_parts)
        if train and self.adapt_alpha and not self.channel_flows and np.all(np.isfinite(moments)) and moments.sum() > 0:
            a = alpha_used * np.sqrt(np.maximum(moments, 0.0))
            a = a / a.sum()
            a = np.maximum(a, self.alpha_min)
            a = a / a.sum()
            if learned:
                self.alpha_learned = float(a[-1])
                self.alpha = a[:-1] / a[:-1].sum()
            else:
                self.alpha = a
        self.history_alpha.append(alpha_used)
        out = dict(I=I, var=var, sigma=math.sqrt(var), wmax=wmax, nonzero=nonzero, npass=npass_tot, n=int(n_events),
                   sum_f=I * N, alpha=alpha_used.tolist(), counts=counts.tolist(),
                   var_parts=[float(v) for v in var_parts])
        if keep_f:
            out["f"] = torch.cat(fs)
        return out
