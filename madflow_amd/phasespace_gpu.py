"""Device-side phase-space helpers over the C ABI (Philox uniforms + RAMBO).

``rambo(xrand, n, sqrts, masses)`` mirrors ``python_package/madflow/phasespace.py:145-212``
(returns momenta (nev, n, 4) and weights (nev,)); ``ramboflow`` and
``PhaseSpaceGenerator`` mirror phasespace.py:255-317 and 359-520 (the hadronic mapping,
register_cut's cuts, compaction and the lab-frame boost); ``sample_momenta`` builds full
``all_ps`` tensors (incoming gluons along +-z at fixed partonic sqrt(s)) for the
generated processes.
"""

import ctypes

import numpy as np
import torch

from . import _lib

PROCESS_KIN = {
    # name: (sqrt(s) of the benchmark configuration, outgoing masses)
    "gg_ttx": (1000.0, [173.0, 173.0]),
    "gg_ttxg": (2000.0, [173.0, 173.0, 0.0]),
    "gg_ttxgg": (3000.0, [173.0, 173.0, 0.0, 0.0]),
    "gg_ttxggg": (5000.0, [173.0, 173.0, 0.0, 0.0, 0.0]),
    "uux_ttx": (1000.0, [173.0, 173.0]),
    "uux_ttxg": (2000.0, [173.0, 173.0, 0.0]),
    "uux_ttxgg": (3000.0, [173.0, 173.0, 0.0, 0.0]),
    "ug_ttxu": (2000.0, [173.0, 173.0, 0.0]),
    "uxg_ttxux": (2000.0, [173.0, 173.0, 0.0]),
}


def philox_uniforms(nev, ndim, seed, iteration, ev0=0, device="cuda", dtype=torch.float64):
    """(nev, ndim) uniforms of the Philox stream (seed, iteration) for events ev0..ev0+nev-1,
    float64 or (dtype=torch.float32) the same values rounded to float32."""
    out = torch.empty((nev, ndim), dtype=dtype, device=device)
    if nev:
        fn = _lib.load().mfx_philox_uniforms_f32 if dtype == torch.float32 else _lib.load().mfx_philox_uniforms
        if dtype not in (torch.float32, torch.float64):
            raise TypeError("philox_uniforms: dtype must be float32 or float64")
        _lib.check(fn(ev0, nev, ndim, seed, iteration, ctypes.c_void_p(out.data_ptr()), _lib.stream_ptr(out.device)))
    return out


def rambo(xrand, n_particles, sqrts, masses=None, check_physical=False, return_status=False):
    """RAMBO on device: xrand (nev, >= 4n) float64 GPU tensor -> (p (nev,n,4), wt (nev,)).

    ``sqrts`` is a number (every event) or a device tensor of per-event values.  A number
    at or below the sum of the masses raises ValueError, as ``rambo(check_physical=True)``
    does (phasespace.py:152-159; the library rejects it whatever check_physical says).  With
    per-event values, ``check_physical`` raises when any event lacks the energy, and
    ``return_status`` adds the per-event status (0 physical, 1 not enough energy, 2 the
    massive Newton iteration did not converge; weight 0 for both)."""
    m = None
    if masses is not None and np.sum(masses) != 0:
        m = np.ascontiguousarray(masses, dtype=np.float64)
        if m.shape != (n_particles,):
            raise ValueError("masses must have one entry per outgoing particle")
    s_ev, s_val = None, 0.0
    if isinstance(sqrts, torch.Tensor) and sqrts.numel() > 1:
        s_ev = _lib.require_device_tensor(sqrts.reshape(-1), torch.float64, "sqrts")
    else:
        s_val = float(sqrts.reshape(()).item() if isinstance(sqrts, torch.Tensor) else sqrts)
        if not s_val > (m.sum() if m is not None else 0.0):
            raise ValueError("Not enough energy (%s) to generate particles of mass: %s" % (
                s_val, list(masses) if masses is not None else []))
    xrand = _lib.require_device_tensor(xrand, torch.float64, "xrand")
    nev = xrand.shape[0]
    dev = xrand.device
    p = torch.empty((nev, n_particles, 4), dtype=torch.float64, device=dev)
    wt = torch.empty(nev, dtype=torch.float64, device=dev)
    status = torch.empty(nev, dtype=torch.int32, device=dev)
    if nev:
        _lib.check(_lib.load().mfx_rambo(
            ctypes.c_void_p(xrand.data_ptr()), nev, xrand.shape[1], n_particles,
            ctypes.c_void_p(s_ev.data_ptr()) if s_ev is not None else None, s_val,
            m.ctypes.data_as(ctypes.c_void_p) if m is not None else None, ctypes.c_void_p(p.data_ptr()),
            ctypes.c_void_p(wt.data_ptr()), ctypes.c_void_p(status.data_ptr()), _lib.stream_ptr(dev)))
    if check_physical and nev and bool((status == 1).any()):
        raise ValueError("Not enough energy to generate particles of mass %s for %d events" % (
            (list(masses) if masses is not None else []), int((status == 1).sum())))
    if return_status:
        return p, wt, status
    return p, wt


def with_incoming(p_out, sqrts):
    """Prepend the two incoming massless partons along +z / -z (phasespace.py:309-313)."""
    nev = p_out.shape[0]
    e = float(sqrts) / 2.0
    pin = torch.zeros((nev, 2, 4), dtype=torch.float64, device=p_out.device)
    pin[:, 0, 0] = e
    pin[:, 0, 3] = e
    pin[:, 1, 0] = e
    pin[:, 1, 3] = -e
    return torch.cat([pin, p_out], dim=1).contiguous()


def sample_momenta(name, nev, seed=1, iteration=0, device="cuda"):
    """All-particle momenta for ``name`` at its benchmark sqrt(s) from Philox + RAMBO."""
    sqrts, masses = PROCESS_KIN[name]
    n = len(masses)
    x = philox_uniforms(nev, 4 * n, seed, iteration, device=device)
    p, _ = rambo(x, n, sqrts, masses)
    return with_incoming(p, sqrts)


class PhaseSpaceGenerator:
    """madflow.phasespace.PhaseSpaceGenerator (phasespace.py:359-520) on the device: the
    ramboflow mapping (x1, x2 from the first two coordinates, RAMBO at sqrt(x1 x2 s) on the
    rest), the cuts registered with register_cut, and the compaction of the passing events,
    in one generation kernel (mfx_ps_generate_xrand).

    ``nparticles`` counts the two incoming partons; ``masses`` are the outgoing masses
    (massless when None); ``com_output`` False boosts the momenta to the lab frame after
    the cuts, as the reference does (phasespace.py:517-518).  Called on xrand, a device
    float64 tensor (nevents, (nparticles - 2) * 4 + 2), it returns (momenta (npass,
    nparticles, 4), weights (npass,), x1, x2, idx): the weights carry the flux 1/(2 shat)
    and the GeV^-2 -> pb factor, and idx holds the indices of the passing events
    (npass, 1) int32 when cuts are registered, else the number 1, as in the reference.
    Events whose massive RAMBO did not converge are dropped like cut ones.  One outgoing
    particle (the reference's on-shell x1, x2 sampling, phasespace.py:291-296) is not
    built."""

    def __init__(self, nparticles, com_sqrts, masses=None, com_output=True, algorithm="ramboflow"):
        from .vegas import GEV2PB, PsParams, _sig
        if masses is None:
            masses = [0.0] * (nparticles - 2)
        if len(masses) != (nparticles - 2):
            raise ValueError("Missmatch in PhaseSpaceGenerator between particles and masses"
                             f" {len(masses)} given for {nparticles - 2} outgoing particles")
        if algorithm != "ramboflow":
            raise ValueError(f"PS algorithm {algorithm} not understood")
        n_out = nparticles - 2
        if not 2 <= n_out <= 8:
            raise NotImplementedError("PhaseSpaceGenerator: between 2 and 8 outgoing particles")
        self._nparticles = nparticles
        self._sqrts = float(com_sqrts)
        self._masses = [float(m) for m in masses]
        self._com_output = bool(com_output)
        self.lib = _lib.load()
        _sig(self.lib)
        p = PsParams()
        p.n_out, p.mapping, p.ndim, p.nbins = n_out, 2, 4 * n_out + 2, 2
        p.sqrts = self._sqrts
        for k, m in enumerate(self._masses):
            p.masses[k] = m
        p.gev2pb = GEV2PB
        p.cut_first_leg = n_out + 2            # only the registered cuts
        p.shat_min = float(sum(self._masses) ** 2)
        p.lab_frame = 0 if self._com_output else 1
        self.p = p

    @property
    def ndim(self):
        return self.p.ndim

    def register_cut(self, variable, particle=None, min_val=None, max_val=None):
        """Keep the events with min_val < variable(particle) < max_val (phasespace.py:424-478):
        variable "pt", "mt" or "mt2" of an outgoing particle (2 .. nparticles - 1), either
        bound None; cuts are ANDed and act in the partonic c.m. frame."""
        from .vegas import add_cut
        add_cut(self.p, self._nparticles, variable, particle, min_val, max_val)

    def clear_cuts(self):
        """Remove every registered cut (phasespace.py:396-402)."""
        self.p.ncuts = 0

    @staticmethod
    def pt(ps_point):
        """Transverse momentum of (..., 4) momenta (phasespace.py:416-421)."""
        return torch.sqrt(ps_point[..., 1] ** 2 + ps_point[..., 2] ** 2)

    @staticmethod
    def mt2(ps_point):
        """Transverse mass squared m^2 + pt^2 of (..., 4) momenta (phasespace.py:405-409)."""
        m2 = ps_point[..., 0] ** 2 - (ps_point[..., 1:] ** 2).sum(dim=-1)
        return m2 + PhaseSpaceGenerator.pt(ps_point) ** 2

    @staticmethod
    def mt(ps_point):
        """Transverse mass (phasespace.py:411-414)."""
        return torch.sqrt(PhaseSpaceGenerator.mt2(ps_point))

    def generate(self, xrand):
        """Every event of xrand: (momenta (nev, nparticles, 4), weights (0 for the events
        that are cut or whose massive RAMBO did not converge), x1, x2, flags (1 = pass))."""
        return self._run(xrand)[:5]

    def _run(self, xrand):
        """generate's outputs, then the passing events' momenta in order and their count."""
        xrand = _lib.require_device_tensor(xrand, torch.float64, "xrand")
        if xrand.dim() != 2 or xrand.shape[1] != self.p.ndim:
            raise ValueError("xrand must be (nevents, %d)" % self.p.ndim)
        nev = xrand.shape[0]
        dev = xrand.device
        all_ps = torch.empty((nev, self._nparticles, 4), dtype=torch.float64, device=dev)
        dense = torch.empty_like(all_ps)
        wps = torch.empty(nev, dtype=torch.float64, device=dev)
        flag = torch.empty(nev, dtype=torch.uint8, device=dev)
        slot = torch.empty(nev, dtype=torch.int32, device=dev)
        x12 = torch.empty((nev, 2), dtype=torch.float64, device=dev)
        ws = torch.empty(self.lib.mfx_integrand_workspace_size(nev, self.p.ndim, self.p.nbins), dtype=torch.uint8,
                         device=dev)
        npass = ctypes.c_int64(0)
        p = lambda t: ctypes.c_void_p(t.data_ptr())
        _lib.check(self.lib.mfx_ps_generate_xrand(
            ctypes.byref(self.p), p(xrand), nev, p(all_ps), p(wps), p(flag), p(slot), p(dense), p(x12), p(ws),
            ws.numel(), ctypes.byref(npass), _lib.stream_ptr(dev)))
        return all_ps, wps, x12[:, 0], x12[:, 1], flag, dense, npass.value

    def __call__(self, xrand):
        all_ps, wps, x1, x2, flag, dense, npass = self._run(xrand)
        ok = flag == 1
        if self.p.ncuts == 0 and bool(ok.all()):
            return all_ps, wps, x1, x2, 1
        idx = torch.nonzero(ok).to(torch.int32)
        return dense[:npass], wps[ok], x1[ok], x2[ok], idx


def ramboflow(xrand, nparticles, com_sqrts, masses=None):
    """madflow.phasespace.ramboflow (phasespace.py:255-317): xrand (nevents,
    (nparticles - 2) * 4 + 2) float64 device tensor -> (momenta (nevents, nparticles, 4) in
    the partonic c.m. frame, weights with the flux 1/(2 shat) and GeV^-2 -> pb, x1, x2),
    for every event (weight 0 where the massive RAMBO did not converge)."""
    gen = PhaseSpaceGenerator(nparticles, com_sqrts, masses)
    all_ps, wps, x1, x2, _ = gen.generate(xrand)
    return all_ps, wps, x1, x2
