"""Device-side phase-space helpers over the C ABI (Philox uniforms + RAMBO).

``rambo(xrand, n, sqrts, masses)`` mirrors ``python_package/madflow/phasespace.py:145-212``
(returns momenta (nev, n, 4) and weights (nev,)); ``sample_momenta`` builds full
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
