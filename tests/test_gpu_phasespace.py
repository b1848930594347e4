"""GPU RAMBO and Philox against the oracle."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O
from madflow_amd import _lib


def _rambo_gpu(x, n, sqrts, masses, dev):
    lib = _lib.load()
    xt = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    nev = x.shape[0]
    p = torch.empty((nev, n, 4), dtype=torch.float64, device=dev)
    w = torch.empty(nev, dtype=torch.float64, device=dev)
    st = torch.full((nev,), -1, dtype=torch.int32, device=dev)
    m = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
    _lib.check(lib.mfx_rambo(ctypes.c_void_p(xt.data_ptr()), nev, x.shape[1], n, None, float(sqrts),
                             m.ctypes.data_as(ctypes.c_void_p) if m is not None else None,
                             ctypes.c_void_p(p.data_ptr()), ctypes.c_void_p(w.data_ptr()),
                             ctypes.c_void_p(st.data_ptr()), _lib.stream_ptr(dev)))
    assert int(st.abs().sum()) == 0
    return p.cpu().numpy(), w.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("n,masses", [(2, [173.0, 173.0]), (3, [173.0, 173.0, 0.0]), (4, [173.0, 173.0, 0, 0]),
                                      (5, None), (7, None)])
def test_rambo_parity(dev, n, masses):
    print("case rambo", n, masses, flush=True)
    rng = np.random.default_rng(n)
    x = rng.random((1000, 4 * n))
    sqrts = 3000.0
    p_ref, w_ref = O.rambo(x, n, sqrts, masses)
    p, w = _rambo_gpu(x, n, sqrts, masses, dev)
    assert np.max(np.abs(p - p_ref)) < 1e-9 * sqrts
    assert np.max(np.abs(w / w_ref - 1)) < 1e-12


@pytest.mark.gpu
def test_philox_bitwise(dev):
    lib = _lib.load()
    nev, ndim = 4099, 13
    out = torch.empty((nev, ndim), dtype=torch.float64, device=dev)
    _lib.check(lib.mfx_philox_uniforms(123456789, nev, ndim, 3, 7, ctypes.c_void_p(out.data_ptr()),
                                       _lib.stream_ptr(dev)))
    ref = O.philox_uniforms(nev, ndim, 3, 7, ev0=123456789)
    assert np.array_equal(out.cpu().numpy(), ref)


@pytest.mark.gpu
def test_rambo_per_event_energy_guard(dev):
    """Per-event sqrt(s): events below the sum of the masses get status 1 with zero
    momenta and weight (the per-event form of phasespace.py:152-159); the others equal
    the oracle's RAMBO at their own sqrt(s)."""
    from madflow_amd.phasespace_gpu import rambo
    print("case rambo per-event energy guard", flush=True)
    n, masses = 3, [173.0, 173.0, 10.0]
    rng = np.random.default_rng(5)
    x = rng.random((512, 4 * n))
    sq = rng.uniform(200.0, 800.0, 512)
    sq[::7] = 356.0                     # exactly the sum of the masses: unphysical too
    p, w, st = rambo(torch.from_numpy(x).to(dev), n, torch.from_numpy(sq).to(dev), masses, return_status=True)
    p, w, st = p.cpu().numpy(), w.cpu().numpy(), st.cpu().numpy()
    bad = sq <= sum(masses)
    assert bad.any() and not bad.all()
    assert np.array_equal(st, np.where(bad, 1, 0))
    assert np.all(w[bad] == 0) and np.all(p[bad] == 0)
    for i in np.flatnonzero(~bad)[:64]:
        pr, wr = O.rambo(x[i:i + 1], n, sq[i], masses)
        assert np.max(np.abs(p[i] - pr[0])) < 1e-9 * sq[i]
        assert abs(w[i] / wr[0] - 1) < 1e-12
    with pytest.raises(ValueError):
        rambo(torch.from_numpy(x).to(dev), n, torch.from_numpy(sq).to(dev), masses, check_physical=True)
