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
    s = torch.full((1,), float(sqrts), dtype=torch.float64, device=dev)
    p = torch.empty((nev, n, 4), dtype=torch.float64, device=dev)
    w = torch.empty(nev, dtype=torch.float64, device=dev)
    m = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
    _lib.check(lib.mfx_rambo(ctypes.c_void_p(xt.data_ptr()), nev, x.shape[1], n, ctypes.c_void_p(s.data_ptr()), 0,
                             m.ctypes.data_as(ctypes.c_void_p) if m is not None else None,
                             ctypes.c_void_p(p.data_ptr()), ctypes.c_void_p(w.data_ptr()), _lib.stream_ptr(dev)))
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
