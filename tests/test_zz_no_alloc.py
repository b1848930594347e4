"""mfx_smatrix_noalloc (the smatrix entry point that allocates nothing: its jamp staging
lives in a fixed device array compiled into the library) against mfx_smatrix with a
caller-allocated workspace, through the C ABI.

Two calls in a row on each case, the second with a smaller batch, so a stale scratch
array from the first call would show.  The first batch is larger than one pass through
the 256 MiB scratch (several sub-chunks and helicity groups).  Bar: 1e-12 relative per
event (both entry points run the same kernels; only the order in which helicity groups
are added in FP64 can differ)."""

import ctypes

import numpy as np
import pytest
import torch

from madflow_amd import _lib
from oracle import oracle as O
from tests.helpers import phase_space_points

CASES = [("gg_ttxgg", 32, 200_000, 3_000), ("gg_ttxgg", 64, 90_000, 1_000), ("gg_ttxggg", 32, 20_000, 500),
         ("gg_ttx", 64, 5_000, 100)]


def _call(lib, fn, proc, ps, mw, gc, precision, ws=None):
    nev = ps.shape[0]
    out = torch.full((nev,), float("nan"), dtype=torch.float64, device=ps.device)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    stream = _lib.stream_ptr(ps.device)
    if fn == "noalloc":
        rc = lib.mfx_smatrix_noalloc(proc, ctypes.c_void_p(ps.data_ptr()), nev, vp(mw), vp(gc), None,
                                     ctypes.c_void_p(out.data_ptr()), precision, 1, stream)
    else:
        rc = lib.mfx_smatrix(proc, ctypes.c_void_p(ps.data_ptr()), nev, vp(mw), vp(gc), None,
                             ctypes.c_void_p(out.data_ptr()), precision, 1,
                             ctypes.c_void_p(ws.data_ptr()) if ws is not None else None,
                             ws.numel() if ws is not None else 0, stream)
    _lib.check(rc)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name,precision,n1,n2", CASES)
def test_noalloc_matches_smatrix(dev, name, precision, n1, n2):
    lib = _lib.load()
    assert lib.mfx_smatrix_noalloc_scratch_bytes() == 256 << 20
    proc = lib.mfx_process_index(name.encode())
    assert proc >= 0
    rp, coup = O.default_params(ref_gs=True)
    mw = np.array([rp[0], rp[1]])
    gc = np.zeros(6)
    for slot, c in enumerate(("GC_10", "GC_11", "GC_12")):
        gc[2 * slot], gc[2 * slot + 1] = coup[c].real, coup[c].imag
    ps_all, _ = phase_space_points(name, n1, seed=41)
    for call, n in enumerate((n1, n2)):
        print("case noalloc %s fp%d call %d nev %d" % (name, precision, call + 1, n), flush=True)
        ps = torch.from_numpy(ps_all[-n:].copy()).to(dev)
        need = int(lib.mfx_smatrix_workspace_size(proc, n, precision, 0))
        ws = torch.empty(min(need, 2 << 30), dtype=torch.uint8, device=dev) if need > 0 else None
        ref = _call(lib, "smatrix", proc, ps, mw, gc, precision, ws).cpu().numpy()
        got = _call(lib, "noalloc", proc, ps, mw, gc, precision).cpu().numpy()
        assert np.all(np.isfinite(got))
        rel = np.max(np.abs(got / ref - 1))
        print("noalloc %s fp%d nev %d: max rel %.3e" % (name, precision, n, rel), flush=True)
        assert rel < 1e-12
    if precision == 64:
        # and the FP64 result is the oracle's
        sub = ps_all[-n2:][:64]
        orc = O.smatrix(name, sub, rp, coup, ref_compat=True)
        np.testing.assert_allclose(got[:64], orc, rtol=1e-12, atol=0)
