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


@pytest.mark.gpu
def test_phase_space_generator_cuts(dev):
    """PhaseSpaceGenerator and register_cut checked as the reference's own test checks them
    (test_ps.py:40-64): 5 particles at 7 TeV, pT(particle 3) in (60, 300); the generator's
    pt equals numpy's on the uncut momenta, and the cut output equals the uncut momenta
    under the numpy mask, with idx the indices of the passing events."""
    from madflow_amd.phasespace_gpu import PhaseSpaceGenerator, philox_uniforms
    print("case PhaseSpaceGenerator cuts", flush=True)
    gen = PhaseSpaceGenerator(5, 7e3)
    gen.register_cut("pt", particle=3, min_val=60, max_val=300.0)
    xrand = philox_uniforms(10000, gen.ndim, 11, 0)
    all_ps, w, x1, x2, idx = gen(xrand)
    gen.clear_cuts()
    full_ps, full_w, fx1, fx2, one = gen(xrand)
    assert one == 1
    internal_pt = gen.pt(full_ps[:, 3, :]).cpu().numpy()
    full_np = full_ps.cpu().numpy()
    numpy_pt = np.sqrt(full_np[:, 3, 1] ** 2 + full_np[:, 3, 2] ** 2)
    np.testing.assert_allclose(numpy_pt, internal_pt)
    mask = np.all([numpy_pt > 60.0, numpy_pt < 300.0], axis=0)
    assert 0.05 < mask.mean() < 0.95
    np.testing.assert_allclose(all_ps.cpu().numpy(), full_np[mask])
    np.testing.assert_allclose(w.cpu().numpy(), full_w.cpu().numpy()[mask])
    np.testing.assert_array_equal(idx.cpu().numpy()[:, 0], np.nonzero(mask)[0])
    np.testing.assert_allclose(x1.cpu().numpy(), fx1.cpu().numpy()[mask])


@pytest.mark.gpu
@pytest.mark.parametrize("com_output", [True, False])
def test_ramboflow_matches_oracle(dev, com_output):
    """ramboflow / PhaseSpaceGenerator on the caller's FP64 uniforms equal the oracle's
    hadronic generation (mfx_ps_generate's restatement) on the same uniforms through an
    identity VEGAS grid: momenta, weights (flux and GeV^-2 -> pb included) and x1, x2
    within 1e-12, in the c.m. frame and boosted to the lab (com_output=False); the
    reference's checks of test_ps.py:67-78 hold (massless incoming legs, outgoing masses
    50 and 125 GeV on shell within 1e-8)."""
    from madflow_amd.phasespace_gpu import PhaseSpaceGenerator, ramboflow
    print("case ramboflow parity com_output", com_output, flush=True)
    masses = [50.0, 125.0]
    n, nb = 2000, 50
    cfg = dict(n_out=2, ndim=10, nbins=nb, mapping=2, sqrts=7000.0, masses=masses, gev2pb=389379365.6, pt_min=0.0,
               abs_eta_max=0.0, dr_min=0.0, cut_first_leg=4, seed=13, shat_min=175.0 ** 2,
               lab_frame=0 if com_output else 1)
    edges = np.tile(np.linspace(0.0, 1.0, nb + 1), (10, 1))
    ref_ps, ref_w, ref_flag, ref_x12 = O.generate(cfg, edges, 0, n, 2)
    y = O.philox_uniforms(n, 10, 13, 2)
    xrand = torch.from_numpy(y).to(dev)
    gen = PhaseSpaceGenerator(4, 7000.0, masses=masses, com_output=com_output)
    ps, w, x1, x2, idx = gen(xrand)
    assert idx == 1 and np.all(ref_flag == 1)
    np.testing.assert_allclose(ps.cpu().numpy(), ref_ps, rtol=1e-12, atol=1e-12 * 7000.0)
    np.testing.assert_allclose(w.cpu().numpy(), ref_w, rtol=1e-12)
    np.testing.assert_allclose(x1.cpu().numpy(), ref_x12[:, 0], rtol=1e-12)
    np.testing.assert_allclose(x2.cpu().numpy(), ref_x12[:, 1], rtol=1e-12)
    m2 = lambda q: q[..., 0] ** 2 - np.sum(q[..., 1:] ** 2, axis=-1)
    p = ps.cpu().numpy()
    # on shell to the massive RAMBO's Newton residual (worst 1.4e-9 here; the reference's
    # test allows 1e-4)
    np.testing.assert_allclose(m2(p[:, 0:2]), 0.0, atol=1e-6)
    np.testing.assert_allclose(m2(p[:, 2]), 50.0 ** 2, rtol=1e-8)
    np.testing.assert_allclose(m2(p[:, 3]), 125.0 ** 2, rtol=1e-8)
    if com_output:
        cps, cw, cx1, cx2 = ramboflow(xrand, 4, 7000.0, masses)
        np.testing.assert_array_equal(cps.cpu().numpy(), p)
        np.testing.assert_array_equal(cw.cpu().numpy(), w.cpu().numpy())
