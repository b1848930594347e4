"""GPU VEGAS integrand pipeline against the oracle: generation (Philox -> VEGAS map ->
RAMBO / RAMBO-on-diet -> cuts), compaction, compensated accumulation, refinement, and
the integrated g g > t t~ cross section against the closed form."""

import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from madflow_amd.vegas import MEIntegrand, VegasFlow, GEV2PB


def _toy_pdf(x, q2):
    """A gluon-like x g(x) for tests (no PDF grid is available): 1.7 x^-0.2 (1 - x)^5."""
    return 1.7 * x ** -0.2 * (1.0 - x) ** 5


CASES = [
    ("gg_ttxg", dict(sqrts=2000.0, pt_min=30.0, abs_eta_max=2.5), "rambo"),
    ("gg_ttxgg", dict(sqrts=3000.0, pt_min=30.0, dr_min=0.4), "rambo"),
    ("gg_ttxgg", dict(sqrts=3000.0, pt_min=30.0, dr_min=0.4), "diet"),
    ("gg_ttx", dict(sqrts=1000.0), "diet"),
    ("gg_ttxggg", dict(sqrts=5000.0, pt_min=30.0, dr_min=0.4), "diet"),
    ("gg_ttxggg", dict(sqrts=5000.0, pt_min=30.0, dr_min=0.4), "rambo"),
    ("gg_ttx", dict(sqrts=13000.0, pdf=_toy_pdf), "hadronic"),
    ("gg_ttxgg", dict(sqrts=13000.0, pt_min=30.0, dr_min=0.4, pdf=_toy_pdf, lab_frame=True), "hadronic"),
]


def _cfg(ig):
    p = ig.p
    return dict(n_out=p.n_out, ndim=p.ndim, nbins=p.nbins, mapping=p.mapping, sqrts=p.sqrts,
                masses=list(p.masses), gev2pb=p.gev2pb, pt_min=p.pt_min, abs_eta_max=p.abs_eta_max,
                dr_min=p.dr_min, cut_first_leg=p.cut_first_leg, seed=p.seed, shat_min=p.shat_min,
                lab_frame=p.lab_frame)



def _warp_grid(ig, rng):
    """A non-uniform monotone grid to exercise the VEGAS map."""
    g = np.sort(rng.random((ig.ndim, ig.nbins - 1)), axis=1)
    g = np.concatenate([np.zeros((ig.ndim, 1)), g, np.ones((ig.ndim, 1))], axis=1)
    ig.grid.copy_(torch.from_numpy(g))
    return g


@pytest.mark.gpu
@pytest.mark.parametrize("name,kw,mapping", CASES)
def test_generate_parity(dev, name, kw, mapping):
    print("case generate", name, kw, mapping, flush=True)
    ig = MEIntegrand(name, precision=64, mapping=mapping, seed=5, **kw)
    edges = _warp_grid(ig, np.random.default_rng(1))
    n, ev0, it = 4096, 12345, 3
    ig._alloc(n)
    ig.chunk(ev0, n, it)
    b = ig._bufs
    ref = O.generate(_cfg(ig), edges, ev0, n, it)
    ps_ref, wps_ref, flag_ref = ref[:3]
    if mapping == "hadronic":
        np.testing.assert_allclose(b["x12"][:n].cpu().numpy(), ref[3], rtol=1e-13)
    flag = b["flag"][:n].cpu().numpy()
    assert np.array_equal(flag, flag_ref)
    assert np.max(np.abs(b["all_ps"][:n].cpu().numpy() - ps_ref)) < 1e-9 * kw["sqrts"]
    w = b["wps"][:n].cpu().numpy()
    nz = wps_ref != 0
    assert np.array_equal(w != 0, nz)
    assert np.max(np.abs(w[nz] / wps_ref[nz] - 1)) < 1e-11
    # compaction keeps event order
    slot = b["slot"][:n].cpu().numpy()
    idx = np.nonzero(flag)[0]
    assert np.array_equal(slot[idx], np.arange(len(idx)))
    assert np.all(slot[flag == 0] == -1)
    np.testing.assert_array_equal(b["dense"][:len(idx)].cpu().numpy(), b["all_ps"][:n].cpu().numpy()[idx])


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 63, 65, 4173])
def test_compaction_ragged_and_sparse(dev, n):
    """Compaction at sizes that end inside a wave and a block, under cuts hard enough that
    many waves hold few or no passing events: the flags equal the oracle's, the passing
    events get slots 0, 1, ... in event order, the others -1, and the dense rows are the
    passing rows of all_ps, bit for bit."""
    print("case compaction n=%d" % n, flush=True)
    ig = MEIntegrand("gg_ttxgg", precision=64, mapping="diet", seed=9, sqrts=3000.0, pt_min=400.0, dr_min=0.4)
    edges = _warp_grid(ig, np.random.default_rng(2))
    ev0, it = 777, 2
    ig._alloc(n)
    ig.chunk(ev0, n, it)
    b = ig._bufs
    flag = b["flag"][:n].cpu().numpy()
    assert np.array_equal(flag, O.generate(_cfg(ig), edges, ev0, n, it)[2])
    idx = np.nonzero(flag == 1)[0]
    if n == 4173:
        assert 0.05 * n < len(idx) < 0.8 * n
    slot = b["slot"][:n].cpu().numpy()
    assert np.array_equal(slot[idx], np.arange(len(idx)))
    assert np.all(slot[flag != 1] == -1)
    np.testing.assert_array_equal(b["dense"][:len(idx)].cpu().numpy(), b["all_ps"][:n].cpu().numpy()[idx])


@pytest.mark.gpu
def test_accumulate_and_refine(dev):
    print("case accumulate", flush=True)
    ig = MEIntegrand("gg_ttxgg", precision=32, mapping="diet", seed=9, sqrts=3000.0, pt_min=30.0, dr_min=0.4)
    n = 300001
    rng = np.random.default_rng(3)
    f = rng.lognormal(size=n) * (rng.random(n) > 0.3)
    f[17] = np.nan
    ft = torch.from_numpy(f).to(dev)
    b = ig._alloc(n)
    import ctypes
    from madflow_amd import _lib
    res = []
    for _ in range(2):
        _lib.check(ig.lib.mfx_vegas_accumulate(ctypes.byref(ig.p), 7, n, 2, None, None, None, None,
                                               ctypes.c_void_p(ft.data_ptr()), None, ctypes.c_void_p(b["ws"].data_ptr()),
                                               b["ws"].numel(), ctypes.c_void_p(b["result"].data_ptr()),
                                               _lib.stream_ptr(dev)))
        res.append(b["result"].cpu().numpy().copy())
    assert np.array_equal(res[0], res[1])            # bitwise reproducible
    r = res[0]
    fc = np.where(np.isfinite(f), f, 0.0)
    assert abs(r[0] / math.fsum(fc) - 1) < 1e-14
    assert abs(r[1] / math.fsum(fc * fc) - 1) < 1e-14
    assert r[2] == fc.max() and r[3] == np.count_nonzero(fc)
    y = O.philox_uniforms(n, ig.ndim, ig.p.seed, 2, ev0=7)
    bins = np.minimum((y * ig.nbins).astype(int), ig.nbins - 1)
    hist = np.stack([np.bincount(bins[:, d], weights=fc * fc, minlength=ig.nbins) for d in range(ig.ndim)])
    np.testing.assert_allclose(r[6:].reshape(ig.ndim, ig.nbins), hist, rtol=1e-12)
    # refinement on the device equals the oracle's
    edges0 = ig.grid.cpu().numpy().copy()
    ref = O.vegas_refine(edges0, hist, alpha=1.5)
    hist_t = torch.from_numpy(r[6:].copy()).to(dev)
    _lib.check(ig.lib.mfx_vegas_refine(ctypes.c_void_p(ig.grid.data_ptr()), ctypes.c_void_p(hist_t.data_ptr()),
                                       ig.ndim, ig.nbins, 1.5, _lib.stream_ptr(dev)))
    np.testing.assert_allclose(ig.grid.cpu().numpy(), ref, rtol=1e-12, atol=1e-15)


@pytest.mark.gpu
def test_device_refine_hand_worked(dev):
    """The device refinement on the two cases worked by hand from Lepage's steps (the
    derivations are in tests/test_oracle_vegas.py): 5 bins with histogram (0, 0, 1, 0, 0)
    and 3 bins with histogram (0, 0, 6) at alpha 1.5."""
    import ctypes
    from madflow_amd import _lib
    lib = _lib.load()
    m1 = (0.6 / math.log(2.5)) ** 1.5
    m2 = (0.4 / math.log(5.0 / 3.0)) ** 1.5
    cases = [(np.linspace(0.0, 1.0, 6), [0.0, 0.0, 1.0, 0.0, 0.0], [0.0, 0.32, 0.44, 0.56, 0.68, 1.0]),
             (np.array([0.0, 1 / 3, 2 / 3, 1.0]), [0.0, 0.0, 6.0],
              [0.0, 1 / 3 + (1 + m2 / m1) / 9, 2 / 3 + (2 * (m1 + m2) / 3 - m1) / m2 / 3, 1.0])]
    for edges, hist, want in cases:
        g = torch.from_numpy(np.array([edges])).to(dev)
        h = torch.tensor([hist], dtype=torch.float64, device=dev)
        _lib.check(lib.mfx_vegas_refine(ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(h.data_ptr()), 1, len(hist),
                                        ctypes.c_double(1.5), _lib.stream_ptr(dev)))
        np.testing.assert_allclose(g.cpu().numpy()[0], want, rtol=0, atol=1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("mapping", ["rambo", "diet"])
def test_ttbar_cross_section(dev, mapping):
    """gg -> tt~ at sqrt(s) = 1 TeV, alpha_s = 0.118, m_t = 173, Gamma_t = 0: the closed
    form gives 11.3683 pb; VEGAS with 1e6 points x 5 iterations must agree within 3 sigma
    (the tolerance of the reference's own integration test, test_integration.py:22; a
    fixed-seed 2-sigma bar fails on one seed in twenty by construction)."""
    print("case sigma ttbar", mapping, flush=True)
    ig = MEIntegrand("gg_ttx", precision=64, mapping=mapping, seed=1, sqrts=1000.0, ref_compat=False, width=0.0)
    vf = VegasFlow(ig.ndim, 1000000)
    vf.compile(ig)
    vf.run_integration(2)
    vf.freeze_grid()
    res, err = vf.run_integration(5)
    s, m, a = 1e6, 173.0, 0.118
    rho = 4 * m * m / s
    b = math.sqrt(1 - rho)
    closed = math.pi * a * a / (3 * s) * ((1 + rho + rho * rho / 16) * math.log((1 + b) / (1 - b))
                                          - b * (7 / 4 + 31 * rho / 16)) * GEV2PB
    print("sigma", res, "+-", err, "closed", closed, flush=True)
    assert abs(res - closed) < 3 * err
    assert err / res < 2e-3


@pytest.mark.gpu
def test_unweighting(dev):
    """Device hit-or-miss unweighting: the keep mask equals the oracle's bit for bit, the
    kept indices are the mask's in order, and the kept count matches sum min(1, w/wmax)."""
    print("case unweight", flush=True)
    ig = MEIntegrand("gg_ttxgg", precision=32, mapping="diet", seed=13, sqrts=3000.0, pt_min=30.0, dr_min=0.4)
    rng = np.random.default_rng(5)
    n = 1_000_003
    w = rng.lognormal(sigma=1.5, size=n) * (rng.random(n) > 0.2)
    wmax = np.quantile(w, 0.999)       # some events are over-weight
    idx, keep = ig.unweight(torch.from_numpy(w).to(dev), wmax, iteration=4, ev0=77)
    ref = O.unweight(w, wmax, 13, 4, ev0=77)
    assert np.array_equal(keep.cpu().numpy(), ref)
    assert np.array_equal(idx.cpu().numpy(), np.nonzero(ref)[0])
    expect = np.minimum(1.0, w / wmax).sum()
    assert abs(len(idx) - expect) < 5 * np.sqrt(expect)
    assert np.all(ref[w > wmax] == 1) and np.all(ref[w == 0] == 0)


@pytest.mark.gpu
def test_ttggg_vegas_and_flow_proposal(dev):
    """g g > t t~ g g g at sqrt(s) = 5 TeV with pT(g) > 30 GeV and dR(g,g) > 0.4 on the
    11-dim diet cube: plain VEGAS (FP32 matrix elements from the colour-ordered kernels,
    FP64 accumulation) and, on the frozen grid, an 11-dim RQ-spline flow proposal trained
    on uniform (point, f) pairs, every sample reweighted by f / q'.  With fixed seeds the
    two estimates agree within 3 combined standard deviations."""
    from madflow_amd.flow import RQSFlow
    from madflow_amd.phasespace_gpu import philox_uniforms
    print("case ttggg vegas + flow", flush=True)
    ig = MEIntegrand("gg_ttxggg", precision=32, mapping="diet", seed=5, sqrts=5000.0, pt_min=30.0, dr_min=0.4)
    assert ig.ndim == 11
    vf = VegasFlow(ig.ndim, 1 << 16)
    vf.compile(ig)
    vf.run_integration(3)
    vf.freeze_grid()
    plain, plain_err = vf.run_integration(2)
    ntr = 1 << 17
    y = philox_uniforms(ntr, ig.ndim, 99, 0).float().contiguous()
    _, _, f = ig.chunk(0, ntr, 0, keep_f=True, y_in=y)
    f = f.clone()
    assert torch.isfinite(f).all() and float(f.max()) > 0
    flow = RQSFlow(ig.ndim, layers=8, seed=1, device=dev)
    flow.train(y, f, steps=300, batch=4096, lr=1e-3)
    res = [ig.proposal_iteration(flow, 1 << 16, 100 + k, mix=0.2) for k in range(2)]
    from madflow_amd.vegas import combine
    learned, learned_err = combine(res)
    rel_plain = math.sqrt(vf.history[-1]["var"]) / vf.history[-1]["I"]
    rel_flow = math.sqrt(res[-1]["var"]) / res[-1]["I"]
    print("ttggg sigma plain %.6g +- %.3g  flow %.6g +- %.3g  (per-iteration rel. error %.3g vs %.3g)"
          % (plain, plain_err, learned, learned_err, rel_plain, rel_flow), flush=True)
    assert plain > 0 and learned > 0 and learned_err > 0
    assert abs(learned - plain) < 3 * math.hypot(plain_err, learned_err)


@pytest.mark.gpu
def test_ttgg_flow_proposal_c4(dev):
    """Configuration C4: g g > t t~ g g at sqrt(s) = 3 TeV, pT(g) > 30 GeV, dR(g,g) > 0.4,
    FP32 matrix elements (the drop-in Matrix_1_gg_ttxgg, colour-ordered kernels) on the
    8-dim diet VEGAS cube, plain VEGAS against VEGAS + the learned proposal: 8 RQ-spline
    couplings with K = 16 bins and a 2 x 128 MLP, bf16 weights with fp32 accumulation,
    trained by weighted maximum likelihood on |M|^2 J pairs from the frozen cube.

    Sizes are reduced from the configuration's (1.6e7 points per iteration, 5e6 training
    pairs, 2e4 steps) to fit a test: 4 warm-up iterations of 1M points, 3 iterations of 2M
    points per method, 1M training pairs plus 100k held out, at most 1,500 steps of batch
    8,192 (held-out check every 250 steps, stop after 2 without a new best).  Asserts:
    the two estimates agree within 2 combined standard deviations, and the flow cuts the
    per-iteration variance at equal points at least 2x (it is 12-15x with the full
    schedule)."""
    from madflow_amd.flow import RQSFlow
    from madflow_amd.phasespace_gpu import philox_uniforms
    from madflow_amd.vegas import combine
    print("case C4: ttgg vegas vs flow proposal", flush=True)
    ig = MEIntegrand("gg_ttxgg", precision=32, mapping="diet", seed=21, sqrts=3000.0, pt_min=30.0, dr_min=0.4)
    assert ig.ndim == 8 and ig.matrix.registry_name == "gg_ttxgg"
    vf = VegasFlow(ig.ndim, 1 << 20)
    vf.compile(ig)
    vf.run_integration(4)
    vf.freeze_grid()
    plain_res = [ig.iteration(2 << 20, 50 + k, train=False) for k in range(3)]
    plain, plain_err = combine(plain_res)
    ntr, nval = 1 << 20, 1 << 17
    y = philox_uniforms(ntr + nval, ig.ndim, 99, 0).float().contiguous()
    f = torch.empty(ntr + nval, dtype=torch.float64, device=dev)
    for a in range(0, ntr + nval, 1 << 19):
        n = min(1 << 19, ntr + nval - a)
        _, _, fb = ig.chunk(a, n, 0, keep_f=True, y_in=y[a:a + n].contiguous())
        f[a:a + n] = fb
    assert torch.isfinite(f).all() and float(f.max()) > 0
    flow = RQSFlow(ig.ndim, layers=8, seed=1, device=dev)
    assert flow.NPmax == 16 * ((4 * 49 + 15) // 16)       # K = 16: 3K + 1 = 49 raw values per dimension
    flow.train(y[:ntr], f[:ntr], steps=1500, batch=8192, lr=1e-3, seed=7, val=(y[ntr:], f[ntr:]),
               eval_every=250, mix=0.2, patience=2)
    res = [ig.proposal_iteration(flow, 2 << 20, 100 + k, mix=0.2) for k in range(3)]
    learned, learned_err = combine(res)
    var_plain = np.mean([r["var"] / r["I"] ** 2 for r in plain_res])
    var_flow = np.mean([r["var"] / r["I"] ** 2 for r in res])
    eff = lambda rs: sum(r["sum_f"] for r in rs) / (sum(r["n"] for r in rs) * max(r["wmax"] for r in rs))
    pull = (learned - plain) / math.hypot(plain_err, learned_err)
    print("C4 sigma plain %.6g +- %.3g  flow %.6g +- %.3g  pull %.2f; per-iteration rel. variance %.3g vs %.3g "
          "(gain %.1fx); unweighting efficiency %.3g vs %.3g; %d steps trained, best at %d"
          % (plain, plain_err, learned, learned_err, pull, var_plain, var_flow, var_plain / var_flow,
             eff(plain_res), eff(res), flow.steps_run, flow.best_step), flush=True)
    assert plain > 0 and learned > 0
    assert abs(pull) < 2.0
    assert var_flow < 0.5 * var_plain


@pytest.mark.gpu
def test_unweighted_events_to_lhe(dev, tmp_path):
    """Unweighted events on the frozen grid: the kept count matches sum min(1, w/wmax)
    over the same (reproducible) points within its binomial spread, the kept events are
    physical (momentum conservation, on-shell tops), a second call returns the same
    events, and the Les Houches file holds exactly them."""
    from madflow_amd.lhe import write_lhe, read_lhe
    print("case unweighted events -> LHE", flush=True)
    ig = MEIntegrand("gg_ttxgg", precision=32, mapping="diet", seed=17, sqrts=3000.0, pt_min=30.0, dr_min=0.4)
    vf = VegasFlow(ig.ndim, 1 << 18)
    vf.compile(ig)
    vf.run_integration(3)
    vf.freeze_grid()
    sigma, err = vf.run_integration(1)
    wmax = 0.1 * vf.history[-1]["wmax"]                  # a tenth: thousands of kept events
    n = 1 << 18
    ps, w, over = ig.unweighted_events(n, 7, wmax)
    _, _, f = ig.chunk(0, n, 7, keep_f=True)
    expect = float(torch.clamp(f / wmax, max=1.0).sum())
    assert abs(len(ps) - expect) < 5 * np.sqrt(expect)
    assert over == int((f > wmax).sum())
    assert np.all(w > 0)
    pin = ps[:, 0] + ps[:, 1]
    pout = ps[:, 2:].sum(axis=1)
    np.testing.assert_allclose(pout, pin, atol=1e-9 * 3000.0)
    for leg in (2, 3):
        m2 = ps[:, leg, 0] ** 2 - np.sum(ps[:, leg, 1:] ** 2, axis=1)
        np.testing.assert_allclose(m2, 173.0 ** 2, rtol=1e-9)
    ps2, w2, _ = ig.unweighted_events(n, 7, wmax)
    np.testing.assert_array_equal(ps2, ps)
    path = tmp_path / "unweighted_events.lhe.gz"
    assert write_lhe(path, "gg_ttxgg", ps, sigma, err, 3000.0, weights=w, wmax=wmax) == len(ps)
    init, events = read_lhe(path)
    assert len(events) == len(ps)
    # over-weight events (w > wmax, always kept) carry sigma * w / wmax, the rest sigma
    xw = np.array([ev[0][2] for ev in events])
    np.testing.assert_allclose(xw, sigma * np.maximum(1.0, w / wmax), rtol=1e-9)
    assert init[0][8] == (4.0 if over else 3.0)


@pytest.mark.gpu
def test_hadronic_ttbar_cross_section(dev):
    """Hadronic mode (ramboflow x1, x2 sampling, per-event flux, parton luminosity
    x1 g(x1) x2 g(x2) / (x1 x2)): pp -> gg -> tt~ at sqrt(s) = 13 TeV with a toy gluon
    density and Gamma_t = 0 must reproduce the direct 2D integral of the closed-form
    partonic cross section, int dx1 dx2 g(x1) g(x2) sigma_hat(x1 x2 s), within 3 sigma."""
    from scipy import integrate
    print("case hadronic sigma ttbar", flush=True)
    ig = MEIntegrand("gg_ttx", precision=64, mapping="hadronic", seed=2, sqrts=13000.0, ref_compat=False,
                     width=0.0, pdf=_toy_pdf, scale=173.0)
    assert ig.ndim == 10
    vf = VegasFlow(ig.ndim, 1000000)
    vf.compile(ig)
    vf.run_integration(3)
    vf.freeze_grid()
    res, err = vf.run_integration(5)
    m, a, s = 173.0, 0.118, 13000.0 ** 2

    def sig_hat(sh):
        rho = 4 * m * m / sh
        b = math.sqrt(max(1 - rho, 0.0))
        return math.pi * a * a / (3 * sh) * ((1 + rho + rho * rho / 16) * math.log((1 + b) / (1 - b))
                                            - b * (7 / 4 + 31 * rho / 16)) * GEV2PB

    t = 4 * m * m / s
    g = lambda x: _toy_pdf(x, 0.0) / x
    # substitute x = exp(u) to tame the small-x growth of g
    direct, derr = integrate.dblquad(
        lambda u2, u1: g(math.exp(u1)) * g(math.exp(u2)) * math.exp(u1 + u2) * sig_hat(math.exp(u1 + u2) * s),
        math.log(t), 0.0, lambda u1: math.log(t) - u1, lambda u1: 0.0, epsabs=0, epsrel=1e-7)
    print("hadronic sigma", res, "+-", err, "direct", direct, flush=True)
    assert abs(res - direct) < 3 * err + 1e-6 * direct
    assert err / res < 1e-2


@pytest.mark.gpu
def test_hadronic_dynamic_scale_and_running_alphas(dev):
    """The dynamic scale Q = sum of outgoing m_T / 2 (madflow_exec.py:427-430) reaches the
    couplings of each event: with alpha_s(Q^2) = 0.118 (Q / 173 GeV)^-0.2 every event's
    result equals the frozen-coupling result scaled by (alpha_s(Q^2) / 0.118)^2, the power
    of alpha_s in gg -> tt~, with Q computed independently from the momenta."""
    print("case hadronic running alpha_s", flush=True)
    kw = dict(precision=64, mapping="hadronic", seed=4, sqrts=13000.0, ref_compat=False, pdf=_toy_pdf)
    run = MEIntegrand("gg_ttx", alphas=lambda q2: 0.118 * (q2 / 173.0 ** 2) ** -0.1, **kw)
    frozen = MEIntegrand("gg_ttx", **kw)
    n = 20000
    _, _, f_run = run.chunk(0, n, 0, keep_f=True)
    f_run = f_run.cpu().numpy().copy()
    _, _, f_fix = frozen.chunk(0, n, 0, keep_f=True)
    f_fix = f_fix.cpu().numpy()
    ps = frozen._bufs["all_ps"][:n].cpu().numpy()
    mt = np.sqrt(np.maximum(ps[:, 2:, 0] ** 2 - ps[:, 2:, 3] ** 2, 0.0)).sum(axis=1)
    q2 = (0.5 * mt) ** 2
    ratio = (0.118 * (q2 / 173.0 ** 2) ** -0.1 / 0.118) ** 2
    nz = f_fix > 0
    assert nz.sum() > 0.9 * n
    np.testing.assert_allclose(f_run[nz], f_fix[nz] * ratio[nz], rtol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("variable,particle,lo,hi", [("pt", 3, 60.0, 300.0), ("mt", 2, 400.0, None),
                                                     ("mt2", 4, None, 250.0 ** 2)])
def test_register_cut_matches_numpy_mask(dev, variable, particle, lo, hi):
    """register_cut (phasespace.py:424-478) checked as the reference's own test checks it
    (test_ps.py:40-64): on the same Philox points, the events that pass equal the uncut
    ones with lo < variable(particle) < hi computed in numpy from their momenta, the
    momenta of those events are unchanged, and clear_cuts restores the uncut flags."""
    print("case register_cut", variable, particle, flush=True)
    ig = MEIntegrand("gg_ttxg", 2000.0, precision=64, mapping="rambo", seed=4)
    n = 20000
    ig.chunk(0, n, 3)
    full = ig._bufs["all_ps"][:n].cpu().numpy().copy()
    flag0 = ig._bufs["flag"][:n].cpu().numpy().copy()
    ig.register_cut(variable, particle=particle, min_val=lo, max_val=hi)
    ig.chunk(0, n, 3)
    flag = ig._bufs["flag"][:n].cpu().numpy().copy()
    cut_ps = ig._bufs["all_ps"][:n].cpu().numpy()
    q = full[:, particle]
    pt = np.sqrt(q[:, 1] ** 2 + q[:, 2] ** 2)
    mt2 = (q[:, 0] ** 2 - q[:, 1] ** 2 - q[:, 2] ** 2 - q[:, 3] ** 2) + pt ** 2
    v = {"pt": pt, "mt": np.sqrt(mt2), "mt2": mt2}[variable]
    mask = flag0 == 1
    if lo is not None:
        mask &= v > lo
    if hi is not None:
        mask &= v < hi
    print("pass fraction", mask.mean(), "of", (flag0 == 1).mean(), flush=True)
    assert 0.05 < mask.mean() < 0.95
    assert np.array_equal(flag == 1, mask)
    np.testing.assert_array_equal(cut_ps[mask], full[mask])
    ig.clear_cuts()
    ig.chunk(0, n, 3)
    assert np.array_equal(ig._bufs["flag"][:n].cpu().numpy(), flag0)
    with pytest.raises(ValueError):
        ig.register_cut("eta", particle=3, min_val=0.0)
    with pytest.raises(ValueError):
        ig.register_cut("pt", particle=5, min_val=10.0)


@pytest.mark.gpu
@pytest.mark.parametrize("hadronic", [False, True])
def test_integrand_sums_subprocesses(dev, hadronic):
    """Several processes in one integrand, as the reference's integrand sums its matrices
    (madflow_exec.py:453-465): on the same points, the per-event f of g g > t t~ g,
    u u~ > t t~ g, u g > t t~ u and u~ g > t t~ u~ together equals the sum of the four
    single-process integrands' f within 1e-12, at fixed energy and in hadronic mode
    (each process with its own flavour luminosity)."""
    print("case integrand subprocess sum hadronic", hadronic, flush=True)
    procs = ["gg_ttxg", "uux_ttxg", "ug_ttxu", "uxg_ttxux"]

    def pdf(x, q2, pid=21):
        return _toy_pdf(x, q2) * (1.0 if pid == 21 else (0.6 if pid > 0 else 0.25))
    kw = dict(precision=64, seed=6, pt_min=30.0, abs_eta_max=2.5)
    if hadronic:
        kw.update(mapping="hadronic", pdf=pdf, scale=173.0)
    else:
        kw.update(mapping="rambo")
    sqrts = 13000.0 if hadronic else 2000.0
    n = 20000
    ig = MEIntegrand(procs, sqrts, **kw)
    _, npass, f = ig.chunk(0, n, 3, keep_f=True)
    f = f.cpu().numpy().copy()
    tot = np.zeros(n)
    for q in procs:
        _, _, fq = MEIntegrand(q, sqrts, **kw).chunk(0, n, 3, keep_f=True)
        fq = fq.cpu().numpy()
        assert np.any(fq > 0), q
        tot += fq
    print("passing", npass, "sum of f", f.sum(), flush=True)
    np.testing.assert_allclose(f, tot, rtol=1e-12, atol=0.0)
    with pytest.raises(ValueError):
        MEIntegrand(["gg_ttxg", "gg_ttxgg"], sqrts, **kw)
