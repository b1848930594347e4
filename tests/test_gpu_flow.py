"""Device RQ-spline flow against the CPU torch restatement: forward, inverse, log-det,
gradients (autograd of the restatement), and the training gain over plain sampling."""

import math

import numpy as np
import pytest
import torch

from madflow_amd.flow import RQSFlow
from tests import flow_reference as R


def _perturbed(flow, seed=1, scale=0.3):
    """Random non-identity parameters so that every spline bin and MLP path is exercised."""
    g = torch.Generator().manual_seed(seed)
    p = flow.params.cpu().clone()
    p += scale * torch.randn(p.shape, generator=g) / math.sqrt(128)
    flow.params.copy_(p)
    flow._cast()
    return p


@pytest.mark.gpu
@pytest.mark.parametrize("D", [8, 11])
def test_flow_forward_inverse(dev, D):
    print("case flow fwd/inv D=%d" % D, flush=True)
    flow = RQSFlow(D, layers=8, seed=3, device=dev)
    p = _perturbed(flow)
    x = torch.rand((1000, D), generator=torch.Generator().manual_seed(2), dtype=torch.float32)
    u, ld = flow.forward(x.to(dev))
    u_ref, ld_ref = R.apply(flow, p, x)
    np.testing.assert_allclose(u.cpu().numpy(), u_ref.numpy(), atol=2e-4)
    np.testing.assert_allclose(ld.cpu().numpy(), ld_ref.numpy(), atol=2e-3, rtol=1e-3)
    xb, ldi = flow.sample(u)
    np.testing.assert_allclose(xb.cpu().numpy(), x.numpy(), atol=2e-4)
    np.testing.assert_allclose(ldi.cpu().numpy(), ld.cpu().numpy(), atol=2e-3)


@pytest.mark.gpu
def test_flow_logdet_is_jacobian(dev):
    """log q from the kernels equals log|det du/dx| of the restated map (autograd Jacobian)."""
    D = 8
    flow = RQSFlow(D, layers=8, seed=4, device=dev)
    p = _perturbed(flow, seed=5)
    x = torch.rand((6, D), generator=torch.Generator().manual_seed(6), dtype=torch.float64)
    ld = flow.log_prob(x.float().to(dev)).cpu().numpy()
    for i in range(6):
        J = torch.autograd.functional.jacobian(lambda v: R.apply(flow, p.double(), v[None])[0][0], x[i])
        assert abs(ld[i] - torch.linalg.slogdet(J)[1].item()) < 5e-3


@pytest.mark.gpu
@pytest.mark.parametrize("n", [512, 1000])
def test_flow_gradients_match_autograd(dev, n):
    """n = 1000: the last 32-sample block of the forward pass holds one tile of the batch."""
    print("case flow gradients n=%d" % n, flush=True)
    D = 8
    flow = RQSFlow(D, layers=8, seed=7, device=dev)
    p = _perturbed(flow, seed=8)
    x = torch.rand((n, D), generator=torch.Generator().manual_seed(9), dtype=torch.float32)
    w = torch.rand(n, generator=torch.Generator().manual_seed(10), dtype=torch.float64) ** 4
    loss, grad = flow.grad(x.to(dev), w.to(dev))
    pr = p.clone().requires_grad_(True)
    _, ldr = R.apply(flow, pr, x)
    lr = -(w.float() * ldr).sum() / w.float().sum()
    lr.backward()
    g_ref = pr.grad.numpy()
    g = grad.cpu().numpy()
    assert abs(float(loss) - float(lr)) < 1e-3 * max(1.0, abs(float(lr)))
    cos = float(np.dot(g, g_ref) / (np.linalg.norm(g) * np.linalg.norm(g_ref)))
    print("flow grad cosine", cos, "rel norm err", np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref), flush=True)
    assert cos > 0.999
    assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 3e-2


@pytest.mark.gpu
@pytest.mark.parametrize("beta", [1.0, 2.0])
def test_flow_alpha_divergence_gradient(dev, beta):
    """beta > 0: the device gradient is that of the alpha-divergence sum_i w_i (f_i / q_i)^beta
    of the restated flow (autograd through log q), divided by the positive factor
    beta * sum_i w_i (f_i / q_i)^beta."""
    print("case flow alpha-divergence beta=%g" % beta, flush=True)
    D = 8
    flow = RQSFlow(D, layers=8, seed=7, device=dev)
    p = _perturbed(flow, seed=12)
    n = 512
    x = torch.rand((n, D), generator=torch.Generator().manual_seed(13), dtype=torch.float32)
    w = torch.rand(n, generator=torch.Generator().manual_seed(14), dtype=torch.float64) ** 2
    f = torch.rand(n, generator=torch.Generator().manual_seed(15), dtype=torch.float64) ** 4 + 1e-3
    _, grad = flow.grad(x.to(dev), w.to(dev), f.to(dev), beta)
    pr = p.clone().double().requires_grad_(True)
    _, ldr = R.apply(flow, pr, x.double())
    terms = w * (f * torch.exp(-ldr)) ** beta
    terms.sum().backward()
    g_ref = pr.grad.numpy() / (beta * float(terms.sum()))
    g = grad.cpu().numpy()
    cos = float(np.dot(g, g_ref) / (np.linalg.norm(g) * np.linalg.norm(g_ref)))
    print("alpha-divergence grad cosine", cos, "rel norm err", np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref),
          flush=True)
    assert cos > 0.999
    assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 3e-2


def _target(y):
    """Peaked, correlated test density on [0,1]^D (unnormalised)."""
    z = (y - 0.3) / 0.08
    return torch.exp(-0.5 * (z[:, 0] ** 2 + (z[:, 1] - 0.8 * z[:, 0]) ** 2)) + 0.05


@pytest.mark.gpu
def test_flow_training_reduces_variance(dev):
    """Trained on (y, f) pairs from uniform points, the flow's importance weights f/q have
    a much smaller relative variance than plain uniform sampling, and the reweighted
    estimate stays unbiased."""
    D = 4
    torch.manual_seed(0)
    y = torch.rand((200000, D), device=dev)
    f = _target(y).to(torch.float64)
    flow = RQSFlow(D, layers=4, seed=11, device=dev)
    flow.train(y, f, steps=400, batch=4096, lr=3e-3)
    u = torch.rand((200000, D), device=dev)
    ys, logq = flow.sample(u)
    wgt = _target(ys).double() / torch.exp(logq.double())
    plain = f
    rv_flow = (wgt.std() / wgt.mean()).item()
    rv_plain = (plain.std() / plain.mean()).item()
    print("relative std plain %.3f flow %.3f" % (rv_plain, rv_flow), flush=True)
    assert rv_flow < 0.5 * rv_plain
    exact = f.mean().item()
    err = math.hypot(wgt.std().item() / math.sqrt(len(wgt)), plain.std().item() / math.sqrt(len(plain)))
    assert abs(wgt.mean().item() - exact) < 4 * err


@pytest.mark.gpu
def test_minibatch_draw_and_seed(dev):
    """mfx_flow_minibatch picks the first index whose normalised running weight reaches
    the uniform (numpy searchsorted, side='left'), zero-weight points are never drawn,
    and mfx_flow_seed gives -v_i / sum v with v_i = w_i (f_i / q_i)^beta."""
    import ctypes
    from madflow_amd import _lib
    print("case flow minibatch / seed", flush=True)
    flow = RQSFlow(3, layers=2, seed=1, device=dev)
    lib = flow.lib
    rng = np.random.default_rng(21)
    n, batch = 5000, 4096
    w = rng.random(n) ** 3 * (rng.random(n) > 0.3)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    r = rng.random(batch)
    r[:16] = cdf[rng.integers(0, n, 16)]           # uniforms that hit a running sum exactly
    y = rng.random((n, 3)).astype(np.float32)
    f = rng.random(n) + 0.1
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    cdf_t, r_t, y_t, f_t = t(cdf), t(r), t(y), t(f)
    yb = torch.empty((batch, 3), dtype=torch.float32, device=dev)
    fb = torch.empty(batch, dtype=torch.float64, device=dev)
    vp = lambda x: ctypes.c_void_p(x.data_ptr())
    _lib.check(lib.mfx_flow_minibatch(vp(cdf_t), n, vp(r_t), batch, vp(y_t), 3, vp(f_t), vp(yb), vp(fb),
                                      _lib.stream_ptr(dev)))
    idx = np.minimum(np.searchsorted(cdf, r, side="left"), n - 1)
    assert np.all(w[idx] > 0)
    np.testing.assert_array_equal(yb.cpu().numpy(), y[idx])
    np.testing.assert_array_equal(fb.cpu().numpy(), f[idx])
    ld = rng.normal(size=batch).astype(np.float32)
    wb = rng.random(batch)
    fi_t, ld_t = t(f[idx]), t(ld)              # kept referenced until the kernels ran
    for beta, wt in ((0.0, None), (0.0, wb), (2.0, wb)):
        gl = torch.empty(batch, dtype=torch.float32, device=dev)
        w_t = t(wt) if wt is not None else None
        _lib.check(lib.mfx_flow_seed(vp(w_t) if w_t is not None else None, vp(fi_t), vp(ld_t), batch,
                                     beta, vp(gl), _lib.stream_ptr(dev)))
        v = (wt if wt is not None else np.ones(batch)) * (f[idx] * np.exp(-ld.astype(np.float64))) ** beta
        np.testing.assert_allclose(gl.cpu().numpy(), -v / v.sum(), rtol=1e-6, atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("beta", [0.0, 2.0])
def test_draw_forward_and_fused_step(dev, beta):
    """mfx_flow_draw_forward draws sample e as the first pool index whose normalised running
    weight reaches the Philox uniform of counter (e, 0, 0, draw) (the oracle's stream),
    returns the point's integrand value, and its log q and saved state equal those of a
    forward pass over the drawn points: the gradient from its save equals flow.grad of the
    drawn batch bitwise.  The draw through a guide table (mfx_flow_cdf_guide) picks the same
    indices as the bisection.  mfx_flow_adam_cast equals mfx_flow_adam then mfx_flow_cast
    bitwise."""
    import ctypes
    from madflow_amd import _lib
    from oracle import oracle as O
    print("case flow draw-forward / fused step beta=%g" % beta, flush=True)
    D, npool, batch, seed, draw = 8, 20000, 3000, 77, 5
    flow = RQSFlow(D, layers=8, seed=3, device=dev)
    _perturbed(flow, seed=4)
    lib = flow.lib
    rng = np.random.default_rng(31)
    w = rng.random(npool) ** 4 * (rng.random(npool) > 0.2)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    y = rng.random((npool, D)).astype(np.float32)
    f = rng.random(npool) + 0.05
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    cdf_t, y_t, f_t = t(cdf), t(y), t(f)
    yo = torch.empty((batch, D), dtype=torch.float32, device=dev)
    fo = torch.empty(batch, dtype=torch.float64, device=dev)
    ld = torch.empty(batch, dtype=torch.float32, device=dev)
    save = torch.empty(lib.mfx_flow_save_bytes(D, 8, flow._m(), batch) // 4 + 1, dtype=torch.float32, device=dev)
    vp = lambda x: ctypes.c_void_p(x.data_ptr())
    u = O.philox_uniforms(batch, 2, seed, draw)[:, 0]
    idx = np.minimum(np.searchsorted(cdf, u, side="left"), npool - 1)
    assert np.all(w[idx] > 0)
    # the bisection, then guide tables of several sizes (a table as long as the pool, and
    # coarse ones whose buckets hold many entries), all drawing the same indices
    for m_guide in (0, npool, 64, 1):
        guide = None
        if m_guide:
            guide = torch.empty(m_guide, dtype=torch.int32, device=dev)
            _lib.check(lib.mfx_flow_cdf_guide(vp(cdf_t), npool, vp(guide), m_guide, _lib.stream_ptr(dev)))
            b = np.arange(m_guide) / m_guide
            np.testing.assert_array_equal(guide.cpu().numpy(), np.minimum(np.searchsorted(cdf, b, side="left"), npool - 1))
        _lib.check(lib.mfx_flow_draw_forward(D, 8, flow._m(), vp(flow.params), vp(flow.pbf16), vp(cdf_t),
                                             vp(guide) if guide is not None else None, m_guide, npool, vp(y_t),
                                             vp(f_t), seed, draw, batch, vp(yo), vp(fo), vp(ld), vp(save),
                                             _lib.stream_ptr(dev)))
        np.testing.assert_array_equal(yo.cpu().numpy(), y[idx])
        np.testing.assert_array_equal(fo.cpu().numpy(), f[idx])
    assert torch.equal(ld, flow.log_prob(yo))
    # the gradient from the draw's saved state equals that of a separate forward pass
    gl = torch.empty(batch, dtype=torch.float32, device=dev)
    _lib.check(lib.mfx_flow_seed(None, vp(fo), vp(ld), batch, beta, vp(gl), _lib.stream_ptr(dev)))
    partial = torch.empty(lib.mfx_flow_partial_bytes(D, 8, flow._m(), batch) // 4 + 1, dtype=torch.float32,
                          device=dev)
    grad = torch.empty_like(flow.params)
    _lib.check(lib.mfx_flow_backward(D, 8, flow._m(), vp(flow.params), vp(flow.pbf16), batch, vp(save), vp(gl), None,
                                     vp(partial), vp(grad), _lib.stream_ptr(dev)))
    _, grad_ref = flow.grad(yo, None, fo if beta else None, beta)
    assert torch.equal(grad, grad_ref)
    # Adam with the bf16 refresh in one launch
    p0, m1, m2 = flow.params.clone(), flow.m1.clone(), flow.m2.clone()
    flow.step = 0
    flow.adam_step(grad, lr=1e-2)
    p_fused, pb_fused = flow.params.clone(), flow.pbf16.clone()
    flow.params.copy_(p0)
    flow.m1.copy_(m1)
    flow.m2.copy_(m2)
    _lib.check(lib.mfx_flow_adam(vp(flow.params), vp(grad), vp(flow.m1), vp(flow.m2), flow.nparams, 1e-2, 0.9, 0.999,
                                 1e-8, 1, 10.0, _lib.stream_ptr(dev)))
    flow._cast()
    assert torch.equal(flow.params, p_fused)
    assert torch.equal(flow.pbf16, pb_fused)


@pytest.mark.gpu
def test_flow_wide_layers(dev):
    """A flow whose layers transform 9..15 of 16 dimensions (the kernels sized for 16):
    forward, inverse and the gradient against the restatement."""
    print("case flow wide layers D=16", flush=True)
    D, L = 16, 4
    masks = np.zeros((L, D), dtype=np.int32)
    for l, nb in enumerate((9, 15, 12, 10)):
        masks[l, (np.arange(D) + 3 * l) % D < nb] = 1
    flow = RQSFlow(D, layers=L, seed=5, device=dev, masks=masks)
    p = _perturbed(flow, seed=6)
    x = torch.rand((700, D), generator=torch.Generator().manual_seed(7), dtype=torch.float32)
    u, ld = flow.forward(x.to(dev))
    u_ref, ld_ref = R.apply(flow, p, x)
    np.testing.assert_allclose(u.cpu().numpy(), u_ref.numpy(), atol=2e-4)
    np.testing.assert_allclose(ld.cpu().numpy(), ld_ref.numpy(), atol=4e-3, rtol=1e-3)
    xb, _ = flow.sample(u)
    np.testing.assert_allclose(xb.cpu().numpy(), x.numpy(), atol=3e-4)
    w = torch.rand(700, generator=torch.Generator().manual_seed(8), dtype=torch.float64) ** 2
    _, grad = flow.grad(x.to(dev), w.to(dev))
    pr = p.clone().requires_grad_(True)
    _, ldr = R.apply(flow, pr, x)
    (-(w.float() * ldr).sum() / w.float().sum()).backward()
    g, g_ref = grad.cpu().numpy(), pr.grad.numpy()
    cos = float(np.dot(g, g_ref) / (np.linalg.norm(g) * np.linalg.norm(g_ref)))
    print("wide-layer grad cosine", cos, flush=True)
    assert cos > 0.999


@pytest.mark.gpu
def test_cdf_moment_mixture_weights(dev):
    """mfx_flow_cdf equals the normalised running sum of max(w, 0) (a non-finite weight
    counts as 0) with its total in the workspace's last double; mfx_flow_moment equals
    sum f^2 / ((1 - mix) e^logq + mix); mfx_flow_mixture_winv equals 1 / ((1 - mix) q + mix)
    for forward and inverse log-dets; the fp32 Philox uniforms are the fp64 ones rounded."""
    import ctypes
    from madflow_amd import _lib
    from madflow_amd.phasespace_gpu import philox_uniforms
    print("case flow cdf / moment / mixture weights", flush=True)
    flow = RQSFlow(3, layers=2, seed=1, device=dev)
    lib = flow.lib
    rng = np.random.default_rng(41)
    n = 100_003
    w = rng.random(n) ** 3 - 0.1
    w[rng.integers(0, n, 50)] = np.nan
    wc = np.where(np.isfinite(w) & (w > 0), w, 0.0)
    ref = np.cumsum(wc)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    w_t = t(w)
    cdf = torch.empty(n, dtype=torch.float64, device=dev)
    ws = torch.empty(lib.mfx_flow_cdf_workspace(n) // 8, dtype=torch.float64, device=dev)
    vp = lambda x: ctypes.c_void_p(x.data_ptr())
    _lib.check(lib.mfx_flow_cdf(vp(w_t), n, vp(cdf), vp(ws), ws.numel() * 8, _lib.stream_ptr(dev)))
    c = cdf.cpu().numpy()
    assert abs(float(ws[-1]) / ref[-1] - 1) < 1e-12
    np.testing.assert_allclose(c, ref / ref[-1], rtol=1e-12, atol=1e-15)
    assert c[-1] == 1.0 and np.all(np.diff(c) >= 0)
    lq = rng.normal(size=n).astype(np.float32)
    f = rng.random(n) + 0.01
    for mix in (0.0, 0.3):
        work = torch.empty(257, dtype=torch.float64, device=dev)
        lq_t, f_t = t(lq), t(f)
        _lib.check(lib.mfx_flow_moment(vp(lq_t), vp(f_t), n, mix, vp(work), _lib.stream_ptr(dev)))
        ref_m = np.sum(f * f / ((1 - mix) * np.exp(lq.astype(np.float64)) + mix))
        assert abs(float(work[256]) / ref_m - 1) < 1e-12
        for inv in (0, 1):
            winv = torch.empty(n, dtype=torch.float64, device=dev)
            _lib.check(lib.mfx_flow_mixture_winv(vp(lq_t), n, inv, mix, vp(winv), _lib.stream_ptr(dev)))
            sgn = -1.0 if inv else 1.0
            np.testing.assert_allclose(winv.cpu().numpy(), 1.0 / ((1 - mix) * np.exp(sgn * lq.astype(np.float64)) + mix),
                                       rtol=1e-14)
    u64 = philox_uniforms(1000, 7, 5, 2, ev0=33)
    u32 = philox_uniforms(1000, 7, 5, 2, ev0=33, dtype=torch.float32)
    assert torch.equal(u32, u64.float())
