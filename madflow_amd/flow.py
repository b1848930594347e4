"""Learned importance-sampling proposal: rational-quadratic spline coupling flow on the
VEGAS-mapped unit hypercube (device kernels in csrc/mfx_flow.hip).

    flow = RQSFlow(D=8, layers=8)
    flow.train(y, w, steps=..., batch=...)      # weighted maximum likelihood on (point, weight) pairs
    y, logq = flow.sample(u)                    # u ~ U(0,1)^D  ->  y, log q(y)
    logq = flow.log_prob(y)

Training maximises sum_i w_i log q(y_i) / sum_i w_i with w_i = |M|^2 x (phase-space and
VEGAS jacobians) at points y_i drawn uniformly on the VEGAS cube, i.e. it minimises
KL(p || q) for p proportional to the integrand pulled back to the cube.  Every sample
of the proposal is reweighted by f(y)/q(y), so the estimate stays unbiased whatever the
quality of the fit; plain VEGAS (q = 1) is the fallback.
"""

import ctypes
import math

import numpy as np
import torch

from . import _lib

H = 128
K = 16
NPD = 3 * K + 1
DMAX = 16
MIND = 1e-3


def default_masks(D, L):
    """Layer l transforms the dimensions whose bit (l//2 mod nbits) equals (l+1) % 2."""
    nbits = max(1, int(math.ceil(math.log2(D))))
    masks = np.zeros((L, D), dtype=np.int32)
    for l in range(L):
        bit = (l // 2) % nbits
        want = (l + 1) % 2
        for d in range(D):
            masks[l, d] = 1 if ((d >> bit) & 1) == want else 0
        if masks[l].sum() in (0, D):
            masks[l] = np.arange(D) % 2 == (l % 2)
    return masks


def _sig(lib):
    vp, i64, i, f = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_float
    lib.mfx_flow_sizes.argtypes = [i, i, vp, vp]
    lib.mfx_flow_sizes.restype = i
    lib.mfx_flow_cast.argtypes = [i, i, vp, vp, vp, vp]
    lib.mfx_flow_cast.restype = i
    lib.mfx_flow_apply.argtypes = [i, i, vp, vp, vp, i64, vp, vp, vp, i, vp, vp]
    lib.mfx_flow_apply.restype = i
    lib.mfx_flow_save_bytes.argtypes = [i, i, vp, i64]
    lib.mfx_flow_save_bytes.restype = i64
    lib.mfx_flow_partial_bytes.argtypes = [i, i, vp, i64]
    lib.mfx_flow_partial_bytes.restype = i64
    lib.mfx_flow_backward.argtypes = [i, i, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp]
    lib.mfx_flow_backward.restype = i
    lib.mfx_flow_adam.argtypes = [vp, vp, vp, vp, i64, f, f, f, f, i, f, vp]
    lib.mfx_flow_adam.restype = i
    lib.mfx_flow_minibatch.argtypes = [vp, i64, vp, i64, vp, i, vp, vp, vp, vp]
    lib.mfx_flow_minibatch.restype = i
    lib.mfx_flow_seed.argtypes = [vp, vp, vp, i64, f, vp, vp]
    lib.mfx_flow_seed.restype = i
    u32 = ctypes.c_uint32
    lib.mfx_flow_draw_forward.argtypes = [i, i, vp, vp, vp, vp, vp, i64, i64, vp, vp, u32, u32, i64, vp, vp, vp, vp,
                                          vp]
    lib.mfx_flow_draw_forward.restype = i
    lib.mfx_flow_adam_cast.argtypes = [i, i, vp, vp, vp, vp, vp, vp, f, f, f, f, i, f, vp]
    lib.mfx_flow_adam_cast.restype = i
    d = ctypes.c_double
    lib.mfx_flow_cdf_workspace.argtypes = [i64]
    lib.mfx_flow_cdf_workspace.restype = i64
    lib.mfx_flow_cdf.argtypes = [vp, i64, vp, vp, i64, vp]
    lib.mfx_flow_cdf.restype = i
    lib.mfx_flow_cdf_guide.argtypes = [vp, i64, vp, i64, vp]
    lib.mfx_flow_cdf_guide.restype = i
    lib.mfx_flow_moment.argtypes = [vp, vp, i64, d, vp, vp]
    lib.mfx_flow_moment.restype = i
    lib.mfx_flow_mixture_winv.argtypes = [vp, i64, i, d, vp, vp]
    lib.mfx_flow_mixture_winv.restype = i


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class RQSFlow:
    def __init__(self, D, layers=8, seed=0, device="cuda", masks=None):
        self.D = D
        self.L = layers
        self.device = torch.device(device)
        self.masks = np.ascontiguousarray(masks if masks is not None else default_masks(D, layers), dtype=np.int32)
        self.lib = _lib.load()
        _sig(self.lib)
        sizes = np.zeros(3 + 2 * layers, dtype=np.int64)
        _lib.check(self.lib.mfx_flow_sizes(D, layers, self._m(), sizes.ctypes.data_as(ctypes.c_void_p)))
        self.nparams, self.nbf16, self.NPmax = int(sizes[0]), int(sizes[1]), int(sizes[2])
        self.layer_off = [(int(sizes[3 + 2 * l]), int(sizes[4 + 2 * l])) for l in range(layers)]
        self.params = torch.from_numpy(self.init_params(seed)).to(self.device)
        self.pbf16 = torch.empty(self.nbf16, dtype=torch.int16, device=self.device)
        self.m1 = torch.zeros_like(self.params)
        self.m2 = torch.zeros_like(self.params)
        self.step = 0
        self._cast()

    def _m(self):
        return self.masks.ctypes.data_as(ctypes.c_void_p)

    def layout(self, l):
        """Offsets of W1 [DMAX][H], b1, W2 [H][H], b2, W3 [H][NP], b3 inside layer l."""
        off, n = self.layer_off[l]
        nB = int(self.masks[l].sum())
        NP = ((nB * NPD + 15) // 16) * 16
        o = {}
        o["W1"] = off
        o["b1"] = o["W1"] + DMAX * H
        o["W2"] = o["b1"] + H
        o["b2"] = o["W2"] + H * H
        o["W3"] = o["b2"] + H
        o["b3"] = o["W3"] + H * NP
        o["NP"] = NP
        o["nB"] = nB
        return o

    def init_params(self, seed):
        rng = np.random.default_rng(seed)
        p = np.zeros(self.nparams, dtype=np.float32)
        for l in range(self.L):
            o = self.layout(l)
            nA = self.D - o["nB"]
            w1 = rng.uniform(-1, 1, size=(DMAX, H)) / math.sqrt(nA)
            w1[nA:] = 0
            p[o["W1"]:o["W1"] + DMAX * H] = w1.reshape(-1)
            p[o["W2"]:o["W2"] + H * H] = (rng.uniform(-1, 1, size=(H, H)) / math.sqrt(H)).reshape(-1)
            # last layer zero: uniform bins; derivative logits so that every knot slope is 1
            b3 = np.zeros(o["NP"], dtype=np.float32)
            dlogit = math.log(math.expm1(1.0 - MIND))
            for b in range(o["nB"]):
                b3[b * NPD + 2 * K: b * NPD + 3 * K + 1] = dlogit
            p[o["b3"]:o["b3"] + o["NP"]] = b3
        return p

    def _cast(self):
        _lib.check(self.lib.mfx_flow_cast(self.D, self.L, self._m(), _p(self.params), _p(self.pbf16),
                                          _lib.stream_ptr(self.device)))

    # --------------------------------------------------------------- evaluation
    def _apply(self, x, inverse, save=None):
        x = _lib.require_device_tensor(x, torch.float32, "x")
        n = x.shape[0]
        out = torch.empty_like(x)
        ld = torch.empty(n, dtype=torch.float32, device=x.device)     # written by the kernel
        _lib.check(self.lib.mfx_flow_apply(self.D, self.L, self._m(), _p(self.params), _p(self.pbf16), n, _p(x),
                                           _p(out), _p(ld), int(inverse), _p(save), _lib.stream_ptr(x.device)))
        return out, ld

    def forward(self, y):
        """y -> u and log|det du/dy| (= log q(y) for a uniform base)."""
        return self._apply(y, False)

    def log_prob(self, y):
        return self._apply(y, False)[1]

    def sample(self, u, logq=True):
        """u -> y and log q(y) (None with logq=False)."""
        y, ld_inv = self._apply(u, True)
        return y, (-ld_inv if logq else None)

    # ----------------------------------------------------------------- training
    def grad(self, y, w, f=None, beta=0.0, need_loss=True):
        """Gradient of L = -sum_i v_i log q(y_i) / sum_i v_i (returns L, grad), with
        v_i = w_i for beta = 0 and v_i = w_i (f_i / q(y_i))^beta otherwise (w = None: w_i = 1).

        For points y_i drawn from a density p with w_i = f_i / p_i, beta = 0 is the
        gradient of KL(f || q) and beta = a - 1 that of the alpha-divergence
        int f^a q^(1-a) (up to a positive factor); beta = 1 minimises the second moment
        of the importance weights f/q itself, and larger beta weighs the largest ratios
        f/q, which set the unweighting efficiency, harder still."""
        y = _lib.require_device_tensor(y, torch.float32, "y")
        n = y.shape[0]
        save = torch.empty(self.lib.mfx_flow_save_bytes(self.D, self.L, self._m(), n) // 4 + 1,
                           dtype=torch.float32, device=y.device)
        u, ld = self._apply(y, False, save=save)
        if beta and f is None:
            raise ValueError("grad: beta != 0 needs the integrand values f")
        w64 = None if w is None else _lib.require_device_tensor(w, None, "w").to(torch.float64)
        f64 = None if not beta else _lib.require_device_tensor(f, None, "f").to(torch.float64)
        gl = torch.empty(n, dtype=torch.float32, device=y.device)
        _lib.check(self.lib.mfx_flow_seed(_p(w64), _p(f64), _p(ld), n, float(beta), _p(gl), _lib.stream_ptr(y.device)))
        gx = torch.zeros_like(y)
        partial = torch.empty(self.lib.mfx_flow_partial_bytes(self.D, self.L, self._m(), n) // 4 + 1,
                              dtype=torch.float32, device=y.device)
        grad = torch.empty_like(self.params)     # every entry is written by the backward pass
        _lib.check(self.lib.mfx_flow_backward(self.D, self.L, self._m(), _p(self.params), _p(self.pbf16), n, _p(save), _p(gl),
                                              _p(gx), _p(partial), _p(grad), _lib.stream_ptr(y.device)))
        loss = (gl.to(torch.float64) * ld.to(torch.float64)).sum() if need_loss else None
        return loss, grad

    def adam_step(self, grad, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, clip=10.0):
        """One Adam step on the fp32 parameters; the same launch refreshes their bf16 copies."""
        self.step += 1
        _lib.check(self.lib.mfx_flow_adam_cast(self.D, self.L, self._m(), _p(self.params), _p(self.pbf16), _p(grad),
                                               _p(self.m1), _p(self.m2), lr, betas[0], betas[1], eps, self.step, clip,
                                               _lib.stream_ptr(self.device)))

    def is_second_moment(self, y, f, mix=0.0, chunk=1 << 20):
        """E_uniform[f^2 / q'] over uniform points y with integrand values f, where
        q' = (1 - mix) q + mix: the second moment of the importance weights f/q' of
        samples drawn from q' (the variance up to the fixed squared integral)."""
        f = _lib.require_device_tensor(f, torch.float64, "f")
        work = torch.empty(257, dtype=torch.float64, device=y.device)
        tot = 0.0
        for a in range(0, y.shape[0], chunk):
            lq = self.log_prob(y[a:a + chunk].contiguous())
            ff = f[a:a + chunk]
            _lib.check(self.lib.mfx_flow_moment(_p(lq), _p(ff), ff.shape[0], float(mix), _p(work),
                                                _lib.stream_ptr(y.device)))
            tot += float(work[256])
        return tot / y.shape[0]

    def train(self, y, w, steps=1000, batch=4096, lr=1e-3, seed=0, log=None, lr_decay=True, resample=True,
              val=None, eval_every=1000, mix=0.0, patience=4, f=None, beta=0.0):
        """Minibatch weighted maximum likelihood on device-resident pairs (y, w).

        resample=True draws each minibatch with probability proportional to w (inverse-CDF
        lookup inside the forward kernel, Philox uniforms keyed on ``seed`` and the step) and
        gives every drawn point unit weight: the same expected
        loss as the w-weighted uniform minibatch, with far less gradient noise when the
        weights are heavy-tailed (matrix elements near collinear/soft limits).

        val=(y_val, f_val): held-out uniform points.  Every eval_every steps the second
        moment of the importance weights on them (is_second_moment, with the sampling
        mixture ``mix``) is measured and the parameters with the smallest one are kept:
        maximum likelihood on heavy-tailed weights over-fits the few largest weights
        after enough steps, which the held-out moment sees and the training loss does not.
        Training stops early after ``patience`` evaluations without a new best.

        beta > 0 trains on the alpha-divergence with a = 1 + beta instead of KL (see grad);
        f holds the integrand values at y (w = f / p for points drawn from p; f = w for
        uniform points, the default)."""
        n = y.shape[0]
        if beta and f is None:
            f = w
        g = torch.Generator(device=y.device)
        g.manual_seed(seed)
        losses = []
        best = None
        if resample:
            w64 = _lib.require_device_tensor(w, None, "w")
            if w64.dtype != torch.float64:
                w64 = w64.to(torch.float64)
            cdf = torch.empty(n, dtype=torch.float64, device=w64.device)
            ws = torch.empty(self.lib.mfx_flow_cdf_workspace(n) // 8, dtype=torch.float64, device=w64.device)
            _lib.check(self.lib.mfx_flow_cdf(_p(w64), n, _p(cdf), _p(ws), ws.numel() * 8, _lib.stream_ptr(w64.device)))
            if not float(ws[-1]) > 0:
                raise ValueError("train: the weights sum to zero")
            # guide table of the cdf: a draw then takes two or three loads, not a bisection
            guide = torch.empty(n, dtype=torch.int32, device=w64.device)
            _lib.check(self.lib.mfx_flow_cdf_guide(_p(cdf), n, _p(guide), n, _lib.stream_ptr(w64.device)))
            y = _lib.require_device_tensor(y, torch.float32, "y")
            fc = _lib.require_device_tensor(f, None, "f").to(torch.float64) if beta else None
            # device workspace of the fused step, allocated once: the minibatch is drawn
            # inside the forward kernel (Philox counter (sample, step), key seed)
            dev = y.device
            fb = torch.empty(batch, dtype=torch.float64, device=dev) if beta else None
            ld = torch.empty(batch, dtype=torch.float32, device=dev)
            gl = torch.empty(batch, dtype=torch.float32, device=dev)
            save = torch.empty(self.lib.mfx_flow_save_bytes(self.D, self.L, self._m(), batch) // 4 + 1,
                               dtype=torch.float32, device=dev)
            partial = torch.empty(self.lib.mfx_flow_partial_bytes(self.D, self.L, self._m(), batch) // 4 + 1,
                                  dtype=torch.float32, device=dev)
            gbuf = torch.empty_like(self.params)
            P = dict(par=_p(self.params), pb=_p(self.pbf16), cdf=_p(cdf), guide=_p(guide), y=_p(y), fc=_p(fc), fb=_p(fb),
                     ld=_p(ld),
                     gl=_p(gl), save=_p(save), partial=_p(partial), grad=_p(gbuf), st=_lib.stream_ptr(dev))
            seed32 = int(seed) & 0xFFFFFFFF
        for it in range(steps + 1):
            if val is not None and (it % eval_every == 0 or it == steps):
                m2 = self.is_second_moment(val[0], val[1], mix)
                if best is None or m2 < best[0]:
                    best = (m2, it, self.params.clone(), self.m1.clone(), self.m2.clone())
                if log:
                    log("flow step %d held-out second moment %.6g (best at step %d)" % (it, m2, best[1]))
                if patience and it - best[1] >= patience * eval_every:
                    break
            if it == steps:
                break
            cur_lr = lr * (0.5 * (1 + math.cos(math.pi * it / steps)) if lr_decay else 1.0)
            show = log and val is None and (it % max(1, steps // 10) == 0 or it == steps - 1)
            need_loss = show
            if resample:
                m = self._m()
                _lib.check(self.lib.mfx_flow_draw_forward(self.D, self.L, m, P["par"], P["pb"], P["cdf"], P["guide"], n,
                                                          n, P["y"], P["fc"], seed32, it, batch, None, P["fb"],
                                                          P["ld"], P["save"], P["st"]))
                _lib.check(self.lib.mfx_flow_seed(None, P["fb"], P["ld"], batch, float(beta), P["gl"], P["st"]))
                _lib.check(self.lib.mfx_flow_backward(self.D, self.L, m, P["par"], P["pb"], batch, P["save"], P["gl"],
                                                      None, P["partial"], P["grad"], P["st"]))
                loss = (gl.to(torch.float64) * ld.to(torch.float64)).sum() if need_loss else None
                grad = gbuf
            else:
                idx = torch.randint(0, n, (batch,), generator=g, device=y.device)
                yb = y[idx].contiguous()
                wb = w[idx].contiguous()
                fb = f[idx].contiguous() if beta else None
                if float(wb.sum()) <= 0:
                    continue
                loss, grad = self.grad(yb, wb, fb, beta, need_loss=need_loss)
            self.adam_step(grad, lr=cur_lr)
            if show:
                log("flow step %d loss %.5f" % (it, float(loss)))
                losses.append(loss)
        self.steps_run = it          # optimiser steps taken (early stopping may end before steps)
        if best is not None:
            self.params.copy_(best[2])
            self.m1.copy_(best[3])
            self.m2.copy_(best[4])
            self._cast()
            self.best_step = best[1]
            self.best_second_moment = best[0]
        return [float(x) for x in losses]

    def sample_mixture(self, u, u_plain, mix):
        """Stratified defensive mixture q' = (1 - mix) q + mix on the unit cube: the first
        round(mix * n) points are u_plain as they are, the rest come from the flow applied
        to u.  Returns (y, 1/q'(y)); the estimate sum f(y)/q'(y) / n stays unbiased."""
        n = u.shape[0]
        nu = int(round(mix * n))
        y = torch.empty_like(u)
        winv = torch.empty(n, dtype=torch.float64, device=u.device)
        st = _lib.stream_ptr(u.device)
        if nu > 0:
            yu = u_plain[:nu].contiguous()
            y[:nu] = yu
            _, ld = self._apply(yu, False)
            _lib.check(self.lib.mfx_flow_mixture_winv(_p(ld), nu, 0, float(mix), _p(winv[:nu]), st))
        if nu < n:
            yf, ld_inv = self._apply(u[nu:].contiguous(), True)
            y[nu:] = yf
            _lib.check(self.lib.mfx_flow_mixture_winv(_p(ld_inv), n - nu, 1, float(mix), _p(winv[nu:]), st))
        return y, winv
