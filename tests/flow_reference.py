"""CPU torch restatement of the RQ-spline coupling flow (test infrastructure): the checker
for the device forward / inverse / backward kernels.  Same parameter layout; the
hidden/output weights and the hidden activations are rounded to bf16 exactly where the
device kernels use bf16 (straight-through for gradients), everything else is fp32."""

import torch

H, K, DMAX, MIND, MINW, MINH = 128, 16, 16, 1e-3, 1e-3, 1e-3
NPD = 3 * K + 1


def bf(x):
    """bf16 rounding with a straight-through gradient."""
    return x + (x.to(torch.bfloat16).to(x.dtype) - x).detach()


def _bins(raw, minv):
    return minv + (1 - K * minv) * torch.softmax(raw, dim=-1)


def spline(raw, x, inverse=False):
    """raw (n, NPD), x (n,) -> (y, log|dy/dx|) for the forward map, (x, log|dx/du|) inverse."""
    w = _bins(raw[:, :K], MINW)
    h = _bins(raw[:, K:2 * K], MINH)
    d = MIND + torch.nn.functional.softplus(raw[:, 2 * K:])
    cw = torch.cat([torch.zeros_like(w[:, :1]), torch.cumsum(w, -1)], -1)
    ch = torch.cat([torch.zeros_like(h[:, :1]), torch.cumsum(h, -1)], -1)
    cw = torch.cat([cw[:, :K], torch.ones_like(cw[:, :1])], -1)
    ch = torch.cat([ch[:, :K], torch.ones_like(ch[:, :1])], -1)
    src = ch if inverse else cw
    v = x.clamp(0, 1)
    k = (v[:, None] >= src[:, 1:K]).sum(-1)
    g = lambda t: t.gather(1, k[:, None])[:, 0]
    X, Y = g(cw), g(ch)
    wk = g(cw[:, 1:]) - X
    hk = g(ch[:, 1:]) - Y
    a = g(d)
    c = d.gather(1, (k + 1)[:, None])[:, 0]
    s = hk / wk
    if not inverse:
        xi = ((v - X) / wk).clamp(0, 1)
        t = xi * (1 - xi)
        y = Y + hk * (s * xi * xi + a * t) / (s + (a + c - 2 * s) * t)
        P = c * xi * xi + 2 * s * t + a * (1 - xi) ** 2
        ld = 2 * torch.log(s) + torch.log(P) - 2 * torch.log(s + (a + c - 2 * s) * t)
        return y, ld
    dd = v - Y
    gg = a + c - 2 * s
    qa = hk * (s - a) + dd * gg
    qb = hk * a - dd * gg
    qc = -s * dd
    xi = (2 * qc / (-qb - torch.sqrt((qb * qb - 4 * qa * qc).clamp(min=0)))).clamp(0, 1)
    t = xi * (1 - xi)
    P = c * xi * xi + 2 * s * t + a * (1 - xi) ** 2
    ld = 2 * torch.log(s) + torch.log(P) - 2 * torch.log(s + gg * t)
    return X + xi * wk, -ld


def layer_params(flow, params, l):
    o = flow.layout(l)
    NP = o["NP"]
    W1 = params[o["W1"]:o["W1"] + DMAX * H].reshape(DMAX, H)
    b1 = params[o["b1"]:o["b1"] + H]
    W2 = params[o["W2"]:o["W2"] + H * H].reshape(H, H)
    b2 = params[o["b2"]:o["b2"] + H]
    W3 = params[o["W3"]:o["W3"] + H * NP].reshape(H, NP)
    b3 = params[o["b3"]:o["b3"] + NP]
    return W1, b1, W2, b2, W3, b3, o


def apply(flow, params, x, inverse=False):
    """Full flow on CPU: returns (out, logdet); ``params`` may require grad."""
    D, L = flow.D, flow.L
    x = x.clone()
    ld = torch.zeros(x.shape[0], dtype=x.dtype)
    order = range(L - 1, -1, -1) if inverse else range(L)
    for l in order:
        W1, b1, W2, b2, W3, b3, o = layer_params(flow, params, l)
        m = torch.as_tensor(flow.masks[l]).bool()
        A = torch.nonzero(~m)[:, 0]
        B = torch.nonzero(m)[:, 0]
        xa = x[:, A]
        h1 = bf(torch.relu(xa @ W1[:len(A)] + b1))
        h2 = bf(torch.relu(h1 @ bf(W2) + b2))
        raw = h2 @ bf(W3) + b3
        cols = []
        new = x.clone()
        for j, dim in enumerate(B.tolist()):
            y, l1 = spline(raw[:, j * NPD:(j + 1) * NPD], x[:, dim], inverse)
            new = new.index_put((torch.arange(x.shape[0]), torch.full((x.shape[0],), dim)), y)
            ld = ld + l1
        x = new
    return x, ld
