"""TEST INFRASTRUCTURE ONLY: numpy FP64 reference of the hedge-ratio MLP and of
its training loss, used to check the device training kernels' forward pass,
gradients and Adam update.

Network (no reference counterpart; the reference has no control variates):
    x (11) -> Linear 64 -> SiLU -> Linear 64 -> SiLU -> Linear 64 -> SiLU -> Linear 5
Loss of one minibatch of B paths with J hedge dates each:
    CV_p = sum_j f(x_{p,j}) . dM_{p,j},   L = mean_p (P_p - CV_p)^2
Adam (Kingma & Ba 2015): m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
    w -= lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps)
"""
import numpy as np

SIZES = [(11, 64), (64, 64), (64, 64), (64, 5)]


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, x):
    """params: list of (W [in,out], b [out]); x [N, 11] -> (f [N, 5], cache)."""
    h = x
    cache = []
    for li, (W, b) in enumerate(params):
        z = h @ W + b
        cache.append((h, z))
        h = silu(z) if li < len(params) - 1 else z
    return h, cache


def loss_and_grads(params, x, dm, payoff):
    """x [B, J, 11], dm [B, J, 5], payoff [B] -> (loss, grads like params)."""
    B, J, _ = x.shape
    f, cache = forward(params, x.reshape(B * J, -1))
    cv = (f.reshape(B, J, 5) * dm).sum(axis=(1, 2))
    r = payoff - cv
    loss = np.mean(r * r)
    df = (-2.0 / B) * r[:, None, None] * dm
    g = df.reshape(B * J, 5)
    grads = [None] * len(params)
    for li in range(len(params) - 1, -1, -1):
        h, z = cache[li]
        if li < len(params) - 1:
            g = g * dsilu(z)
        W, _ = params[li]
        grads[li] = (h.T @ g, g.sum(axis=0))
        g = g @ W.T
    return loss, grads


def adam_step(params, grads, state, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    out = []
    for i, ((W, b), (gW, gb)) in enumerate(zip(params, grads)):
        new = []
        for j, (p, g) in enumerate(((W, gW), (b, gb))):
            m, v = state.setdefault((i, j), (np.zeros_like(p), np.zeros_like(p)))
            m = b1 * m + (1 - b1) * g
            v = b2 * v + (1 - b2) * g * g
            state[(i, j)] = (m, v)
            mh = m / (1 - b1 ** t)
            vh = v / (1 - b2 ** t)
            new.append(p - lr * mh / (np.sqrt(vh) + eps))
        out.append(tuple(new))
    return out
