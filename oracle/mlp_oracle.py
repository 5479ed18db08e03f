"""TEST INFRASTRUCTURE ONLY: numpy FP64 reference of the hedge-ratio MLP and of
its training loss, used to check the device training kernels' forward pass,
gradients and Adam update.

Network (no reference counterpart; the reference has no control variates):
    x (11) -> Linear 64 -> a -> Linear 64 -> a -> Linear 64 -> a -> Linear 5
    a = SiLU, a(z) = z / (1 + e^-z)   ("silu", the specified network), or
    a(z) = z s(t), t = clip(z/6 + 1/2, 0, 1), s(t) = t^2 (3 - 2t)   ("gate")
Loss of one minibatch of B paths with J hedge dates each:
    CV_p = sum_j f(x_{p,j}) . dM_{p,j},   L = mean_p (P_p - CV_p)^2
Adam (Kingma & Ba 2015): m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
    w -= lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps)
"""
import numpy as np

SIZES = [(11, 64), (64, 64), (64, 64), (64, 5)]
# flat weight vector of the C-ABI (thoth_cv_weights): per layer W^T [out_pad][in_pad]
# row-major, then the biases padded to [64][64][64][16]
PAD = [(64, 16), (64, 64), (64, 64), (16, 64)]
W_OFF = [0, 1024, 1024 + 4096, 1024 + 8192]
B_OFF = [10240, 10240 + 64, 10240 + 128, 10240 + 192]
N_WEIGHTS = 10448


def unpack(flat):
    """flat float vector (thoth_cv_weights layout) -> [(W [in, out], b [out])]."""
    flat = np.asarray(flat, np.float64)
    params = []
    for (n_in, n_out), (rows, cols), wo, bo in zip(SIZES, PAD, W_OFF, B_OFF):
        WT = flat[wo:wo + rows * cols].reshape(rows, cols)
        params.append((WT[:n_out, :n_in].T.copy(), flat[bo:bo + n_out].copy()))
    return params


def pack(params):
    flat = np.zeros(N_WEIGHTS)
    for (W, b), (rows, cols), wo, bo in zip(params, PAD, W_OFF, B_OFF):
        WT = np.zeros((rows, cols))
        WT[:W.shape[1], :W.shape[0]] = W.T
        flat[wo:wo + rows * cols] = WT.ravel()
        flat[bo:bo + b.shape[0]] = b
    return flat


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def act(z, kind="gate"):
    """Hidden activation: "silu" a(z) = z / (1 + e^-z); "gate" the
    smoothstep-gated linear unit a(z) = z s(t), t = clip(z/6 + 1/2, 0, 1),
    s = t^2 (3 - 2t)."""
    if kind == "silu":
        return z * _sigmoid(z)
    t = np.clip(z / 6.0 + 0.5, 0.0, 1.0)
    return z * t * t * (3.0 - 2.0 * t)


def dact(z, kind="gate"):
    if kind == "silu":
        s = _sigmoid(z)
        return s * (1.0 + z * (1.0 - s))
    t = np.clip(z / 6.0 + 0.5, 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t) + z * t * (1.0 - t)


def _sigmoid(z):
    # 1 / (1 + e^-z) without overflow for large |z|
    e = np.exp(-np.abs(z))
    return np.where(z >= 0, 1.0 / (1.0 + e), e / (1.0 + e))


ACTIVATIONS = {0: "silu", 1: "gate"}  # THOTH_CV_SILU, THOTH_CV_GATE


def forward(params, x, matrix_cores=False, activation="silu"):
    """params: list of (W [in,out], b [out]); x [N, 11] -> (f [N, 5], cache).

    matrix_cores=True rounds where the BF16 matrix-core kernels round: the
    weights and every hidden activation a(z) to bfloat16 (products and sums
    stay exact here, FP32 on the device)."""
    h = x
    cache = []
    rnd = bf16 if matrix_cores else (lambda a: a)
    for li, (W, b) in enumerate(params):
        z = h @ rnd(W) + b
        cache.append((h, z))
        h = rnd(act(z, activation)) if li < len(params) - 1 else z
    return h, cache


def loss_and_grads(params, x, dm, payoff, matrix_cores=False, activation="silu"):
    """x [B, J, 11], dm [B, J, 5], payoff [B] -> (loss, grads like params).

    matrix_cores=True also rounds the back-propagated error to bfloat16 where
    it enters a matrix product (weight gradient, error of the layer below);
    the bias gradients use the unrounded error, as on the device."""
    B, J, _ = x.shape
    f, cache = forward(params, x.reshape(B * J, -1), matrix_cores, activation)
    cv = (f.reshape(B, J, 5) * dm).sum(axis=(1, 2))
    r = payoff - cv
    loss = np.mean(r * r)
    df = (-2.0 / B) * r[:, None, None] * dm
    g = df.reshape(B * J, 5)
    rnd = bf16 if matrix_cores else (lambda a: a)
    grads = [None] * len(params)
    for li in range(len(params) - 1, -1, -1):
        h, z = cache[li]
        if li < len(params) - 1:
            g = g * dact(z, activation)
        W, _ = params[li]
        grads[li] = (h.T @ rnd(g), g.sum(axis=0))
        g = rnd(g) @ rnd(W).T
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
