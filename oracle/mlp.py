"""ORACLE — test infrastructure only (see oracle/__init__.py).

FP64 forward and backward pass of the surrogate network n_in -> width (SiLU) x depth -> n_out under a mean
squared error, with the GPU kernels' BF16 rounding points reproduced (round to nearest even): the input, the
weights, every hidden activation SiLU(z), the output-loss gradient 2 (y - t)/(n n_out) and every back-propagated
pre-activation gradient are rounded to BF16; everything else is exact. The kernels then differ from this only
by FP32 accumulation order and the hardware exp2/rcp in the sigmoid.

Textbook back-propagation: for z_l = h_{l-1} W_lᵀ + b_l, h_l = SiLU(z_l) = z σ(z),
    dW_l = dz_lᵀ h_{l-1},  db_l = Σ_n dz_l,  dz_{l-1} = (dz_l W_l) ⊙ SiLU'(z_{l-1}),  SiLU'(z) = σ(z)(1 + z(1 - σ(z))).
"""
import numpy as np


def bf16(x):
    """Round FP32 values to BF16 (nearest, ties to even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def unpack(params, layers):
    """layers: (W offset, b offset, out, in) per layer -> [(W [out][in], b [out])]."""
    return [(params[w:w + o * i].reshape(o, i), params[b:b + o]) for w, b, o, i in layers]


def forward(params, layers, x, rounding=True):
    """x [n][in_pad] -> (y [n][out_pad], activations (h list, z list)); rounding=False is exact FP64."""
    r = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    p = unpack(np.asarray(params, np.float64), layers)
    h = [r(x)]
    zs = []
    for l, (W, b) in enumerate(p):
        z = h[-1] @ r(W).T + b
        zs.append(z)
        if l + 1 < len(p):
            h.append(r(z * sigmoid(z)))
    return zs[-1], (h, zs)


def loss(params, layers, x, target, n_out):
    y, _ = forward(params, layers, x, rounding=False)
    return np.mean((y[:, :n_out] - target[:, :n_out]) ** 2)


def backward_mse(params, layers, x, target, n_out, rounding=True):
    """Gradient (flat, FP64) of mean((y - t)²) over the first n_out outputs."""
    r = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    p = unpack(np.asarray(params, np.float64), layers)
    y, (h, zs) = forward(params, layers, x, rounding)
    n = x.shape[0]
    dz = np.zeros_like(y)
    if rounding:
        dz[:, :n_out] = bf16(2.0 * (np.float32(y[:, :n_out]) - np.float32(target[:, :n_out])) / np.float32(n * n_out))
    else:
        dz[:, :n_out] = 2.0 * (y[:, :n_out] - target[:, :n_out]) / (n * n_out)
    grad = np.zeros(len(params))
    for l in range(len(p) - 1, -1, -1):
        W, b = p[l]
        w_off, b_off, o, i = layers[l]
        grad[w_off:w_off + o * i] = (dz.T @ h[l]).ravel()
        grad[b_off:b_off + o] = dz.sum(axis=0)
        if l > 0:
            z = zs[l - 1]
            s = sigmoid(z)
            dz = r((dz @ r(W)) * s * (1.0 + z * (1.0 - s)))
    return grad, y
