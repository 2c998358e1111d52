"""FP64 CPU restatement of the warm-start network (no reference counterpart:
the reference has no learned component).  TEST INFRASTRUCTURE ONLY: tests/
use it to check the GPU kernels of libbo_hip.so (bo_mlp_*), never as the
thing measured or shipped.

Network: layers l = 0..L-1, Z_l = H_{l-1} W_l + b_l, H_l = relu(Z_l) for
l < L-1, output Y = Z_{L-1}; loss = mean((Y - T)^2).  With quant=True,
`forward` and `gradients` quantise to BF16 exactly where the GPU kernels do (inputs,
weights, hidden activations, output and hidden gradients), and compute
everything else in FP64, so the GPU's FP32-accumulated results must agree up
to FP32 summation rounding (and rare BF16 rounding-boundary flips).  The
exact (quant=False) versions are pinned by central finite differences in
tests/test_mlp_oracle.py.
"""
import numpy as np


def bf16_round(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    u = np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def unpack(params, shapes):
    """Flat parameters (per layer: W row-major d_in x d_out, then b) -> [(W, b)]."""
    out, o = [], 0
    for din, dout in shapes:
        W = params[o:o + din * dout].reshape(din, dout)
        o += din * dout
        b = params[o:o + dout]
        o += dout
        out.append((W, b))
    assert o == params.size
    return out


def pack(grads):
    return np.concatenate([np.concatenate([gW.ravel(), gb.ravel()]) for gW, gb in grads])


def _q(a, on):
    return bf16_round(a) if on else np.asarray(a, np.float64)


def forward(params, shapes, x, quant=False):
    layers = unpack(np.asarray(params, np.float64), shapes)
    h = _q(x, quant)
    hs = [h]
    for l, (W, b) in enumerate(layers):
        z = h @ _q(W, quant) + b
        if l + 1 < len(layers):
            h = _q(np.maximum(z, 0.0), quant)
            hs.append(h)
        else:
            return z, hs


def gradients(params, shapes, x, t, quant=False):
    """loss and d loss / d params (flat, same layout)."""
    layers = unpack(np.asarray(params, np.float64), shapes)
    y, hs = forward(params, shapes, x, quant)
    t = np.asarray(t, np.float64)
    n = y.size
    loss = float(np.sum((y - t) ** 2) / n)
    dz = _q(2.0 * (y - t) / n, quant)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        hin = hs[l]
        grads[l] = (hin.T @ dz, dz.sum(axis=0))
        if l == 0:
            break
        dh = dz @ _q(W, quant).T
        dz = _q(np.where(hs[l] > 0.0, dh, 0.0), quant)
    return loss, pack(grads)


def adam_steps(params, grads, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """Adam (Kingma & Ba 2015) in FP64 from zero moments: for t = 1, 2, ...
    with gradient grads[t-1], m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
    w -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps).  Returns the
    parameters after every step (a list)."""
    w = np.asarray(params, np.float64).copy()
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    out = []
    for t, g in enumerate(grads, start=1):
        g = np.asarray(g, np.float64)
        m = beta1 * m + (1.0 - beta1) * g
        v = beta2 * v + (1.0 - beta2) * g * g
        w = w - lr * (m / (1.0 - beta1 ** t)) / (np.sqrt(v / (1.0 - beta2 ** t)) + eps)
        out.append(w.copy())
    return out
