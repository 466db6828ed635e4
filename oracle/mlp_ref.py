"""CPU oracle of the surrogate network (test infrastructure only: imported by
tests/ and never by the product).

A fully connected network y = W_L g(... g(W_1 x + b_1) ...) + b_L with
g = GELU (erf form), mean-squared-error loss and bias-corrected Adam, written
out in numpy so that the GPU kernels of liblymph_mlp.so can be checked
against it.  No reference counterpart: LYMPH3D has no learned component; the
formulas are the textbook ones and are pinned in tests/test_mlp_oracle.py by
finite differences and by torch's CPU autograd/Adam.

`emulate=True` reproduces the GPU's storage roundings: GEMM operands (input,
weights, hidden activations, output and hidden-layer gradients) rounded to
bfloat16 (round to nearest even), everything else kept in float64 here
(FP32 on the GPU), so GPU and oracle differ only by accumulation order.
"""
import numpy as np
from scipy.special import erf


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float32 values."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(a.shape)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def split(params, widths):
    """Parameter vector -> [(W_l [n_out, n_in], b_l [n_out])], the ABI layout."""
    out, o = [], 0
    for k, n in zip(widths[:-1], widths[1:]):
        W = params[o:o + n * k].reshape(n, k)
        o += n * k
        b = params[o:o + n]
        o += n
        out.append((W, b))
    assert o == params.size
    return out


def forward(params, widths, x, emulate=True):
    """Returns (y, cache)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    layers = split(np.asarray(params, np.float64), widths)
    H = [np.asarray(r(x), np.float64)]
    Z = []
    for l, (W, b) in enumerate(layers):
        Wr = np.asarray(r(W), np.float64)
        z = H[-1] @ Wr.T + b
        if l + 1 < len(layers):
            Z.append(z)
            H.append(np.asarray(r(gelu(z)), np.float64))
        else:
            y = z
    return y, (H, Z, layers)


def loss_and_grads(params, widths, x, t, emulate=True):
    """Mean-squared error and its gradient w.r.t. the parameter vector."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    y, (H, Z, layers) = forward(params, widths, x, emulate)
    d = y - t
    loss = float(np.mean(d * d))
    dZ = np.asarray(r(2.0 * d / d.size), np.float64)
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads.append((dZ.T @ H[l], dZ.sum(axis=0)))
        if l > 0:
            Wr = np.asarray(r(W), np.float64)
            dZ = np.asarray(r((dZ @ Wr) * gelu_grad(Z[l - 1])), np.float64)
    g = []
    for gW, gb in reversed(grads):
        g.append(gW.ravel())
        g.append(gb)
    return loss, np.concatenate(g)


def adam(params, grads, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (t counts from 1); returns (p, m, v)."""
    m = b1 * m + (1 - b1) * grads
    v = b2 * v + (1 - b2) * grads * grads
    p = params - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
