"""The surrogate network's CPU oracle (oracle/mlp_ref.py) pinned against
what fixes it independently: torch's bfloat16 conversion (round to nearest
even), torch's GELU and autograd in float64, finite differences, and one step
of torch.optim.Adam.  No reference counterpart (LYMPH3D has no learned
part)."""
import numpy as np
import pytest
import torch

from oracle import mlp_ref as R

WIDTHS = [8, 24, 16, 32]


def _params(widths, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for k, n in zip(widths[:-1], widths[1:]):
        s = 1.0 / np.sqrt(k)
        out.append(rng.uniform(-s, s, n * k))
        out.append(rng.uniform(-s, s, n))
    return np.concatenate(out)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    a = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 10000),
                        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -0.0, np.inf, -np.inf, 3.4e38,
                                  1e-40, 2 ** -133], np.float32)])
    ours = R.bf16(a)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours, ref)
    assert R.bf16(np.float32(1.0 + 2 ** -8)) == 1.0  # tie -> even
    assert np.isnan(R.bf16(np.array([np.nan], np.float32)))[0]


def test_gelu_matches_torch():
    z = np.linspace(-8, 8, 2001)
    zt = torch.tensor(z, dtype=torch.float64, requires_grad=True)
    y = torch.nn.functional.gelu(zt)
    y.sum().backward()
    assert np.allclose(R.gelu(z), y.detach().numpy(), rtol=1e-14, atol=1e-15)
    assert np.allclose(R.gelu_grad(z), zt.grad.numpy(), rtol=1e-13, atol=1e-14)


def test_forward_and_gradients_match_torch_autograd():
    p = _params(WIDTHS)
    rng = np.random.default_rng(2)
    x, t = rng.standard_normal((16, WIDTHS[0])), rng.standard_normal((16, WIDTHS[-1]))
    loss, g = R.loss_and_grads(p, WIDTHS, x, t, emulate=False)
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    h = torch.tensor(x)
    o = 0
    for l, (k, n) in enumerate(zip(WIDTHS[:-1], WIDTHS[1:])):
        W = pt[o:o + n * k].reshape(n, k)
        o += n * k
        b = pt[o:o + n]
        o += n
        h = h @ W.T + b
        if l + 2 < len(WIDTHS):
            h = torch.nn.functional.gelu(h)
    lt = torch.mean((h - torch.tensor(t)) ** 2)
    lt.backward()
    assert abs(loss - lt.item()) < 1e-13 * max(1.0, abs(loss))
    assert np.allclose(g, pt.grad.numpy(), rtol=1e-12, atol=1e-14)


def test_gradients_match_finite_differences():
    p = _params(WIDTHS, seed=3)
    rng = np.random.default_rng(4)
    x, t = rng.standard_normal((8, WIDTHS[0])), rng.standard_normal((8, WIDTHS[-1]))
    _, g = R.loss_and_grads(p, WIDTHS, x, t, emulate=False)
    for i in rng.choice(p.size, 40, replace=False):
        e = np.zeros_like(p)
        e[i] = 1e-6
        lp = R.loss_and_grads(p + e, WIDTHS, x, t, emulate=False)[0]
        lm = R.loss_and_grads(p - e, WIDTHS, x, t, emulate=False)[0]
        assert abs((lp - lm) / 2e-6 - g[i]) < 1e-7 + 1e-6 * abs(g[i])


def test_emulated_rounding_is_a_small_perturbation():
    p = _params(WIDTHS, seed=5)
    rng = np.random.default_rng(6)
    x, t = rng.standard_normal((16, WIDTHS[0])), rng.standard_normal((16, WIDTHS[-1]))
    l0, g0 = R.loss_and_grads(p, WIDTHS, x, t, emulate=False)
    l1, g1 = R.loss_and_grads(p, WIDTHS, x, t, emulate=True)
    assert abs(l1 - l0) < 2e-2 * l0
    assert np.linalg.norm(g1 - g0) < 3e-2 * np.linalg.norm(g0)
    assert not np.array_equal(g1, g0)


def test_adam_matches_torch():
    rng = np.random.default_rng(7)
    p = rng.standard_normal(50)
    pt = torch.tensor(p, requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    m = v = np.zeros_like(p)
    q = p.copy()
    for step in range(1, 4):
        g = rng.standard_normal(50)
        opt.zero_grad()
        pt.grad = torch.tensor(g)
        opt.step()
        q, m, v = R.adam(q, g, m, v, step, 1e-2)
    assert np.allclose(q, pt.detach().numpy(), rtol=1e-12, atol=1e-14)
