"""GPU parity of the surrogate network's BF16 MFMA kernels (liblymph_mlp.so)
against the CPU oracle oracle/mlp_ref.py with the same bfloat16 roundings:
forward outputs, loss, every parameter gradient, and the Adam update.  GPU
and oracle differ only by accumulation order (FP32 on the matrix cores,
float64 in the oracle) and by the rare bfloat16 rounding of an activation
that lands on the other side of a tie, hence the tolerances below."""
import sys

import numpy as np
import pytest

from lymph3d_apsc_amd import device_count
from oracle import mlp_ref as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if device_count() == 0:
        pytest.skip("no HIP device")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - b) / max(np.linalg.norm(b), 1e-300))


# shapes that exercise full and partial 64x64 tiles, k tails, width 8 and the
# split-k path (output layers of 1536 and 38,400 outputs)
SHAPES = [([8, 64, 64, 128], 64), ([8, 72, 136, 40], 24), ([8, 512, 512, 512, 1536], 128),
          ([8, 64, 64, 38400], 32)]


@pytest.mark.parametrize("widths,batch", SHAPES)
def test_forward_loss_and_gradients_match_oracle(widths, batch):
    from lymph3d_apsc_amd.mlp import Mlp
    say(f"[mlp grads] {widths} batch {batch}")
    rng = np.random.default_rng(len(widths) + batch)
    x = rng.standard_normal((batch, widths[0])).astype(np.float32)
    t = rng.standard_normal((batch, widths[-1])).astype(np.float32)
    with Mlp(widths, max_batch=max(batch, 64), seed=3) as net:
        p = net.params().astype(np.float64)
        y_ref, _ = R.forward(p, widths, x, emulate=True)
        y = net.predict(x)
        assert rel(y, y_ref) < 5e-3
        loss, g = net.gradients(x, t)
        loss_ref, g_ref = R.loss_and_grads(p, widths, x, t, emulate=True)
        assert abs(loss - loss_ref) < 1e-3 * loss_ref
        # per layer: weights and biases
        o = 0
        for k, n in zip(widths[:-1], widths[1:]):
            for size in (n * k, n):
                assert rel(g[o:o + size], g_ref[o:o + size]) < 2e-2, (k, n, size)
                o += size


def test_adam_step_matches_oracle_and_training_reduces_loss():
    from lymph3d_apsc_amd.mlp import Mlp
    say("[mlp train]")
    widths, batch = [8, 64, 64, 32], 32
    rng = np.random.default_rng(11)
    x = rng.standard_normal((256, 8)).astype(np.float32)
    A = rng.standard_normal((8, 32)) / 3
    y = np.tanh(x @ A).astype(np.float32)  # smooth target
    with Mlp(widths, max_batch=64, seed=5) as net:
        net.set_data(x, y)
        p0 = net.params().astype(np.float64)
        order = np.arange(batch, dtype=np.int32)
        _, g = net.gradients(x[:batch], y[:batch])
        net.train(order, batch, lr=1e-3)
        p1 = net.params()
        q, _, _ = R.adam(p0, g.astype(np.float64), 0.0, 0.0, 1, 1e-3)
        assert np.allclose(p1, q, rtol=1e-5, atol=1e-7)
        losses = []
        for epoch in range(60):
            order = rng.permutation(256).astype(np.int32)
            losses.append(net.train(order, batch, lr=3e-3).mean())
        assert losses[-1] < 0.1 * losses[0]
        yp = net.predict(x)
        assert rel(yp, y) < 0.25
