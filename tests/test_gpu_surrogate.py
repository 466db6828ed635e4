"""The learned surrogate end to end at a small size: training data from the
batched FP32 ensemble, networks trained on the GPU (BF16 MFMA kernels), held
out instances judged against the FP64 solver.  Shows what the network gains
over the method without it: predictions far faster than solver runs, more
accurate than the best parameter-free guess (the training-set mean trace),
and the spread-based fallback never makes the answer worse."""
import numpy as np
import pytest

from lymph3d_apsc_amd import device_count

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if device_count() == 0:
        pytest.skip("no HIP device")


def test_surrogate_small_end_to_end():
    from lymph3d_apsc_amd import surrogate as S
    print("[surrogate small]", flush=True)
    r = S.run(n_train=512, n_eval=32, n=4, p=2, steps=200, every=5, n_receivers=8, n_nets=2,
              width=128, epochs=150, batch=64, chunk=256, plain_sample=2, threshold=5e-2)
    print(r, flush=True)
    e = r["rel_l2_vs_fp64"]
    th = r["throughput_solves_per_s"]
    assert e["surrogate"]["mean"] < 0.8 * e["training_mean_predictor"]["mean"]
    assert e["ensemble_fp32"]["max"] < 1e-4
    assert e["hybrid_with_fallback"]["mean"] <= e["surrogate"]["mean"]
    assert th["surrogate"] > 10 * th["plain_fp64"]
    assert 0.0 <= r["fallback_fraction"] <= 1.0
