"""Learned surrogate of the batched solver: instance parameters -> receiver
traces, with the solver as fallback and as the judge of its error.

Instance parameters (8): source position (x, y, z) ~ U[0.3, 0.7]^3, Ricker
peak frequency f0 ~ U[4, 8] (t0 = 1.2/f0), and four coefficients
xi ~ N(0, 1) (clipped to +-2.5) of a smooth lognormal material field on the
n^3-cube mesh (6 tets per cube agglomerated per cube):
    lambda = exp(sigma (xi1 + xi2 (2x - 1))),  mu = exp(sigma (xi3 + xi4 (2z - 1))),
rho = 1, sigma = 0.5, evaluated at tet barycentres; Dirichlet on all faces; a
z-directed point force; 64 fixed seeded receivers recording u every `every`
steps.

Pipeline: the training and held-out sets are solved by the batched ensemble
path (ensemble.Ensemble: 32-bit apply, HIP-graph captured); the held-out
truth by the same path with the FP64 apply.  An ensemble of networks (widths
8-512-512-512-out, GELU) is trained on the GPU by liblymph_mlp.so (BF16 MFMA,
FP32 master weights, Adam) on standardised inputs and centred traces
scaled by one global factor.  At evaluation the spread of the networks' predictions
is the error estimate: a prediction whose relative spread exceeds the
threshold is replaced by a solver run.

No reference counterpart: LYMPH3D has no learned component.
"""
import time

import numpy as np

from . import ensemble as E
from .mlp import Mlp

SIGMA = 0.5
CLIP = 2.5
IN_MEAN = np.array([0.5, 0.5, 0.5, 6.0, 0, 0, 0, 0], np.float64)
IN_STD = np.array([0.4 / np.sqrt(12)] * 3 + [4.0 / np.sqrt(12)] + [1.0] * 4, np.float64)


def sample_params(seed):
    rng = np.random.default_rng(seed)
    src = rng.uniform(0.3, 0.7, 3)
    f0 = rng.uniform(4.0, 8.0)
    xi = np.clip(rng.standard_normal(4), -CLIP, CLIP)
    return np.concatenate([src, [f0], xi])


def receivers(n_receivers=64, seed=20240):
    return np.random.default_rng(seed).uniform(0.1, 0.9, (n_receivers, 3))


def instance(theta, n, p, rec):
    m = E._cube(n, True)
    v = m.xyz[m.tet_vert]
    bc = (((v[:, 0] + v[:, 1]) + v[:, 2]) + v[:, 3]) / 4
    xi = theta[4:]
    m.lam = np.exp(SIGMA * (xi[0] + xi[1] * (2 * bc[:, 0] - 1)))
    m.mu = np.exp(SIGMA * (xi[2] + xi[3] * (2 * bc[:, 2] - 1)))
    m.rho = np.ones(m.n_tet)
    return dict(n=n, p=p, mesh=m, source=theta[None, :3], force=np.array([[0.0, 0.0, 1.0]]),
                receivers=rec, f0=float(theta[3]))


def cfl_dt(n, p):
    """dt = 0.2 h / (c_p p^2) with the largest P-wave speed any parameter
    set can produce (|xi| <= CLIP)."""
    lam_max = np.exp(SIGMA * 2 * CLIP)
    return 0.2 * (1.0 / n) / (np.sqrt(3 * lam_max) * p * p)


def solve(thetas, n, p, steps, every, rec, apply_bits=32, chunk=512, log=None):
    """Receiver traces (len(thetas), steps//every, n_receivers, 3) of every
    parameter set, solved as batched ensembles of `chunk` instances; also
    returns (seconds in replays, seconds in set-up)."""
    dt = cfl_dt(n, p)
    out = np.zeros((len(thetas), steps // every, len(rec), 3),
                   np.float32 if apply_bits == 32 else np.float64)
    t_run = t_setup = 0.0
    for c0 in range(0, len(thetas), chunk):
        t0 = time.perf_counter()
        inst = [instance(th, n, p, rec) for th in thetas[c0:c0 + chunk]]
        ens = E.Ensemble(inst, dt, steps, every=every, apply_bits=apply_bits, autotune=False)
        t1 = time.perf_counter()
        tr = ens.run()
        t2 = time.perf_counter()
        for k, a in enumerate(tr):
            out[c0 + k] = a
        ens.close()
        t_setup += t1 - t0
        t_run += t2 - t1
        if log:
            log(f"  solved {min(c0 + chunk, len(thetas))}/{len(thetas)} ({apply_bits}-bit apply): "
                f"set-up {t1 - t0:.1f} s, solve {t2 - t1:.1f} s")
    return out, t_run, t_setup


class Surrogate:
    """Ensemble of networks on standardised inputs/outputs."""

    def __init__(self, n_out, n_nets=4, width=512, depth=4, max_batch=256, seed=0):
        self.widths = [8] + [width] * (depth - 1) + [n_out]
        self.nets = [Mlp(self.widths, max_batch=max_batch, seed=seed + 17 * i) for i in range(n_nets)]
        self.seed = seed

    def fit(self, thetas, traces, epochs=200, batch=256, lr=1e-3, log=None):
        x = ((thetas - IN_MEAN) / IN_STD).astype(np.float32)
        y = traces.reshape(len(traces), -1).astype(np.float64)
        # per-output mean, one global scale: the loss then weighs every
        # output as the relative L2 error of a trace does
        self.y_mean = y.mean(axis=0)
        self.y_std = max(float(np.sqrt(np.mean((y - self.y_mean) ** 2))), 1e-300)
        yn = ((y - self.y_mean) / self.y_std).astype(np.float32)
        n = len(x)
        steps_per_epoch = n // batch
        hist = []
        for i, net in enumerate(self.nets):
            net.set_data(x, yn)
            rng = np.random.default_rng(self.seed + 1000 + i)
            losses = []
            for ep in range(epochs):
                frac = ep / max(1, epochs - 1)
                rate = lr if frac < 0.6 else (lr * 0.3 if frac < 0.85 else lr * 0.1)
                order = rng.permutation(n)[:steps_per_epoch * batch].astype(np.int32)
                losses.append(float(net.train(order, batch, rate).mean()))
            hist.append(losses)
            if log:
                log(f"net {i}: loss {losses[0]:.3e} -> {losses[-1]:.3e}")
        return hist

    def predict_all(self, thetas):
        x = ((thetas - IN_MEAN) / IN_STD).astype(np.float32)
        return np.stack([net.predict(x).astype(np.float64) * self.y_std + self.y_mean
                         for net in self.nets])

    def predict(self, thetas):
        """(mean prediction, relative spread per sample)."""
        P = self.predict_all(thetas)
        mean = P.mean(axis=0)
        spread = np.linalg.norm(P.std(axis=0), axis=1) / np.maximum(np.linalg.norm(mean, axis=1), 1e-300)
        return mean, spread

    def time_predict(self, thetas, reps=10):
        """Seconds for one prediction of every sample by every network
        (inputs resident on the device)."""
        x = ((thetas - IN_MEAN) / IN_STD).astype(np.float32)
        dummy = np.zeros((len(x), self.widths[-1]), np.float32)
        total = 0.0
        for net in self.nets:
            net.set_data(x, dummy)
            total += net.time_predict(len(x), reps) * 1e-3
        return total

    def close(self):
        for net in self.nets:
            net.close()


def rel_err(a, b):
    a = a.reshape(len(a), -1)
    b = b.reshape(len(b), -1)
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)


def run(n_train=4096, n_eval=512, n=16, p=2, steps=1000, every=5, n_receivers=64, n_nets=4,
        width=512, epochs=200, batch=256, lr=1e-3, threshold=1e-2, chunk=512, plain_sample=8,
        seed0=10_000, log=None):
    """Whole pipeline; returns a dict of results (errors, throughputs)."""
    from . import _lib
    from .matrix_free import Context
    log = log or (lambda *a: None)
    rec = receivers(n_receivers)
    th_train = np.stack([sample_params(seed0 + i) for i in range(n_train)])
    th_eval = np.stack([sample_params(seed0 + n_train + i) for i in range(n_eval)])
    t0 = time.perf_counter()
    y_train, run32, set32 = solve(th_train, n, p, steps, every, rec, 32, chunk, log)
    log(f"training set: {n_train} FP32 ensemble solves in {run32:.1f} s (+{set32:.1f} s set-up)")
    y_eval32, run32e, _ = solve(th_eval, n, p, steps, every, rec, 32, chunk, log)
    y_eval64, run64e, set64 = solve(th_eval, n, p, steps, every, rec, 64, chunk, log)
    log(f"held-out set: FP32 {run32e:.1f} s, FP64 {run64e:.1f} s")
    t_data = time.perf_counter() - t0
    # plain FP64 path, one instance at a time, uncaptured
    dt = cfl_dt(n, p)
    t_plain = 0.0
    for k in range(plain_sample):
        ins = instance(th_eval[k], n, p, rec)
        with Context(ins["mesh"], p) as ctx:
            ctx.set_keep_adg(False)
            ctx.set_build("factored")
            t = ctx.start_from_rest(dt, ins["source"], ins["force"], f0=ins["f0"])
            ctx.set_receivers(rec, max_records=steps // every)
            ctx.sync()
            s1 = time.perf_counter()
            ctx.step_n_record(_lib.PROB_RICKER, steps, t, dt, every=every)
            ctx.sync()
            t_plain += time.perf_counter() - s1
            ref = ctx.traces()
        assert np.allclose(ref, y_eval64[k], rtol=0, atol=1e-12 * np.abs(ref).max())
    plain_rate = plain_sample / t_plain
    # networks
    t0 = time.perf_counter()
    sur = Surrogate(y_train[0].size, n_nets=n_nets, width=width, max_batch=max(batch, 64))
    hist = sur.fit(th_train, y_train, epochs=epochs, batch=batch, lr=lr, log=log)
    t_train = time.perf_counter() - t0
    w = sur.widths
    # GEMM flops of one training step: forward, weight gradient, input gradient
    # (none for the first layer's input)
    step_flops = 2.0 * batch * sum(k * n for k, n in zip(w[:-1], w[1:])) * 3 - 2.0 * batch * w[0] * w[1]
    n_steps = n_nets * epochs * (n_train // batch)
    mean, spread = sur.predict(th_eval)
    t_pred = sur.time_predict(th_eval)
    err_sur = rel_err(mean, y_eval64)
    err_32 = rel_err(y_eval32, y_eval64)
    err_meanpred = rel_err(np.broadcast_to(y_train.mean(axis=0), y_eval64.shape), y_eval64)
    fallback = spread > threshold
    hybrid_err = np.where(fallback, 0.0, err_sur)
    n_fb = int(fallback.sum())
    hybrid_time = t_pred + n_fb / plain_rate
    corr = float(np.corrcoef(spread, err_sur)[0, 1]) if n_eval > 2 else None
    # what each spread threshold buys: share answered by the networks, their
    # error, and the hybrid's throughput and error (fallbacks exact)
    tradeoff = []
    for tau in (threshold, 0.02, 0.05, 0.1, 0.2, 0.3, 0.5, 1.0):
        acc = spread <= tau
        t_h = t_pred + int((~acc).sum()) / plain_rate
        tradeoff.append({"threshold": tau, "accepted": float(acc.mean()),
                         "accepted_mean_err": float(err_sur[acc].mean()) if acc.any() else None,
                         "hybrid_mean_err": float(np.where(acc, err_sur, 0.0).mean()),
                         "hybrid_solves_per_s": n_eval / t_h})
    sur.close()
    q = lambda a: {"mean": float(np.mean(a)), "median": float(np.median(a)),  # noqa: E731
                   "p90": float(np.quantile(a, 0.9)), "max": float(np.max(a))}
    return {
        "config": {"n_train": n_train, "n_eval": n_eval, "mesh": f"{n}^3 cubes x 6 tets, agglomerated",
                   "p": p, "steps": steps, "every": every, "receivers": n_receivers,
                   "outputs": int(y_train[0].size), "widths": sur.widths, "nets": n_nets,
                   "epochs": epochs, "batch": batch, "lr": lr, "threshold": threshold, "dt": dt},
        "throughput_solves_per_s": {
            "plain_fp64": plain_rate,
            "ensemble_fp32": (n_train + n_eval) / (run32 + run32e),
            "ensemble_fp64": n_eval / run64e,
            "surrogate": n_eval / t_pred,
            "hybrid_with_fallback": n_eval / hybrid_time,
        },
        "rel_l2_vs_fp64": {
            "surrogate": q(err_sur), "ensemble_fp32": q(err_32),
            "training_mean_predictor": q(err_meanpred),
            "hybrid_with_fallback": q(hybrid_err),
        },
        "fallback_fraction": n_fb / n_eval,
        "threshold_tradeoff": tradeoff,
        "spread_vs_error_correlation": corr,
        "final_train_loss": [h[-1] for h in hist],
        "seconds": {"data": t_data, "train": t_train, "predict_all": t_pred},
        "training": {"steps": n_steps, "ms_per_step": 1e3 * t_train / n_steps,
                     "gemm_tflops": step_flops * n_steps / t_train / 1e12,
                     "bf16_mfma_peak_tflops": 2500.0},
    }


if __name__ == "__main__":
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="train and evaluate the learned surrogate")
    ap.add_argument("--n-train", type=int, default=4096)
    ap.add_argument("--n-eval", type=int, default=512)
    ap.add_argument("--n", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--nets", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    a = ap.parse_args()
    res = run(n_train=a.n_train, n_eval=a.n_eval, n=a.n, epochs=a.epochs, n_nets=a.nets, lr=a.lr,
              log=lambda *m: print(*m, file=sys.stderr, flush=True))
    print(json.dumps(res))
