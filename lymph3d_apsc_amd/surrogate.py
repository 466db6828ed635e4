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
truth by the same path with the FP64 apply.  An ensemble of networks (GELU)
is trained on the GPU by liblymph_mlp.so (BF16 MFMA, FP32 master weights,
Adam).  Two models:
- "receiver" (default): one example per (instance, receiver).  Inputs are
  the parameters together with what the physics fixes for that pair
  (receiver position, source-receiver direction and distance, P and S travel
  times along the straight path through the instance's field, the local wave
  speeds); the output is the receiver's trace, times its distance (geometric
  spreading), in the leading SVD modes of the training traces.
- "global": the parameters to all 64 receivers' traces at once (widths
  8-512-512-512-38400), centred and scaled by one global factor.
At evaluation the spread of the networks' predictions is the error
estimate: a prediction whose relative spread exceeds the threshold is
replaced by a solver run.

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


def train_networks(nets, n, epochs, batch, lr, seed, log=None):
    """Each network trained on its n examples (set_data done) for `epochs`
    epochs of shuffled batches (its own seeded order), Adam at lr (x0.3 from
    60% of the epochs, x0.1 from 85%).  The networks run on their own
    streams from one host thread each, so their small GEMMs overlap on the
    GPU; each network's arithmetic is that of a run on its own.  Returns the
    mean loss per epoch of every network."""
    from concurrent.futures import ThreadPoolExecutor
    steps_per_epoch = n // batch

    def one(i):
        net = nets[i]
        rng = np.random.default_rng(seed + 1000 + i)
        losses = []
        for ep in range(epochs):
            frac = ep / max(1, epochs - 1)
            rate = lr if frac < 0.6 else (lr * 0.3 if frac < 0.85 else lr * 0.1)
            order = rng.permutation(n)[:steps_per_epoch * batch].astype(np.int32)
            losses.append(float(net.train(order, batch, rate).mean()))
            if log and (ep + 1) % max(1, epochs // 10) == 0:
                log(f"net {i}: epoch {ep + 1}/{epochs}, loss {losses[-1]:.3e}")
        return losses

    with ThreadPoolExecutor(max_workers=len(nets)) as ex:
        hist = list(ex.map(one, range(len(nets))))
    if log:
        for i, losses in enumerate(hist):
            log(f"net {i}: loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    return hist


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
        for net in self.nets:
            net.set_data(x, yn)
        return train_networks(self.nets, len(x), epochs, batch, lr, self.seed, log)

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


N_FEAT = 32
R_MIN = 0.05  # distance floor of the geometric-spreading factor


def _gauss01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1), 0.5 * w


def receiver_features(thetas, rec, n_cells, n_q=8):
    """(N, R, N_FEAT) inputs of the receiver model: source xyz, receiver xyz,
    unit source-receiver direction, distance, log distance, f0, the four
    field coefficients, P and S travel times along the straight segment
    (distance times the segment's mean slowness, Gauss-Legendre in n_q
    points; rho = 1 so c_P = sqrt(lambda + 2 mu), c_S = sqrt(mu)), the same
    times f0, c_P and c_S at the source and at the receiver, and the source's
    and the receiver's positions inside their cells of the n_cells^3 cube
    mesh (a point force's discrete response near the source depends on
    them), and two zeros."""
    th = np.asarray(thetas, np.float64)
    src, f0, xi = th[:, :3], th[:, 3], th[:, 4:8]
    d = rec[None, :, :] - src[:, None, :]
    dist = np.linalg.norm(d, axis=2)
    gam = d / np.maximum(dist, 1e-12)[..., None]
    sq, wq = _gauss01(n_q)
    pts = src[:, None, None, :] + sq[None, None, :, None] * d[:, :, None, :]

    def speeds(x, z, k):
        lam = np.exp(SIGMA * (xi[:, 0].reshape(k) + xi[:, 1].reshape(k) * (2 * x - 1)))
        mu = np.exp(SIGMA * (xi[:, 2].reshape(k) + xi[:, 3].reshape(k) * (2 * z - 1)))
        return np.sqrt(lam + 2 * mu), np.sqrt(mu)

    cp, cs = speeds(pts[..., 0], pts[..., 2], (-1, 1, 1))
    tp = dist * (wq / cp).sum(axis=2)
    ts = dist * (wq / cs).sum(axis=2)
    cps, css = speeds(src[:, 0], src[:, 2], (-1,))
    cpr, csr = speeds(rec[None, :, 0], rec[None, :, 2], (-1, 1))
    N, R = dist.shape
    f = np.zeros((N, R, N_FEAT))
    f[..., 0:3] = src[:, None, :]
    f[..., 3:6] = rec[None, :, :]
    f[..., 6:9] = gam
    f[..., 9] = dist
    f[..., 10] = np.log(np.maximum(dist, R_MIN))
    f[..., 11] = f0[:, None]
    f[..., 12:16] = xi[:, None, :]
    f[..., 16] = tp
    f[..., 17] = ts
    f[..., 18] = tp * f0[:, None]
    f[..., 19] = ts * f0[:, None]
    f[..., 20] = cps[:, None]
    f[..., 21] = css[:, None]
    f[..., 22] = cpr
    f[..., 23] = csr
    f[..., 24:27] = (n_cells * src - np.floor(n_cells * src))[:, None, :]
    f[..., 27:30] = (n_cells * rec - np.floor(n_cells * rec))[None, :, :]
    return f


class ReceiverSurrogate:
    """Networks from (parameters, receiver) features to one receiver's trace
    in n_modes SVD modes; rec is the fixed receiver array."""

    def __init__(self, rec, n_t, n_cells, n_nets=4, width=512, depth=6, n_modes=64, max_batch=256,
                 seed=0):
        self.rec = np.asarray(rec, np.float64)
        self.n_cells = n_cells
        self.n_t, self.n_modes = n_t, n_modes
        self.widths = [N_FEAT] + [width] * (depth - 1) + [n_modes]
        self.nets = [Mlp(self.widths, max_batch=max_batch, seed=seed + 17 * i) for i in range(n_nets)]
        self.seed = seed

    def _x(self, thetas):
        f = receiver_features(thetas, self.rec, self.n_cells)
        return ((f - self.f_mean) / self.f_std).reshape(-1, N_FEAT).astype(np.float32), f[..., 9]

    def fit(self, thetas, traces, epochs=100, batch=256, lr=1e-3, log=None):
        N, T, R, _ = traces.shape
        f = receiver_features(thetas, self.rec, self.n_cells)
        self.f_mean = f.reshape(-1, N_FEAT).mean(axis=0)
        self.f_std = np.maximum(f.reshape(-1, N_FEAT).std(axis=0), 1e-12)
        x, dist = self._x(thetas)
        # one example per (instance, receiver): the trace times max(distance, R_MIN)
        y = np.ascontiguousarray(traces.transpose(0, 2, 1, 3)).reshape(N * R, T * 3).astype(np.float64)
        y *= np.maximum(dist, R_MIN).reshape(-1, 1)
        self.y_mean = y.mean(axis=0)
        rng = np.random.default_rng(self.seed + 7)
        sub = rng.choice(len(y), min(len(y), 20000), replace=False)
        _, sv, vt = np.linalg.svd(y[sub] - self.y_mean, full_matrices=False)
        self.basis = vt[:self.n_modes].T  # (T*3, n_modes), orthonormal columns
        self.basis_energy = float((sv[:self.n_modes] ** 2).sum() / (sv ** 2).sum())
        c = (y - self.y_mean) @ self.basis
        self.c_std = max(float(np.sqrt(np.mean(c ** 2))), 1e-300)
        cn = (c / self.c_std).astype(np.float32)
        for net in self.nets:
            net.set_data(x, cn)
        return train_networks(self.nets, len(x), epochs, batch, lr, self.seed, log)

    def _decode(self, c, dist, N):
        y = (c.astype(np.float64) * self.c_std) @ self.basis.T + self.y_mean
        y /= np.maximum(dist, R_MIN).reshape(-1, 1)
        R = len(self.rec)
        return y.reshape(N, R, self.n_t, 3).transpose(0, 2, 1, 3).reshape(N, -1)

    def project(self, thetas, traces):
        """The traces as the SVD basis represents them (the floor of the
        model's error)."""
        N, T, R, _ = traces.shape
        _, dist = self._x(thetas)
        y = np.ascontiguousarray(traces.transpose(0, 2, 1, 3)).reshape(N * R, T * 3).astype(np.float64)
        y *= np.maximum(dist, R_MIN).reshape(-1, 1)
        return self._decode(((y - self.y_mean) @ self.basis) / self.c_std, dist, N)

    def predict_all(self, thetas):
        x, dist = self._x(thetas)
        return np.stack([self._decode(net.predict(x), dist, len(thetas)) for net in self.nets])

    def predict(self, thetas):
        """(mean prediction, relative spread per sample)."""
        P = self.predict_all(thetas)
        mean = P.mean(axis=0)
        spread = np.linalg.norm(P.std(axis=0), axis=1) / np.maximum(np.linalg.norm(mean, axis=1), 1e-300)
        return mean, spread

    def time_predict(self, thetas, reps=10):
        """Seconds for one prediction of every sample by every network
        (inputs resident on the device; the modes' decoding is not timed)."""
        x, _ = self._x(thetas)
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
        width=512, epochs=300, batch=256, lr=1e-3, threshold=5e-2, chunk=512, plain_sample=8,
        seed0=10_000, model="receiver", n_modes=64, log=None):
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
    floor = None
    if model == "receiver":
        sur = ReceiverSurrogate(rec, steps // every, n, n_nets=n_nets, width=width, n_modes=n_modes,
                                max_batch=max(batch, 64))
        hist = sur.fit(th_train, y_train, epochs=epochs, batch=batch, lr=lr, log=log)
        n_examples = n_train * n_receivers
        floor = rel_err(sur.project(th_eval, y_eval64), y_eval64)
    else:
        sur = Surrogate(y_train[0].size, n_nets=n_nets, width=width, max_batch=max(batch, 64))
        hist = sur.fit(th_train, y_train, epochs=epochs, batch=batch, lr=lr, log=log)
        n_examples = n_train
    t_train = time.perf_counter() - t0
    w = sur.widths
    # GEMM flops of one training step: forward, weight gradient, input gradient
    # (none for the first layer's input)
    step_flops = 2.0 * batch * sum(k * n for k, n in zip(w[:-1], w[1:])) * 3 - 2.0 * batch * w[0] * w[1]
    n_steps = n_nets * epochs * (n_examples // batch)
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
    for tau in sorted({threshold, 0.01, 0.02, 0.05, 0.1, 0.2, 0.3, 0.5, 1.0}):
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
                   "outputs": int(y_train[0].size), "model": model, "widths": sur.widths,
                   "nets": n_nets, "epochs": epochs, "batch": batch, "lr": lr,
                   "threshold": threshold, "dt": dt,
                   "n_modes": n_modes if model == "receiver" else None},
        "basis": None if floor is None else {
            "energy_kept": sur.basis_energy, "projection_rel_l2_vs_fp64": q(floor)},
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
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--nets", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--model", choices=["receiver", "global"], default="receiver")
    ap.add_argument("--modes", type=int, default=64)
    ap.add_argument("--batch", type=int, default=256)
    a = ap.parse_args()
    res = run(n_train=a.n_train, n_eval=a.n_eval, n=a.n, epochs=a.epochs, n_nets=a.nets, lr=a.lr,
              model=a.model, n_modes=a.modes, batch=a.batch,
              log=lambda *m: print(*m, file=sys.stderr, flush=True))
    print(json.dumps(res))
