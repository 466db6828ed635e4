"""Batched ensembles (the config of 256 instances of 10^3-20^3 cubes, p = 2-3,
smooth lognormal lambda/mu, Ricker point force, 64 receivers, 1000 steps, FP32
apply, whole solve captured in a HIP graph).

1. Against the oracle: an FP64 ensemble of small uneven instances (mixed
   degree, size and wavelet frequency, joined into block-diagonal contexts,
   per-group Ricker factors computed on the device, graph replay) gives each
   instance's receiver traces within 1e-11 of the oracle's own step loop with
   the same point source and receivers.
2. At full size: the 256-instance ensemble, one replay of the captured solve,
   every instance's traces finite and nonzero, and a sample of instances
   within 1e-4 (relative L2 over the whole trace) of their standalone FP64
   runs.  The FP32 apply's error against FP64 is stated per result; 1e-4 is
   thirty times the 3e-6 measured on this workload.
3. The 48-bit apply in the ensemble (FP32 operator plus 16-bit corrections,
   FP64 sums, FP64 blocks released after the split): the small uneven
   instances' traces within 1e-7 of the FP64 ensemble's and a hundredth of
   the 32-bit ensemble's distance from it.
"""
import time

import numpy as np
import pytest

import lymph3d_apsc_amd as L
from lymph3d_apsc_amd import _lib, ensemble as E
from oracle import oracle

from tests.test_gpu_parity import oracle_for, rel, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if L.device_count() == 0:
        pytest.skip("no HIP device")


def _oracle_traces(ins, dt, nsteps, every):
    """The oracle's step loop from rest with the instance's point force and
    wavelet; receiver values of u^{n+1} after every `every`-th step."""
    m, p, f0 = ins["mesh"], ins["p"], ins["f0"]
    o = oracle_for(m, p)
    o.build_operator()
    rs = o.rhs(oracle.PROB_NULL)
    rd = o.rhs(oracle.PROB_NULL) + o.point_sources(ins["source"], o.locate(ins["source"]), ins["force"])
    rtet = o.locate(ins["receivers"])
    z = np.zeros(o.n_poly * o.n3)
    t = 0.0
    un = o.first_step(oracle.ricker(f0, 1.2 / f0, t), dt, z, z, rs, rd)
    up = z
    t = t + dt
    out = []
    for s in range(1, nsteps + 1):
        up, un = un, o.step(oracle.ricker(f0, 1.2 / f0, t), dt, up, un, rs, rd)
        t = t + dt
        if s % every == 0:
            out.append(o.eval_points(ins["receivers"], rtet, un))
    return np.array(out)


def test_fp64_ensemble_matches_oracle():
    say("[ensemble] fp64 vs oracle")
    inst = [E.make_instance(100 + s, n_range=(2, 4), n_receivers=12, f0=4.0 + 0.7 * s) for s in range(6)]
    assert len({i["p"] for i in inst}) == 2 and len({i["n"] for i in inst}) >= 2
    dt, nsteps, every = 2e-3, 60, 4
    ens = E.Ensemble(inst, dt, nsteps, every=every, apply_bits=64, autotune=False)
    tr = ens.run()
    ens.close()
    for k, ins in enumerate(inst):
        ref = _oracle_traces(ins, dt, nsteps, every)
        assert np.abs(ref).max() > 0
        assert tr[k].shape == ref.shape
        assert rel(tr[k], ref) < 1e-11, (k, rel(tr[k], ref))


def test_full_size_fp32_ensemble():
    """256 instances at the configured sizes, 1000 steps, one replay."""
    n_inst, steps, f0 = 256, 1000, 5.0
    t0 = time.time()
    inst = [E.make_instance(1000 + s) for s in range(n_inst)]
    dt = min(0.2 * (1.0 / i["n"]) / (np.sqrt(((i["mesh"].lam + 2 * i["mesh"].mu) / i["mesh"].rho).max())
                                     * i["p"] ** 2) for i in inst)
    say(f"[ensemble] 256 instances generated in {time.time() - t0:.1f} s")
    ens = E.Ensemble(inst, dt, steps, every=1, f0=f0, apply_bits=32, autotune=False)
    say(f"[ensemble] set up in {time.time() - t0:.1f} s")
    tr = ens.run()
    say(f"[ensemble] one replay done at {time.time() - t0:.1f} s")
    assert {g["p"] for g in ens.groups} == {2, 3}
    ens.close()
    del ens
    for k, t in enumerate(tr):
        assert t.shape == (steps, 64, 3), k
        assert np.isfinite(t).all(), k
        assert np.abs(t).max() > 0, k
    for k in (0, 77, 201):  # sampled standalone FP64 runs
        ins = inst[k]
        with L.Context(ins["mesh"], ins["p"]) as ctx:
            ctx.set_keep_adg(False)
            ctx.set_build("factored")
            t = ctx.start_from_rest(dt, ins["source"], ins["force"], f0=f0)
            ctx.set_receivers(ins["receivers"], max_records=steps)
            ctx.step_n_record(L.PROB_RICKER, steps, t, dt, every=1)
            ref = ctx.traces()
        err = np.linalg.norm(tr[k] - ref) / np.linalg.norm(ref)
        say(f"[ensemble] instance {k} (n={ins['n']}, p={ins['p']}): FP32 ensemble vs FP64 {err:.2e}")
        assert err < 1e-4, (k, err)


def test_48bit_ensemble_matches_fp64_ensemble():
    """The same small uneven instances (mixed degree and size, per-instance
    wavelets) solved as an ensemble with the 48-bit apply, the 32-bit apply
    and the FP64 apply: per instance, the 48-bit traces within 1e-7 relative
    L2 of the FP64 ones and within a hundredth of the 32-bit traces' distance
    (the operator's rounding is 2^-36 against 2^-25)."""
    say("[ensemble] 48-bit vs fp64")
    inst = [E.make_instance(200 + s, n_range=(2, 4), n_receivers=12, f0=4.0 + 0.5 * s) for s in range(6)]
    dt, nsteps = 2e-3, 200
    out = {}
    for bits in (64, 48, 32):
        ens = E.Ensemble(inst, dt, nsteps, every=4, apply_bits=bits, autotune=False)
        out[bits] = ens.run()
        ens.close()
    for k in range(len(inst)):
        ref = out[64][k]
        assert np.abs(ref).max() > 0
        e48 = np.linalg.norm(out[48][k] - ref) / np.linalg.norm(ref)
        e32 = np.linalg.norm(out[32][k] - ref) / np.linalg.norm(ref)
        say(f"  instance {k}: 48-bit {e48:.2e}, 32-bit {e32:.2e}")
        assert e48 < 1e-7, (k, e48)
        assert e48 < 1e-2 * e32, (k, e48, e32)
