#!/usr/bin/env python3
"""Benchmark of the FP64 matrix-free PolyDG leapfrog step on one MI355X.

Workload (default): the unit cube split into 40^3 cubes x 6 tets (384,000
tets), each cube's 6 tets agglomerated into one polyhedron (64,000
polyhedra), p = 3 (Np = 20, 3Np = 60 -> 3,840,000 DOFs), standing-wave data
u_i = sin(pi x)sin(pi y)sin(pi z)cos(sqrt3 pi t), Dirichlet on all faces,
rho = lambda = mu = 1, SIPG alpha = 10, dt = 0.2 h / (c_p p^2) with
c_p = sqrt(3), 1000 timed leapfrog steps.  Synthetic mesh built in-process.

One "step" = one leapfrog step over the whole mesh (one launch of the fused
step kernel).  value = DOF-updates/s = 3 Np N_poly / t_step with all data
resident in HBM.  The timed region is bracketed by device synchronisation
and measured with HIP events recorded on the stream the steps run on.

roofline: algorithmic bytes per step B = 8 [n_blk (3Np)^2 + N_poly Np(Np+1)/2
+ 5 N_poly 3Np] (operator blocks, packed mass factor, five vectors) divided by
the average step-kernel time, against the MI355X HBM peak of 8 TB/s.
cpu_baseline: the C oracle (a scalar restatement of the reference's
algorithm, OpenMP over elements) on the reference's own per-tet layout of a
mesh with the same DOF count (22^3 cubes x 6 tets at p = 3), timed for a
bounded number of steps on every core this job may use and on one thread,
with the GPU's FP64 step on that same mesh beside it.

Usage: python bench.py [--gpus 1] [--steps 1000] [--warmup 20] [--n 40] [--p 3]
       python bench.py --degree-sweep     (p = 1..4 step table instead)
       python bench.py --uneven           (uneven polyhedra, p = 4, heterogeneous, instead)
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DOF-updates/s + %HBM roofline at p=3; ensemble solves/s vs plain at fixed L2"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stream_probe_gbps():
    """Achievable HBM bandwidth of this box: (copy GB/s, read GB/s) of
    STREAM-style kernels (liblymph_bwprobe.so), or (None, None).  Copy: 2 GiB.
    Read: the best of plain loads (k_read, 4 per thread) and non-temporal
    loads (k_read_nt, 8-32 per thread; k_read_persist, long-lived readers
    with 16 or 32 per thread: the way the step kernel streams its blocks)
    over 2 GiB and over 12 GiB (the step's size)."""
    path = os.path.join(ROOT, "lymph3d_apsc_amd", "liblymph_bwprobe.so")
    if not os.path.exists(path):
        return None, None
    L = C.CDLL(path)
    out = []
    for fn, size, reps in ((L.lymph_bw_copy, 2 << 30, 20), (L.lymph_bw_read, 2 << 30, 20),
                           (L.lymph_bw_read_nt, 2 << 30, 20), (L.lymph_bw_read_nt, 12 << 30, 10)):
        fn.argtypes = [C.c_size_t, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        g, ms = C.c_double(), C.c_double()
        out.append(g.value if fn(size, reps, C.byref(g), C.byref(ms)) == 0 else None)
    reads = [r for r in out[1:] if r]
    return out[0], (max(reads) if reads else None)


def pmc_traffic(n, p, kernel="k_leapfrog"):
    """HBM bytes per step-kernel launch measured by the FETCH_SIZE/WRITE_SIZE
    counter runs summarised in profiles/ (calibrated against the copy and
    read probes), for the workload they were taken on; else None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_cube{n}_p{p}_pmc_traffic.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("- traffic_bytes_per_launch:"):
            return float(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def host_threads():
    """The cores this job may use: the CPU affinity set, capped by
    OMP_NUM_THREADS when the environment sets it."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n, (os.cpu_count() or 1)


def cpu_baseline(p, n_sample, target_s):
    """The C oracle's FP64 leapfrog step (the reference's per-element loops,
    OpenMP over elements) on the reference's own layout: n_sample^3 cubes x 6
    tets, one block row per tet (the reference's stepper does not agglomerate;
    22^3 x 6 = 63,888 tets at p = 3 has the bench workload's 3.83 M DOFs),
    timed for about target_s on every core this job may use and again on one
    thread.  The GPU's FP64 step on the same mesh is timed beside it."""
    from lymph3d_apsc_amd import mesh as M
    from oracle import oracle
    import lymph3d_apsc_amd as L
    from lymph3d_apsc_amd import _lib
    m = M.cube_mesh(n_sample, agglomerate=False)
    kinds = np.zeros(8, np.int32)
    kinds[[2, 3, 4]] = -1
    kinds[[5, 6, 7]] = -2
    space = np.zeros(8, np.int32)
    space[2] = space[5] = 1
    threads, visible = host_threads()
    oracle.lib().o_num_threads(threads)
    o = oracle.Oracle(m.xyz, m.tet_vert, m.tria_vert, m.tria_tag, kinds, space, p)
    t0 = time.perf_counter()
    o.build_operator()
    t_build = time.perf_counter() - t0
    rs = o.rhs(oracle.PROB_NULL)
    rd = o.rhs(oracle.PROB_CUBE_WAVE)
    u0 = o.project(oracle.FN_CUBE_U)
    dt = 0.2 * (1.0 / n_sample) / (np.sqrt(3.0) * p * p)
    dofs = 3 * o.Np * o.n_poly
    B = 8 * (o.n_blk * o.n3 ** 2 + o.n_poly * o.Np * (o.Np + 1) // 2 + 5 * o.n_poly * o.n3)

    def timed(nthreads, seconds):
        oracle.lib().o_num_threads(nthreads)
        a, b = u0.copy(), u0.copy()
        o.step(0.5, dt, a, b, rs, rd)  # one untimed step
        k, t1 = 0, time.perf_counter()
        while True:
            a, b = b, o.step(0.5, dt, a, b, rs, rd)
            k += 1
            el = time.perf_counter() - t1
            if el >= seconds:
                return k, el, b

    k_all, el_all, u_cpu = timed(threads, target_s)
    k_one, el_one, _ = timed(1, target_s)
    oracle.lib().o_num_threads(threads)
    # the GPU's FP64 step on the same per-tet mesh and state
    gpu = None
    if L.device_count() > 0:
        with L.Context(m, p) as ctx:
            ctx.set_keep_adg(False)
            ctx.set_build("factored")
            ctx.make_matrices_free()
            ctx.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_STAT)
            ctx.make_rhs_free(L.PROB_CUBE_WAVE, _lib.SLOT_RHS_DYN)
            ctx.set_slot(_lib.SLOT_U_PREV, u0)
            ctx.set_slot(_lib.SLOT_U_CUR, u0)
            ctx.autotune(reps=5)
            ctx.time_steps(L.PROB_CUBE_WAVE, 10, 0.0, dt)
            ms = ctx.time_steps(L.PROB_CUBE_WAVE, 200, 0.0, dt) / 200
        gpu = {"value": dofs / (ms * 1e-3), "ms_per_step": ms, "roofline_frac": B / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    sample = (f"C oracle on {n_sample}^3 cubes x 6 tets, one block row per tet as the reference stores "
              f"it ({o.n_poly} tets, p={p}, {dofs} DOFs, {o.n_blk} blocks, {B / 1e9:.2f} GB per step): "
              f"{k_all} FP64 leapfrog steps in {el_all:.1f} s on {threads} OpenMP threads "
              f"(the cores this job may use: affinity set capped by OMP_NUM_THREADS; the host shows "
              f"{visible}); {k_one} steps in {el_one:.1f} s on 1 thread; operator build {t_build:.1f} s")
    return dict(value=dofs * k_all / el_all, unit="DOF-updates/s", cores=threads, kind="port",
                sample=sample, steps=k_all, seconds=el_all, ms_per_step=el_all / k_all * 1e3,
                value_1thread=dofs * k_one / el_one, ms_per_step_1thread=el_one / k_one * 1e3,
                host_cpus_visible=visible, operator_build_s=t_build, bytes_per_step=B,
                gpu_fp64_same_mesh=gpu,
                gpu_over_cpu_same_mesh=(gpu["value"] / (dofs * k_all / el_all)) if gpu else None,
                _unused=float(u_cpu[0]))


def ensemble_leg(n_inst, steps, plain_sample, f0=5.0, seed0=1000):
    """Batched ensemble: n_inst seeded instances (n^3 cubes, n uniform in
    10..20, p in {2, 3}, lognormal lambda/mu, Ricker point force, 64
    receivers recorded every step), one context per degree with the 32-bit
    apply and the whole solve captured in a HIP graph, against the plain path
    (FP64, one instance at a time, no graph) on a sample of the same
    instances; accuracy = relative L2 difference of the receiver traces."""
    from lymph3d_apsc_amd import ensemble as E
    from lymph3d_apsc_amd import _lib
    import lymph3d_apsc_amd as L
    t0 = time.perf_counter()
    inst = [E.make_instance(seed0 + s) for s in range(n_inst)]
    # one dt for every instance: CFL 0.2 with the fastest P-wave speed
    dt = min(0.2 * (1.0 / i["n"]) / (np.sqrt(((i["mesh"].lam + 2 * i["mesh"].mu) / i["mesh"].rho).max())
                                     * i["p"] ** 2) for i in inst)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    ens = E.Ensemble(inst, dt, steps, every=1, f0=f0, apply_bits=32, autotune=True)
    t_setup = time.perf_counter() - t0
    setup_parts = dict(ens.setup_seconds)
    ens.run()  # first replay (untimed)
    t0 = time.perf_counter()
    tr = ens.run()
    t_ens = time.perf_counter() - t0
    groups = [{"p": g["p"], "instances": len(g["members"]), "n_poly": g["ctx"].n_poly,
               "n_blk": g["ctx"].n_blk, "device_GB": g["ctx"].device_bytes / 1e9,
               "step_bytes": g["ctx"].step_bytes_for(32)} for g in ens.groups]
    ens.close()
    del ens
    # the same ensemble with the FP64 apply (graph, no 32-bit operator): what
    # batching alone gains over the plain path, apart from the halved bytes
    t0 = time.perf_counter()
    ens64 = E.Ensemble(inst, dt, steps, every=1, f0=f0, apply_bits=64, autotune=True)
    t_setup64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    tr64 = ens64.run()  # one replay
    t_ens64 = time.perf_counter() - t0
    ens64.close()
    del ens64
    # and with the 48-bit apply (FP32 blocks plus 16-bit corrections, FP64 sums)
    t0 = time.perf_counter()
    ens48 = E.Ensemble(inst, dt, steps, every=1, f0=f0, apply_bits=48, autotune=True)
    t_setup48 = time.perf_counter() - t0
    t0 = time.perf_counter()
    tr48 = ens48.run()  # one replay
    t_ens48 = time.perf_counter() - t0
    ens48.close()
    del ens48
    errs48 = [float(np.linalg.norm(tr48[k] - tr64[k]) / max(np.linalg.norm(tr64[k]), 1e-300))
              for k in range(n_inst)]
    # plain path on a sample: FP64, one instance at a time, uncaptured
    sample = list(range(0, n_inst, max(1, n_inst // plain_sample)))[:plain_sample]
    t_plain, t_plain_setup, errs = 0.0, 0.0, []
    for k in sample:
        ins = inst[k]
        s0 = time.perf_counter()
        with L.Context(ins["mesh"], ins["p"]) as ctx:
            ctx.set_keep_adg(False)
            ctx.set_build("factored")
            t = ctx.start_from_rest(dt, ins["source"], ins["force"], f0=f0)
            ctx.set_receivers(ins["receivers"], max_records=steps)
            ctx.sync()
            s1 = time.perf_counter()
            ctx.step_n_record(L.PROB_RICKER, steps, t, dt, every=1)
            ctx.sync()
            s2 = time.perf_counter()
            ref = ctx.traces()
        t_plain += s2 - s1
        t_plain_setup += s1 - s0
        errs.append(float(np.linalg.norm(tr[k] - ref) / max(np.linalg.norm(ref), 1e-300)))
    plain_rate = len(sample) / t_plain
    ens_rate = n_inst / t_ens
    errs64 = [float(np.linalg.norm(tr[k] - tr64[k]) / max(np.linalg.norm(tr64[k]), 1e-300))
              for k in range(n_inst)]
    return {
        "value": ens_rate, "unit": "ensemble solves/s",
        "instances": n_inst, "steps": steps, "dt": dt, "receivers_per_instance": 64,
        "dofs_total": int(sum(g["n_poly"] * 3 * (g["p"] + 1) * (g["p"] + 2) * (g["p"] + 3) // 6
                              for g in groups)),
        "groups": groups, "seconds_per_ensemble_solve": t_ens,
        "step_GBps": sum(g["step_bytes"] for g in groups) * (steps + 1) / t_ens / 1e9,
        "setup_s": t_setup, "setup_seconds_by_phase": setup_parts, "instance_generation_s": t_gen,
        "plain_fp64": {"value": plain_rate, "unit": "solves/s", "sample_instances": len(sample),
                       "seconds": t_plain, "setup_seconds": t_plain_setup},
        "speedup_vs_plain": ens_rate / plain_rate,
        "solves_per_s_including_setup": n_inst / (t_setup + t_ens),
        "trace_rel_l2_vs_fp64_max": max(errs), "trace_rel_l2_vs_fp64_mean": float(np.mean(errs)),
        "fp64_ensemble": {
            "value": n_inst / t_ens64, "unit": "ensemble solves/s", "seconds_per_ensemble_solve": t_ens64,
            "setup_s": t_setup64, "solves_per_s_including_setup": n_inst / (t_setup64 + t_ens64),
            "speedup_vs_plain": (n_inst / t_ens64) / plain_rate,
            "fp32_trace_rel_l2_vs_this_max": max(errs64),
            "fp32_trace_rel_l2_vs_this_mean": float(np.mean(errs64)),
            "note": "the same instances, graph and launches with the FP64 apply: the gain of "
                    "batching alone; the FP32 ensemble's further gain is its halved bytes"},
        "fp48_ensemble": {
            "value": n_inst / t_ens48, "unit": "ensemble solves/s", "seconds_per_ensemble_solve": t_ens48,
            "setup_s": t_setup48, "solves_per_s_including_setup": n_inst / (t_setup48 + t_ens48),
            "speedup_vs_plain": (n_inst / t_ens48) / plain_rate,
            "trace_rel_l2_vs_fp64_ensemble_max": max(errs48),
            "trace_rel_l2_vs_fp64_ensemble_mean": float(np.mean(errs48)),
            "note": "the same instances with the 48-bit apply (FP32 blocks plus a 16-bit "
                    "correction per entry, FP64 sums), against the FP64 ensemble's traces"},
        "note": "solve = zero state, first step, then `steps` leapfrog steps recording 64 "
                "receivers per step; value excludes the one-time setup (operator build, "
                "transforms, RHS, capture), which solves_per_s_including_setup adds for one "
                "ensemble solve",
    }


def surrogate_leg(n_train, n_eval, epochs, threshold=5e-2):
    """Learned surrogate on a bounded sample of its workload: n_train + n_eval
    parameter sets on 16^3 cubes at p=2 (1000 steps, 64 receivers every 5
    steps), solved by the 32-bit ensemble path (truth: the FP64 ensemble), four
    per-receiver networks (32 physics features, 5 hidden layers of 512, 64 SVD modes
    of the traces) trained on the device; throughput and error against the
    plain FP64 path one instance at a time."""
    from lymph3d_apsc_amd import surrogate as S
    r = S.run(n_train=n_train, n_eval=n_eval, epochs=epochs, threshold=threshold, plain_sample=4,
              log=log)
    thr = r["throughput_solves_per_s"]
    err = r["rel_l2_vs_fp64"]
    return {
        "value": thr["surrogate"], "unit": "surrogate predictions/s (4 networks, inputs resident)",
        "sample": f"{n_train} training / {n_eval} held-out instances, {epochs} epochs "
                  "(the full-size run uses 4096 / 512 and 300 epochs)",
        "plain_fp64_solves_per_s": thr["plain_fp64"],
        "ensemble_fp32_solves_per_s": thr["ensemble_fp32"],
        "hybrid_solves_per_s": thr["hybrid_with_fallback"], "threshold": threshold,
        "fallback_fraction": r["fallback_fraction"],
        "rel_l2_vs_fp64_mean": err["surrogate"]["mean"],
        "rel_l2_vs_fp64_median": err["surrogate"]["median"],
        "mean_trace_predictor_rel_l2": err["training_mean_predictor"]["mean"],
        "ensemble_fp32_rel_l2_max": err["ensemble_fp32"]["max"],
        "spread_error_correlation": r["spread_vs_error_correlation"],
        "threshold_tradeoff": r["threshold_tradeoff"],
        "training_ms_per_step": r["training"]["ms_per_step"],
        "training_gemm_tflops": r["training"]["gemm_tflops"],
        "seconds": r["seconds"],
    }


def uneven_leg(n=30, p=4, seed=3, steps=200, run_steps=2000):
    """Heterogeneous uneven polyhedra: n^3 cubes x 6 tets grown into polyhedra
    of 4-40 tets (seeded), smooth lognormal lambda and mu, rho ~ U[0.8, 1.2],
    Neumann on z=1 and Dirichlet elsewhere, a z-directed Ricker point force
    (f0 = 5) at a seeded interior point, 64 seeded receivers, p = 4, dt from
    CFL 0.2 with the fastest P-wave speed.  Times `steps` steps of the FP64
    and of the 32-bit apply (each autotuned) from the same start, the
    difference of the two states, and a `run_steps`-step 32-bit run
    recording the receivers every step."""
    import lymph3d_apsc_amd as L
    from lymph3d_apsc_amd import _lib, mesh as M, ensemble as E
    m = M.cube_mesh(n, dirichlet_faces={"x0", "x1", "y0", "y1", "z0"}, poly_size=(4, 40), seed=seed)
    rng = np.random.default_rng(seed)
    v = m.xyz[m.tet_vert]
    bc = (((v[:, 0] + v[:, 1]) + v[:, 2]) + v[:, 3]) / 4
    m.lam = E.lognormal_field(rng, bc)
    m.mu = E.lognormal_field(rng, bc)
    m.rho = rng.uniform(0.8, 1.2, m.n_tet)
    dt = 0.2 * (1 / n) / (np.sqrt(((m.lam + 2 * m.mu) / m.rho).max()) * p * p)
    src = rng.uniform(0.3, 0.7, (1, 3))
    rec = rng.uniform(0.05, 0.95, (64, 3))
    res = {"mesh": f"{n}^3 cubes x 6 tets grown into polyhedra of 4-40 tets (seed {seed})",
           "n_tet": m.n_tet, "n_poly": m.n_poly, "p": p, "dt": dt}
    with L.Context(m, p) as ctx:
        ctx.set_keep_adg(False)
        ctx.set_build("factored")
        t0 = time.perf_counter()
        ctx.make_matrices_free()
        res["build_s"] = time.perf_counter() - t0
        res["n_blk"] = ctx.n_blk
        out = {}
        for bits in (64, 32):
            ctx.set_apply_precision(bits)
            t = ctx.start_from_rest(dt, src, [[0.0, 0.0, 1.0]], f0=5.0)
            chosen, _ = ctx.autotune(reps=10)
            t = ctx.start_from_rest(dt, src, [[0.0, 0.0, 1.0]], f0=5.0)
            ms = ctx.time_steps(L.PROB_RICKER, steps, t, dt) / steps
            B = ctx.step_bytes_for(bits)
            out[bits] = ctx.get_slot(_lib.SLOT_U_CUR)
            res[f"fp{bits}"] = {"ms_per_step": ms, "variant": list(chosen), "GBps": B / ms / 1e6,
                                "frac": B / ms / 1e6 / HBM_PEAK_GBS, "bytes_per_step": B}
        res["fp32"]["rel_l2_diff_vs_fp64"] = float(np.linalg.norm(out[32] - out[64]) /
                                                   np.linalg.norm(out[64]))
        t = ctx.start_from_rest(dt, src, [[0.0, 0.0, 1.0]], f0=5.0)
        ctx.set_receivers(rec, max_records=run_steps)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step_n_record(L.PROB_RICKER, run_steps, t, dt, every=1)
        ctx.sync()
        res["fp32_run_with_receivers_s"] = time.perf_counter() - t0
        res["run_steps"] = run_steps
        res["traces_finite"] = bool(np.isfinite(ctx.traces()).all())
    log(json.dumps(res))
    return res


def pairs_once_bytes(ctx):
    """Algorithmic bytes of one FP64 step with face pairs stored once: the
    pair blocks and the self blocks' upper triangles (column pair k's rows
    0..2k+1) read once, the per-block products written and read back, the
    mass factor's inverse, five vectors and three int32 indices per block."""
    n3, npk = ctx.n3, ctx.Np * (ctx.Np + 1) // 2
    n_pair = (ctx.n_blk - ctx.n_poly) // 2
    h = n3 // 2
    return 8 * (n_pair * n3 * n3 + ctx.n_poly * 2 * h * (h + 1) + 2 * ctx.n_blk * n3
                + ctx.n_poly * npk + 5 * ctx.n_poly * n3) + 12 * ctx.n_blk


def pairs_once_tilde_bytes(ctx, bits):
    """Algorithmic bytes of one 32- or 48-bit step with face pairs stored once
    (p = 3): the pair blocks (FP32, and for 48 bits their 16-bit corrections)
    and the self blocks' packed upper triangles read once -- quad q's rows
    0..4q+3, and per pair of quads its rows 0..8k+7 (4 halves per row for the
    last quad of an odd count) -- the pair blocks' transposed products (FP32
    for 32 bits, FP64 for 48) and each polyhedron's summed row products (FP64)
    written and read back, five FP64 vectors, and four int32 indices per
    block (three read by the products pass, one by the update)."""
    n3 = ctx.n3
    nq = n3 // 4
    n_pair = (ctx.n_blk - ctx.n_poly) // 2
    self32 = sum(4 * (4 * q + 4) for q in range(nq))
    self16 = sum((8 if 2 * k + 1 < nq else 4) * min(8 * k + 8, n3) for k in range((nq + 1) // 2))
    blocks = 4 * (n_pair * n3 * n3 + ctx.n_poly * self32)
    if bits == 48:
        blocks += 2 * (n_pair * n3 * n3 + ctx.n_poly * self16)
    prod = (4 if bits == 32 else 8) * 2 * n_pair * n3 + 8 * 2 * ctx.n_poly * n3
    return blocks + prod + 8 * 5 * ctx.n_poly * n3 + 16 * ctx.n_blk


def degree_sweep(steps=200):
    """Step throughput at p = 1..4 on agglomerated cubes of about 4 M DOFs
    (70^3, 52^3, 40^3, 33^3 cubes), FP64, 32- and 48-bit apply, each
    autotuned, FP64 with face pairs stored once at p <= 3; and the time of
    both operator builds (keep_adg off, K and M only)."""
    import lymph3d_apsc_amd as L
    from lymph3d_apsc_amd import _lib, mesh as M
    rows = []
    for p, n in ((1, 70), (2, 52), (3, 40), (4, 33)):
        m = M.cube_mesh(n, agglomerate=True)
        dt = 0.2 * (1 / n) / (np.sqrt(3.0) * p * p)
        with L.Context(m, p) as ctx:
            ctx.set_keep_adg(False)
            build_s = {}
            for build in ("reference", "factored"):  # the factored build is kept
                ctx.set_build(build)
                ctx.sync()
                t0 = time.perf_counter()
                ctx.make_matrices_free()
                ctx.sync()
                build_s[build] = time.perf_counter() - t0
            ctx.make_rhs_free(L.PROB_NULL, _lib.SLOT_RHS_STAT)
            ctx.make_rhs_free(L.PROB_CUBE_WAVE, _lib.SLOT_RHS_DYN)
            row = dict(p=p, n=n, n_poly=m.n_poly, dofs=m.n_poly * ctx.n3, n_blk=ctx.n_blk,
                       operator_build_s=build_s)
            for bits in (64, 32, 48):
                ctx.set_apply_precision(bits)
                chosen, _ = ctx.autotune(reps=10)
                ms = ctx.time_steps(L.PROB_CUBE_WAVE, steps, dt, dt) / steps
                B = ctx.step_bytes_for(bits)
                row[f"fp{bits}"] = dict(ms_per_step=ms, dof_updates_per_s=row["dofs"] / (ms * 1e-3),
                                        variant=list(chosen), GBps=B / ms / 1e6,
                                        frac=B / ms / 1e6 / HBM_PEAK_GBS)
            if p <= 3:  # FP64 with face pairs stored once (p <= 3)
                ctx.set_apply_precision(64)
                ctx.set_symmetric_storage(True)
                ms = ctx.time_steps(L.PROB_CUBE_WAVE, steps, dt, dt) / steps
                B = pairs_once_bytes(ctx)
                row["fp64_pairs_once"] = dict(ms_per_step=ms, dof_updates_per_s=row["dofs"] / (ms * 1e-3),
                                              speedup_vs_full_rows=row["fp64"]["ms_per_step"] / ms,
                                              GBps=B / ms / 1e6, frac=B / ms / 1e6 / HBM_PEAK_GBS)
                ctx.set_symmetric_storage(False)
            log(json.dumps(row))
            rows.append(row)
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=40, help="cubes per side")
    ap.add_argument("--p", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=22,
                    help="cubes per side of the CPU sample (per-tet layout)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="seconds per thread count")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-apply measurement")
    ap.add_argument("--no-fp48", action="store_true",
                    help="skip the 48-bit apply (FP32 blocks plus 16-bit corrections) measurement")
    ap.add_argument("--no-pairs-once", action="store_true",
                    help="skip the FP64 step with face pairs stored once")
    ap.add_argument("--no-autotune", action="store_true", help="use the default step variant")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--ens-instances", type=int, default=256)
    ap.add_argument("--ens-steps", type=int, default=1000)
    ap.add_argument("--ens-plain-sample", type=int, default=8)
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-surrogate leg")
    ap.add_argument("--sur-train", type=int, default=1024)
    ap.add_argument("--sur-eval", type=int, default=128)
    ap.add_argument("--sur-epochs", type=int, default=150)
    ap.add_argument("--uneven", action="store_true",
                    help="instead of the bench line: heterogeneous uneven polyhedra at p=4 (one JSON line)")
    ap.add_argument("--degree-sweep", action="store_true",
                    help="instead of the bench line: step throughput at p = 1..4 (one JSON line)")
    a = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the whole mesh lives on one "
                                   "MI355X and nothing is sharded; run with --gpus 1"}))
        return 2

    import lymph3d_apsc_amd as L
    from lymph3d_apsc_amd import _lib, mesh as M

    if L.device_count() == 0:
        print(json.dumps({"error": "no HIP device visible"}))
        return 1
    if a.uneven:
        print(json.dumps({"uneven": uneven_leg()}))
        return 0
    if a.degree_sweep:
        copy, read = stream_probe_gbps()
        print(json.dumps({"degree_sweep": degree_sweep(), "stream_copy_GBps": copy,
                          "stream_read_GBps": read}))
        return 0
    t_setup = time.perf_counter()
    m = M.cube_mesh(a.n, agglomerate=True)
    log(f"mesh: {m.n_tet} tets, {m.n_poly} polyhedra ({time.perf_counter() - t_setup:.1f}s)")
    p = a.p
    h = 1.0 / a.n
    dt = 0.2 * h / (np.sqrt(3.0) * p * p)
    prob = L.PROB_CUBE_WAVE
    ctx = L.Context(m, p)
    ctx.set_build("factored")  # same operator to rounding, ~5x faster set-up at p=3
    t0 = time.perf_counter()
    # operator, RHS, IC projections, first step and the warm-up steps
    r0 = ctx.run(prob, dt, 1e30, max_steps=1 + a.warmup)
    t_build = time.perf_counter() - t0
    log(f"setup + {1 + a.warmup} steps: {t_build:.2f}s; device bytes {ctx.device_bytes / 1e9:.2f} GB")
    t_start = r0["t_final"]
    chosen, tune = ((None, None, None), {}) if a.no_autotune else ctx.autotune(reps=20)
    log(f"autotune: chosen (group size, broadcast mode, waves per polyhedron) {chosen}; ms/step {tune}")
    # timed region: K device-resident steps, HIP events on the context stream
    w0 = time.perf_counter()
    ms = ctx.time_steps(prob, a.steps, t_start, dt)
    wall = time.perf_counter() - w0
    # final time and errors vs the projected exact solution
    t = t_start
    for _ in range(a.steps):
        t = t + dt
    u = ctx.get_slot(_lib.SLOT_U_CUR)
    uex = ctx.compute_modal_coefficients_free(L.FN_CUBE_U) * L.time_function(prob, t)
    e_l2 = np.sqrt(ctx.compute_error_l2_matrix_free(u, uex))
    e_dg = np.sqrt(ctx.compute_error_dg_matrix_free(u, uex))
    ms_step = ms / a.steps
    dofs = 3 * ctx.Np * ctx.n_poly
    value = dofs / (ms_step * 1e-3)
    B = ctx.step_bytes
    achieved = B / (ms_step * 1e-3) / 1e9
    copy, read = stream_probe_gbps()
    traffic, traffic_src = pmc_traffic(a.n, p)
    res = {
        "metric": METRIC, "value": value, "unit": "DOF-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": f"unit cube {a.n}^3 cubes x 6 tets ({m.n_tet} tets) agglomerated per cube "
                        f"to {m.n_poly} polyhedra, p={p} ({dofs} DOFs), FP64 matrix-free PolyDG "
                        f"leapfrog, SIPG alpha=10, standing-wave data, dt=CFL 0.2 ({dt:.3e})",
            "n_poly": m.n_poly, "n_tet": m.n_tet, "p": p, "dofs": dofs, "n_blk": ctx.n_blk,
            "dt": dt, "t_final": t, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
            "traffic_over_algorithmic": (traffic / B) if traffic else None,
            "bytes_per_step": B, "kernel": f"k_leapfrog<{p},0>",
            "stream_copy_GBps": copy,
            "frac_of_stream_copy": (achieved / copy) if copy else None,
            "stream_read_GBps": read,
            "frac_of_stream_read": (achieved / read) if read else None,
            "fp64_tflops": ctx.step_flops / (ms_step * 1e-3) / 1e12,
            "fp64_frac": ctx.step_flops / (ms_step * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
        },
        "wall_ms_per_step": wall * 1e3 / a.steps,
        "autotune": {"chosen": list(chosen),
                     "ms_per_step": {f"unr{k[0]}_x{k[1]}_w{k[2]}": v for k, v in tune.items()}},
        "err_l2": e_l2, "err_dg": e_dg,
        "setup_s": t_build, "operator_build": "factored",
    }
    if not a.no_pairs_once:
        # face pairs stored once (the operator is symmetric here: SIPG, one material): same start
        # state and steps as the FP64 timed run
        ctx.run(prob, dt, 1e30, max_steps=1 + a.warmup)
        asym = ctx.set_symmetric_storage(True)
        mss = ctx.time_steps(prob, a.steps, t_start, dt) / a.steps
        us = ctx.get_slot(_lib.SLOT_U_CUR)
        n_st = ctx.n_poly + (ctx.n_blk - ctx.n_poly) // 2
        Bs = pairs_once_bytes(ctx)
        traffic_s, traffic_s_src = pmc_traffic(a.n, p, "k_sym")
        res["fp64_pairs_once"] = {
            "value": dofs / (mss * 1e-3), "unit": "DOF-updates/s", "ms_per_step": mss,
            "speedup_vs_full_rows": ms_step / mss, "stored_blocks": n_st, "max_pair_asymmetry": asym,
            "roofline": {"bound": "hbm", "achieved": Bs / (mss * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": Bs / (mss * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "bytes_per_step": Bs, "kernel": f"k_sym_products<{p}> + k_sym_update<{p},0>",
                         "traffic": traffic_s, "traffic_source": traffic_s_src,
                         "traffic_over_algorithmic": (traffic_s / Bs) if traffic_s else None},
            "rel_diff_vs_full_rows": float(np.abs(us - u).max() / np.abs(u).max()),
            "err_l2": float(np.sqrt(ctx.compute_error_l2_matrix_free(us, uex))),
        }
        ctx.set_symmetric_storage(False)
        log(f"pairs stored once: {mss:.3f} ms/step ({ms_step / mss:.2f}x the full rows), "
            f"rel diff {res['fp64_pairs_once']['rel_diff_vs_full_rows']:.2e}")
    if not a.no_fp32:
        # 32-bit operator apply (mass-orthonormal basis; FP64 RHS, transforms, update): same start state
        # and the same number of steps as the FP64 timed run
        ctx.run(prob, dt, 1e30, max_steps=1 + a.warmup)
        ctx.set_apply_precision(32)
        ch32, tune32 = ((None, None, None), {}) if a.no_autotune else ctx.autotune(reps=20)
        ms32 = ctx.time_steps(prob, a.steps, t_start, dt)
        u32 = ctx.get_slot(_lib.SLOT_U_CUR)
        ms32_step = ms32 / a.steps
        # every 32-bit stepping call transforms the states and RHS into the
        # mass-orthonormal basis and back once: twice the steps from the same
        # start give the per-step time without that per-call cost
        ctx.set_apply_precision(64)
        ctx.run(prob, dt, 1e30, max_steps=1 + a.warmup)
        ctx.set_apply_precision(32)
        if not a.no_autotune:
            ctx.set_step_variant(ctx.chosen_variant)
        ms32_2 = ctx.time_steps(prob, 2 * a.steps, t_start, dt)
        ms32_kernel = (ms32_2 - ms32) / a.steps
        l2_32_vs_64 = float(np.sqrt(ctx.compute_error_l2_matrix_free(u32, u)))
        B32 = ctx.step_bytes_for(32)
        traffic32, traffic32_src = pmc_traffic(a.n, p, "k_leapfrog32")
        res["fp32_apply"] = {
            "value": dofs / (ms32_step * 1e-3), "unit": "DOF-updates/s", "ms_per_step": ms32_step,
            "speedup_vs_fp64": ms_step / ms32_step,
            "roofline": {"bound": "hbm", "achieved": B32 / (ms32_step * 1e-3) / 1e9,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": B32 / (ms32_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "traffic": traffic32, "traffic_source": traffic32_src,
                         "traffic_over_algorithmic": (traffic32 / B32) if traffic32 else None,
                         "bytes_per_step": B32, "kernel": f"k_leapfrog32<{p},0>"},
            "ms_per_step_includes": f"the basis transforms of one stepping call of {a.steps} steps",
            "ms_per_step_without_transforms": ms32_kernel,
            "rel_l2_diff_vs_fp64": float(np.linalg.norm(u32 - u) / np.linalg.norm(u)),
            "err_l2": float(np.sqrt(ctx.compute_error_l2_matrix_free(u32, uex))),
            "l2_diff_vs_fp64": l2_32_vs_64,
            "l2_diff_vs_fp64_over_fp64_err_l2": l2_32_vs_64 / e_l2,
            "steps_from_rest": 1 + a.warmup + a.steps,
            "autotune": {"chosen": list(ch32),
                         "ms_per_step": {f"unr{k[0]}_w{k[2]}": v for k, v in tune32.items()}},
        }
        if not a.no_pairs_once and p == 3:
            # the same 32-bit steps with face pairs stored once, from the same start state
            ctx.set_apply_precision(64)
            ctx.run(prob, dt, 1e30, max_steps=1 + a.warmup)
            ctx.set_apply_precision(32)
            ctx.set_symmetric_storage(True, fp32=True)
            ms32s = ctx.time_steps(prob, a.steps, t_start, dt) / a.steps
            u32s = ctx.get_slot(_lib.SLOT_U_CUR)
            B32s = pairs_once_tilde_bytes(ctx, 32)
            traffic32s, traffic32s_src = pmc_traffic(a.n, p, "k_sym32")
            res["fp32_apply"]["pairs_once"] = {
                "value": dofs / (ms32s * 1e-3), "unit": "DOF-updates/s", "ms_per_step": ms32s,
                "speedup_vs_full_rows": ms32_step / ms32s,
                "roofline": {"bound": "hbm", "achieved": B32s / (ms32s * 1e-3) / 1e9,
                             "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": B32s / (ms32s * 1e-3) / 1e9 / HBM_PEAK_GBS, "bytes_per_step": B32s,
                             "kernel": f"k_sym_products32<{p}> + k_sym_update32<{p},0>",
                             "traffic": traffic32s, "traffic_source": traffic32s_src,
                             "traffic_over_algorithmic": (traffic32s / B32s) if traffic32s else None},
                "rel_diff_vs_full_rows": float(np.abs(u32s - u32).max() / np.abs(u32).max()),
                "rel_l2_diff_vs_fp64": float(np.linalg.norm(u32s - u) / np.linalg.norm(u)),
                "err_l2": float(np.sqrt(ctx.compute_error_l2_matrix_free(u32s, uex))),
            }
            ctx.set_symmetric_storage(False)
        log(f"fp32 apply: {ms32_step:.3f} ms/step, rel diff vs fp64 {res['fp32_apply']['rel_l2_diff_vs_fp64']:.2e}")
    if not a.no_fp48:
        # 48-bit apply (FP32 blocks plus a 16-bit correction per entry, FP64 sums): same start
        # state and steps as the FP64 timed run
        ctx.set_apply_precision(64)
        ctx.run(prob, dt, 1e30, max_steps=1 + a.warmup)
        ctx.set_apply_precision(48)
        ch48, tune48 = ((None, None, None), {}) if a.no_autotune else ctx.autotune(reps=20)
        ms48 = ctx.time_steps(prob, a.steps, t_start, dt) / a.steps
        u48 = ctx.get_slot(_lib.SLOT_U_CUR)
        B48 = ctx.step_bytes_for(48)
        traffic48, traffic48_src = pmc_traffic(a.n, p, "k_leapfrog48")
        l2_48 = float(np.sqrt(ctx.compute_error_l2_matrix_free(u48, u)))
        res["fp48_apply"] = {
            "value": dofs / (ms48 * 1e-3), "unit": "DOF-updates/s", "ms_per_step": ms48,
            "speedup_vs_fp64": ms_step / ms48,
            "roofline": {"bound": "hbm", "achieved": B48 / (ms48 * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": B48 / (ms48 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "bytes_per_step": B48, "kernel": f"k_leapfrog48<{p},0>",
                         "traffic": traffic48, "traffic_source": traffic48_src,
                         "traffic_over_algorithmic": (traffic48 / B48) if traffic48 else None},
            "ms_per_step_includes": f"the basis transforms of one stepping call of {a.steps} steps",
            "rel_l2_diff_vs_fp64": float(np.linalg.norm(u48 - u) / np.linalg.norm(u)),
            "err_l2": float(np.sqrt(ctx.compute_error_l2_matrix_free(u48, uex))),
            "l2_diff_vs_fp64": l2_48,
            "l2_diff_vs_fp64_over_fp64_err_l2": l2_48 / e_l2,
            "steps_from_rest": 1 + a.warmup + a.steps,
            "autotune": {"chosen": list(ch48),
                         "ms_per_step": {f"unr{k[0]}_w{k[2]}": v for k, v in tune48.items()}},
        }
        if not a.no_pairs_once and p == 3:
            # the same 48-bit steps with face pairs stored once, from the same start state
            ctx.set_apply_precision(64)
            ctx.run(prob, dt, 1e30, max_steps=1 + a.warmup)
            ctx.set_apply_precision(48)
            ctx.set_symmetric_storage(True, fp32=True)
            ms48s = ctx.time_steps(prob, a.steps, t_start, dt) / a.steps
            u48s = ctx.get_slot(_lib.SLOT_U_CUR)
            B48s = pairs_once_tilde_bytes(ctx, 48)
            traffic48s, traffic48s_src = pmc_traffic(a.n, p, "k_sym48")
            l2_48s = float(np.sqrt(ctx.compute_error_l2_matrix_free(u48s, u)))
            res["fp48_apply"]["pairs_once"] = {
                "value": dofs / (ms48s * 1e-3), "unit": "DOF-updates/s", "ms_per_step": ms48s,
                "speedup_vs_full_rows": ms48 / ms48s, "speedup_vs_fp64": ms_step / ms48s,
                "roofline": {"bound": "hbm", "achieved": B48s / (ms48s * 1e-3) / 1e9,
                             "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": B48s / (ms48s * 1e-3) / 1e9 / HBM_PEAK_GBS, "bytes_per_step": B48s,
                             "kernel": f"k_sym_products48<{p}> + k_sym_update32<{p},0,double>",
                             "traffic": traffic48s, "traffic_source": traffic48s_src,
                             "traffic_over_algorithmic": (traffic48s / B48s) if traffic48s else None},
                "rel_diff_vs_full_rows": float(np.abs(u48s - u48).max() / np.abs(u48).max()),
                "rel_l2_diff_vs_fp64": float(np.linalg.norm(u48s - u) / np.linalg.norm(u)),
                "err_l2": float(np.sqrt(ctx.compute_error_l2_matrix_free(u48s, uex))),
                "l2_diff_vs_fp64_over_fp64_err_l2": l2_48s / e_l2,
            }
            ctx.set_symmetric_storage(False)
            log(f"fp48 apply, pairs stored once: {ms48s:.3f} ms/step")
        ctx.set_apply_precision(64)
        log(f"fp48 apply: {ms48:.3f} ms/step, rel diff vs fp64 {res['fp48_apply']['rel_l2_diff_vs_fp64']:.2e}")
    del ctx
    if not a.no_ensemble:
        try:
            res["ensemble"] = ensemble_leg(a.ens_instances, a.ens_steps, a.ens_plain_sample)
            e = res["ensemble"]
            log(f"ensemble: {e['value']:.2f} solves/s ({e['instances']} instances), plain FP64 "
                f"{e['plain_fp64']['value']:.2f} solves/s, trace rel L2 {e['trace_rel_l2_vs_fp64_max']:.2e}")
        except Exception as ex:  # reported, never fatal
            res["ensemble"] = {"error": repr(ex)}
    if not a.no_surrogate:
        try:
            res["surrogate"] = surrogate_leg(a.sur_train, a.sur_eval, a.sur_epochs)
            sg = res["surrogate"]
            log(f"surrogate: {sg['value']:.3g} predictions/s, rel L2 vs FP64 mean "
                f"{sg['rel_l2_vs_fp64_mean']:.3f} (mean trace {sg['mean_trace_predictor_rel_l2']:.3f}); "
                f"fallback {sg['fallback_fraction']:.2f} at {sg['threshold']}")
        except Exception as ex:  # reported, never fatal
            res["surrogate"] = {"error": repr(ex)}
    if not a.no_cpu_baseline:
        try:
            cb = cpu_baseline(p, a.cpu_sample, a.cpu_seconds)
            cb.pop("_unused", None)
            res["cpu_baseline"] = cb
        except Exception as e:  # reported, never fatal
            res["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
