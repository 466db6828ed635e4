"""Batched ensembles: many independent elastodynamics instances (own mesh
size, own lognormal material field, own Ricker point-source location, own
receivers) advanced together.

Design: instances of the same polynomial degree are joined into one mesh with
disjoint components (`union`); the block-row operator of the union is block
diagonal by instance, so one launch of the step kernel per time step advances
every instance of the group, and each instance performs exactly the
arithmetic of a standalone run.  The time factor (Ricker wavelet, common f0)
is shared; the spatial source term of each instance lives in its own rows of
the dynamic RHS.  A whole solve (zero state, first step, N steps with
receiver sampling) is captured once into a HIP graph and replayed.

No reference counterpart: the reference runs one instance per MPI job.
"""
import copy

import numpy as np

from . import _lib
from .matrix_free import Context
from .mesh import Mesh, cube_mesh


def union(meshes):
    """One Mesh whose connected components are `meshes`: node, tet and
    polyhedron ids offset, per-tet and per-polyhedron data concatenated (no
    face matching is redone: no face is shared between components)."""
    no = np.cumsum([0] + [m.xyz.shape[0] for m in meshes])
    to = np.cumsum([0] + [m.n_tet for m in meshes])
    po = np.cumsum([0] + [m.n_poly for m in meshes])
    u = Mesh.__new__(Mesh)
    cat = np.concatenate
    u.xyz = cat([m.xyz for m in meshes])
    u.tet_vert = cat([m.tet_vert + no[k] for k, m in enumerate(meshes)]).astype(np.int32)
    u.tet_poly = cat([m.tet_poly + po[k] for k, m in enumerate(meshes)]).astype(np.int32)
    u.n_tet, u.n_poly = int(to[-1]), int(po[-1])
    u.tria_vert = cat([m.tria_vert + no[k] for k, m in enumerate(meshes)]).astype(np.int32)
    u.tria_tag = cat([m.tria_tag for m in meshes])
    u.tag_space = {}
    for m in meshes:
        u.tag_space.update(m.tag_space)
    for f in ("rho", "lam", "mu", "neigh_face", "bc_space", "normal", "area", "bbox", "hk"):
        setattr(u, f, np.ascontiguousarray(cat([getattr(m, f) for m in meshes])))
    u.neigh = cat([np.where(m.neigh >= 0, m.neigh + to[k], m.neigh)
                   for k, m in enumerate(meshes)]).astype(np.int32)
    u.poly_tets = cat([m.poly_tets + to[k] for k, m in enumerate(meshes)]).astype(np.int32)
    u.poly_ptr = cat([[0]] + [m.poly_ptr[1:] + to[k] for k, m in enumerate(meshes)]).astype(np.int32)
    u.degree = u.dt = u.stop_time = None
    return u


_CUBES = {}


def _cube(n, agglomerate):
    """Cube meshes are reused across instances of the same size; each
    instance gets its own copy of the material arrays."""
    if (n, agglomerate) not in _CUBES:
        m = cube_mesh(n, agglomerate=agglomerate)
        m.locate(np.full((1, 3), 0.5))  # builds the point locator shared by the copies
        _CUBES[(n, agglomerate)] = m
    return copy.copy(_CUBES[(n, agglomerate)])


def lognormal_field(rng, points, sigma=0.5, modes=4):
    """Smooth lognormal random field at `points` (n, 3): exp(sigma s(x)) with
    s = sqrt(2/modes) sum_k a_k cos(pi k.x + phase_k), a_k ~ N(0, 1), wave
    vectors k in {1, 2}^3 -- pointwise median 1 and log-standard deviation
    sigma on average."""
    kv = rng.integers(1, 3, (modes, 3))
    ph = rng.uniform(0.0, 2.0 * np.pi, modes)
    a = rng.standard_normal(modes)
    s = np.zeros(points.shape[0])
    for k in range(modes):
        s += a[k] * np.cos(np.pi * (points @ kv[k]) + ph[k])
    return np.exp(sigma * np.sqrt(2.0 / modes) * s)


def make_instance(seed, n_range=(10, 20), p_choices=(2, 3), n_receivers=64, sigma=0.5,
                  agglomerate=True, f0=5.0):
    """One seeded instance: n^3 cubes (n uniform in n_range), degree from
    p_choices, smooth lognormal lambda and mu fields (lognormal_field at the
    tet barycentres), rho ~ U[0.8, 1.2] per tet, Dirichlet on all faces, a
    z-directed Ricker point force at a seeded interior point and seeded
    receivers.

    The fields are smooth on purpose: the face terms take the material of the
    element that owns the face on both sides (src/assemble_element.f90:446),
    so with independent per-tet lambda/mu the operator is far from symmetric
    and the leapfrog iteration grows without bound (a 6^3 mesh at p=2 grows
    tenfold per unit time whatever dt); smooth fields keep the jumps small and
    the iteration stable."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(n_range[0], n_range[1] + 1))
    p = int(rng.choice(p_choices))
    m = _cube(n, agglomerate)
    v = m.xyz[m.tet_vert]
    bc = (((v[:, 0] + v[:, 1]) + v[:, 2]) + v[:, 3]) / 4
    m.lam = lognormal_field(rng, bc, sigma)
    m.mu = lognormal_field(rng, bc, sigma)
    m.rho = rng.uniform(0.8, 1.2, m.n_tet)
    src = rng.uniform(0.25, 0.75, (1, 3))
    rec = rng.uniform(0.05, 0.95, (n_receivers, 3))
    return dict(seed=seed, n=n, p=p, mesh=m, source=src, force=np.array([[0.0, 0.0, 1.0]]),
                receivers=rec, f0=f0)


def _locate_each(meshes, points):
    """meshes[k].locate(points[k]) for every k, with one query per distinct
    geometry (instances of one cube size share it): the same tets, one
    locator call per size instead of one per instance."""
    out = [None] * len(meshes)
    by_geom = {}
    for k, m in enumerate(meshes):
        by_geom.setdefault((id(m.xyz), id(m.tet_vert)), []).append(k)
    for ks in by_geom.values():
        pts = [np.asarray(points[k], np.float64).reshape(-1, 3) for k in ks]
        tets = meshes[ks[0]].locate(np.concatenate(pts))
        cut = np.cumsum([0] + [len(q) for q in pts])
        for j, k in enumerate(ks):
            out[k] = tets[cut[j]:cut[j + 1]]
    return out


class Ensemble:
    """Instances grouped by degree; one context per group on the GPU.  Each
    instance's Ricker wavelet uses its own f0 (instance["f0"], default `f0`)
    through per-group time factors computed on the device each step."""

    def __init__(self, instances, dt, nsteps, every=1, f0=5.0, apply_bits=32, autotune=True):
        import time
        self.instances = list(instances)
        self.dt, self.nsteps, self.every, self.f0 = dt, nsteps, every, f0
        self.groups = []
        # seconds per set-up phase, summed over the groups (device work is
        # synchronised at each phase's end)
        self.setup_seconds = dict(union=0.0, context=0.0, operator=0.0, precision=0.0,
                                  sources_receivers=0.0, autotune=0.0, capture=0.0)
        tick = [time.perf_counter()]

        def lap(phase, ctx=None):
            if ctx is not None:
                ctx.sync()
            now = time.perf_counter()
            self.setup_seconds[phase] += now - tick[0]
            tick[0] = now

        for p in sorted({ins["p"] for ins in self.instances}):
            members = [i for i, ins in enumerate(self.instances) if ins["p"] == p]
            meshes = [self.instances[i]["mesh"] for i in members]
            um = union(meshes)
            lap("union")
            ctx = Context(um, p)
            ctx.set_keep_adg(False)  # traces only: no DG-norm error, half the build memory
            ctx.set_build("factored")
            lap("context", ctx)
            ctx.make_matrices_free()
            lap("operator", ctx)
            if apply_bits != 64:  # 32- or 48-bit apply: the FP64 blocks are not needed after it
                ctx.set_apply_precision(apply_bits)
                ctx.release_fp64_operator()
            lap("precision", ctx)
            # sources and receivers located in each instance, ids offset into the union
            tet_off = np.cumsum([0] + [m.n_tet for m in meshes])
            src = np.concatenate([self.instances[i]["source"] for i in members])
            frc = np.concatenate([self.instances[i]["force"] for i in members])
            stet = np.concatenate([t + tet_off[k] for k, t in enumerate(
                _locate_each(meshes, [self.instances[i]["source"] for i in members]))])
            rec = np.concatenate([self.instances[i]["receivers"] for i in members])
            rtet = np.concatenate([t + tet_off[k] for k, t in enumerate(
                _locate_each(meshes, [self.instances[i]["receivers"] for i in members]))])
            ctx.set_ricker(f0)
            # each instance keeps its own wavelet (f0, t0 = 1.2/f0)
            f0s = np.array([self.instances[i].get("f0", f0) for i in members], np.float64)
            ctx.set_source_groups(np.repeat(np.arange(len(members)), [m.n_poly for m in meshes]), f0s)
            ctx.make_rhs_free(_lib.PROB_NULL, _lib.SLOT_RHS_STAT)
            ctx.make_rhs_free(_lib.PROB_RICKER, _lib.SLOT_RHS_DYN)
            ctx.add_point_sources(src, frc, tets=stet.astype(np.int32))
            ctx.set_receivers(rec, max_records=nsteps // every, tets=rtet.astype(np.int32))
            lap("sources_receivers", ctx)
            if autotune:
                ctx.autotune(reps=2)
            lap("autotune", ctx)
            ctx.capture_solve(_lib.PROB_RICKER, dt, nsteps, every)
            lap("capture", ctx)
            nrec = [len(self.instances[i]["receivers"]) for i in members]
            self.groups.append(dict(p=p, members=members, ctx=ctx, rec_split=np.cumsum([0] + nrec)))
        self.n_dofs = sum(g["ctx"].n_poly * g["ctx"].n3 for g in self.groups)

    def run(self):
        """One ensemble solve (every instance, all steps); returns traces per
        instance: list of (records, receivers, 3) arrays in instance order."""
        for g in self.groups:
            g["ctx"].replay()
        for g in self.groups:
            g["ctx"].sync()
        return self.traces()

    def traces(self):
        out = [None] * len(self.instances)
        for g in self.groups:
            tr = g["ctx"].traces()
            for k, i in enumerate(g["members"]):
                out[i] = tr[:, g["rec_split"][k]:g["rec_split"][k + 1], :]
        return out

    def close(self):
        for g in self.groups:
            g["ctx"].close()
