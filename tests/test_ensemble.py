"""Host logic of ensembles (lymph3d_apsc_amd.ensemble) on the CPU: the union
of instance meshes is block diagonal by instance, and the operator and mass
factors the oracle builds for the union are, bit for bit, those of each
instance built alone (so a batched solve performs each instance's own
arithmetic)."""
import numpy as np

from lymph3d_apsc_amd import ensemble as E
from lymph3d_apsc_amd import mesh as M
from oracle import oracle


def _oracle_for(m, p):
    kinds = np.zeros(8, np.int32)
    kinds[[2, 3, 4]] = -1
    kinds[[5, 6, 7]] = -2
    space = np.zeros(8, np.int32)
    for t, s in m.tag_space.items():
        space[t] = s
    return oracle.Oracle(m.xyz, m.tet_vert, m.tria_vert, m.tria_tag, kinds, space, p,
                         tet_poly=m.tet_poly, rho=m.rho, lam=m.lam, mu=m.mu)


def test_make_instance_is_seeded():
    a, b = E.make_instance(7, n_range=(2, 3)), E.make_instance(7, n_range=(2, 3))
    assert a["n"] == b["n"] and a["p"] == b["p"]
    assert np.array_equal(a["mesh"].lam, b["mesh"].lam)
    assert np.array_equal(a["receivers"], b["receivers"])
    c = E.make_instance(8, n_range=(2, 3))
    assert not np.array_equal(a["mesh"].lam, c["mesh"].lam)
    assert (a["mesh"].lam > 0).all() and (a["mesh"].mu > 0).all()


def test_union_is_block_diagonal_and_bitwise_per_instance():
    inst = [E.make_instance(s, n_range=(2, 3), p_choices=(1,)) for s in (1, 2, 3)]
    meshes = [i["mesh"] for i in inst]
    u = E.union(meshes)
    assert u.n_poly == sum(m.n_poly for m in meshes)
    assert u.n_tet == sum(m.n_tet for m in meshes)
    # connectivity: neighbours offset, never across instances
    to = np.cumsum([0] + [m.n_tet for m in meshes])
    for k, m in enumerate(meshes):
        nb = u.neigh[to[k]:to[k + 1]]
        ref = np.where(m.neigh >= 0, m.neigh + to[k], m.neigh)
        assert np.array_equal(nb, ref)
        assert np.array_equal(u.normal[to[k]:to[k + 1]], m.normal)
    ou = _oracle_for(u, 1)
    ou.build_operator()
    Ku = ou.K()
    rpu, colu = ou.pattern()
    po = 0
    blk = 0
    for m in meshes:
        om = _oracle_for(m, 1)
        om.build_operator()
        rp, col = om.pattern()
        assert np.array_equal(rpu[po:po + m.n_poly + 1] - rpu[po], rp)
        assert np.array_equal(colu[blk:blk + len(col)], col + po)
        assert np.array_equal(Ku[blk:blk + len(col)], om.K())
        Mu = ou.mass()[2][po:po + m.n_poly]
        assert np.array_equal(Mu, om.mass()[2])
        po += m.n_poly
        blk += len(col)
    assert blk == ou.n_blk


def test_union_equals_mesh_built_from_scratch():
    """The concatenating union gives the same arrays as face matching and
    geometry redone on the joined mesh."""
    inst = [E.make_instance(s, n_range=(2, 3)) for s in (1, 2, 3)]
    meshes = [i["mesh"] for i in inst]
    u = E.union(meshes)
    no = np.cumsum([0] + [m.xyz.shape[0] for m in meshes])
    po = np.cumsum([0] + [m.n_poly for m in meshes])
    cat = np.concatenate
    full = M.Mesh(cat([m.xyz for m in meshes]), cat([m.tet_vert + no[k] for k, m in enumerate(meshes)]),
                  cat([m.tet_poly + po[k] for k, m in enumerate(meshes)]),
                  cat([m.tria_vert + no[k] for k, m in enumerate(meshes)]),
                  cat([m.tria_tag for m in meshes]), meshes[0].tag_space,
                  cat([m.rho for m in meshes]), cat([m.lam for m in meshes]), cat([m.mu for m in meshes]))
    assert set(vars(full)) == set(vars(u))
    for k, v in vars(full).items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(v, getattr(u, k)), k
            assert v.dtype == getattr(u, k).dtype, k
        else:
            assert v == getattr(u, k), k


def test_union_point_location_offsets():
    inst = [E.make_instance(s, n_range=(2, 2), n_receivers=5) for s in (4, 5)]
    meshes = [i["mesh"] for i in inst]
    u = E.union(meshes)
    # instances overlap in space, so location is per instance then offset
    t0 = meshes[0].locate(inst[0]["receivers"])
    t1 = meshes[1].locate(inst[1]["receivers"]) + meshes[0].n_tet
    assert t0.max() < meshes[0].n_tet <= t1.min()
    assert isinstance(u, M.Mesh)


def _asymmetry(m, p=2):
    o = _oracle_for(m, p)
    o.build_operator()
    K = o.K()
    rp, col = o.pattern()
    blk = {}
    for P in range(o.n_poly):
        for b in range(rp[P], rp[P + 1]):
            blk[(P, col[b])] = K[b]
    return max(np.abs(A - blk[(Q, P)].T).max() for (P, Q), A in blk.items()) / np.abs(K).max()


def test_face_material_convention_and_field_smoothness():
    """The face terms take the owning element's material on both sides
    (src/assemble_element.f90:446), so the operator is symmetric (SIPG) only
    for uniform material; independent per-tet lognormal fields make it far
    from symmetric, smooth fields (the ensembles' instances) only slightly."""
    m = M.cube_mesh(3, agglomerate=True)
    assert _asymmetry(m) < 1e-12
    rng = np.random.default_rng(0)
    rough = M.cube_mesh(8, agglomerate=True)
    rough.lam = np.exp(rng.normal(0, 0.5, rough.n_tet))
    rough.mu = np.exp(rng.normal(0, 0.5, rough.n_tet))
    smooth = E.make_instance(0, n_range=(8, 8), p_choices=(2,))["mesh"]
    a_rough, a_smooth = _asymmetry(rough), _asymmetry(smooth)
    assert a_rough > 0.1
    assert a_smooth < 0.4 * a_rough


def test_grouped_location_equals_per_instance():
    """Sources and receivers located with one query per cube size give the
    tets that each instance's own query gives (here the grouped queries take
    the k-d tree path and the single ones the exhaustive one)."""
    from lymph3d_apsc_amd import ensemble as E
    inst = [E.make_instance(300 + s, n_range=(8, 9), n_receivers=9) for s in range(12)]
    meshes = [i["mesh"] for i in inst]
    pts = [np.concatenate([i["source"], i["receivers"]]) for i in inst]
    grouped = E._locate_each(meshes, pts)
    for m, q, t in zip(meshes, pts, grouped):
        assert np.array_equal(t, m.locate(q))
