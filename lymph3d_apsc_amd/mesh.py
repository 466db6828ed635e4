"""Host-side mesh pipeline: the part of LYMPH3D's set-up that produces the
flattened mesh the GPU path consumes (struct LymphMesh of include/lymph_gpu.h).

It stands in for the reference's Fortran host (READ_INPUT_FILES,
MAKE_PARTITION_AND_MPI_FILES) in standalone runs:
  * deck reader: Poly.input / .mate / .mesh (src/Poly_readfile.f90:9-187,
    327-707, 715-1049);
  * structured cube generator: n^3 cubes x 6 tets (Kuhn split, positively
    oriented), optional agglomeration of each cube's 6 tets into one
    polyhedron, or seeded region growing into polyhedra of a target size;
  * face matching by sorted vertex triples, vectorised (replaces the O(F^2)
    search of CREATE_NEIGH_EL_TRIA, src/MAKE_PARTITION_AND_MPI_FILES.f90:
    2030-2145) with boundary tags mapped as :2105-2109 (2,3,4 -> Dirichlet -1;
    5,6,7 -> Neumann -2);
  * face normals/areas (CREATE_NORMAL_FACE :1429-1504) and polyhedron
    bounding boxes/diameters (CREATE_BBOX_EL :1510-1580).
Everything is vectorised numpy; nothing here is on the per-step path.
"""
import os

import numpy as np

# tet faces 1..4 of the reference: vertex sets for matching and ordered
# triplets (a, b, c) for n = (c-b) x (a-b)
FACE_SETS = np.array([[0, 1, 2], [0, 1, 3], [1, 2, 3], [2, 0, 3]])
FACE_TRI = np.array([[2, 1, 0], [0, 1, 3], [1, 2, 3], [2, 0, 3]])

DIRICHLET_TAGS = (2, 3, 4)
NEUMANN_TAGS = (5, 6, 7)


class MeshError(RuntimeError):
    """Mesh set-up failure; .code follows the reference's exit codes."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class Mesh:
    """Flattened mesh (0-based).  Arrays are C-contiguous numpy arrays matching
    the fields of LymphMesh."""

    def __init__(self, xyz, tet_vert, tet_poly, tria_vert, tria_tag, tag_space,
                 rho=None, lam=None, mu=None, degree=None, dt=None, stop_time=None):
        self.xyz = np.ascontiguousarray(xyz, np.float64)
        self.tet_vert = np.ascontiguousarray(tet_vert, np.int32)
        nt = self.tet_vert.shape[0]
        if tet_poly is None:
            tet_poly = np.arange(nt)
        self.tet_poly = np.ascontiguousarray(tet_poly, np.int32)
        self.n_tet = nt
        self.n_poly = int(self.tet_poly.max()) + 1
        self.tria_vert = np.ascontiguousarray(tria_vert, np.int32).reshape(-1, 3)
        self.tria_tag = np.ascontiguousarray(tria_tag, np.int32)
        self.tag_space = dict(tag_space)
        ones = np.ones(nt)
        self.rho = np.ascontiguousarray(ones if rho is None else rho, np.float64)
        self.lam = np.ascontiguousarray(ones if lam is None else lam, np.float64)
        self.mu = np.ascontiguousarray(ones if mu is None else mu, np.float64)
        self.degree, self.dt, self.stop_time = degree, dt, stop_time
        self._build()

    # --- connectivity and geometry --------------------------------------
    def _build(self):
        xyz, tv = self.xyz, self.tet_vert
        nt = self.n_tet
        v = xyz[tv]  # (nt, 4, 3)
        J = np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]], axis=2)
        det = np.linalg.det(J)
        if np.any(det <= 0):
            raise MeshError(11, f"{int((det <= 0).sum())} tetrahedra are degenerate or "
                                "negatively oriented (EXIT_ELEM_ORIENT)")
        nn = np.int64(xyz.shape[0])

        def key(tri):
            s = np.sort(tri, axis=-1).astype(np.int64)
            return (s[..., 0] * nn + s[..., 1]) * nn + s[..., 2]

        fk = key(tv[:, FACE_SETS]).reshape(-1)  # (4 nt,)
        order = np.argsort(fk, kind="stable")
        sk = fk[order]
        neigh = np.full(4 * nt, -100, np.int32)
        nface = np.full(4 * nt, -1, np.int32)
        same = sk[1:] == sk[:-1]
        if np.any(same[1:] & same[:-1]):
            raise MeshError(18, "a face is shared by more than two tetrahedra")
        i = np.nonzero(same)[0]
        a, b = order[i], order[i + 1]
        neigh[a], neigh[b] = b // 4, a // 4
        nface[a], nface[b] = b % 4, a % 4
        bc = np.zeros(4 * nt, np.int32)
        free = np.nonzero(neigh == -100)[0]
        if free.size:
            if self.tria_vert.shape[0] == 0:
                raise MeshError(18, "boundary faces without boundary triangles")
            tk = key(self.tria_vert)
            to = np.argsort(tk, kind="stable")
            pos = np.searchsorted(tk[to], fk[free])
            pos = np.minimum(pos, tk.size - 1)
            hit = tk[to][pos] == fk[free]
            if not np.all(hit):
                raise MeshError(18, f"{int((~hit).sum())} boundary faces have no tagged triangle")
            tag = self.tria_tag[to][pos]
            kind = np.where(np.isin(tag, DIRICHLET_TAGS), -1,
                            np.where(np.isin(tag, NEUMANN_TAGS), -2, 0))
            if np.any(kind == 0):
                raise MeshError(18, "boundary triangle tag is neither Dirichlet nor Neumann")
            neigh[free] = kind
            nface[free] = free % 4
            bc[free] = np.array([self.tag_space.get(int(t), 0) for t in tag], np.int32) \
                if len(self.tag_space) else 0
        self.neigh = neigh.reshape(nt, 4)
        self.neigh_face = nface.reshape(nt, 4)
        self.bc_space = bc.reshape(nt, 4)
        # normals / areas
        A = xyz[tv[:, FACE_TRI[:, 0]]]
        B = xyz[tv[:, FACE_TRI[:, 1]]]
        Cc = xyz[tv[:, FACE_TRI[:, 2]]]
        Q = A - B
        P = Cc - B
        n1 = P[..., 1] * Q[..., 2] - P[..., 2] * Q[..., 1]
        n2 = -(P[..., 0] * Q[..., 2]) + P[..., 2] * Q[..., 0]
        n3 = P[..., 0] * Q[..., 1] - P[..., 1] * Q[..., 0]
        norm = np.sqrt(n1 * n1 + n2 * n2 + n3 * n3)
        self.normal = np.ascontiguousarray(np.stack([n1 / norm, n2 / norm, n3 / norm], axis=-1))
        self.area = np.ascontiguousarray(norm)
        self._poly_geometry(v)

    def _poly_geometry(self, v):
        npoly = self.n_poly
        order = np.argsort(self.tet_poly, kind="stable")
        counts = np.bincount(self.tet_poly, minlength=npoly)
        if np.any(counts == 0):
            raise MeshError(4, "empty polyhedron id")
        self.poly_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        self.poly_tets = order.astype(np.int32)
        bbox = np.zeros((npoly, 6))
        hk = np.zeros(npoly)
        for size in np.unique(counts):
            polys = np.nonzero(counts == size)[0]
            tets = order[self.poly_ptr[polys][:, None] + np.arange(size)[None, :]]
            pts = v[tets].reshape(len(polys), 4 * size, 3)  # vertices with duplicates
            bbox[polys, 0::2] = pts.min(axis=1)
            bbox[polys, 1::2] = pts.max(axis=1)
            # diameter over distinct vertices only: a repeated vertex repeats
            # distances (same squares, same bits), so the maximum is unchanged
            vid = np.sort(self.tet_vert[tets].reshape(len(polys), 4 * size), axis=1)
            first = np.ones(vid.shape, bool)
            first[:, 1:] = vid[:, 1:] != vid[:, :-1]
            nv = int(first.sum(axis=1).max())
            sel = np.argsort(~first, axis=1, kind="stable")[:, :nv]
            cols = np.take_along_axis(vid, sel, axis=1)
            cols = np.where(np.take_along_axis(first, sel, axis=1), cols, vid[:, :1])
            pts = self.xyz[cols]
            ii, jj = np.triu_indices(nv, 1)
            h = np.zeros(len(polys))
            step = max(1, int(4e6 // max(1, len(ii))))
            for s in range(0, len(polys), step):
                d = pts[s:s + step, ii] - pts[s:s + step, jj]
                dist = np.sqrt(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2])
                h[s:s + step] = dist.max(axis=1)
            hk[polys] = h
        self.bbox = np.ascontiguousarray(bbox)
        self.hk = np.ascontiguousarray(hk)

    def locate(self, points):
        """Tet of each point as FIND_ELEM_FROM_POINT picks it (src/find_tet.f90:
        82-124): nearest barycentre ((x1+x2+x3+x4)/4), first minimum on ties.
        Large queries take 32 candidates from a k-d tree and decide among them
        with the same squared-distance expression; a point whose candidates
        are too close to call is searched exhaustively."""
        pts = np.asarray(points, np.float64).reshape(-1, 3)
        loc = getattr(self, "_locator", None)
        if loc is None or loc[0] is not self.tet_vert:
            v = self.xyz[self.tet_vert]
            loc = [self.tet_vert, (((v[:, 0] + v[:, 1]) + v[:, 2]) + v[:, 3]) / 4, None]
            self._locator = loc  # shared by shallow copies of this mesh
        bc = loc[1]

        def brute(q):
            d = q[:, None, :] - bc[None, :, :]
            d2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]
            return np.argmin(d2, axis=1)  # first minimum on ties

        out = np.zeros(pts.shape[0], np.int32)
        if pts.shape[0] * self.n_tet <= 200_000 or self.n_tet <= 64:
            step = max(1, int(2e7 // max(1, self.n_tet)))
            for k in range(0, pts.shape[0], step):
                out[k:k + step] = brute(pts[k:k + step])
            return out
        if loc[2] is None:
            from scipy.spatial import cKDTree
            loc[2] = cKDTree(bc)
        tree = loc[2]
        dist, cand = tree.query(pts, k=32)
        d = pts[:, None, :] - bc[cand]
        d2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]
        m = d2.min(axis=1)
        out[:] = np.where(d2 == m[:, None], cand, self.n_tet).min(axis=1)
        # the exhaustive rule could pick a tet outside the candidates only if
        # the 32nd candidate ties the nearest up to rounding
        unsure = np.nonzero(dist[:, -1] * dist[:, -1] <= m * (1 + 1e-9) + 1e-300)[0]
        if unsure.size:
            out[unsure] = brute(pts[unsure])
        return out

    def hmax(self):
        return float(self.hk.max())  # compute_hmax, src/solution_processing.f90:197-214

    def n_dof(self, p):
        return 3 * self.n_poly * (p + 1) * (p + 2) * (p + 3) // 6

    def write_flat(self, path, p, dt, stop_time):
        """Write the flattened mesh as the stream file read by the Fortran
        drop-in driver (fortran/drive_mesh.f90): int32 n_node, n_tet, n_poly,
        p, 0; float64 dt, stop_time; then the LymphMesh arrays in field order."""
        with open(path, "wb") as f:
            np.array([self.xyz.shape[0], self.n_tet, self.n_poly, p, 0], np.int32).tofile(f)
            np.array([dt, stop_time], np.float64).tofile(f)
            for a, dt_ in ((self.xyz, np.float64), (self.tet_vert, np.int32),
                           (self.tet_poly, np.int32), (self.neigh, np.int32),
                           (self.bc_space, np.int32), (self.normal, np.float64),
                           (self.area, np.float64), (self.bbox, np.float64),
                           (self.hk, np.float64), (self.rho, np.float64),
                           (self.lam, np.float64), (self.mu, np.float64)):
                np.ascontiguousarray(a, dt_).tofile(f)

    def to_c(self):
        """Return (LymphMesh struct, keep-alive list)."""
        from ._lib import LymphMesh, ptr
        arrs = [self.xyz, self.tet_vert, self.tet_poly,
                np.ascontiguousarray(self.neigh, np.int32),
                np.ascontiguousarray(self.bc_space, np.int32),
                self.normal, self.area, self.bbox, self.hk, self.rho, self.lam, self.mu]
        m = LymphMesh(self.xyz.shape[0], self.n_tet, self.n_poly, *(ptr(a) for a in arrs))
        return m, arrs


# ----------------------------------------------------------------------------
def read_deck(folder, force_unit_material=True):
    """Read a LYMPH3D deck (Poly.input + .mate + .mesh).  With
    force_unit_material the driver's override rho=lambda=mu=1 is applied
    (src/Lymph3D.f90:196-203); otherwise lambda=rho(vp^2-2vs^2), mu=rho vs^2
    from the MATE line (src/Poly_readfile.f90:408-412)."""
    hdr = {}
    with open(os.path.join(folder, "Poly.input")) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line or line[0] == " ":
                continue  # comment lines start with a blank
            parts = line.split()
            if len(parts) < 2:
                raise MeshError(8, f"no argument given for command {parts[0]!r} (EXIT_SYNTAX_ERROR)")
            hdr[parts[0]] = parts[1]
    for k in ("GRIDFILE", "MATFILE"):
        if k not in hdr:
            raise MeshError(9, f"Poly.input lacks {k}")
    mate = {}
    tag_space = {}
    matfile = os.path.join(folder, hdr["MATFILE"] + ".mate")
    if not os.path.exists(matfile):
        raise MeshError(9, f"missing file {matfile} (EXIT_MISSING_FILE)")
    with open(matfile) as f:
        for line in f:
            key, parts = line[:4], line.split()
            if key == "MATE":
                tag, deg = int(parts[1]), int(parts[2])
                rho, vs, vp = float(parts[3]), float(parts[4]), float(parts[5])
                mate[tag] = (deg, rho, rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2)
            elif key in ("DIRI", "NEUM"):
                tag_space[int(parts[1])] = int(parts[2])
    gridfile = os.path.join(folder, hdr["GRIDFILE"] + ".mesh")
    if not os.path.exists(gridfile):
        raise MeshError(9, f"missing file {gridfile} (EXIT_MISSING_FILE)")
    with open(gridfile) as f:
        lines = f.read().splitlines()
    i = 0
    while lines[i].startswith("#"):
        i += 1
    nnode, nelem = (int(x) for x in lines[i].split()[:2])
    i += 1
    xyz = np.loadtxt(lines[i:i + nnode], usecols=(1, 2, 3), ndmin=2)
    i += nnode
    tets, tmat, trias, ttag, polys = [], [], [], [], []
    for line in lines[i:i + nelem]:
        parts = line.split()
        tag, kind = int(parts[1]), parts[2].lower()
        if kind == "tetra" and tag in mate:
            tets.append([int(x) - 1 for x in parts[3:7]])
            tmat.append(tag)
        elif kind == "tria" and tag in tag_space:
            trias.append([int(x) - 1 for x in parts[3:6]])
            ttag.append(tag)
        elif kind == "poly":
            polys.append(int(parts[3]) - 1)
    if not tets:
        raise MeshError(13, "no tetrahedra (EXIT_NO_ELEMENTS)")
    tmat = np.asarray(tmat)
    degs = {mate[t][0] for t in tmat}
    deg = mate[int(tmat[0])][0]  # Elem_loc(1)%Degree (src/Lymph3D.f90:171-172)
    if force_unit_material:
        rho = lam = mu = np.ones(len(tets))
    else:
        rho = np.array([mate[t][1] for t in tmat])
        lam = np.array([mate[t][2] for t in tmat])
        mu = np.array([mate[t][3] for t in tmat])
    tet_poly = np.asarray(polys) if len(polys) == len(tets) else None
    m = Mesh(xyz, np.asarray(tets), tet_poly, np.asarray(trias).reshape(-1, 3),
             np.asarray(ttag), tag_space, rho, lam, mu, degree=deg,
             dt=float(hdr.get("TIMESTEP", "nan")), stop_time=float(hdr.get("STOPTIME", "nan")))
    m.mixed_degree = len(degs) > 1
    m.monitor_every = int(hdr.get("TMONITOR", "0"))  # steps between SAVEOUT exports
    return m


# ----------------------------------------------------------------------------
# Kuhn (Freudenthal) split of the unit cube along its main diagonal: one tet
# per permutation of the axes, made positively oriented.
_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def _kuhn_tets():
    tets = []
    for p in _PERMS:
        c = [0, 0, 0]
        path = [tuple(c)]
        for ax in p:
            c[ax] = 1
            path.append(tuple(c))
        ids = [a + 2 * b + 4 * cc for a, b, cc in path]  # corner id in the cube
        pts = np.array(path, float)
        det = np.linalg.det(np.stack([pts[1] - pts[0], pts[2] - pts[0], pts[3] - pts[0]], axis=1))
        if det < 0:
            ids[1], ids[2] = ids[2], ids[1]
        tets.append(ids)
    return np.array(tets)


def cube_mesh(n, agglomerate=False, dirichlet_faces="all", poly_size=None, seed=0,
              lengths=(1.0, 1.0, 1.0)):
    """Unit cube split into n^3 cubes x 6 tets.

    agglomerate: each cube's 6 tets form one polyhedron (n^3 polyhedra).
    poly_size=(lo, hi): instead, seeded region growing over the tet adjacency
      into polyhedra of lo..hi tets (uneven elements).
    dirichlet_faces: "all" (tag 2 on all six faces) or a set of face names
      among x0,x1,y0,y1,z0,z1 that get Dirichlet tag 2; others get Neumann tag 5.
    """
    L = np.asarray(lengths, float)
    g = np.arange(n + 1)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1) / n * L
    nid = lambda i, j, k: (i * (n + 1) + j) * (n + 1) + k  # noqa: E731
    ci, cj, ck = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    corners = np.stack([nid(ci + a, cj + b, ck + c) for c in (0, 1) for b in (0, 1) for a in (0, 1)],
                       axis=1)  # corner id a + 2b + 4c
    kt = _kuhn_tets()
    tets = corners[:, kt].reshape(-1, 4)
    cube_of_tet = np.repeat(np.arange(n ** 3), 6)
    # boundary triangles: every tet face on the box boundary
    keys = np.sort(tets[:, FACE_SETS], axis=2).reshape(-1, 3)
    fx = xyz[keys]  # (F, 3, 3)
    names = ["x0", "x1", "y0", "y1", "z0", "z1"]
    tri, tag = [], []
    if dirichlet_faces == "all":
        dset = set(names)
    else:
        dset = set(dirichlet_faces)
    for d in range(3):
        for side, val in ((0, 0.0), (1, L[d])):
            on = np.all(np.isclose(fx[:, :, d], val), axis=1)
            tri.append(keys[on])
            tag.append(np.full(on.sum(), 2 if names[2 * d + side] in dset else 5))
    tri = np.concatenate(tri)
    tag = np.concatenate(tag)
    if poly_size is not None:
        tet_poly = grow_polyhedra(tets, n, poly_size, seed)
    elif agglomerate:
        tet_poly = cube_of_tet
    else:
        tet_poly = None
    return Mesh(xyz, tets, tet_poly, tri, tag, {2: 1, 5: 1})


def grow_polyhedra(tets, n, size, seed):
    """Seeded region growing over face adjacency into polyhedra whose tet
    counts are drawn uniformly in [lo, hi] (leftover small regions merge into
    a neighbour)."""
    lo, hi = size
    rng = np.random.default_rng(seed)
    nt = tets.shape[0]
    fk = np.sort(tets[:, FACE_SETS], axis=2).reshape(-1, 3).astype(np.int64)
    m = int(tets.max()) + 1
    key = (fk[:, 0] * m + fk[:, 1]) * m + fk[:, 2]
    order = np.argsort(key, kind="stable")
    sk = key[order]
    i = np.nonzero(sk[1:] == sk[:-1])[0]
    a, b = order[i] // 4, order[i + 1] // 4
    adj = [[] for _ in range(nt)]
    for x, y in zip(a.tolist(), b.tolist()):
        adj[x].append(y)
        adj[y].append(x)
    poly = np.full(nt, -1, np.int64)
    pid = 0
    for start in rng.permutation(nt):
        if poly[start] >= 0:
            continue
        target = int(rng.integers(lo, hi + 1))
        front = [start]
        poly[start] = pid
        cnt = 1
        k = 0
        while k < len(front) and cnt < target:
            for nb in adj[front[k]]:
                if poly[nb] < 0 and cnt < target:
                    poly[nb] = pid
                    front.append(nb)
                    cnt += 1
            k += 1
        pid += 1
    # merge regions smaller than lo into a neighbouring region
    counts = np.bincount(poly, minlength=pid)
    for t in range(nt):
        if counts[poly[t]] < lo:
            for nb in adj[t]:
                if poly[nb] != poly[t] and counts[poly[nb]] + counts[poly[t]] <= hi:
                    old = poly[t]
                    poly[poly == old] = poly[nb]
                    counts[poly[nb]] += counts[old]
                    counts[old] = 0
                    break
    _, dense = np.unique(poly, return_inverse=True)
    return dense.astype(np.int32)


# ----------------------------------------------------------------------------
# Writers of the reference's input and partition files: a mesh generated or
# agglomerated here can be handed to the reference's Fortran host, and decks
# and partitions round-trip through read_deck / read_elem_in_poly.

def _mate_tags(mesh):
    """Material tag per tet: tets with equal (rho, lambda, mu) share a tag,
    numbered 1, 8, 9, ... (2-7 are the boundary tags of the BC mapping)."""
    trip = np.stack([mesh.rho, mesh.lam, mesh.mu], axis=1)
    uniq, inv = np.unique(trip, axis=0, return_inverse=True)
    tags = np.array([1] + list(range(8, 8 + len(uniq) - 1)), np.int64)
    return uniq, tags, tags[inv.ravel()]


def write_deck(folder, mesh, degree, dt, stop_time, grid="grid", mat="mat", monitor_every=100,
               save_output=0):
    """Write `mesh` as a LYMPH3D deck: Poly.input (src/Poly_readfile.f90:9-187),
    <mat>.mate (:327-707) and <grid>.mesh (:715-1049).

    * Materials: one MATE line per distinct (rho, lambda, mu), with
      vs = sqrt(mu/rho), vp = sqrt((lambda + 2 mu)/rho), the inverse of the
      reader's lambda = rho (vp^2 - 2 vs^2), mu = rho vs^2 (:408-412).
    * Boundary triangles keep their tags; one DIRI/NEUM line per tag with its
      space-function tag.
    * Agglomerated meshes: one `poly` element line per tet, in tet order, with
      the 1-based polyhedron id (read as elem_in_poly, :1017-1032).
    * Nodes are written with 17 significant digits (lines stay under the
      reader's 100 characters), so coordinates round-trip exactly."""
    os.makedirs(folder, exist_ok=True)
    uniq, tags, tet_tag = _mate_tags(mesh)
    with open(os.path.join(folder, "Poly.input"), "w") as f:
        f.write(f"GRIDFILE  {grid}\nMATFILE   {mat}\nMPIFILE   FILES_MPI\nMONFILE   MONITORS\n\n")
        f.write(f"TIMESTEP {float(dt)!r}\nTMONITOR {int(monitor_every)}\nSTOPTIME {float(stop_time)!r}\n"
                f"SAVEOUT {int(save_output)}\n")
    with open(os.path.join(folder, mat + ".mate"), "w") as f:
        f.write("      mat_id   degree    rho    vs    vp    qs    qp\n")
        for (rho, lam, mu), t in zip(uniq, tags):
            vs, vp = np.sqrt(mu / rho), np.sqrt((lam + 2 * mu) / rho)
            f.write(f"MATE {int(t):4d} {int(degree):4d} {float(rho)!r} {float(vs)!r} {float(vp)!r} 30 60\n")
        f.write("\n")
        for t in sorted({int(x) for x in mesh.tria_tag}):
            key = "DIRI" if t in DIRICHLET_TAGS else "NEUM"
            f.write(f"{key} {t:4d} {int(mesh.tag_space.get(t, 0)):4d} 0.0 0.0 0.0 0.0\n")
    agg = mesh.n_poly != mesh.n_tet or np.any(mesh.tet_poly != np.arange(mesh.n_tet))
    n_elem = mesh.tria_vert.shape[0] + mesh.n_tet + (mesh.n_tet if agg else 0)
    with open(os.path.join(folder, grid + ".mesh"), "w") as f:
        f.write(f"{mesh.xyz.shape[0]:8d} {n_elem:8d}   0   0   0\n")
        for i, (x, y, z) in enumerate(mesh.xyz, 1):
            f.write(f"{i}  {x:+.16e}  {y:+.16e}  {z:+.16e}\n")
        k = 0
        for (a, b, c), t in zip(mesh.tria_vert + 1, mesh.tria_tag):
            k += 1
            f.write(f"{k} {int(t)} tria {a} {b} {c}\n")
        for (a, b, c, d), t in zip(mesh.tet_vert + 1, tet_tag):
            k += 1
            f.write(f"{k} {int(t)} tetra {a} {b} {c} {d}\n")
        if agg:
            for pid in mesh.tet_poly + 1:
                k += 1
                f.write(f"{k} 0 poly {int(pid)}\n")


def partition_file_name(rank=0):
    """elem_in_poly_NNNNNN.mpi with the rank in the last digits (the name the
    reference gives rank `rank` < 10, src/MAKE_PARTITION_AND_MPI_FILES.f90:401-404)."""
    return f"elem_in_poly_{rank:06d}.mpi"


def write_elem_in_poly(folder, mesh, rank=0):
    """Tet -> polyhedron map in the reference's partition-file format (written
    by MESH_AGGLOMERATION / WRITE_POLY_INFO, :463-476, :551-556; read by
    CREATE_LOCAL_MESH, :1163-1172): the tet count, then one line per tet with
    its 1-based global id and 1-based polyhedron id."""
    os.makedirs(folder, exist_ok=True)
    path = os.path.join(folder, partition_file_name(rank))
    with open(path, "w") as f:
        f.write(f" {mesh.n_tet:12d}\n")
        for i, pid in enumerate(mesh.tet_poly + 1, 1):
            f.write(f" {i:12d} {int(pid):12d}\n")
    return path


def write_poly_global(folder, mesh):
    """poly_global.mpi (CREATE_GLOBAL_POLY_MAP, :618-630): the tet count, then
    the 1-based polyhedron id of every tet, one per line."""
    os.makedirs(folder, exist_ok=True)
    path = os.path.join(folder, "poly_global.mpi")
    with open(path, "w") as f:
        f.write(f" {mesh.n_tet:12d}\n")
        for pid in mesh.tet_poly + 1:
            f.write(f" {int(pid):12d}\n")
    return path


def read_elem_in_poly(path, n_poly=None):
    """0-based polyhedron id per tet from a partition file (second column,
    positional, as CREATE_LOCAL_MESH reads it).  With n_poly given, the
    reference's MESH_CORRECTION (src/MESH_CORRECTION.f90, single rank) is
    applied: every id 1..n_poly left without tets takes the first tet of a
    polyhedron with more than one, scanning polyhedra from 1 upwards and
    filling the missing ids from the last."""
    with open(path) as f:
        n = int(f.readline().split()[0])
        ids = np.array([int(f.readline().split()[1]) for _ in range(n)], np.int64)
    if n_poly is not None:
        counts = np.bincount(ids, minlength=n_poly + 1)
        missing = [q for q in range(1, n_poly + 1) if counts[q] == 0]
        i = 1
        while missing and i <= n_poly:
            tets = np.nonzero(ids == i)[0]
            if tets.size > 1:
                ids[tets[0]] = missing.pop()
            i += 1
        if missing:
            raise MeshError(4, f"{len(missing)} polyhedra left without tetrahedra")
    return (ids - 1).astype(np.int32)
