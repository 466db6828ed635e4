"""Independent reader for LYMPH3D input decks (Poly.input + .mate + .mesh).

TEST INFRASTRUCTURE ONLY (part of the oracle): used by tests/ and bench.py's
cpu_baseline leg to feed the C oracle; the product has its own reader.

Follows the reference readers:
  Poly.input keywords  src/Poly_readfile.f90:60-170 (lines starting with a
                       blank are comments; keyword = text up to first blank)
  .mate lines          src/Poly_readfile.f90:380-470 (keyword = first 4 chars;
                       MATE tag degree rho vs vp qs qp; DIRI/NEUM face_tag
                       space_fun_tag v1..v4)
  .mesh                src/Poly_readfile.f90:847-1049 (header "nnode nelem";
                       node lines "id x y z"; element lines "id tag kind v...";
                       tetra kept if tag is a MATE tag, tria kept if tag is a
                       DIRI/NEUM tag)
  BC kind by tag       src/MAKE_PARTITION_AND_MPI_FILES.f90:2105-2109
                       (2,3,4 -> Dirichlet -1; 5,6,7 -> Neumann -2)
"""
import os

import numpy as np


def read_deck(folder):
    hdr = {}
    with open(os.path.join(folder, "Poly.input")) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line or line[0] == " ":
                continue
            parts = line.split()
            if len(parts) >= 2:
                hdr[parts[0]] = parts[1]
    mate_tags, degrees, diri, neum = [], [], {}, {}
    with open(os.path.join(folder, hdr["MATFILE"] + ".mate")) as f:
        for line in f:
            key = line[:4]
            parts = line.split()
            if key == "MATE":
                mate_tags.append(int(parts[1]))
                degrees.append(int(parts[2]))
            elif key == "DIRI":
                diri[int(parts[1])] = int(parts[2])
            elif key == "NEUM":
                neum[int(parts[1])] = int(parts[2])
    with open(os.path.join(folder, hdr["GRIDFILE"] + ".mesh")) as f:
        lines = f.read().splitlines()
    i = 0
    while lines[i].startswith("#"):
        i += 1
    nnode, nelem = (int(v) for v in lines[i].split()[:2])
    i += 1
    xyz = np.array([[float(v) for v in lines[i + k].split()[1:4]] for k in range(nnode)])
    i += nnode
    tets, tmat, trias, ttag = [], [], [], []
    for k in range(nelem):
        parts = lines[i + k].split()
        tag, kind = int(parts[1]), parts[2].lower()
        if kind == "tetra" and tag in mate_tags:
            tets.append([int(v) - 1 for v in parts[3:7]])
            tmat.append(tag)
        elif kind == "tria" and (tag in diri or tag in neum):
            trias.append([int(v) - 1 for v in parts[3:6]])
            ttag.append(tag)
    n_tags = max([8] + [t + 1 for t in ttag])
    tag_kind = np.zeros(n_tags, np.int32)
    tag_space = np.zeros(n_tags, np.int32)
    for t in range(n_tags):
        if t in (2, 3, 4):
            tag_kind[t] = -1
        elif t in (5, 6, 7):
            tag_kind[t] = -2
        if t in diri:
            tag_space[t] = diri[t]
        if t in neum:
            tag_space[t] = neum[t]
    return {
        "dt": float(hdr["TIMESTEP"]),
        "stop_time": float(hdr["STOPTIME"]),
        "degree": degrees[0],
        "xyz": np.ascontiguousarray(xyz, dtype=np.float64),
        "tet_vert": np.ascontiguousarray(tets, dtype=np.int32),
        "tet_mat": np.asarray(tmat, np.int32),
        "tria_vert": np.ascontiguousarray(trias, dtype=np.int32).reshape(-1, 3),
        "tria_tag": np.asarray(ttag, np.int32),
        "tag_kind": tag_kind,
        "tag_space": tag_space,
    }
