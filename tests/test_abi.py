"""The C-ABI library loads and exports every symbol include/lymph_gpu.h
declares; without a GPU it refuses to create a context (no CPU fallback)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from lymph3d_apsc_amd import _lib, mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "lymph_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lymph_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert sorted(_lib.EXPORTS) == names
    # the .so's dynamic symbol table has them as real exports
    out = os.popen(f"nm -D --defined-only {_lib.LIB_PATH}").read()
    for n in names:
        assert re.search(rf"\bT {n}\b", out), n


def test_built_for_gfx950():
    out = os.popen(f"strings {_lib.LIB_PATH} | grep -c gfx950").read().strip()
    assert int(out or 0) > 0


def test_no_device_means_loud_failure():
    L = _lib.lib()
    assert L.lymph_version().decode().startswith("lymph3d_apsc_amd")
    if L.lymph_device_count() > 0:
        pytest.skip("a GPU is visible")
    m = mesh.cube_mesh(1)
    cm, keep = m.to_c()
    h = C.c_void_p()
    rc = L.lymph_ctx_create(C.byref(cm), 2, 10.0, -1.0, C.byref(h))
    assert rc == _lib.ERR_HIP
    assert b"no HIP device" in L.lymph_last_error()
    assert not h.value


def test_argument_errors_before_device():
    L = _lib.lib()
    m = mesh.cube_mesh(1)
    cm, keep = m.to_c()
    h = C.c_void_p()
    assert L.lymph_ctx_create(C.byref(cm), 7, 10.0, -1.0, C.byref(h)) == _lib.ERR_ARG
    assert L.lymph_ctx_create(None, 2, 10.0, -1.0, C.byref(h)) == _lib.ERR_ARG
    bad = m.neigh.copy()
    bad[0, 0] = 10 ** 6
    arrs = list(keep)
    arrs[3] = np.ascontiguousarray(bad, np.int32)
    cm.neigh = arrs[3].ctypes.data
    rc = L.lymph_ctx_create(C.byref(cm), 2, 10.0, -1.0, C.byref(h))
    assert rc == _lib.EXIT_NEIGHBOUR
    assert L.lymph_build_operator(None) == _lib.ERR_ARG
    assert abs(L.lymph_time_function(1, 0.25) - np.sin(np.sqrt(2) * np.pi * 0.25)) < 1e-15


def test_fortran_dropin_links_and_fails_loudly_without_gpu(tmp_path):
    """The ISO_C_BINDING module and the reference-loop driver
    (lymph3d_apsc_amd/fortran) link against liblymph_gpu.so; with no device
    the driver stops at lymph_ctx_create with the library's message."""
    import subprocess
    exe = os.path.join(ROOT, "lymph3d_apsc_amd", "fortran", "drive_mesh")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built (no amdflang)")
    if _lib.lib().lymph_device_count() > 0:
        pytest.skip("a GPU is visible")
    m = mesh.read_deck(os.path.join(ROOT, "tests", "golden", "decks", "tet_5"))
    f = tmp_path / "t5.bin"
    m.write_flat(str(f), m.degree, m.dt, m.stop_time)
    r = subprocess.run([exe, str(f)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3, r.stdout + r.stderr
    assert "no HIP device" in r.stdout


def test_fortran_module_binds_every_entry_point():
    """lymph3d_apsc_amd/fortran/lymph_gpu.f90 declares an interface, bound by
    name, for every function of include/lymph_gpu.h."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "lymph_gpu.h")).read()
    declared = set(re.findall(r"^[A-Za-z_][A-Za-z0-9_ ]*\*?\s*\b(lymph_\w+)\s*\(", hdr, re.M))
    f90 = open(os.path.join(root, "lymph3d_apsc_amd", "fortran", "lymph_gpu.f90")).read()
    bound = set(re.findall(r"bind\(C, name='(lymph_\w+)'\)", f90))
    assert declared and declared <= bound, sorted(declared - bound)
