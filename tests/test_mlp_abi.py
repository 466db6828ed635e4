"""liblymph_mlp.so loads and exports every symbol include/lymph_mlp.h
declares; without a GPU it fails loudly (no CPU fallback)."""
import ctypes as C
import os
import re

import pytest

from lymph3d_apsc_amd import device_count, mlp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_and_exports_agree():
    hdr = open(os.path.join(ROOT, "include", "lymph_mlp.h")).read()
    declared = set(re.findall(r"\b(lymph_mlp_\w+)\s*\(", hdr))
    assert declared == set(mlp.EXPORTS)
    L = mlp.lib()
    for name in declared:
        assert hasattr(L, name), name


@pytest.mark.skipif(device_count() > 0, reason="needs a machine without a GPU")
def test_no_gpu_fails_loudly():
    with pytest.raises(mlp.MlpError) as e:
        mlp.Mlp([8, 16, 8])
    assert "no HIP device" in str(e.value)


def test_argument_errors_need_no_gpu():
    h = C.c_void_p()
    import numpy as np
    w = np.array([8, 12, 8], np.int32)
    assert mlp.lib().lymph_mlp_create(2, C.c_void_p(w.ctypes.data), 64, C.byref(h)) == 101
    assert b"multiples of 8" in mlp.lib().lymph_mlp_last_error()
