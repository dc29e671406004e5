"""The C-ABI library: loads without a GPU and exports every function that
include/c3d.h declares (no compute calls here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "c3d.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|long long|char|void)\s*\*?\s*(c3d_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("c3d_create", "c3d_destroy", "c3d_getup", "c3d_step", "c3d_divcheck", "c3d_cfl",
              "c3d_chkmf", "c3d_upload_state", "c3d_download_state", "c3d_rhs", "c3d_getuh",
              "c3d_dpcalc", "c3d_upcalc", "c3d_sgs", "c3d_wall_model", "c3d_apply_op"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from compact3d_amd import c3d
    lib = c3d.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_no_gpu_creation_fails_loudly():
    """on a machine without a GPU, creating a channel raises (no CPU fallback)"""
    import pytest
    from compact3d_amd import c3d
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except Exception:
        pass
    with pytest.raises(c3d.C3DError):
        c3d.Channel(33, 33, 33, 4200.0, 6.28, 2.0, 3.14, 1e-3)


def test_footprint_is_positive_and_scales():
    from compact3d_amd import c3d
    a = c3d.footprint_bytes(n1=65, n2=65, n3=65, re=1.0, alx=1.0, aly=2.0, alz=1.0, dt=1.0)
    b = c3d.footprint_bytes(n1=129, n2=65, n3=129, re=1.0, alx=1.0, aly=2.0, alz=1.0, dt=1.0)
    assert a > 0 and 3.5 < b / a < 4.5
