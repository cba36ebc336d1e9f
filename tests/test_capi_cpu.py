"""CPU checks of the C ABI: libscalehip.so loads and exports every entry point
include/scale_hip.h declares; host-side argument validation rejects what the
library does not support (no GPU call is made)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "scale_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(scale_hip_\w+)\s*\(", src)))


def test_header_declares_the_interfaces():
    names = declared_functions()
    for need in ("scale_hip_init", "scale_hip_finalize", "scale_hip_tstep_short_hevi", "scale_hip_tstep_tracer_heve",
                 "scale_hip_numfilter_coef", "scale_hip_model_large_step", "scale_hip_model_short_step"):
        assert need in names


def test_library_exports_every_declared_symbol():
    from scale_amd import hip
    lib = hip.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_error_strings():
    from scale_amd import hip
    lib = hip.load_library()
    assert b"gfx950" in lib.scale_hip_version()
    assert isinstance(lib.scale_hip_last_error(), bytes)


def test_init_rejects_bad_grids():
    from scale_amd import hip
    lib = hip.load_library()
    h = C.c_void_p()
    # halo other than 2
    g = hip.HipGrid(20, 14, 14, 3, 18, 3, 12, 3, 12, 3, 3, 3, 3, 3, 0)
    assert lib.scale_hip_init(C.byref(g), C.byref(h)) != 0
    assert b"HALO" in lib.scale_hip_last_error()
    # UD5/CD6 tracers need a lateral halo of 3 and KMAX >= 5
    for s in (5, 6):
        g = hip.HipGrid(20, 14, 14, 3, 18, 3, 12, 3, 12, 2, 2, 2, 3, s, 0)
        assert lib.scale_hip_init(C.byref(g), C.byref(h)) != 0
        assert b"UD5" in lib.scale_hip_last_error()
        g = hip.HipGrid(8, 16, 16, 3, 6, 4, 13, 4, 13, 2, 3, 3, 3, s, 0)  # KMAX = 4
        assert lib.scale_hip_init(C.byref(g), C.byref(h)) != 0
        assert b"KMAX >= 5" in lib.scale_hip_last_error()
    # unsupported flux scheme
    g = hip.HipGrid(20, 14, 14, 3, 18, 3, 12, 3, 12, 2, 2, 2, 5, 3, 0)
    assert lib.scale_hip_init(C.byref(g), C.byref(h)) != 0
    assert b"UD3" in lib.scale_hip_last_error()
    # inconsistent extents
    g = hip.HipGrid(21, 14, 14, 3, 18, 3, 12, 3, 12, 2, 2, 2, 3, 3, 0)
    assert lib.scale_hip_init(C.byref(g), C.byref(h)) != 0


def test_null_context_is_an_error():
    from scale_amd import hip
    lib = hip.load_library()
    assert lib.scale_hip_synchronize(None) != 0
    assert b"null" in lib.scale_hip_last_error()


def test_python_front_end_raises_without_library(tmp_path):
    from scale_amd import hip
    with pytest.raises(hip.ScaleHipError):
        hip._lib_backup = hip._lib
        hip._lib = None
        try:
            hip.load_library(str(tmp_path / "missing.so"))
        finally:
            hip._lib = hip._lib_backup


@pytest.mark.parametrize("height", [-1.0, 8000.0, 15000.0, 1.0e9])
def test_wdamp_setup_matches_oracle(height):
    """The host Rayleigh-coefficient setup equals the oracle's bit for bit, on a
    stretched grid, including the 'off' cases (negative height, height above
    the top)."""
    import numpy as np
    import pyoracle as po
    from scale_amd import hip, inputs
    case = inputs.make_case(kmax=40, imax=8, jmax=8, stretched=(50.0, 500.0), ntracer=0)
    g = case.grid
    fz = case.fz
    got = hip.ATMOS_DYN_wdamp_setup(g, 300.0, height, fz)
    ref = po.wdamp_setup(g, 300.0, height, fz)
    assert np.array_equal(got, ref)
    if 0.0 <= height < fz[g.KE + 1]:
        assert got.max() > 0.0 and got[g.KS] == 0.0
