"""C ABI checks that need no GPU: the library loads, exports every entry point
declared in include/channel3d.h, and its host-only routines agree with the
oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "channel3d.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(c3d_\w+)\s*\(", hdr)))


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("c3d_create", "c3d_clcRhsX", "c3d_clcU1Hat", "c3d_clcPPE", "c3d_clcPrSrc",
                 "c3d_FluidVelUpdate", "c3d_PressureUpdate", "c3d_step"):
        assert must in names


def test_library_exports_all_declared():
    from channel3d_amd import lib
    L = lib.load()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing
    # and the python binding covers them all
    assert set(declared_symbols()) == set(lib.EXPORTS)


def test_mesh_coords_matches_oracle(oracle_mod):
    from channel3d_amd import channel
    for args in [
        (64, 2.0, [1.0, 1.0], [32, 32], [1, 1], [0, 1], [1.5, 1.5]),
        (50, 2.0, [1.0, 3.0], [20, 30], [1, 2], [0, 1], [1.4, 0.7]),
        (30, 1.0, [1.0], [30], [3], [0], [1.05]),
        (16, 2.0, [1.0, 1.0], [8, 8], [0, 3], [0, 1], [1.0, 1.1]),
    ]:
        a = channel.mesh_coords(*args)
        b = oracle_mod.mesh_coords(*args)
        assert np.array_equal(a[1:], b[1:])


def test_missing_library_fails_loudly(tmp_path):
    from channel3d_amd import lib
    with pytest.raises(lib.C3DError):
        saved = lib._lib
        lib._lib = None
        try:
            lib.load(str(tmp_path / "nope.so"))
        finally:
            lib._lib = saved


def test_create_rejects_unsupported(oracle_mod):
    # argument validation runs on the host before any device call
    from channel3d_amd import channel, lib
    import tests.cases as cases
    bad = [dict(cases.SMALL, nxc=31), dict(cases.SMALL, nzc=42),
           dict(cases.SMALL, bc=(-1, -1, -1, -1, 0, 0)), dict(cases.SMALL, dt=0.0)]
    for case in bad:
        p = channel.make_params(case)
        yp = np.zeros(case["nyc"] + 2)
        h = ctypes.c_void_p()
        rc = lib.load().c3d_create(ctypes.byref(p), yp.ctypes.data_as(lib._DP), ctypes.byref(h))
        assert rc in (lib_err for lib_err in (-1, -4)), rc
        assert not h.value
