"""The Fortran binding (channel3d_c, m_channel3d_gpu) driven as Channel3d's
ChannelIterate drives its time-step routines, against the CPU oracle.

The check program (channel3d_amd/fortran/c3d_fortran_check.f90, built by
__graft_entry__.build() with amdflang) runs one RK substep through the
per-routine wrappers on host arrays, in the reference's call order and
scratch aliasing, then several time steps on the device-resident state,
takes a statistics sample and writes the stats and restart files, and
reads the restart file into a fresh context, and finally runs the LES
scheme (WALE) the same two ways.
"""
import os
import subprocess

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "channel3d_amd", "fortran", "c3d_fortran_check")


def _ymid(a):
    nx, nz = a.shape[0] - 3, a.shape[2] - 3
    return a[1:nx + 1, :, 1:nz + 1]


def test_fortran_call_sites_match_oracle(oracle_mod, tmp_path):
    if not os.path.exists(PROG):
        pytest.skip("Fortran binding not built (amdflang missing at build time)")
    from channel3d_amd import synth
    O = oracle_mod
    case = dict(cases.SMALL)
    nsteps = 3
    o = O.Oracle(case)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    ux, uy, uz = synth.laminar_noise(nx, ny, nz, o.yp, yly=case["yly"], ubulk=1.0, seed=7)
    o.field(O.UX)[...] = ux
    o.field(O.UY)[...] = uy
    o.field(O.UZ)[...] = uz
    o.project()
    init = [o.field(f).copy(order="F") for f in (O.UX, O.UY, O.UZ, O.PR)]
    (tmp_path / "case.txt").write_text(
        f"{nx} {ny} {nz}\n{case['xlx']!r} {case['yly']!r} {case['zlz']!r}\n"
        f"{case['xnu']!r} {case['dt']!r}\n{nsteps}\n2 1\n")
    with open(tmp_path / "in.bin", "wb") as f:
        np.asarray(o.yp, dtype=np.float64).tofile(f)
        for a in init:
            a.ravel(order="F").tofile(f)
    print(f"fortran check: {nx}x{ny}x{nz}, 1 substep per routine + {nsteps} steps", flush=True)
    r = subprocess.run([PROG, str(tmp_path / "case.txt"), str(tmp_path / "in.bin"),
                        str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    shp = (nx + 3, ny + 3, nz + 3)
    n = int(np.prod(shp))
    out = np.fromfile(tmp_path / "out.bin", dtype=np.float64)
    assert out.size == 21 * n
    got = [out[q * n:(q + 1) * n].reshape(shp, order="F") for q in range(21)]
    # 1. one substep through the per-routine wrappers
    o.substep(1)
    for q, fid in enumerate((O.UX, O.UY, O.UZ, O.PR)):
        ref = _ymid(o.field(fid))
        err = np.abs(_ymid(got[q]) - ref).max() / np.abs(ref).max()
        assert err < 1e-12, (q, err)
    # 2. nsteps device-resident steps
    o2 = O.Oracle(case)
    for fid, a in zip((O.UX, O.UY, O.UZ, O.PR), init):
        o2.field(fid)[...] = a
    for _ in range(nsteps):
        o2.iterate()
    for q, fid in enumerate((O.UX, O.UY, O.UZ, O.PR)):
        ref = _ymid(o2.field(fid))
        err = np.abs(_ymid(got[4 + q]) - ref).max() / np.abs(ref).max()
        assert err < 1e-12, (q, err)
    # 3. the restart file holds the interiors; read back (and the ghosts
    # rebuilt as ChannelInitialize does) it gives the state of 2. exactly
    rst = np.fromfile(str(tmp_path / "out.bin.") + "RestartForCheck%010d" % nsteps,
                      dtype=np.float64)
    assert rst.size == 4 * nx * ny * nz
    for q in range(4):
        a = rst[q * nx * ny * nz:(q + 1) * nx * ny * nz].reshape((nx, ny, nz), order="F")
        assert np.array_equal(a, got[4 + q][1:nx + 1, 1:ny + 1, 1:nz + 1])
        assert np.array_equal(_ymid(got[8 + q])[:, :ny + 2], _ymid(got[4 + q])[:, :ny + 2]), q
    # the stats file of one sample: nyc+1 rows of 35 values, mean u > 0
    lines = open(str(tmp_path / "out.bin.") + "stats%010d" % nsteps).read().split("\n")
    rows = [ln for ln in lines[5:] if ln.strip()]
    assert len(rows) == ny + 1 and all(len(ln) == 35 * 24 for ln in rows)
    assert float(rows[ny // 2][:24]) > 0.5
    # 4. LES (WALE, Simpson filter): nut, one per-routine substep, nsteps steps
    les = dict(case, les_type=2, filter_type=1)
    o3 = O.Oracle(les)
    for fid, a in zip((O.UX, O.UY, O.UZ, O.PR), init):
        o3.field(fid)[...] = a
    o3.call("Clc_SGS_Vis", o3.field(O.UX), o3.field(O.UY), o3.field(O.UZ), o3.field(O.NUT))
    ref = _ymid(o3.field(O.NUT))[:, :ny + 2]
    assert np.abs(_ymid(got[12])[:, :ny + 2] - ref).max() / np.abs(ref).max() < 1e-14
    o3.substep(1)
    for q, fid in enumerate((O.UX, O.UY, O.UZ, O.PR)):
        ref = _ymid(o3.field(fid))
        err = np.abs(_ymid(got[13 + q]) - ref).max() / np.abs(ref).max()
        assert err < 1e-12, ("les substep", q, err)
    o4 = O.Oracle(les)
    for fid, a in zip((O.UX, O.UY, O.UZ, O.PR), init):
        o4.field(fid)[...] = a
    for _ in range(nsteps):
        o4.iterate()
    for q, fid in enumerate((O.UX, O.UY, O.UZ, O.PR)):
        ref = _ymid(o4.field(fid))
        err = np.abs(_ymid(got[17 + q]) - ref).max() / np.abs(ref).max()
        assert err < 1e-12, ("les steps", q, err)
    o.close()
    o2.close()
    o3.close()
    o4.close()
