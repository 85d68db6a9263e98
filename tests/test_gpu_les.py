"""Parity of the HIP LES layer (WALE / MTS eddy viscosity and the partially
implicit LES scheme, src/CFD_2nd/m_LESModels.f90 and the *_PIMP_LES
routines of m_TSchemePIMP_inc.f90) with the CPU oracle, through the C ABI.

The stencils, solves and pressure update use the reference's operation order
without contraction, so given the same nut they reproduce the oracle bit for
bit; the MTS viscosity uses only correctly rounded operations (sqrt and
division) and is bitwise too.  WALE evaluates x^1.5, x^2.5 and x^1.25 with
the device's pow, which may differ from the host's in the last bit.  Whole
substeps and steps are held to the 1e-12 relative max-norm bar of the plain
path.
"""
import numpy as np
import pytest

from tests import cases
from tests.test_gpu_parity import (TOL, interior, make_pair, periodic_fill, random_state,
                                   relerr, ymid)

pytestmark = pytest.mark.gpu

LES = {
    "wale": dict(cases.SMALL, les_type=2, filter_type=1),
    "mts_trap": dict(cases.SMALL, les_type=3, filter_type=0),
    "mts_simpson": dict(cases.ODD, les_type=3, filter_type=1),
    "wale_freeslip": dict(cases.FREESLIP, les_type=2, filter_type=1),
}


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def _nut_pair(O, o, g):
    """nut from both sides' Clc_SGS_Vis on the current velocity"""
    from channel3d_amd import lib as L
    o.call("Clc_SGS_Vis", o.field(O.UX), o.field(O.UY), o.field(O.UZ), o.field(O.NUT))
    g.Clc_SGS_Vis(L.UX, L.UY, L.UZ, L.NUT)
    return g.download(L.NUT), o.field(O.NUT)


def _share_nut(O, o, g):
    """the oracle's nut in the device context, so routines see identical nut"""
    from channel3d_amd import lib as L
    o.call("Clc_SGS_Vis", o.field(O.UX), o.field(O.UY), o.field(O.UZ), o.field(O.NUT))
    g.upload(L.NUT, o.field(O.NUT))


def test_delta_cell_identical(O):
    from channel3d_amd import lib as L
    o, g = make_pair(O, LES["wale"], laminar=False)
    ny = o.nyc
    assert np.array_equal(g.metric(L.M_DELTA, 0, ny + 1), o.metric(O.M_DELTA, 0, ny + 1))
    g.close()


@pytest.mark.parametrize("name", list(LES))
def test_sgs_viscosity(O, name):
    o, g = make_pair(O, LES[name], seed=3)
    got, ref = _nut_pair(O, o, g)
    # interior x, z; every stored y plane the reference sets (0 .. nyc+1)
    sl = (slice(1, o.nxc + 1), slice(0, o.nyc + 2), slice(1, o.nzc + 1))
    err = relerr(got[sl], ref[sl])
    if LES[name]["les_type"] == 3:
        assert err == 0.0, err
    else:
        assert err < 1e-14, err
    # the model's nut is at least the molecular viscosity in every cell (the
    # no-slip ghosts 2 nu - nut(1) may be negative)
    assert ref[1:o.nxc + 1, 1:o.nyc + 1, 1:o.nzc + 1].min() >= LES[name]["xnu"]
    g.close()


@pytest.mark.parametrize("name", ["wale", "mts_simpson", "wale_freeslip"])
@pytest.mark.parametrize("ns", [1, 2, 3])
def test_les_rhs_components(O, name, ns):
    from channel3d_amd import lib as L
    o, g = make_pair(O, LES[name], laminar=False)
    random_state(O, o, g, seed=ns)
    _share_nut(O, o, g)
    o.call("PMcoeUpdate", ns)
    g.PMcoeUpdate(ns)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    nut = o.field(O.NUT)
    rx, ry = np.zeros((nx, ny, nz), order="F"), np.zeros((nx, ny, nz), order="F")
    rz = np.zeros((nx + 3, ny + 3, nz + 3), order="F")
    u = (o.field(O.UX), o.field(O.UY), o.field(O.UZ))
    o.call("clcRhsX_les", *u, rx, o.field(O.HISTX), o.field(O.PR), nut)
    o.call("clcRhsY_les", *u, ry, o.field(O.HISTY), o.field(O.PR), nut)
    o.call("clcRhsZ_les", *u, rz, o.field(O.HISTZ), o.field(O.PR), nut)
    g.clcRhsX(L.UX, L.UY, L.UZ, L.ARR1, L.HISTX, L.PR)
    g.clcRhsY(L.UX, L.UY, L.UZ, L.ARR2, L.HISTY, L.PR)
    g.clcRhsZ(L.UX, L.UY, L.UZ, L.HALOARR, L.HISTZ, L.PR)
    assert relerr(g.download(L.ARR1), rx) == 0.0
    assert relerr(g.download(L.ARR2), ry) == 0.0
    assert relerr(interior(g.download(L.HALOARR)), interior(rz)) == 0.0
    for fid, gid in ((O.HISTX, L.HISTX), (O.HISTY, L.HISTY), (O.HISTZ, L.HISTZ)):
        assert relerr(g.download(gid), o.field(fid)) == 0.0
    g.close()


@pytest.mark.parametrize("name", ["wale", "mts_simpson", "wale_freeslip"])
@pytest.mark.parametrize("ns", [1, 3])
def test_les_solves(O, name, ns):
    from channel3d_amd import lib as L, synth
    o, g = make_pair(O, LES[name], laminar=False)
    random_state(O, o, g, seed=40 + ns)
    _share_nut(O, o, g)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    o.call("PMcoeUpdate", ns)
    g.PMcoeUpdate(ns)
    r1 = synth.random_interior(nx, ny, nz, 5, 1, 0.01)
    r2 = synth.random_interior(nx, ny, nz, 5, 2, 0.01)
    r3 = periodic_fill(synth.random_halo(nx, ny, nz, 5, 3, 0.01))
    duy = synth.uniform_pm1(5, 4, (nx, nz)) * 1e-3
    g.upload(L.ARR1, r1)
    g.upload(L.ARR2, r2)
    g.upload(L.HALOARR, r3)
    g.upload(L.DUY_YM, duy)
    nut = o.field(O.NUT)
    o.call("clcU1Hat_les", o.field(O.UX), np.asfortranarray(r1.copy()), nut)
    o.call("clcU2Hat_les", o.field(O.UY), np.asfortranarray(r2.copy()), np.asfortranarray(duy), nut)
    o.call("clcU3Hat_les", o.field(O.UZ), np.asfortranarray(r3.copy()), nut)
    g.clcU1Hat(L.UX, L.ARR1)
    g.clcU2Hat(L.UY, L.ARR2, L.DUY_YM)
    g.clcU3Hat(L.UZ, L.HALOARR)
    for fid, gid in ((O.UX, L.UX), (O.UY, L.UY), (O.UZ, L.UZ)):
        assert relerr(interior(g.download(gid)), interior(o.field(fid))) == 0.0
    g.close()


@pytest.mark.parametrize("name", ["wale", "mts_simpson"])
def test_les_prsrc_and_pressure_update(O, name):
    from channel3d_amd import lib as L, synth
    o, g = make_pair(O, LES[name], laminar=False)
    random_state(O, o, g, seed=9)
    _share_nut(O, o, g)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    o.call("PMcoeUpdate", 2)
    g.PMcoeUpdate(2)
    prsrc = np.zeros((nx, ny, nz), order="F")
    dm = np.zeros(1)
    pr0 = o.field(O.PR).copy()
    o.call("clcPrSrcOther", o.field(O.UX), o.field(O.UY), o.field(O.UZ), prsrc, dm)
    dmg = g.clcPrSrc(L.UX, L.UY, L.UZ, L.ARR1, L.PR)
    assert relerr(g.download(L.ARR1), prsrc) == 0.0
    assert dmg == dm[0]
    # clcPrSrcOther leaves the pressure alone
    assert np.array_equal(interior(g.download(L.PR)), interior(pr0))
    phi = periodic_fill(synth.random_halo(nx, ny, nz, 9, 7))
    o.call("SetBC_and_UpdateHalo_pr", phi)
    g.upload(L.HALOARR, phi)
    o.call("PressureUpdate_les", o.field(O.PR), phi, o.field(O.NUT))
    g.PressureUpdate(L.PR, L.HALOARR)
    assert relerr(interior(g.download(L.PR)), interior(o.field(O.PR))) == 0.0
    g.close()


@pytest.mark.parametrize("name", list(LES))
def test_les_fused_substeps(O, name):
    o, g = make_pair(O, LES[name], seed=1)
    _nut_pair(O, o, g)
    from tests.test_gpu_parity import compare_state
    for ns in (1, 2, 3):
        o.substep(ns)
        g.substep(ns)
        compare_state(O, o, g, TOL)
    g.close()


def test_les_per_routine_sequence_matches_fused(O):
    """ChannelIterate's LES call sequence through the per-routine entries
    gives the fused substeps' result"""
    from channel3d_amd import Channel, lib as L
    case = LES["mts_trap"]
    o, g = make_pair(O, case, seed=5)
    h = Channel(case, yp=o.yp)
    for gid in (L.UX, L.UY, L.UZ, L.PR):
        h.upload(gid, g.download(gid))
    g.Clc_SGS_Vis()
    h.Clc_SGS_Vis(L.UX, L.UY, L.UZ, L.NUT)
    for ns in (1, 2, 3):
        g.substep(ns)
        h.PMcoeUpdate(ns)
        h.Update_uy_ym(L.UY_YM, L.DUY_YM)
        h.clcRhsX(L.UX, L.UY, L.UZ, L.ARR1, L.HISTX, L.PR)
        h.clcRhsY(L.UX, L.UY, L.UZ, L.ARR2, L.HISTY, L.PR)
        h.clcRhsZ(L.UX, L.UY, L.UZ, L.HALOARR, L.HISTZ, L.PR)
        h.clcU1Hat(L.UX, L.ARR1)
        h.clcU2Hat(L.UY, L.ARR2, L.DUY_YM)
        h.clcU3Hat(L.UZ, L.HALOARR)
        h.SetBC_and_UpdateHaloForPrSrc(L.UX, L.UY, L.UZ, L.UY_YM)
        h.clcPrSrc(L.UX, L.UY, L.UZ, L.ARR1, L.PR)
        h.clcPPE(L.ARR1, L.HALOARR)
        h.SetBC_and_UpdateHalo_pr(L.HALOARR)
        h.FluidVelUpdate(L.HALOARR, L.UX, L.UY, L.UZ)
        h.PressureUpdate(L.PR, L.HALOARR)
        h.SetBC_and_UpdateHalo(L.UX, L.UY, L.UZ, L.UY_YM)
        h.SetBC_and_UpdateHalo_pr(L.PR)
        for gid in (L.UX, L.UY, L.UZ, L.PR):
            assert relerr(ymid(h.download(gid)), ymid(g.download(gid))) == 0.0
    g.close()
    h.close()


@pytest.mark.parametrize("name,nsteps", [("wale", 10), ("mts_simpson", 10)])
def test_les_multistep_parity(O, name, nsteps):
    from channel3d_amd import lib as L
    from tests.test_gpu_parity import compare_state
    o, g = make_pair(O, LES[name], seed=4)
    for n in range(nsteps):
        o.iterate()
        g.step()
    compare_state(O, o, g, TOL)
    sl = (slice(1, o.nxc + 1), slice(0, o.nyc + 2), slice(1, o.nzc + 1))
    assert relerr(g.download(L.NUT)[sl], o.field(O.NUT)[sl]) < TOL
    d = g.diagnostics()
    assert d["divmax2"] < 1e-12
    assert abs(d["PrGradAver"] - o.scalar(O.S_PRGRADAVER)) < 1e-10
    g.close()


def test_set_les_rejects(O):
    from channel3d_amd import Channel, lib as L
    g = Channel(cases.SMALL)
    for lt, ft, msg in ((1, 1, "empty case"), (4, 1, "LES_type"), (2, 2, "FilterType")):
        with pytest.raises(L.C3DError, match=msg):
            g.set_les(lt, ft)
    with pytest.raises(L.C3DError, match="LES is off"):
        g.Clc_SGS_Vis(L.UX, L.UY, L.UZ, L.UX)
    with pytest.raises(L.C3DError, match="nut exists only"):
        import ctypes
        v = ctypes.c_void_p()
        L.check(g.lib.c3d_state_field(g._h, L.NUT, ctypes.byref(v)), "c3d_state_field")
    g.close()
    g4 = Channel(cases.SMALL4)
    with pytest.raises(L.C3DError, match="2nd-order"):
        g4.set_les(2, 1)
    g4.close()


def test_les_off_restores_the_full_implicit_scheme(O):
    """switching LES off returns every entry to the FIMP routines"""
    from tests.test_gpu_parity import compare_state
    case = dict(cases.SMALL)
    o, g = make_pair(O, case, seed=2)
    g.set_les(2, 1)
    g.set_les(0, 1)
    for ns in (1, 2, 3):
        o.substep(ns)
        g.substep(ns)
    compare_state(O, o, g, TOL)
    g.close()
