"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Each test builds the same seeded inputs for both, runs one reference routine
(or a fused substep / step) and compares in the relative max norm,
max|gpu - oracle| / max|oracle|.  Bar: 1e-12 for FP64 (the plain path).
"""
import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu

TOL = 1e-12
CASES = {"small": cases.SMALL, "odd": cases.ODD, "stiff": cases.STIFF,
         "freeslip": cases.FREESLIP, "c1": cases.C1, "tiles": cases.TILES}


def relerr(a, b):
    d = np.abs(a - b).max()
    s = np.abs(b).max()
    return d / s if s > 0 else d


def periodic_fill(a):
    """x then z periodic halo layers as myupdate_halo on one rank leaves them"""
    nx, nz = a.shape[0] - 3, a.shape[2] - 3
    a[0] = a[nx]
    a[nx + 1] = a[1]
    a[nx + 2] = 0.0
    a[:, :, 0] = a[:, :, nz]
    a[:, :, nz + 1] = a[:, :, 1]
    a[:, :, nz + 2] = 0.0
    return a


def interior(a):
    nx, ny, nz = a.shape[0] - 3, a.shape[1] - 3, a.shape[2] - 3
    return a[1:nx + 1, 1:ny + 1, 1:nz + 1]


def ymid(a):
    """halo array restricted to interior x,z, all stored y planes"""
    nx, nz = a.shape[0] - 3, a.shape[2] - 3
    return a[1:nx + 1, :, 1:nz + 1]


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def make_pair(O, case, seed=0, laminar=True):
    from channel3d_amd import Channel, synth
    o = O.Oracle(case)
    g = Channel(case, yp=o.yp)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    if laminar:
        ux, uy, uz = synth.laminar_noise(nx, ny, nz, o.yp, yly=case["yly"],
                                         ubulk=case.get("ubulk", 1.0), seed=seed)
        o.field(O.UX)[...] = ux
        o.field(O.UY)[...] = uy
        o.field(O.UZ)[...] = uz
        o.project()
        for fid, gid in ((O.UX, 0), (O.UY, 1), (O.UZ, 2), (O.PR, 3)):
            g.upload(gid, o.field(fid))
    return o, g


def random_state(O, o, g, seed):
    """random, periodic-consistent u, v, w, p and histories in both"""
    from channel3d_amd import synth
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    for n, (fid, gid) in enumerate(((O.UX, 0), (O.UY, 1), (O.UZ, 2), (O.PR, 3))):
        a = periodic_fill(synth.random_halo(nx, ny, nz, seed, 10 + n))
        o.field(fid)[...] = a
        g.upload(gid, a)
    for n, (fid, gid) in enumerate(((O.HISTX, 4), (O.HISTY, 5), (O.HISTZ, 6))):
        a = synth.random_interior(nx, ny, nz, seed, 20 + n)
        o.field(fid)[...] = a
        g.upload(gid, a)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["small", "odd", "freeslip"])
def test_metrics_identical(O, name):
    case = CASES[name]
    o, g = make_pair(O, case, laminar=False)
    ny = o.nyc
    pairs = [(O.M_YP, 1, ny + 1), (O.M_YC, 0, ny + 1), (O.M_DYP, 0, ny + 1),
             (O.M_DYC, 1, ny + 1), (O.M_RDYP, 0, ny + 1), (O.M_RDYC, 1, ny + 1),
             (O.M_YINTERP, 1, ny + 1), (O.M_AM2PR, 1, ny), (O.M_AC2PR, 1, ny),
             (O.M_AP2PR, 1, ny), (O.M_AM2C, 1, ny), (O.M_AC2C, 1, ny), (O.M_AP2C, 1, ny),
             (O.M_AM2P, 1, ny), (O.M_AC2P, 1, ny), (O.M_AP2P, 1, ny), (O.M_AM2CCN, 1, ny),
             (O.M_AP2CCN, 1, ny)]
    for mid, lo, hi in pairs:
        assert np.array_equal(g.metric(mid, lo, hi), o.metric(mid, lo, hi)), mid
    wx = o.metric(O.M_WAVEX, 1, o.nxc)
    assert relerr(g.metric(O.M_WAVEX, 0, o.nxc // 2), wx[: o.nxc // 2 + 1]) < 1e-15
    g.close()


@pytest.mark.parametrize("name", ["small", "odd", "freeslip", "tiles", "c1"])
@pytest.mark.parametrize("ns", [1, 2, 3])
def test_rhs_components(O, name, ns):
    case = CASES[name]
    o, g = make_pair(O, case, laminar=False)
    random_state(O, o, g, seed=ns)
    o.call("PMcoeUpdate", ns)
    g.PMcoeUpdate(ns)
    from channel3d_amd import lib as L
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    rx, ry = np.zeros((nx, ny, nz), order="F"), np.zeros((nx, ny, nz), order="F")
    rz = np.zeros((nx + 3, ny + 3, nz + 3), order="F")
    o.call("clcRhsX", o.field(O.UX), o.field(O.UY), o.field(O.UZ), rx, o.field(O.HISTX), o.field(O.PR))
    o.call("clcRhsY", o.field(O.UX), o.field(O.UY), o.field(O.UZ), ry, o.field(O.HISTY), o.field(O.PR))
    o.call("clcRhsZ", o.field(O.UX), o.field(O.UY), o.field(O.UZ), rz, o.field(O.HISTZ), o.field(O.PR))
    g.clcRhsX(L.UX, L.UY, L.UZ, L.ARR1, L.HISTX, L.PR)
    g.clcRhsY(L.UX, L.UY, L.UZ, L.ARR2, L.HISTY, L.PR)
    g.clcRhsZ(L.UX, L.UY, L.UZ, L.HALOARR, L.HISTZ, L.PR)
    # same operation order, no contraction, reproducible bulk integral: the
    # right-hand sides and histories reproduce the oracle bit for bit
    assert relerr(g.download(L.ARR1), rx) == 0.0
    assert relerr(g.download(L.ARR2), ry) == 0.0
    assert relerr(interior(g.download(L.HALOARR)), interior(rz)) == 0.0
    for fid, gid in ((O.HISTX, L.HISTX), (O.HISTY, L.HISTY), (O.HISTZ, L.HISTZ)):
        assert relerr(g.download(gid), o.field(fid)) == 0.0
    d = g.diagnostics()
    assert abs(d["PrGradNow"] - o.scalar(O.S_PRGRADNOW)) <= 1e-12 * max(1.0, abs(o.scalar(O.S_PRGRADNOW)))
    g.close()


@pytest.mark.parametrize("name", ["small", "odd", "stiff", "freeslip"])
@pytest.mark.parametrize("ns", [1, 3])
def test_adi_solves(O, name, ns):
    case = CASES[name]
    o, g = make_pair(O, case, laminar=False)
    random_state(O, o, g, seed=40 + ns)
    from channel3d_amd import lib as L, synth
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
    o.call("clcU1Hat", o.field(O.UX), np.asfortranarray(r1.copy()))
    o.call("clcU2Hat", o.field(O.UY), np.asfortranarray(r2.copy()), np.asfortranarray(duy))
    o.call("clcU3Hat", o.field(O.UZ), np.asfortranarray(r3.copy()))
    g.clcU1Hat(L.UX, L.ARR1)
    g.clcU2Hat(L.UY, L.ARR2, L.DUY_YM)
    g.clcU3Hat(L.UZ, L.HALOARR)
    for fid, gid in ((O.UX, L.UX), (O.UY, L.UY), (O.UZ, L.UZ)):
        ref = o.field(fid)
        got = g.download(gid)
        # compare the increments (u is O(1), du is O(1e-2))
        assert relerr(interior(got), interior(ref)) < TOL
    g.close()


@pytest.mark.parametrize("name", ["small", "odd", "c1"])
def test_prsrc(O, name):
    case = CASES[name]
    o, g = make_pair(O, case, laminar=False)
    random_state(O, o, g, seed=7)
    from channel3d_amd import lib as L
    o.call("PMcoeUpdate", 2)
    g.PMcoeUpdate(2)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    src = np.zeros((nx, ny, nz), order="F")
    import ctypes
    dm = ctypes.c_double()
    o.call("clcPrSrc", o.field(O.UX), o.field(O.UY), o.field(O.UZ), src, o.field(O.PR), ctypes.byref(dm))
    gd = g.clcPrSrc(L.UX, L.UY, L.UZ, L.ARR1, L.PR)
    assert relerr(g.download(L.ARR1), src) < TOL
    assert relerr(interior(g.download(L.PR)), interior(o.field(O.PR))) < TOL
    assert abs(gd - dm.value) <= 1e-12 * dm.value
    g.close()


@pytest.mark.parametrize("name", ["small", "odd", "stiff", "freeslip", "c1"])
def test_ppe(O, name):
    case = CASES[name]
    o, g = make_pair(O, case, laminar=False)
    from channel3d_amd import lib as L, synth
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    dyp = o.metric(O.M_DYP, 1, ny)
    f = synth.ppe_rhs(nx, ny, nz, seed=3)
    w = np.broadcast_to(dyp[None, :, None], f.shape)
    f = np.asfortranarray(f - (f * w).sum() / w.sum())
    phi = np.zeros((nx + 3, ny + 3, nz + 3), order="F")
    o.call("clcPPE", np.asfortranarray(f.copy()), phi)
    g.upload(L.ARR1, f)
    g.clcPPE(L.ARR1, L.HALOARR)
    got = interior(g.download(L.HALOARR))
    assert relerr(got, interior(phi)) < TOL
    g.close()


@pytest.mark.parametrize("name", ["small", "odd", "c1"])
def test_tridiag_lane_forms_bitwise(O, name, monkeypatch):
    """the checkpointed PPE sweep with one lane per wavenumber pair and with
    the real and imaginary parts on two lanes: bit for bit the same, and at
    the oracle within the bar"""
    from channel3d_amd import Channel, lib as L
    from channel3d_amd import synth
    case = CASES[name]
    o, g0 = make_pair(O, case, laminar=False)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    dyp = o.metric(O.M_DYP, 1, ny)
    src = synth.ppe_rhs(nx, ny, nz, seed=7)
    w = np.broadcast_to(dyp[None, :, None], src.shape)
    src = np.asfortranarray(src - (src * w).sum() / w.sum())
    res = []
    for lanes in ("1", "2"):
        monkeypatch.setenv("C3D_TRIDIAG_LANES", lanes)
        g = Channel(case, yp=o.yp)
        monkeypatch.delenv("C3D_TRIDIAG_LANES")
        g.upload(L.ARR1, src)
        g.ppe(L.ARR1, L.HALOARR)
        res.append(interior(g.download(L.HALOARR)))
        g.close()
    assert np.array_equal(res[0], res[1])
    phi = np.zeros_like(o.field(O.HALO))
    o.call("clcPPE", np.asfortranarray(src.copy()), phi)
    assert relerr(res[0], interior(phi)) < TOL
    g0.close()


def test_ppe_twenty_seeded_sources(O):
    """config 1's standalone solves: 20 sources ~ N(0,1), seeds 0..19, with
    the dyp-weighted mean removed, at 64^3, each against the oracle's clcPPE"""
    from channel3d_amd import lib as L, synth
    o, g = make_pair(O, cases.C1, laminar=False)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    dyp = o.metric(O.M_DYP, 1, ny)
    w = np.broadcast_to(dyp[None, :, None], (nx, ny, nz))
    worst = 0.0
    for seed in range(20):
        f = synth.ppe_rhs(nx, ny, nz, seed=seed)
        f = np.asfortranarray(f - (f * w).sum() / w.sum())
        phi = np.zeros((nx + 3, ny + 3, nz + 3), order="F")
        o.call("clcPPE", np.asfortranarray(f.copy()), phi)
        g.upload(L.ARR1, f)
        g.clcPPE(L.ARR1, L.HALOARR)
        worst = max(worst, relerr(interior(g.download(L.HALOARR)), interior(phi)))
    print(f"20 PPE solves at 64^3: worst rel err {worst:.2e}", flush=True)
    assert worst < TOL
    g.close()


def test_ppe_single_mode_and_gauge(O):
    # a source made of one (kx,kz) mode only: both Re/Im parts, kx = nx/2 Nyquist
    case = cases.SMALL
    o, g = make_pair(O, case, laminar=False)
    from channel3d_amd import lib as L
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    x = np.arange(nx)[:, None, None]
    z = np.arange(nz)[None, None, :]
    yv = np.linspace(-1, 1, ny)[None, :, None]
    for m, q in [(0, 0), (1, 0), (0, 3), (nx // 2, 0), (nx // 2, nz // 2), (5, 7)]:
        f = np.cos(2 * np.pi * m * x / nx + 0.3) * np.cos(2 * np.pi * q * z / nz + 0.1) * (yv ** 2 + 0.5)
        f = np.asfortranarray(np.broadcast_to(f, (nx, ny, nz)))
        dyp = o.metric(O.M_DYP, 1, ny)
        w = np.broadcast_to(dyp[None, :, None], f.shape)
        f = np.asfortranarray(f - (f * w).sum() / w.sum())
        phi = np.zeros((nx + 3, ny + 3, nz + 3), order="F")
        o.call("clcPPE", np.asfortranarray(f.copy()), phi)
        g.upload(L.ARR1, f)
        g.clcPPE(L.ARR1, L.HALOARR)
        got = interior(g.download(L.HALOARR))
        assert relerr(got, interior(phi)) < TOL, (m, q)
        assert abs(got[:, 0, :].mean()) <= 1e-12 * np.abs(got).max()
    g.close()


@pytest.mark.parametrize("name", ["small", "odd"])
def test_projection_routines(O, name):
    case = CASES[name]
    o, g = make_pair(O, case, laminar=False)
    random_state(O, o, g, seed=9)
    from channel3d_amd import lib as L, synth
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    o.call("PMcoeUpdate", 1)
    g.PMcoeUpdate(1)
    phi = periodic_fill(synth.random_halo(nx, ny, nz, 9, 50, 1e-3))
    g.upload(L.HALOARR, phi)
    o.call("SetBC_and_UpdateHalo_pr", phi)
    g.SetBC_and_UpdateHalo_pr(L.HALOARR)
    assert relerr(ymid(g.download(L.HALOARR)), ymid(phi)) == 0.0
    o.call("FluidVelUpdate", phi, o.field(O.UX), o.field(O.UY), o.field(O.UZ))
    o.call("PressureUpdate", o.field(O.PR), phi)
    g.FluidVelUpdate(L.HALOARR, L.UX, L.UY, L.UZ)
    g.PressureUpdate(L.PR, L.HALOARR)
    for fid, gid in ((O.UX, L.UX), (O.UY, L.UY), (O.UZ, L.UZ), (O.PR, L.PR)):
        assert relerr(interior(g.download(gid)), interior(o.field(fid))) < TOL
    g.close()


@pytest.mark.parametrize("name", ["small", "freeslip"])
def test_boundary_ghosts(O, name):
    case = dict(CASES[name], uxbc=(0, 0, 0.1, -0.2, 0, 0), uybc=(0, 0, 0.0, 0.05, 0, 0),
                uzbc=(0, 0, 0.3, 0.4, 0, 0))
    o, g = make_pair(O, case, laminar=False)
    random_state(O, o, g, seed=11)
    from channel3d_amd import lib as L, synth
    nx, nz = o.nxc, o.nzc
    uyym = synth.uniform_pm1(11, 60, (nx, nz)) * 0.01
    o.field(O.UY_YM)[...] = uyym
    g.upload(L.UY_YM, uyym)
    o.call("SetBC_and_UpdateHaloForPrSrc", o.field(O.UX), o.field(O.UY), o.field(O.UZ), o.field(O.UY_YM))
    g.SetBC_and_UpdateHaloForPrSrc(L.UX, L.UY, L.UZ, L.UY_YM)
    assert np.array_equal(ymid(g.download(L.UY)), ymid(o.field(O.UY)))
    o.call("SetBC_and_UpdateHalo", o.field(O.UX), o.field(O.UY), o.field(O.UZ), o.field(O.UY_YM))
    g.SetBC_and_UpdateHalo(L.UX, L.UY, L.UZ, L.UY_YM)
    # no-slip ghosts (2 ubc - u) are exact; free-slip ones (u +- ubc dyp) may
    # differ by the rounding of a fused multiply-add
    for fid, gid in ((O.UX, L.UX), (O.UY, L.UY), (O.UZ, L.UZ)):
        e = relerr(ymid(g.download(gid)), ymid(o.field(fid)))
        assert e <= 2.3e-16, f"field {fid}: rel err {e:.3e}"
    g.close()


@pytest.mark.parametrize("name", ["small", "c1"])
def test_diagnostics(O, name):
    case = CASES[name]
    o, g = make_pair(O, case, seed=2)
    ux, uy, uz = o.field(O.UX), o.field(O.UY), o.field(O.UZ)
    assert relerr(np.array([g.CalcMaxCFL()]), np.array([o.call("CalcMaxCFL", ux, uy, uz)])) < 1e-14
    assert abs(g.CheckDivergence() - o.call("CheckDivergence", ux, uy, uz)) < 1e-14
    v = np.zeros(3)
    o.call("CalcVmax", ux, uy, uz, v)
    assert relerr(np.array(g.CalcVmax()), v) < 1e-15
    assert abs(g.CalcUxAver() - o.call("CalcUxAver", ux)) < 1e-13
    g.close()


def compare_state(O, o, g, tol):
    from channel3d_amd import lib as L
    errs = {}
    for fid, gid, nm in ((O.UX, L.UX, "ux"), (O.UY, L.UY, "uy"), (O.UZ, L.UZ, "uz"),
                         (O.PR, L.PR, "p")):
        errs[nm] = relerr(ymid(g.download(gid)), ymid(o.field(fid)))
    for fid, gid, nm in ((O.HISTX, L.HISTX, "hx"), (O.HISTY, L.HISTY, "hy"),
                         (O.HISTZ, L.HISTZ, "hz")):
        errs[nm] = relerr(g.download(gid), o.field(fid))
    bad = {k: v for k, v in errs.items() if not v < tol}
    assert not bad, errs
    return errs


@pytest.mark.parametrize("name", ["small", "odd", "stiff", "freeslip", "c1"])
def test_fused_substeps(O, name):
    case = CASES[name]
    o, g = make_pair(O, case, seed=1)
    for ns in (1, 2, 3):
        o.substep(ns)
        g.substep(ns)
        compare_state(O, o, g, TOL)
    g.close()


@pytest.mark.parametrize("name", ["small", "odd", "c1"])
def test_ysweep_workgroup_forms_bitwise(O, name, monkeypatch):
    """the checkpointed wall-normal sweeps of clcU1Hat..clcU3Hat with 64- and
    256-line workgroups: bit for bit the same substeps, and at the oracle
    within the bar"""
    from channel3d_amd import lib as L
    case = CASES[name]
    res = []
    for nth in ("64", "256"):
        monkeypatch.setenv("C3D_YSOLVE_THREADS", nth)
        o, g = make_pair(O, case, seed=1)
        monkeypatch.delenv("C3D_YSOLVE_THREADS")
        for ns in (1, 2, 3):
            o.substep(ns)
            g.substep(ns)
        compare_state(O, o, g, TOL)
        res.append([g.download(gid) for gid in (L.UX, L.UY, L.UZ, L.PR)])
        g.close()
    for a, b in zip(*res):
        assert np.array_equal(ymid(a), ymid(b))


@pytest.mark.parametrize("name", ["small", "odd", "c1"])
def test_fftz_tile_forms_bitwise(O, name, monkeypatch):
    """the PPE's z transforms with 8- and 16-line tiles (16 is the default
    for FP64 lengths from 256 up, forced here on the small grids): bit for
    bit the same substeps, and at the oracle within the bar"""
    from channel3d_amd import lib as L
    case = CASES[name]
    res = []
    for tile in ("8", "16"):
        monkeypatch.setenv("C3D_FFTZ_TILE", tile)
        o, g = make_pair(O, case, seed=2)
        monkeypatch.delenv("C3D_FFTZ_TILE")
        for ns in (1, 2, 3):
            o.substep(ns)
            g.substep(ns)
        compare_state(O, o, g, TOL)
        res.append([g.download(gid) for gid in (L.UX, L.UY, L.UZ, L.PR)])
        g.close()
    for a, b in zip(*res):
        assert np.array_equal(ymid(a), ymid(b))


@pytest.mark.parametrize("form", ["fp32_context", "fp32_refine1"])
def test_fftz_tile_forms_bitwise_fp32(O, form, monkeypatch):
    """the FP32 z transforms (an FP32 context's, and the FP32 passes of the
    FP64 context's refined pressure solve) with 8- and 16-line tiles: bit
    for bit the same pressure solution"""
    from channel3d_amd import Channel, lib as L
    from channel3d_amd import synth
    case = cases.C1
    nx, ny, nz = case["nxc"], case["nyc"], case["nzc"]
    src = np.asfortranarray(synth.ppe_rhs(nx, ny, nz, seed=11))
    res = []
    for tile in ("8", "16"):
        monkeypatch.setenv("C3D_FFTZ_TILE", tile)
        if form == "fp32_context":
            g = Channel(case, precision=L.FP32)
        else:
            g = Channel(case)
            g.set_ppe_refine(1)
        monkeypatch.delenv("C3D_FFTZ_TILE")
        g.upload(L.ARR1, src)
        g.ppe(L.ARR1, L.HALOARR)
        res.append(interior(g.download(L.HALOARR)))
        g.close()
    assert np.all(np.isfinite(res[0]))
    assert np.array_equal(res[0], res[1])


def test_per_routine_sequence_matches_fused(O):
    """ChannelIterate's call sequence through the per-routine entries gives the
    fused substep's result"""
    from channel3d_amd import Channel, lib as L
    case = cases.SMALL
    o, g = make_pair(O, case, seed=5)
    h = Channel(case, yp=o.yp)
    for gid in (L.UX, L.UY, L.UZ, L.PR):
        h.upload(gid, g.download(gid))
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
            assert relerr(ymid(h.download(gid)), ymid(g.download(gid))) < 1e-14
    g.close()
    h.close()


# c1 runs config 1's whole length: 100 RK3 steps of dt 0.005 (the reference's
# own x1 and z1 PPE routes differ by 7.5e-13 in p and 1e-14 in u after it)
@pytest.mark.parametrize("name,nsteps", [("c1", 100), ("odd", 10)])
def test_multistep_parity(O, name, nsteps):
    case = CASES[name]
    o, g = make_pair(O, case, seed=4)
    for n in range(nsteps):
        o.iterate()
        g.step()
    errs = compare_state(O, o, g, TOL)
    d = g.diagnostics()
    assert d["divmax2"] < 1e-12
    assert abs(d["divmax2"] - o.scalar(O.S_DIVMAX2)) < 1e-13
    assert relerr(np.array([d["cfl"]]), np.array([o.scalar(O.S_CFL)])) < 1e-12
    assert relerr(np.array([d["vmax_x"], d["vmax_y"], d["vmax_z"]]),
                  np.array([o.scalar(O.S_VMAX1), o.scalar(O.S_VMAX2), o.scalar(O.S_VMAX3)])) < TOL
    assert abs(d["PrGradAver"] - o.scalar(O.S_PRGRADAVER)) < 1e-10
    assert abs(d["SimTime"] - o.scalar(O.S_SIMTIME)) < 1e-13
    g.close()


def test_cfl_follows_uploaded_state(O):
    """a state uploaded between steps replaces the one the last step's closing
    pass computed the next CFL number from: the next step reports the CFL of
    the state it actually starts from (m_ChannelSystem.f90:80-81)"""
    from channel3d_amd import lib as L
    case = cases.SMALL
    o, g = make_pair(O, case, seed=1)
    g.step()
    g.step()
    o2, g2 = make_pair(O, case, seed=7)
    for gid in (L.UX, L.UY, L.UZ, L.PR):
        g.upload(gid, g2.download(gid))
    g2.close()
    o2.iterate()
    g.step()
    d = g.diagnostics()
    assert relerr(np.array([d["cfl"]]), np.array([o2.scalar(O.S_CFL)])) < 1e-12
    g.close()


@pytest.mark.parametrize("name", ["small", "odd"])
def test_nontemporal_variants_bitwise(O, name, monkeypatch):
    """the kernel variants with non-temporal loads and stores (chosen for
    fields of 64 MiB and more) forced on a small grid: bit for bit the cached
    variants, and both at the oracle within the bar"""
    from channel3d_amd import Channel, lib as L
    case = CASES[name]
    o, g0 = make_pair(O, case, seed=3)
    monkeypatch.setenv("C3D_NT", "1")
    g1 = Channel(case, yp=o.yp)
    monkeypatch.setenv("C3D_NT", "0")
    for gid in (L.UX, L.UY, L.UZ, L.PR):
        g1.upload(gid, g0.download(gid))
    src = np.random.default_rng(5).standard_normal((o.nxc, o.nyc, o.nzc))
    src = np.asfortranarray(src - src.mean())
    phis = []
    for g in (g0, g1):
        g.upload(L.ARR1, src)
        g.clcPPE(L.ARR1, L.HALOARR)
        phis.append(g.download(L.HALOARR))
    assert np.array_equal(phis[0], phis[1])
    for _ in range(2):
        o.iterate()
        g0.step()
        g1.step()
    for gid in (L.UX, L.UY, L.UZ, L.PR, L.HISTX, L.HISTY, L.HISTZ):
        assert np.array_equal(g0.download(gid), g1.download(gid)), gid
    compare_state(O, o, g1, TOL)
    assert g0.diagnostics() == g1.diagnostics()
    g0.close()
    g1.close()


@pytest.mark.parametrize("ischeme", [1, 2], ids=["ab2", "rk2"])
def test_ab2_rk2_steps(O, ischeme):
    """the AB2 and RK2 tables (ischeme 1, 2; m_Parameters.f90:135-142):
    free-running steps against the oracle with the same scheme"""
    case = dict(cases.SMALL, ischeme=ischeme)
    o, g = make_pair(O, case, seed=6)
    for _ in range(4):
        o.iterate()
        g.step()
    compare_state(O, o, g, TOL)
    d = g.diagnostics()
    assert d["divmax2"] < 1e-12
    assert abs(d["SimTime"] - o.scalar(O.S_SIMTIME)) < 1e-13
    g.close()


def test_store_y_solver_bitwise(O, monkeypatch):
    """C3D_Y_SOLVER=store (the two-pass Thomas sweeps storing the forward
    pass, also the fallback for lines too long for the checkpoints) runs the
    same operations as the default checkpointed sweeps: equal bit for bit"""
    from channel3d_amd import Channel, lib as L
    case = cases.ODD
    o, g0 = make_pair(O, case, seed=8)
    monkeypatch.setenv("C3D_Y_SOLVER", "store")
    g1 = Channel(case, yp=o.yp)
    monkeypatch.delenv("C3D_Y_SOLVER")
    for gid in (L.UX, L.UY, L.UZ, L.PR):
        g1.upload(gid, g0.download(gid))
    for _ in range(2):
        g0.step()
        g1.step()
    for gid in (L.UX, L.UY, L.UZ, L.PR, L.HISTX, L.HISTY, L.HISTZ):
        assert np.array_equal(g0.download(gid), g1.download(gid)), gid
    g0.close()
    g1.close()
