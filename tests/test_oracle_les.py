"""Pins of the oracle's LES layer (src/CFD_2nd/m_LESModels.f90 and the
*_PIMP_LES routines of m_TSchemePIMP_inc.f90) against what the mathematics
fixes.  The reference ships no vectors for these routines either, so:

* WALE on a linear velocity field must equal the textbook WALE formula of
  Nicoud & Ducros (1999) evaluated on the constant gradient tensor (the
  staggered differences are exact on linear fields), and vanish in pure shear;
* MTS on quadratic fields must equal the closed form built from the second
  moment of the 3x3 test filter (trapezoidal: dx^2/2, Simpson: dx^2/3) and
  the exact strain rate at the cell centre; uniform flow gives nut = nu;
* the LES right-hand sides with a constant eddy viscosity on a divergence-free
  field must coincide with the independently restated full-implicit ones at
  the first substep (where gamma = alpha), because the cross terms of the
  stress divergence are nu * grad(div u) = 0;
* with a variable viscosity they must converge at second order to
  -div(u u_i) + div(nu (grad u_i + d_i u));
* each implicit y-solve must invert 1 - betaT * (the wall-normal part of the
  explicit operator), the latter extracted by probing the right-hand side;
* the pressure update must converge to phi - betaT d/dy(nu dphi/dy).
"""
import math

import numpy as np
import pytest

from tests import cases
from tests.test_oracle_pins import _coords, _fill, _lamb, _uniform_yp

sympy = pytest.importorskip("sympy")


def _case(**kw):
    c = dict(cases.SMALL, is_ux_const=False, les_type=2, filter_type=1)
    c.update(kw)
    return c


def _halo(n3, val=0.0):
    return np.full(n3, val, order="F")


# ------------------------------------------------------------------ WALE --
def _wale_textbook(g, delta, nu, Cw=0.325):
    S = 0.5 * (g + g.T)
    g2 = g @ g
    Sd = 0.5 * (g2 + g2.T) - np.eye(3) * np.trace(g2) / 3.0
    SS, SdSd = (S * S).sum(), (Sd * Sd).sum()
    return (Cw * delta) ** 2 * SdSd ** 1.5 / (SS ** 2.5 + SdSd ** 1.25) + nu


def _linear_fields(o, c, g, nx, ny, nz):
    """u_i = sum_j g_ij x_j at every staggered halo position"""
    def lin(i, xs, ys, zs):
        return _fill(lambda x, y, z: g[i, 0] * x + g[i, 1] * y + g[i, 2] * z + 0.1 * (i + 1),
                     xs, ys, zs)
    ux = lin(0, c["xf"], c["yc"], c["zc"])
    uy = lin(1, c["xc"], c["yn"], c["zc"])
    uz = lin(2, c["xc"], c["yc"], c["zf"])
    return ux, uy, uz


@pytest.mark.parametrize("stretched", [False, True])
def test_wale_equals_textbook_on_linear_fields(oracle_mod, stretched):
    O = oracle_mod
    case = _case(nxc=8, nyc=10, nzc=6)
    o = O.Oracle(case, yp=None if stretched else _uniform_yp(10, 2.0))
    c = _coords(o, O, case)
    rng = np.random.default_rng(3)
    g = rng.normal(size=(3, 3))
    ux, uy, uz = _linear_fields(o, c, g, 8, 10, 6)
    nut = _halo(ux.shape)
    o.call("Clc_SGS_WALE", ux, uy, uz, nut)
    delta = o.metric(O.M_DELTA, 0, 11)
    for j in range(1, 11):
        want = _wale_textbook(g, delta[j], case["xnu"])
        got = nut[1:9, j, 1:7]
        assert np.abs(got - want).max() <= 1e-12 * abs(want), (j, got.max(), want)
    # DeltaCell = (dy dx dz)^(1/3)
    dyp = o.metric(O.M_DYP, 0, 11)
    dx, dz = case["xlx"] / 8, case["zlz"] / 6
    assert np.allclose(delta, (dyp * dx * dz) ** (1.0 / 3.0), rtol=1e-15, atol=0)
    o.close()


def test_wale_vanishes_in_pure_shear_not_in_rotation(oracle_mod):
    O = oracle_mod
    case = _case(nxc=8, nyc=10, nzc=6)
    o = O.Oracle(case)
    c = _coords(o, O, case)
    g = np.zeros((3, 3))
    g[0, 1] = 2.5                       # u = S y: the wall-adapting property
    nut = _halo((11, 13, 9))
    o.call("Clc_SGS_WALE", *_linear_fields(o, c, g, 8, 10, 6), nut)
    assert np.all(nut[1:9, 1:11, 1:7] == case["xnu"])
    g = np.zeros((3, 3))
    g[0, 2], g[2, 0] = 1.5, -1.5       # solid-body rotation about y: S = 0
    o.call("Clc_SGS_WALE", *_linear_fields(o, c, g, 8, 10, 6), nut)
    delta = o.metric(O.M_DELTA, 0, 11)
    # S = 0 leaves CwDelta^2 * (SdSd)^(1/4); for this rotation SdSd = (2/3) w^4
    want = (0.325 * delta[1:11]) ** 2 * ((2.0 / 3.0) * 1.5 ** 4) ** 0.25 + case["xnu"]
    assert np.allclose(nut[1:9, 1:11, 1:7], want[None, :, None], rtol=1e-13, atol=0)
    o.close()


# ------------------------------------------------------------------- MTS --
@pytest.mark.parametrize("filt", [0, 1])
def test_mts_closed_form_on_quadratic_fields(oracle_mod, filt):
    O = oracle_mod
    nx, ny, nz = 8, 10, 6
    case = _case(nxc=nx, nyc=ny, nzc=nz, les_type=3, filter_type=filt)
    o = O.Oracle(case)
    c = _coords(o, O, case)
    rng = np.random.default_rng(11)
    A, B = rng.normal(size=3), rng.normal(size=3)
    g = rng.normal(size=(3, 3))

    a = (0.0, 0.7, 0.0)                 # uy's x^2 term varies linearly in y

    def quad(i, xs, ys, zs):
        return _fill(lambda x, y, z: A[i] * (1 + a[i] * y) * x * x + B[i] * z * z + g[i, 0] * x
                     + g[i, 1] * y + g[i, 2] * z + 0.2, xs, ys, zs)
    ux = quad(0, c["xf"], c["yc"], c["zc"])
    uy = quad(1, c["xc"], c["yn"], c["zc"])
    uz = quad(2, c["xc"], c["yc"], c["zf"])
    nut = _halo(ux.shape)
    o.call("Clc_SGS_MTS", ux, uy, uz, nut)
    dx, dz = case["xlx"] / nx, case["zlz"] / nz
    m2 = 0.5 if filt == 0 else 1.0 / 3.0      # second moment of [1,2,1]/4 or [1,4,1]/6
    # the filter residual of each component is A m2 dx^2 + B m2 dz^2; uy's is
    # taken on the two nodes around the cell and averaged in its square
    r = lambda i, y: A[i] * (1 + a[i] * y) * m2 * dx * dx + B[i] * m2 * dz * dz
    yn, yc = c["yn"][1:ny + 2], c["yc"][1:ny + 1]
    K = (r(0, yc) ** 2 + r(2, yc) ** 2 + 0.5 * (r(1, yn[:-1]) ** 2 + r(1, yn[1:]) ** 2))
    K = K[None, :, None]
    x = c["xc"][1:nx + 1][:, None, None]
    y = yc[None, :, None]
    z = c["zc"][1:nz + 1][None, None, :]
    G = np.empty((3, 3) + (nx, ny, nz))
    for i in range(3):
        G[i, 0] = 2 * A[i] * (1 + a[i] * y) * x + g[i, 0] + 0 * z
        G[i, 1] = A[i] * a[i] * x * x + g[i, 1] + 0 * y * z
        G[i, 2] = 2 * B[i] * z + g[i, 2] + 0 * x * y
    S = 0.5 * (G + G.transpose(1, 0, 2, 3, 4))
    Smag = np.sqrt(2 * (S * S).sum(axis=(0, 1)))
    delta = o.metric(O.M_DELTA, 0, ny + 1)[1:ny + 1][None, :, None]
    want = 0.05 * K / (np.sqrt(K) / delta + Smag / 10.0) + case["xnu"]
    got = nut[1:nx + 1, 1:ny + 1, 1:nz + 1]
    assert np.allclose(got, want, rtol=1e-11, atol=0), np.abs(got / want - 1).max()
    # the strain rate alone
    st = np.zeros((nx, ny, nz), order="F")
    o.call("ClcVelStrainRate", ux, uy, uz, st)
    assert np.allclose(st, Smag, rtol=1e-12, atol=1e-13)
    o.close()


def test_mts_uniform_flow_gives_molecular_viscosity(oracle_mod):
    O = oracle_mod
    case = _case(nxc=8, nyc=10, nzc=6, les_type=3)
    o = O.Oracle(case)
    ux, uy, uz = _halo((11, 13, 9), 1.0), _halo((11, 13, 9), 0.0), _halo((11, 13, 9), -0.5)
    nut = _halo((11, 13, 9))
    o.call("Clc_SGS_MTS", ux, uy, uz, nut)     # K = |S| = 0: the time scale is cut to 0
    assert np.all(nut[1:9, 1:11, 1:7] == case["xnu"])
    o.close()


def test_sgs_ghosts_put_nu_on_the_wall(oracle_mod):
    """no-slip: ghost = 2 nu - first cell, so the wall-face average is nu;
    free-slip: the ghost copies the first cell; x and z wrap."""
    O = oracle_mod
    for bc in ((0, 0, -1, -1, 0, 0), (0, 0, -1, -2, 0, 0)):
        case = _case(nxc=6, nyc=8, nzc=5, bc=bc)
        o = O.Oracle(case)
        nut = np.asfortranarray(np.random.default_rng(2).uniform(1e-4, 1e-3, (9, 11, 8)))
        o.call("SetBC_and_UpdateHalo_SGS", nut)
        nu = case["xnu"]
        assert np.allclose(0.5 * (nut[1:7, 0, 1:6] + nut[1:7, 1, 1:6]), nu, rtol=1e-14)
        if bc[3] == -1:
            assert np.allclose(0.5 * (nut[1:7, 9, 1:6] + nut[1:7, 8, 1:6]), nu, rtol=1e-14)
        else:
            assert np.array_equal(nut[1:7, 9, 1:6], nut[1:7, 8, 1:6])
        assert np.array_equal(nut[0, :10, 1:6], nut[6, :10, 1:6])
        assert np.array_equal(nut[7, :10, 1:6], nut[1, :10, 1:6])
        assert np.array_equal(nut[:8, :10, 0], nut[:8, :10, 5])
        assert np.array_equal(nut[:8, :10, 6], nut[:8, :10, 1])
        o.close()


# ------------------------------------------- LES right-hand sides --
def _projected_state(O, case, yp, seed=5):
    """a divergence-free laminar+noise field in an oracle context"""
    o = O.Oracle(case, yp=yp)
    rng = np.random.default_rng(seed)
    nx, ny, nz = case["nxc"], case["nyc"], case["nzc"]
    ycc = o.metric(O.M_YC, 0, ny + 1)
    ux = o.field(O.UX)
    ux[1:nx + 1, 1:ny + 1, 1:nz + 1] = (1.5 * (1 - (ycc[1:ny + 1] - 1) ** 2))[None, :, None]
    for fid in (O.UX, O.UY, O.UZ):
        f = o.field(fid)
        f[1:nx + 1, 1:ny + 1, 1:nz + 1] += 0.1 * rng.uniform(-1, 1, (nx, ny, nz))
    uy = o.field(O.UY)
    uy[:, 1, :] = 0.0
    o.project()
    return o


def test_les_rhs_with_constant_nut_equals_full_implicit_rhs(oracle_mod):
    O = oracle_mod
    n = 12
    case = _case(nxc=n, nyc=n, nzc=n, les_type=2)
    yp = _uniform_yp(n, 2.0)
    o = _projected_state(O, case, yp)
    ux, uy, uz, pr = (o.field(f).copy(order="F") for f in (O.UX, O.UY, O.UZ, O.PR))
    pr[...] = np.asfortranarray(np.random.default_rng(8).normal(size=pr.shape))
    o.call("SetBC_and_UpdateHalo_pr", pr)
    nut = _halo(ux.shape, case["xnu"])
    o.call("PMcoeUpdate", 1)
    assert o.scalar(O.S_PMGAMMA) == o.scalar(O.S_PMALPHA)
    assert o.scalar(O.S_PMBETAT) * case["xnu"] == pytest.approx(o.scalar(O.S_PMBETA), rel=1e-15)
    for comp in ("X", "Y", "Z"):
        shp = (n + 3,) * 3 if comp == "Z" else (n, n, n)
        r_les, r_fimp = np.zeros(shp, order="F"), np.zeros(shp, order="F")
        h1, h2 = np.zeros((n, n, n), order="F"), np.zeros((n, n, n), order="F")
        o.call("clcRhs" + comp + "_les", ux, uy, uz, r_les, h1, pr, nut)
        o.call("clcRhs" + comp, ux, uy, uz, r_fimp, h2, pr)
        if comp == "Z":
            r_les, r_fimp = (r[1:n + 1, 1:n + 1, 1:n + 1] for r in (r_les, r_fimp))
        j0 = 1 if comp == "Y" else 0          # uy row 1 is the wall
        err = np.abs(r_les - r_fimp)[:, j0:, :].max()
        scale = np.abs(r_fimp).max()
        assert err <= 1e-12 * scale, (comp, err, scale)
    o.close()


def _mms_les(Lx, Lz):
    X, Y, Z = sympy.symbols("x y z")
    kx, kz = 2 * sympy.pi / Lx, 2 * sympy.pi / Lz
    u = sympy.sin(kx * X + 0.3) * sympy.cos(kz * Z) * (1 + 0.3 * Y) + 0.5 * sympy.cos(sympy.pi * Y / 2)
    v = 0.7 * sympy.cos(kx * X) * sympy.sin(kz * Z + 0.2) * sympy.sin(sympy.pi * Y / 2)
    w = sympy.sin(kx * X) * sympy.sin(kz * Z) * (0.4 - 0.2 * Y ** 2) + 0.3
    nu = 0.01 * (1.5 + 0.5 * sympy.sin(kx * X - 0.4) * sympy.cos(kz * Z + 0.7)
                 + 0.6 * sympy.cos(1.1 * Y))
    return (X, Y, Z), u, v, w, nu


@pytest.mark.parametrize("comp", ["x", "y", "z"])
@pytest.mark.parametrize("stretched", [False, True])
def test_les_rhs_converges_second_order(oracle_mod, comp, stretched):
    """RhsX/Y/Z_PIMP_LES of substep 1 divided by pmGamma, split by linearity
    in nut into its convective part (nut = 0) and its viscous part, converge
    at second order to -div(u u_i) and div(nu (grad u_i + d_i u)) with a
    variable nu; the viscous part of the history to the same minus
    d/dy(nu du_i/dy).  Rows next to the walls are excluded."""
    O = oracle_mod
    errs = {}
    for n in (16, 32, 64):
        case = _case(nxc=n, nyc=n, nzc=n, xlx=2 * math.pi, yly=2.0, zlz=math.pi, dt=0.01,
                     gamma=1.2)
        o = O.Oracle(case, yp=None if stretched else _uniform_yp(n, 2.0))
        c = _coords(o, O, case)
        sym, u, v, w, nu = _mms_les(case["xlx"], case["zlz"])
        X, Y, Z = sym
        xyz = (X, Y, Z)
        uu = (u, v, w)
        i = "xyz".index(comp)
        ui = uu[i]
        conv = -sum(sympy.diff(uu[j] * ui, xyz[j]) for j in range(3))
        visc = sum(sympy.diff(nu * (sympy.diff(ui, xyz[j]) + sympy.diff(uu[j], xyz[i])), xyz[j])
                   for j in range(3))
        yimp = sympy.diff(nu * sympy.diff(ui, Y), Y)
        ux = _fill(_lamb(sym, u), c["xf"], c["yc"], c["zc"])
        uy = _fill(_lamb(sym, v), c["xc"], c["yn"], c["zc"])
        uz = _fill(_lamb(sym, w), c["xc"], c["yc"], c["zf"])
        nut = _fill(_lamb(sym, nu), c["xc"], c["yc"], c["zc"])
        pr = np.zeros_like(ux)
        o.call("PMcoeUpdate", 1)
        out = []
        for nf in (np.zeros_like(nut), nut):
            hist = np.zeros((n, n, n), order="F")
            rhs = np.zeros((n + 3,) * 3 if comp == "z" else (n, n, n), order="F")
            o.call("clcRhs" + comp.upper() + "_les", ux, uy, uz, rhs, hist, pr, nf)
            if comp == "z":
                rhs = rhs[1:n + 1, 1:n + 1, 1:n + 1]
            out.append((rhs / o.scalar(O.S_PMGAMMA), hist))
        pos = {"x": (c["xf"], c["yc"], c["zc"]), "y": (c["xc"], c["yn"], c["zc"]),
               "z": (c["xc"], c["yc"], c["zf"])}[comp]
        xs, ys, zs = (q[1:n + 1] for q in pos)
        C = _fill(_lamb(sym, conv), xs, ys, zs)
        V = _fill(_lamb(sym, visc), xs, ys, zs)
        H = _fill(_lamb(sym, visc - yimp), xs, ys, zs)
        sl = (slice(None), slice(1, n - 1), slice(None))
        e_c = np.abs(out[0][0] - C)[sl].max()
        e_v = np.abs(out[1][0] - out[0][0] - V)[sl].max()
        e_h = np.abs(out[1][1] - out[0][1] - H)[sl].max()
        errs[n] = (e_c, e_v, e_h, np.abs(C).max(), np.abs(V).max(), np.abs(H).max())
        o.close()
    for k in (0, 1, 2):
        r1 = errs[16][k] / errs[32][k]
        r2 = errs[32][k] / errs[64][k]
        assert r2 > 3.5 and r1 > 3.0, (comp, k, errs)
        assert errs[64][k] < 1e-2 * errs[64][k + 3], (comp, k, errs)


@pytest.mark.parametrize("stretched", [False, True])
def test_les_pressure_update_converges_second_order(oracle_mod, stretched):
    O = oracle_mod
    errs = {}
    for n in (16, 32, 64):
        case = _case(nxc=n, nyc=n, nzc=n, xlx=2 * math.pi, yly=2.0, zlz=math.pi, dt=0.01,
                     gamma=1.2)
        o = O.Oracle(case, yp=None if stretched else _uniform_yp(n, 2.0))
        c = _coords(o, O, case)
        X, Y, Z = sympy.symbols("x y z")
        _, _, _, _, nu = _mms_les(case["xlx"], case["zlz"])
        phi = sympy.cos(X + 0.2) * sympy.sin(2 * Z) * (Y ** 3 - 0.4 * Y) + sympy.cos(1.7 * Y)
        sym = (X, Y, Z)
        ph = _fill(_lamb(sym, phi), c["xc"], c["yc"], c["zc"])
        nut = _fill(_lamb(sym, nu), c["xc"], c["yc"], c["zc"])
        pr = np.zeros_like(ph)
        o.call("PMcoeUpdate", 2)
        bt = o.scalar(O.S_PMBETAT)
        o.call("PressureUpdate_les", pr, ph, nut)
        D = _fill(_lamb(sym, sympy.diff(nu * sympy.diff(phi, Y), Y)), *(q[1:n + 1] for q in
                                                                         (c["xc"], c["yc"], c["zc"])))
        got = (pr - ph)[1:n + 1, 1:n + 1, 1:n + 1] / (-bt)
        # on a stretched grid the wall rows are first-order (the wall face is
        # centred between the ghost and the first cell, the face above is not)
        sl = slice(1, n - 1) if stretched else slice(None)
        errs[n] = np.abs(got - D)[:, sl, :].max()
        o.close()
    assert errs[16] / errs[32] > 3.0 and errs[32] / errs[64] > 3.6, errs


# ------------------------------------------------- implicit solves --
@pytest.mark.parametrize("comp", ["x", "y", "z"])
@pytest.mark.parametrize("top", [-1, -2])
def test_les_solves_invert_the_wall_normal_operator(oracle_mod, comp, top):
    """clcU{1,2,3}Hat_PIMP_LES return du with (1 - betaT L_y) du = rhs, where
    L_y is the wall-normal viscous operator of the explicit right-hand side
    (extracted from clcRhs*_PIMP_LES on du with homogeneous wall ghosts)."""
    O = oracle_mod
    nx, ny, nz = 6, 14, 5
    case = _case(nxc=nx, nyc=ny, nzc=nz, xnu=0.02, dt=0.05, bc=(0, 0, -1, top, 0, 0))
    o = O.Oracle(case)
    rng = np.random.default_rng(21)
    shp = (nx + 3, ny + 3, nz + 3)
    nut = np.asfortranarray(rng.uniform(0.01, 0.05, shp))
    o.call("SetBC_and_UpdateHalo_SGS", nut)
    du = np.zeros(shp, order="F")
    du[1:nx + 1, 1:ny + 1, 1:nz + 1] = rng.normal(size=(nx, ny, nz))
    z = np.zeros(shp, order="F")
    fields = {"x": [du, z.copy(), z.copy()], "y": [z.copy(), du, z.copy()],
              "z": [z.copy(), z.copy(), du]}[comp]
    if comp == "y":
        du[:, 1, :] = 0.0                                  # the wall row
    uy_ym = np.zeros((nx, nz), order="F")
    o.call("SetBC_and_UpdateHalo", *fields, uy_ym)
    pr = np.zeros(shp, order="F")
    o.call("PMcoeUpdate", 1)
    hist = np.zeros((nx, ny, nz), order="F")
    rhs = np.zeros(shp if comp == "z" else (nx, ny, nz), order="F")
    o.call("clcRhs" + comp.upper() + "_les", *fields, rhs, hist, pr, nut)
    if comp == "z":
        rhs = rhs[1:nx + 1, 1:ny + 1, 1:nz + 1]
    gam, bt = o.scalar(O.S_PMGAMMA), o.scalar(O.S_PMBETAT)
    Ly = (rhs - gam * hist) / (2 * bt)
    d = du[1:nx + 1, 1:ny + 1, 1:nz + 1]
    b = d - bt * Ly
    u = np.zeros(shp, order="F")
    if comp == "z":
        bh = np.zeros(shp, order="F")
        bh[1:nx + 1, 1:ny + 1, 1:nz + 1] = b
        o.call("clcU3Hat_les", u, bh, nut)
    elif comp == "y":
        bb = np.asfortranarray(b)
        o.call("clcU2Hat_les", u, bb, np.zeros((nx, nz), order="F"), nut)
    else:
        o.call("clcU1Hat_les", u, np.asfortranarray(b), nut)
    got = u[1:nx + 1, 1:ny + 1, 1:nz + 1]
    assert np.abs(got - d).max() <= 1e-11 * np.abs(d).max(), np.abs(got - d).max()
    o.close()


# ------------------------------------------------------------ full step --
@pytest.mark.parametrize("les", [2, 3])
def test_les_step_keeps_divergence_free_and_flow_rate(oracle_mod, les):
    O = oracle_mod
    case = _case(nxc=16, nyc=16, nzc=12, les_type=les, is_ux_const=True)
    o = _projected_state(O, case, None)
    ux, uy, uz = o.field(O.UX), o.field(O.UY), o.field(O.UZ)
    o.call("SetBC_and_UpdateHalo", ux, uy, uz, o.field(O.UY_YM))
    for _ in range(3):
        o.iterate()
        assert o.scalar(O.S_DIVMAX2) < 1e-12
    nut = o.field(O.NUT)[1:17, 1:17, 1:13]
    assert np.all(nut >= case["xnu"]) and nut.max() > 1.5 * case["xnu"]
    # the bulk forcing restores ubulk at the start of every RhsX: a further
    # substep's mean-gradient correction is small
    assert abs(o.scalar(O.S_UXAVER) - case["ubulk"]) < 1e-3
    o.close()
