"""Pins of the oracle's time-scheme pieces against what the mathematics fixes.

The reference ships no golden vectors for this path, so every pin below is
independent of the oracle's own formulas:

* the RK3 tables must give a third-order amplification factor;
* the wall-normal difference operators must be exact on the polynomials
  their finite-volume derivation is exact for;
* the explicit right-hand sides, the divergence and the gradient must converge
  at second order to the analytic operators on manufactured fields;
* the implicit solves must return an increment that the factored operator,
  assembled here independently, maps back to the right-hand side;
* the boundary rules must put the wall value exactly on the boundary value.
"""
import math

import numpy as np
import pytest

from tests import cases

sympy = pytest.importorskip("sympy")


def _case(**kw):
    c = dict(cases.SMALL, is_ux_const=False)
    c.update(kw)
    return c


def _uniform_yp(ny, yly):
    yp = np.zeros(ny + 2)
    yp[1:] = np.linspace(0.0, yly, ny + 1)
    return yp


# ------------------------------------------------------------------ RK3 --
def test_rk3_tables_third_order(oracle_mod):
    """du/dt = lam*u through the low-storage substeps
    u_k = u_{k-1} + gamma_k dt f(u_{k-1}) + zeta_k dt f(u_{k-2})
    must reproduce exp(z) to third order: 1 + z + z^2/2 + z^3/6 exactly."""
    O = oracle_mod
    case = _case()
    o = O.Oracle(case)
    dt, nu = case["dt"], case["xnu"]
    g, t, a, b, ac = [], [], [], [], []
    for ns in (1, 2, 3):
        o.call("PMcoeUpdate", ns)
        g.append(o.scalar(O.S_PMGAMMA) / dt)
        t.append(o.scalar(O.S_PMTHETA) / dt)
        a.append(o.scalar(O.S_PMALPHA) / dt)
        b.append(o.scalar(O.S_PMBETA))
        ac.append(o.scalar(O.S_PMALPHAC))
    # alpha = gamma + zeta, beta = alpha*nu*dt/2 (Crank-Nicolson half step)
    assert np.allclose(a, np.add(g, t), rtol=0, atol=1e-15)
    assert np.allclose(b, 0.5 * np.array(a) * nu * dt, rtol=1e-15, atol=0)
    assert np.allclose(ac, a, rtol=0, atol=1e-15)
    assert abs(sum(a) - 1.0) < 1e-15                      # the step spans dt
    assert abs(o.scalar(O.S_SIMTIME) - dt) < 1e-15
    assert t[0] == 0.0
    # amplification polynomial: evaluate at several z and fit the cubic
    zs = np.array([-0.7, -0.3, 0.2, 0.5, 0.9])
    amp = []
    for z in zs:
        um, u = 0.0, 1.0
        for k in range(3):
            un = u + g[k] * z * u + t[k] * z * um
            um, u = u, un
        amp.append(u)
    coef = np.polyfit(zs, amp, 3)[::-1]
    assert np.allclose(coef, [1.0, 1.0, 0.5, 1.0 / 6.0], rtol=0, atol=1e-13)
    o.close()



@pytest.mark.parametrize("ischeme", [1, 2])
def test_ab2_rk2_tables_second_order(oracle_mod, ischeme):
    """the other two schemes of m_Parameters.f90:135-142 through the same
    recurrence u_k = u_{k-1} + gamma_k z u_{k-1} + zeta_k z u_{k-2}: RK2 (two
    substeps) reproduces the midpoint method 1 + z + z^2/2 exactly; AB2 (one
    substep, the history carried from the previous step) is the two-step
    Adams-Bashforth method, whose principal root is exp(z) to second order"""
    O = oracle_mod
    case = dict(_case(), ischeme=ischeme)
    o = O.Oracle(case)
    dt = case["dt"]
    nsub = 2 if ischeme == 2 else 1
    g, t = [], []
    for ns in range(1, nsub + 1):
        o.call("PMcoeUpdate", ns)
        g.append(o.scalar(O.S_PMGAMMA) / dt)
        t.append(o.scalar(O.S_PMTHETA) / dt)
        assert abs(o.scalar(O.S_PMALPHA) / dt - (g[-1] + t[-1])) < 1e-15
    assert abs(sum(gk + tk for gk, tk in zip(g, t)) - 1.0) < 1e-15
    assert abs(o.scalar(O.S_SIMTIME) - dt) < 1e-15
    if ischeme == 2:
        zs = np.array([-0.7, -0.3, 0.2, 0.5, 0.9])
        amp = []
        for z in zs:
            um, u = 0.0, 1.0
            for k in range(2):
                un = u + g[k] * z * u + t[k] * z * um
                um, u = u, un
            amp.append(u)
        coef = np.polyfit(zs, amp, 2)[::-1]
        assert np.allclose(coef, [1.0, 1.0, 0.5], rtol=0, atol=1e-13)
    else:
        # zeta^2 - (1 + gamma z) zeta - theta z = 0, principal root
        for z in (1e-2, -1e-2, 2e-2):
            r = np.roots([1.0, -(1.0 + g[0] * z), -t[0] * z])
            root = r[np.argmin(np.abs(r - 1.0))].real
            assert abs(root - np.exp(z)) < 0.5 * abs(z) ** 3, z
            assert abs(root - (1 + z + z * z / 2)) < 0.5 * abs(z) ** 3
    o.close()

# ----------------------------------------------------- y-operator pins --
def test_wall_normal_operators_exact(oracle_mod):
    """finite-volume y operators on the stretched grid:
    cell-centred (ux/uz) rows are exact for linear data, give
    (yc[j+1]-yc[j-1])/dyp[j] on y^2, and the no-slip wall rows with the
    ghost u0 = -u1 are exact for quadratics vanishing at the wall; the
    node-centred (uy) operator is exact for quadratics everywhere; the
    pressure operator has zero flux through the walls."""
    O = oracle_mod
    o = O.Oracle(_case(nyc=24, gamma=2.0))
    ny = o.nyc
    yp = o.metric(O.M_YP, 0, ny + 1)
    yc = o.metric(O.M_YC, 0, ny + 1)
    dyp = o.metric(O.M_DYP, 0, ny + 1)
    am, ap = o.metric(O.M_AM2C, 0, ny), o.metric(O.M_AP2C, 0, ny)

    def apply_c(u, j):
        return am[j] * (u[j - 1] - u[j]) + ap[j] * (u[j + 1] - u[j])

    lin = 0.3 + 1.7 * yc
    sq = yc ** 2
    for j in range(2, ny):
        assert abs(apply_c(lin, j)) < 1e-9 * am[j]
        want = (yc[j + 1] - yc[j - 1]) / dyp[j]
        assert abs(apply_c(sq, j) - want) < 1e-12 * abs(want) * am[j] * dyp[j] ** 2 + 1e-9
    # bottom wall y=0 and top wall y=yly: u = c1*s + c2*s^2, s = distance to wall
    for c1, c2 in ((1.3, 0.0), (0.0, 2.5), (-0.4, 0.8)):
        s = yc.copy()
        u = c1 * s + c2 * s ** 2
        u[0] = -u[1]
        assert abs(apply_c(u, 1) - 2 * c2) < 1e-9 * max(1.0, am[1] * dyp[1] ** 2)
        s = yp[ny + 1] - yc
        u = c1 * s + c2 * s ** 2
        u[ny + 1] = -u[ny]
        assert abs(apply_c(u, ny) - 2 * c2) < 1e-9 * max(1.0, am[ny] * dyp[ny] ** 2)
    # node-centred operator (uy): exact for quadratics
    amp_, acp, app = (o.metric(m, 0, ny) for m in (O.M_AM2P, O.M_AC2P, O.M_AP2P))
    q = 0.2 - 0.9 * yp + 1.1 * yp ** 2
    for j in range(2, ny + 1):
        v = amp_[j] * q[j - 1] + acp[j] * q[j] + app[j] * q[j + 1]
        assert abs(v - 2.2) < 1e-9 * acp.__abs__()[j] * 1e-3 + 1e-9
    # pressure: Neumann walls, rows sum to zero
    amr, acr, apr = (o.metric(m, 1, ny) for m in (O.M_AM2PR, O.M_AC2PR, O.M_AP2PR))
    assert amr[0] == 0.0 and apr[-1] == 0.0
    assert np.all(np.abs(amr + acr + apr) <= 1e-12 * np.abs(acr))
    o.close()


# --------------------------------------------- manufactured solutions --
def _mms(Lx, Lz):
    X, Y, Z = sympy.symbols("x y z")
    kx, kz = 2 * sympy.pi / Lx, 2 * sympy.pi / Lz
    u = sympy.sin(kx * X + 0.3) * sympy.cos(kz * Z) * (1 + 0.3 * Y) + 0.5 * sympy.cos(sympy.pi * Y / 2)
    v = 0.7 * sympy.cos(kx * X) * sympy.sin(kz * Z + 0.2) * sympy.sin(sympy.pi * Y / 2)
    w = sympy.sin(kx * X) * sympy.sin(kz * Z) * (0.4 - 0.2 * Y ** 2) + 0.3
    p = 0.5 * sympy.cos(kx * X + 0.1) * sympy.sin(kz * Z) * (Y ** 2 + 0.3 * sympy.sin(1.3 * Y))
    return (X, Y, Z), u, v, w, p


def _lamb(sym, expr):
    f = sympy.lambdify(sym, expr, "numpy")
    return lambda x, y, z: np.broadcast_to(f(x, y, z), np.broadcast_shapes(x.shape, y.shape, z.shape)).astype(float)


def _coords(o, O, case):
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    dx, dz = case["xlx"] / nx, case["zlz"] / nz
    ic = np.arange(nx + 3)
    kc = np.arange(nz + 3)
    yn = np.zeros(ny + 3)
    yn[: ny + 2] = o.metric(O.M_YP, 0, ny + 1)
    yn[0] = 2 * yn[1] - yn[2]
    yn[ny + 2] = 2 * yn[ny + 1] - yn[ny]
    ycc = np.zeros(ny + 3)
    ycc[: ny + 2] = o.metric(O.M_YC, 0, ny + 1)
    ycc[ny + 2] = 2 * ycc[ny + 1] - ycc[ny]
    return dict(xf=(ic - 1) * dx, xc=(ic - 0.5) * dx, zf=(kc - 1) * dz, zc=(kc - 0.5) * dz,
                yn=yn, yc=ycc)


def _fill(fn, xs, ys, zs):
    return np.asfortranarray(fn(xs[:, None, None], ys[None, :, None], zs[None, None, :]))


def _mms_setup(O, n, nu=0.01):
    case = _case(nxc=n, nyc=n, nzc=n, xlx=2 * math.pi, yly=2.0, zlz=math.pi, xnu=nu,
                 dt=0.01)
    o = O.Oracle(case, yp=_uniform_yp(n, 2.0))
    c = _coords(o, O, case)
    sym, u, v, w, p = _mms(case["xlx"], case["zlz"])
    return case, o, c, sym, u, v, w, p


@pytest.mark.parametrize("comp", ["x", "y", "z"])
def test_rhs_converges_second_order(oracle_mod, comp):
    """RhsX/Y/Z of substep 1 (zeta=0) divided by pmGamma converges to
    -div(u u_i) + nu lap(u_i) - d_i p, and the history (the convection alone:
    every viscous term is implicit in this scheme) to -div(u u_i), at second
    order on a uniform grid
    (rows adjacent to the walls excluded: they are pinned by the wall-row
    test above)."""
    O = oracle_mod
    errs = {}
    for n in (16, 32, 64):
        case, o, c, sym, u, v, w, p = _mms_setup(O, n)
        X, Y, Z = sym
        nu = case["xnu"]
        ui = {"x": u, "y": v, "z": w}[comp]
        conv = -(sympy.diff(u * ui, X) + sympy.diff(v * ui, Y) + sympy.diff(w * ui, Z))
        hist_t = conv
        lap = sympy.diff(ui, X, 2) + sympy.diff(ui, Y, 2) + sympy.diff(ui, Z, 2)
        rhs_t = conv + nu * lap - sympy.diff(p, {"x": X, "y": Y, "z": Z}[comp])
        ux = _fill(_lamb(sym, u), c["xf"], c["yc"], c["zc"])
        uy = _fill(_lamb(sym, v), c["xc"], c["yn"], c["zc"])
        uz = _fill(_lamb(sym, w), c["xc"], c["yc"], c["zf"])
        pr = _fill(_lamb(sym, p), c["xc"], c["yc"], c["zc"])
        o.call("PMcoeUpdate", 1)
        hist = np.zeros((n, n, n), order="F")
        if comp == "z":
            rhs = np.zeros((n + 3, n + 3, n + 3), order="F")
        else:
            rhs = np.zeros((n, n, n), order="F")
        o.call("clcRhs" + comp.upper(), ux, uy, uz, rhs, hist, pr)
        if comp == "z":
            rhs = rhs[1:n + 1, 1:n + 1, 1:n + 1]
        pos = {"x": (c["xf"], c["yc"], c["zc"]), "y": (c["xc"], c["yn"], c["zc"]),
               "z": (c["xc"], c["yc"], c["zf"])}[comp]
        xs, ys, zs = (q[1:n + 1] for q in pos)
        R = _fill(_lamb(sym, rhs_t), xs, ys, zs)
        H = _fill(_lamb(sym, hist_t), xs, ys, zs)
        sl = (slice(None), slice(1, n - 1), slice(None))
        e_r = np.abs(rhs / o.scalar(O.S_PMGAMMA) - R)[sl].max()
        e_h = np.abs(hist - H)[sl].max()
        errs[n] = (e_r, e_h, np.abs(R).max())
        o.close()
    for k in (0, 1):
        r1 = errs[16][k] / errs[32][k]
        r2 = errs[32][k] / errs[64][k]
        assert r2 > 3.6 and r1 > 3.0, (comp, k, errs)
    assert errs[64][0] < 1e-2 * errs[64][2]


def test_divergence_and_gradient_converge_second_order(oracle_mod):
    """clcPrSrc: prsrc*pmAlpha -> div(u) and p -> p - nu/2 div(u);
    FluidVelUpdate: u -> u - pmAlpha grad(phi); CheckDivergence = max|div|."""
    O = oracle_mod
    errs = {}
    for n in (16, 32, 64):
        case, o, c, sym, u, v, w, p = _mms_setup(O, n)
        X, Y, Z = sym
        nu = case["xnu"]
        div = sympy.diff(u, X) + sympy.diff(v, Y) + sympy.diff(w, Z)
        ux = _fill(_lamb(sym, u), c["xf"], c["yc"], c["zc"])
        uy = _fill(_lamb(sym, v), c["xc"], c["yn"], c["zc"])
        uz = _fill(_lamb(sym, w), c["xc"], c["yc"], c["zf"])
        pr = _fill(_lamb(sym, p), c["xc"], c["yc"], c["zc"])
        p0 = pr.copy(order="F")
        o.call("PMcoeUpdate", 2)
        alpha = o.scalar(O.S_PMALPHA)
        src = np.zeros((n, n, n), order="F")
        dm = np.zeros(1)
        o.call("clcPrSrc", ux, uy, uz, src, pr, dm)
        dm = float(dm[0])
        D = _fill(_lamb(sym, div), c["xc"][1:n + 1], c["yc"][1:n + 1], c["zc"][1:n + 1])
        e_div = np.abs(src * alpha - D).max()
        e_p = np.abs((p0 - pr)[1:n + 1, 1:n + 1, 1:n + 1] - 0.5 * nu * D).max()
        assert abs(dm - np.abs(src * alpha).max()) <= 1e-12 * dm
        assert abs(o.call("CheckDivergence", ux, uy, uz) - dm) <= 1e-15 * dm
        # projection with phi = p
        phi = pr.copy(order="F")
        phi[...] = _fill(_lamb(sym, p), c["xc"], c["yc"], c["zc"])
        u1, v1, w1 = ux.copy(order="F"), uy.copy(order="F"), uz.copy(order="F")
        o.call("FluidVelUpdate", phi, u1, v1, w1)
        eg = []
        for arr0, arr1, dvar, pos in ((ux, u1, X, (c["xf"], c["yc"], c["zc"])),
                                      (uy, v1, Y, (c["xc"], c["yn"], c["zc"])),
                                      (uz, w1, Z, (c["xc"], c["yc"], c["zf"]))):
            G = _fill(_lamb(sym, sympy.diff(p, dvar)), *(q[1:n + 1] for q in pos))
            got = (arr0 - arr1)[1:n + 1, 1:n + 1, 1:n + 1] / alpha
            j0 = 1 if dvar is Y else 0   # uy row 1 is the wall node
            eg.append(np.abs(got - G)[:, j0:, :].max())
        errs[n] = (e_div, e_p, *eg)
        o.close()
    for k in range(5):
        r = errs[32][k] / errs[64][k]
        assert r > 3.6, (k, errs)


# ------------------------------------------------------ implicit solves --
def _periodic_d2(a, axis, r2):
    return r2 * (np.roll(a, 1, axis) - 2 * a + np.roll(a, -1, axis))


@pytest.mark.parametrize("name", ["small", "freeslip"])
def test_implicit_solves_invert_factored_operator(oracle_mod, name):
    """(1 - b Dxx)(1 - b Dzz)(1 - b Dyy) delta = R for clcU1Hat / clcU3Hat
    with Dyy from the raw cell metrics and the wall ghost rule (no-slip:
    delta_0 = -delta_1; free-slip: delta_{n+1} = delta_n); for clcU2Hat the
    node operator on rows 2..n with the wall nodes' increments fixed
    (row 1 = duy_ym, row n+1 = 0)."""
    O = oracle_mod
    case = dict({"small": cases.SMALL, "freeslip": cases.FREESLIP}[name])
    o = O.Oracle(case)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    rdx2 = (nx / case["xlx"]) ** 2
    rdz2 = (nz / case["zlz"]) ** 2
    o.call("PMcoeUpdate", 2)
    b = o.scalar(O.S_PMBETA)
    rng = np.random.default_rng(11)
    am, ap = o.metric(O.M_AM2C, 0, ny + 1), o.metric(O.M_AP2C, 0, ny + 1)
    top_noslip = case.get("bc", (0, 0, -1, -1, 0, 0))[3] == -1

    def axz(d):
        t = d - b * _periodic_d2(d, 0, rdx2)
        return t - b * _periodic_d2(t, 2, rdz2)

    for comp in ("U1", "U3"):
        R = np.asfortranarray(rng.standard_normal((nx, ny, nz)))
        u = np.zeros((nx + 3, ny + 3, nz + 3), order="F")
        if comp == "U1":
            o.call("clcU1Hat", u, R.copy(order="F"))
        else:
            Rh = np.zeros((nx + 3, ny + 3, nz + 3), order="F")
            Rh[1:nx + 1, 1:ny + 1, 1:nz + 1] = R
            o.call("clcU3Hat", u, Rh)
        d = u[1:nx + 1, 1:ny + 1, 1:nz + 1]
        g = np.zeros((nx, ny + 2, nz))
        g[:, 1:ny + 1] = d
        g[:, 0] = -d[:, 0]
        g[:, ny + 1] = -d[:, ny - 1] if top_noslip else d[:, ny - 1]
        j = np.arange(1, ny + 1)
        Dy = (am[j][None, :, None] * (g[:, :-2] - d) + ap[j][None, :, None] * (g[:, 2:] - d))
        if not top_noslip:
            Dy[:, -1] = am[ny] * (g[:, ny - 1] - d[:, -1])   # no flux through the free-slip wall
        res = axz(d - b * Dy) - R
        assert np.abs(res).max() < 1e-12 * np.abs(R).max(), comp
    # uy
    amp_, acp, app = (o.metric(m, 0, ny + 1) for m in (O.M_AM2P, O.M_AC2P, O.M_AP2P))
    R = np.asfortranarray(rng.standard_normal((nx, ny, nz)))
    dym = np.asfortranarray(0.01 * rng.standard_normal((nx, nz)))
    u = np.zeros((nx + 3, ny + 3, nz + 3), order="F")
    o.call("clcU2Hat", u, R.copy(order="F"), dym)
    d = u[1:nx + 1, 1:ny + 1, 1:nz + 1]
    assert np.abs(d[:, 0, :] - dym).max() < 1e-15
    g = np.zeros((nx, ny + 2, nz))
    g[:, 1:ny + 1] = d                               # g[:, ny+1] = 0: top wall node fixed
    j = np.arange(2, ny + 1)
    Ay = d[:, 1:] - b * (amp_[j][None, :, None] * g[:, 1:ny] + acp[j][None, :, None] * d[:, 1:]
                         + app[j][None, :, None] * g[:, 3:ny + 2])
    res = axz(Ay) - R[:, 1:]
    assert np.abs(res).max() < 1e-12 * np.abs(R).max()
    o.close()


# ------------------------------------------------------------ BC rules --
def test_wall_ghosts_put_the_boundary_value_on_the_wall(oracle_mod):
    O = oracle_mod
    for bc in ((0, 0, -1, -1, 0, 0), (0, 0, -1, -2, 0, 0)):
        case = _case(bc=bc, uxbc=(0, 0, 0.3, -0.2, 0, 0), uzbc=(0, 0, -0.1, 0.4, 0, 0),
                     uybc=(0, 0, 0.0, 0.05, 0, 0))
        o = O.Oracle(case)
        nx, ny, nz = o.nxc, o.nyc, o.nzc
        rng = np.random.default_rng(3)
        fields = [np.asfortranarray(rng.standard_normal((nx + 3, ny + 3, nz + 3))) for _ in range(3)]
        ym = np.asfortranarray(0.02 * rng.standard_normal((nx, nz)))
        o.call("SetBC_and_UpdateHalo", *fields, ym)
        ux, uy, uz = fields
        I = slice(1, nx + 1)
        K = slice(1, nz + 1)
        dyp = o.metric(O.M_DYP, 0, ny + 1)
        assert np.abs(0.5 * (ux[I, 0, K] + ux[I, 1, K]) - 0.3).max() < 1e-15
        assert np.abs(0.5 * (uz[I, 0, K] + uz[I, 1, K]) + 0.1).max() < 1e-15
        assert np.array_equal(uy[I, 1, K], ym)
        assert np.abs(0.5 * (uy[I, 0, K] + uy[I, 2, K]) - ym).max() < 1e-15
        assert np.all(uy[I, ny + 1, K] == 0.05)
        if bc[3] == -1:
            assert np.abs(0.5 * (ux[I, ny + 1, K] + ux[I, ny, K]) + 0.2).max() < 1e-15
            assert np.abs(0.5 * (uz[I, ny + 1, K] + uz[I, ny, K]) - 0.4).max() < 1e-15
        else:
            assert np.abs((ux[I, ny + 1, K] - ux[I, ny, K]) / dyp[ny + 1] + 0.2).max() < 1e-12
            assert np.abs((uz[I, ny + 1, K] - uz[I, ny, K]) / dyp[ny + 1] - 0.4).max() < 1e-12
        # periodic layers are the opposite interior planes
        for a in fields:
            assert np.array_equal(a[0], a[nx]) and np.array_equal(a[nx + 1], a[1])
            assert np.array_equal(a[:, :, 0], a[:, :, nz]) and np.array_equal(a[:, :, nz + 1], a[:, :, 1])
        o.close()


def test_diagnostics_closed_forms(oracle_mod):
    O = oracle_mod
    case = _case()
    o = O.Oracle(case)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    U = 1.7
    ux = np.full((nx + 3, ny + 3, nz + 3), U, order="F")
    uy = np.zeros_like(ux)
    uz = np.zeros_like(ux)
    # |u_i + u_{i+1}| rdx / 2 = U rdx
    assert abs(o.call("CalcMaxCFL", ux, uy, uz) - U * nx / case["xlx"]) < 1e-13 * U * nx
    assert abs(o.call("CalcUxAver", ux) - U) < 1e-14
    v = np.zeros(3)
    o.call("CalcVmax", ux, uy, uz, v)
    assert v[0] == U and v[1] == 0 and v[2] == 0
    assert o.call("CheckDivergence", ux, uy, uz) == 0.0
    # laminar profile: the bulk of 1.5(1-(y-1)^2) sampled at cell centres is
    # its midpoint rule, exact integral 2 plus -f''/24 sum dyp^3 = sum dyp^3 / 8
    yc = o.metric(O.M_YC, 1, ny)
    dyp = o.metric(O.M_DYP, 1, ny)
    ux[:, 1:ny + 1, :] = (1.5 * (1 - (yc - 1) ** 2))[None, :, None]
    want = (2.0 + (dyp ** 3).sum() / 8.0) / case["yly"]
    assert abs(o.call("CalcUxAver", ux) - want) < 1e-14
    o.close()


# ----------------------------------------- y-interpolation and bulk forcing --
def test_wall_normal_interpolation_exact_on_stretched_grid(oracle_mod):
    """the y-interpolations inside the convection terms are exact for data
    linear in y on the stretched grid:
    * u = a + b y carried by a uniform wall-normal velocity V has the flux
      divergence V b (history = -V b) in RhsX and RhsZ;
    * in RhsY, the x- and z-fluxes of uy = g(x) or g(z) carried by a velocity
      U(y) linear in y scale exactly with U at the uy nodes."""
    O = oracle_mod
    case = _case(nyc=20, gamma=2.0)
    o = O.Oracle(case)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    shp = (nx + 3, ny + 3, nz + 3)
    yc = np.zeros(ny + 3)
    yc[:ny + 2] = o.metric(O.M_YC, 0, ny + 1)
    yc[ny + 2] = 2 * yc[ny + 1] - yc[ny]
    yn = np.zeros(ny + 3)
    yn[:ny + 2] = o.metric(O.M_YP, 0, ny + 1)
    o.call("PMcoeUpdate", 1)
    pr = np.zeros(shp, order="F")
    a, b, V = 0.4, -1.3, 0.7
    lin = np.asfortranarray(np.broadcast_to((a + b * yc)[None, :, None], shp))
    vel = np.asfortranarray(np.full(shp, V))
    zero = np.zeros(shp, order="F")
    for comp, args in (("X", (lin, vel, zero)), ("Z", (zero, vel, lin))):
        rhs = np.zeros((nx, ny, nz), order="F") if comp == "X" else np.zeros(shp, order="F")
        hist = np.zeros((nx, ny, nz), order="F")
        o.call("clcRhs" + comp, *args, rhs, hist, pr)
        assert np.abs(hist + V * b).max() < 1e-12, comp
    # RhsY: fluxes in x and z scale with U(yp_j)
    dx, dz = case["xlx"] / nx, case["zlz"] / nz
    gx = np.sin(2 * np.pi * (np.arange(nx + 3) - 0.5) * dx / case["xlx"] + 0.4)
    gz = np.cos(2 * np.pi * (np.arange(nz + 3) - 0.5) * dz / case["zlz"] + 0.2)
    U = a + b * yn
    for which in ("x", "z"):
        g = (np.broadcast_to(gx[:, None, None], shp) if which == "x"
             else np.broadcast_to(gz[None, None, :], shp))
        hs = []
        for carrier in (np.ones(ny + 3), a + b * yc):
            c3 = np.asfortranarray(np.broadcast_to(carrier[None, :, None], shp))
            args = (c3, np.asfortranarray(g), zero) if which == "x" else (zero, np.asfortranarray(g), c3)
            rhs = np.zeros((nx, ny, nz), order="F")
            hist = np.zeros((nx, ny, nz), order="F")
            o.call("clcRhsY", *args, rhs, hist, pr)
            hs.append(hist)
        j = np.arange(2, ny + 1)              # rows 2..ny (row 1 is the wall node)
        want = U[j][None, :, None] * hs[0][:, 1:, :]
        assert np.abs(hs[1][:, 1:, :] - want).max() < 1e-12 * np.abs(want).max(), which
    o.close()


def test_constant_flow_rate_forcing(oracle_mod):
    """IsUxConst: RhsX gains pmAlphaC*(ubulk - <ux>) on every cell, <ux> the
    volume mean of ux (cells of dyp weight), and PrGradNow = (ubulk - <ux>)/dt"""
    import math as _m
    O = oracle_mod
    rng = np.random.default_rng(8)
    for forced in (False, True):
        case = dict(cases.SMALL, is_ux_const=forced, ubulk=1.1)
        o = O.Oracle(case)
        nx, ny, nz = o.nxc, o.nyc, o.nzc
        shp = (nx + 3, ny + 3, nz + 3)
        rng = np.random.default_rng(8)
        f = [np.asfortranarray(1.0 + 0.3 * rng.standard_normal(shp)) for _ in range(4)]
        o.call("PMcoeUpdate", 2)
        rhs = np.zeros((nx, ny, nz), order="F")
        hist = np.asfortranarray(rng.standard_normal((nx, ny, nz)))
        o.call("clcRhsX", f[0], f[1], f[2], rhs, hist, f[3])
        if not forced:
            base = rhs.copy()
            o.close()
            continue
        dyp = o.metric(O.M_DYP, 1, ny)
        vals = (f[0][1:nx + 1, 1:ny + 1, 1:nz + 1] * dyp[None, :, None]).ravel()
        mean = _m.fsum(vals) / (nx * nz) / case["yly"]
        shift = o.scalar(O.S_PMALPHAC) * (1.1 - mean)
        assert np.abs((rhs - base) - shift).max() < 1e-13
        assert abs(o.scalar(O.S_PRGRADNOW) - (1.1 - mean) / case["dt"]) < 1e-9
        o.close()


def test_rhs_wall_normal_diffusion_exact_on_stretched_grid(oracle_mod):
    """the viscous part of RhsX/Y/Z (isolated as the difference of two
    viscosities) applies the wall-normal operators exactly as pinned above:
    uy = quadratic in y -> 2 c2 at every node row 2..ny; ux, uz = quadratic
    in y vanishing on both walls (ghosts from the no-slip rule) -> 2 c2 in
    the wall rows, (yc[j+1]-yc[j-1])/dyp[j]-type values inside."""
    O = oracle_mod
    res = {}
    for nu in (1e-3, 3e-3):
        case = _case(nyc=20, gamma=2.0, xnu=nu)
        o = O.Oracle(case)
        nx, ny, nz = o.nxc, o.nyc, o.nzc
        shp = (nx + 3, ny + 3, nz + 3)
        yc = np.zeros(ny + 3)
        yc[:ny + 2] = o.metric(O.M_YC, 0, ny + 1)
        yn = np.zeros(ny + 3)
        yn[:ny + 2] = o.metric(O.M_YP, 0, ny + 1)
        yn[0] = 2 * yn[1] - yn[2]
        yn[ny + 2] = 2 * yn[ny + 1] - yn[ny]
        L = case["yly"]
        quad_c = 0.8 * yc * (L - yc)           # vanishes at y=0 and y=L, f'' = -1.6
        quad_c[0] = -quad_c[1]
        quad_c[ny + 1] = -quad_c[ny]
        quad_n = 0.3 - 0.5 * yn + 0.9 * yn ** 2  # f'' = 1.8
        fc = np.asfortranarray(np.broadcast_to(quad_c[None, :, None], shp))
        fn = np.asfortranarray(np.broadcast_to(quad_n[None, :, None], shp))
        zero = np.zeros(shp, order="F")
        o.call("PMcoeUpdate", 1)
        beta = o.scalar(O.S_PMBETA)
        out = {}
        for comp, args in (("X", (fc, zero, zero)), ("Y", (zero, fn, zero)), ("Z", (zero, zero, fc))):
            rhs = np.zeros(shp, order="F") if comp == "Z" else np.zeros((nx, ny, nz), order="F")
            hist = np.zeros((nx, ny, nz), order="F")
            o.call("clcRhs" + comp, *args, rhs, hist, zero)
            if comp == "Z":
                rhs = rhs[1:nx + 1, 1:ny + 1, 1:nz + 1]
            out[comp] = rhs.copy()
        res[nu] = (out, beta, yc, o.metric(O.M_DYP, 0, ny + 1))
        o.close()
    (o1, b1, yc, dyp), (o2, b2, _, _) = res[1e-3], res[3e-3]
    ny = dyp.size - 2
    for comp in ("X", "Y", "Z"):
        d22 = (o2[comp] - o1[comp]) / (2 * (b2 - b1))
        if comp == "Y":
            assert np.abs(d22[:, 1:, :] - 1.8).max() < 1e-8, comp
        else:
            assert np.abs(d22[:, 0, :] + 1.6).max() < 1e-8, comp
            assert np.abs(d22[:, ny - 1, :] + 1.6).max() < 1e-8, comp
            j = np.arange(2, ny)
            want = -0.8 * (yc[j + 1] - yc[j - 1]) / dyp[j]
            assert np.abs(d22[:, 1:ny - 1, :] - want[None, :, None]).max() < 1e-8, comp


def test_bulk_sum_order_gap(oracle_mod):
    """The library and the oracle sum the bulk velocity (clcRhsX_FIMP's s1tot,
    CalcUxAver) exactly in fixed point; the reference sums in double, in
    kc, jc, ic order (m_TSchemeFIMP_inc.f90:46, m_Tools.f90:262).  The oracle's
    sequential mode reproduces that order: the gap it opens on config 1
    (64^3, laminar + noise) after five steps stays well inside the 1e-12
    parity bar (measured: velocity ~4e-15, pressure ~5e-13 after 1-20 steps,
    the pressure source amplifying the 1-ulp change of dpmdxns)"""
    O = oracle_mod
    from channel3d_amd import synth
    from tests import cases
    os_ = []
    for seq in (0, 1):
        o = O.Oracle(cases.C1)
        ux, uy, uz = synth.laminar_noise(o.nxc, o.nyc, o.nzc, o.yp, seed=4)
        o.field(O.UX)[...] = ux
        o.field(O.UY)[...] = uy
        o.field(O.UZ)[...] = uz
        o.project()
        o.set_sum_order(seq)
        os_.append(o)
    ux_av = [o.call("CalcUxAver", o.field(O.UX)) for o in os_]
    assert abs(ux_av[1] - ux_av[0]) <= 1e-14 * abs(ux_av[0])
    for _ in range(5):
        for o in os_:
            o.iterate()

    def rel(f):
        a, b = os_[1].field(f)[1:-2, :, 1:-2], os_[0].field(f)[1:-2, :, 1:-2]
        return float(np.abs(a - b).max() / np.abs(b).max())
    gaps = {nm: rel(f) for nm, f in (("ux", O.UX), ("uy", O.UY), ("uz", O.UZ), ("p", O.PR))}
    print("sequential vs fixed-point bulk sum, 5 steps:", gaps)
    assert max(gaps["ux"], gaps["uy"], gaps["uz"]) < 1e-13
    assert 0.0 < gaps["p"] < 1e-12
    for o in os_:
        o.close()
