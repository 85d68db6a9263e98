"""Pins of the 4th-order oracle (src/CFD_4th restatement) against what the
mathematics fixes:

* the x/z stencils (staggered derivative (1,-27,27,-1)/24dx, interpolation
  (-1,9,9,-1)/16, second derivative (-1,16,-30,16,-1)/12dx^2) converge at
  fourth order to the analytic operators on manufactured fields whose
  y-dependence the 2nd-order y discretisation represents exactly;
* the modified wavenumber of the Poisson solver is the Fourier symbol of the
  discrete divergence of the discrete gradient, so the solver inverts the
  operator that clcPrSrc / FluidVelUpdate apply (projection leaves round-off
  divergence), checked against a sparse direct solve of that operator
  assembled here from the textbook coefficients;
* the y-implicit Crank-Nicolson solve returns an increment that the
  operator assembled here maps back to the right-hand side;
* the pressure update adds (1 - beta Lyy) phi with the finite-volume Neumann
  Lyy, and the boundary rules put the wall value on the boundary value.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from tests import cases

sympy = pytest.importorskip("sympy")

C4 = dict(nxc=32, nyc=24, nzc=24, xlx=4 * math.pi, yly=2.0, zlz=4 * math.pi / 3,
          xnu=3.4969e-4, dt=0.005, gamma=1.5)


def to4(a, nx, ny, nz):
    """2nd-order halo layout (nx+3, ny+3, nz+3) -> 4th-order (nx+4, ny+2, nz+4)"""
    b = np.zeros((nx + 4, ny + 2, nz + 4), order="F")
    b[2:nx + 2, 0:ny + 2, 2:nz + 2] = a[1:nx + 1, 0:ny + 2, 1:nz + 1]
    return b


def laminar4(O, case, seed=0):
    from channel3d_amd import synth
    o = O.Oracle4(case)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    ux, uy, uz = synth.laminar_noise(nx, ny, nz, o.yp, yly=case["yly"], ubulk=1.0, seed=seed)
    for fid, a in ((O.UX, ux), (O.UY, uy), (O.UZ, uz)):
        o.field(fid)[...] = to4(a, nx, ny, nz)
    o.project()
    return o


def _coords4(o, O, case):
    """staggered coordinates for the 4th-order halo index ranges"""
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    dx, dz = case["xlx"] / nx, case["zlz"] / nz
    ic = np.arange(-1, nx + 3)
    kc = np.arange(-1, nz + 3)
    yn = o.metric(O.M_YP, 0, ny + 1)
    yn[0] = 2 * yn[1] - yn[2]
    yc = o.metric(O.M_YC, 0, ny + 1)
    return dict(xf=(ic - 1) * dx, xc=(ic - 0.5) * dx, zf=(kc - 1) * dz, zc=(kc - 0.5) * dz,
                yn=yn, yc=yc)


def _lamb(sym, expr):
    f = sympy.lambdify(sym, expr, "numpy")
    return lambda x, y, z: np.broadcast_to(f(x, y, z), np.broadcast_shapes(x.shape, y.shape, z.shape)).astype(float)


def _fill(fn, xs, ys, zs):
    return np.asfortranarray(fn(xs[:, None, None], ys[None, :, None], zs[None, None, :]))


def _mms4(Lx, Lz):
    """fields whose y-dependence is linear (u, w) or constant (v): the
    2nd-order y operators are exact on them, so only x/z truncation remains"""
    X, Y, Z = sympy.symbols("x y z")
    kx, kz = 2 * sympy.pi / Lx, 2 * sympy.pi / Lz
    u = (sympy.sin(kx * X + 0.3) * sympy.cos(kz * Z) + 0.4) * (1 + 0.3 * Y)
    v = 0.7 * sympy.cos(kx * X) * sympy.sin(2 * kz * Z + 0.2) + 0.1
    w = (sympy.sin(kx * X) * sympy.sin(kz * Z) + 0.3 * sympy.cos(2 * kx * X)) * (0.4 - 0.2 * Y)
    p = 0.5 * sympy.cos(kx * X + 0.1) * sympy.sin(kz * Z) * (1.0 + 0.5 * Y)
    return (X, Y, Z), u, v, w, p


def _setup4(O, n, ny=8):
    case = dict(C4, nxc=n, nyc=ny, nzc=n, xnu=0.01, dt=0.01, is_ux_const=False,
                xlx=2 * math.pi, zlz=math.pi)
    yp = np.zeros(ny + 2)
    yp[1:] = np.linspace(0.0, 2.0, ny + 1) ** 1.0 + 0.05 * np.sin(np.linspace(0, np.pi, ny + 1))
    yp[1], yp[ny + 1] = 0.0, 2.0 + 0.0
    o = O.Oracle4(case, yp=yp)
    return case, o, _coords4(o, O, case)


@pytest.mark.parametrize("comp", ["x", "y", "z"])
def test_rhs4_converges_fourth_order_in_x_and_z(oracle_mod, comp):
    O = oracle_mod
    errs = {}
    for n in (16, 32, 64):
        case, o, c = _setup4(O, n)
        sym, u, v, w, p = _mms4(case["xlx"], case["zlz"])
        X, Y, Z = sym
        nu = case["xnu"]
        ui = {"x": u, "y": v, "z": w}[comp]
        conv = -(sympy.diff(u * ui, X) + sympy.diff(v * ui, Y) + sympy.diff(w * ui, Z))
        hist_t = conv + nu * (sympy.diff(ui, X, 2) + sympy.diff(ui, Z, 2))
        rhs_t = hist_t + nu * sympy.diff(ui, Y, 2) - sympy.diff(p, {"x": X, "y": Y, "z": Z}[comp])
        ux = _fill(_lamb(sym, u), c["xf"], c["yc"], c["zc"])
        uy = _fill(_lamb(sym, v), c["xc"], c["yn"], c["zc"])
        uz = _fill(_lamb(sym, w), c["xc"], c["yc"], c["zf"])
        pr = _fill(_lamb(sym, p), c["xc"], c["yc"], c["zc"])
        o.call("PMcoeUpdate", 1)
        nx, ny, nz = o.nxc, o.nyc, o.nzc
        r = [np.zeros((nx, ny, nz), order="F"), np.zeros((nx, ny, nz), order="F"),
             np.zeros((nx + 4, ny + 2, nz + 4), order="F")]
        h = [np.zeros((nx, ny, nz), order="F") for _ in range(3)]
        o.call("clcRhsXYZ", ux, uy, uz, *r, *h, pr)
        q = "xyz".index(comp)
        rhs = r[q] if q < 2 else r[2][2:nx + 2, 1:ny + 1, 2:nz + 2]
        pos = {"x": (c["xf"], c["yc"], c["zc"]), "y": (c["xc"], c["yn"], c["zc"]),
               "z": (c["xc"], c["yc"], c["zf"])}[comp]
        xs, ys, zs = pos[0][2:nx + 2], pos[1][1:ny + 1], pos[2][2:nz + 2]
        R = _fill(_lamb(sym, rhs_t), xs, ys, zs)
        H = _fill(_lamb(sym, hist_t), xs, ys, zs)
        # uy row 1 is the wall node; RhsX/RhsZ of the top row take the wall
        # value of uy as zero (m_TScheme.f90:227-230), these fields do not
        rows = slice(1, ny) if comp == "y" else slice(0, ny - 1)
        e_r = np.abs(rhs / o.scalar(O.S_PMGAMMA) - R)[:, rows, :].max()
        e_h = np.abs(h[q] - H)[:, rows, :].max()
        errs[n] = (e_r, e_h)
        o.close()
    for k in (0, 1):
        assert errs[32][k] / errs[64][k] > 13.0, (comp, k, errs)


def test_divergence_and_gradient4_converge_fourth_order(oracle_mod):
    O = oracle_mod
    errs = {}
    for n in (16, 32, 64):
        case, o, c = _setup4(O, n)
        sym, u, v, w, p = _mms4(case["xlx"], case["zlz"])
        X, Y, Z = sym
        nx, ny, nz = o.nxc, o.nyc, o.nzc
        ux = _fill(_lamb(sym, u), c["xf"], c["yc"], c["zc"])
        uy = _fill(_lamb(sym, v), c["xc"], c["yn"], c["zc"])
        uz = _fill(_lamb(sym, w), c["xc"], c["yc"], c["zf"])
        o.call("PMcoeUpdate", 3)
        alpha = o.scalar(O.S_PMALPHA)
        src = np.zeros((nx, ny, nz), order="F")
        dm = np.zeros(1)
        o.call("clcPrSrc", ux, uy, uz, src, dm)
        div = sympy.diff(u, X) + sympy.diff(v, Y) + sympy.diff(w, Z)
        D = _fill(_lamb(sym, div), c["xc"][2:nx + 2], c["yc"][1:ny + 1], c["zc"][2:nz + 2])
        e_div = np.abs(src * alpha - D).max()
        assert abs(dm[0] - np.abs(src * alpha).max()) <= 1e-12 * dm[0]
        assert abs(o.call("CheckDivergence", ux, uy, uz) - dm[0]) <= 1e-15 * dm[0]
        phi = _fill(_lamb(sym, p), c["xc"], c["yc"], c["zc"])
        u1, v1, w1 = ux.copy(order="F"), uy.copy(order="F"), uz.copy(order="F")
        o.call("FluidVelUpdate", phi, u1, v1, w1)
        eg = []
        for a0, a1, dv, pos in ((ux, u1, X, (c["xf"], c["yc"], c["zc"])),
                                (uz, w1, Z, (c["xc"], c["yc"], c["zf"]))):
            G = _fill(_lamb(sym, sympy.diff(p, dv)), pos[0][2:nx + 2], pos[1][1:ny + 1], pos[2][2:nz + 2])
            got = (a0 - a1)[2:nx + 2, 1:ny + 1, 2:nz + 2] / alpha
            eg.append(np.abs(got - G).max())
        errs[n] = (e_div, *eg)
        o.close()
    for k in range(3):
        assert errs[32][k] / errs[64][k] > 13.0, (k, errs)


def _stag_derivative(n, h):
    """4th-order staggered first derivative as a periodic sparse matrix: from
    cell values at i-2, i-1, i, i+1 to the face between i-1 and i
    (coefficients 1/24, -27/24, 27/24, -1/24 divided by h)"""
    D = sp.lil_matrix((n, n))
    for i in range(n):
        for off, cf in ((-2, 1 / 24), (-1, -27 / 24), (0, 27 / 24), (1, -1 / 24)):
            D[i, (i + off) % n] += cf / h
    return D.tocsr()


def test_wavenumber4_is_symbol_of_div_grad(oracle_mod):
    """eigenvalues of D_face->cell o G_cell->face on Fourier modes = WaveNumX"""
    O = oracle_mod
    o = O.Oracle4(C4)
    nx = o.nxc
    h = C4["xlx"] / nx
    G = _stag_derivative(nx, h)                  # cell -> face (face i-1/2 uses i-2..i+1)
    Dv = -G.T                                    # face -> cell: the adjoint stencil
    L = (Dv @ G).toarray()
    w = o.metric(O.M_WAVEX, 1, nx)
    for m in (0, 1, 3, nx // 4, nx // 2):
        mode = np.cos(2 * np.pi * m * np.arange(nx) / nx + 0.2)
        assert np.allclose(L @ mode, w[m] * mode, rtol=0, atol=1e-10 * (1 + abs(w[m])))
    # the closed form of the symbol: -((27 sin(w/2) - sin(3w/2)) / (12 h))^2
    k = np.arange(nx)
    th = 2 * np.pi * k / nx
    assert np.allclose(w, -((27 * np.sin(th / 2) - np.sin(1.5 * th)) / (12 * h)) ** 2,
                       rtol=1e-12, atol=1e-9)
    o.close()


@pytest.mark.parametrize("shape", [(8, 6, 8), (12, 9, 10)])
@pytest.mark.parametrize("route", [0, 1], ids=["x1", "z1"])
def test_ppe4_vs_sparse_direct(oracle_mod, shape, route):
    """both PPE routes of the reference (x-first clcPPE_x1, z-first
    clcPPE_z1, m_Poisson.f90:371-489) against a direct solve of the assembled
    operator with the same gauge"""
    O = oracle_mod
    nx, ny, nz = shape
    case = dict(C4, nxc=nx, nyc=ny, nzc=nz)
    o = O.Oracle4(case)
    o.set_ppe_route(route)
    hx, hz = case["xlx"] / nx, case["zlz"] / nz
    Gx, Gz = _stag_derivative(nx, hx), _stag_derivative(nz, hz)
    Lx, Lz = (-Gx.T @ Gx), (-Gz.T @ Gz)
    am, ac, ap = (o.metric(m, 1, ny) for m in (O.M_AM2PR, O.M_AC2PR, O.M_AP2PR))
    Dy = sp.diags([am[1:], ac, ap[:-1]], [-1, 0, 1])
    Ix, Iy, Iz = sp.identity(nx), sp.identity(ny), sp.identity(nz)
    A = (sp.kron(Iz, sp.kron(Iy, Lx)) + sp.kron(Iz, sp.kron(Dy, Ix)) +
         sp.kron(Lz, sp.kron(Iy, Ix))).tocsr()
    dyp = o.metric(O.M_DYP, 1, ny)
    f = np.random.default_rng(2).standard_normal((nx, ny, nz))
    wgt = dyp[None, :, None] * np.ones_like(f)
    f = f - (f * wgt).sum() / wgt.sum()
    phi_h = np.zeros((nx + 4, ny + 2, nz + 4), order="F")
    o.call("clcPPE", np.asfortranarray(f.copy()), phi_h)
    phi = phi_h[2:nx + 2, 1:ny + 1, 2:nz + 2]
    n = nx * ny * nz
    g = np.zeros(n)
    g[[i + nx * ny * k for i in range(nx) for k in range(nz)]] = 1.0
    M = sp.vstack([sp.hstack([A, sp.csr_matrix(g[:, None])]),
                   sp.hstack([sp.csr_matrix(g[None, :]), sp.csr_matrix((1, 1))])]).tocsc()
    sol = spla.spsolve(M, np.concatenate([f.reshape(-1, order="F"), [0.0]]))
    ref = sol[:n].reshape((nx, ny, nz), order="F")
    assert abs(phi[:, 0, :].mean()) < 1e-12 * np.abs(phi).max()
    assert np.abs(phi - ref).max() <= 1e-11 * np.abs(ref).max()
    o.close()


def test_projection4_is_divergence_free_and_steps_are_stable(oracle_mod):
    O = oracle_mod
    o = laminar4(O, C4, seed=2)
    assert o.call("CheckDivergence", o.field(O.UX), o.field(O.UY), o.field(O.UZ)) < 1e-12
    u0 = o.call("CalcUxAver", o.field(O.UX))
    for _ in range(4):
        o.iterate()
        assert o.scalar(O.S_DIVMAX2) < 1e-12
    # the wall-shear forcing holds the bulk velocity (its change is O(dt * nu))
    assert abs(o.scalar(O.S_UXAVER) - u0) < 1e-6
    assert np.abs(o.field(O.UY)[1:-1, 1, 1:-1]).max() == 0.0
    assert np.abs(o.field(O.UY)[1:-1, o.nyc + 1, 1:-1]).max() == 0.0
    o.close()


@pytest.mark.parametrize("flow_type", [1, 2])
def test_y_implicit_solve4_inverts_operator(oracle_mod, flow_type):
    """(1 - beta Dyy) delta = R with Dyy from the raw cell metrics and the
    ghost rule (no-slip: delta_0 = -delta_1; half channel top: no flux), and
    for uy the node operator on rows 2..ny with row 1 = duy_ym, row ny+1 = 0"""
    O = oracle_mod
    case = dict(C4, flow_type=flow_type, gamma=2.0)
    o = O.Oracle4(case)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    o.call("PMcoeUpdate", 2)
    b = o.scalar(O.S_PMBETA)
    rng = np.random.default_rng(4)
    am, ap = o.metric(O.M_AM2C, 0, ny + 1), o.metric(O.M_AP2C, 0, ny + 1)
    for comp in ("U1", "U3"):
        R = np.asfortranarray(rng.standard_normal((nx, ny, nz)))
        u = np.zeros((nx + 4, ny + 2, nz + 4), order="F")
        if comp == "U1":
            o.call("clcU1Hat", u, R.copy(order="F"))
        else:
            Rh = np.zeros((nx + 4, ny + 2, nz + 4), order="F")
            Rh[2:nx + 2, 1:ny + 1, 2:nz + 2] = R
            o.call("clcU3Hat", u, Rh)
        d = u[2:nx + 2, 1:ny + 1, 2:nz + 2]
        gh = np.zeros((nx, ny + 2, nz))
        gh[:, 1:ny + 1] = d
        gh[:, 0] = -d[:, 0]
        gh[:, ny + 1] = -d[:, ny - 1]
        j = np.arange(1, ny + 1)
        Dy = am[j][None, :, None] * (gh[:, :-2] - d) + ap[j][None, :, None] * (gh[:, 2:] - d)
        res = d - b * Dy - R
        assert np.abs(res).max() < 1e-12 * np.abs(R).max(), comp
    amp_, acp, app = (o.metric(m, 0, ny + 1) for m in (O.M_AM2P, O.M_AC2P, O.M_AP2P))
    R = np.asfortranarray(rng.standard_normal((nx, ny, nz)))
    dym = np.asfortranarray(0.01 * rng.standard_normal((nx, nz)))
    u = np.zeros((nx + 4, ny + 2, nz + 4), order="F")
    o.call("clcU2Hat", u, R.copy(order="F"), dym)
    d = u[2:nx + 2, 1:ny + 1, 2:nz + 2]
    assert np.abs(d[:, 0, :] - dym).max() < 1e-15
    gh = np.zeros((nx, ny + 2, nz))
    gh[:, 1:ny + 1] = d
    j = np.arange(2, ny + 1)
    Ay = d[:, 1:] - b * (amp_[j][None, :, None] * gh[:, 1:ny] + acp[j][None, :, None] * d[:, 1:]
                         + app[j][None, :, None] * gh[:, 3:ny + 2])
    assert np.abs(Ay - R[:, 1:]).max() < 1e-12 * np.abs(R).max()
    o.close()


def test_wall_rows4_exact_for_quadratics(oracle_mod):
    """4th-order code's cell metrics at the no-slip walls, with the ghost
    u0 = -u1, differentiate quadratics vanishing at the wall exactly; the
    half-channel top row has zero flux"""
    O = oracle_mod
    for ft in (1, 2):
        o = O.Oracle4(dict(C4, flow_type=ft, gamma=2.0))
        ny = o.nyc
        yc = o.metric(O.M_YC, 0, ny + 1)
        yp = o.metric(O.M_YP, 0, ny + 1)
        am, ap = o.metric(O.M_AM2C, 0, ny), o.metric(O.M_AP2C, 0, ny)
        dyp = o.metric(O.M_DYP, 0, ny + 1)
        s = yc.copy()
        u = 0.7 * s - 1.1 * s ** 2
        u[0] = -u[1]
        assert abs(am[1] * (u[0] - u[1]) + ap[1] * (u[2] - u[1]) + 2.2) < 1e-9 * am[1] * dyp[1] ** 2
        if ft == 1:
            s = yp[ny + 1] - yc
            u = 0.7 * s - 1.1 * s ** 2
            u[ny + 1] = -u[ny]
            assert abs(am[ny] * (u[ny - 1] - u[ny]) + ap[ny] * (u[ny + 1] - u[ny]) + 2.2) < 1e-9 * am[ny] * dyp[ny] ** 2
        else:
            assert ap[ny] == 0.0
        o.close()


def test_pressure_update4_adds_helmholtz_of_phi(oracle_mod):
    """p += (1 - beta Lyy) phi, Lyy the finite-volume Neumann operator
    (1/dyp_j) [(phi_{j+1}-phi_j)/dyc_{j+1} - (phi_j-phi_{j-1})/dyc_j], no
    flux through the walls"""
    O = oracle_mod
    o = O.Oracle4(dict(C4, gamma=1.8))
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    o.call("PMcoeUpdate", 3)
    b = o.scalar(O.S_PMBETA)
    rng = np.random.default_rng(6)
    phi = np.asfortranarray(rng.standard_normal((nx + 4, ny + 2, nz + 4)))
    o.call("SetBC_and_UpdateHalo_pr", phi)
    p = np.asfortranarray(rng.standard_normal((nx + 4, ny + 2, nz + 4)))
    p0 = p.copy()
    o.call("PressureUpdate", p, phi)
    yp = o.metric(O.M_YP, 0, ny + 1)
    yc = 0.5 * (yp[1:ny + 1] + yp[2:ny + 2])
    dyp = np.diff(yp[1:ny + 2])
    f = phi[2:nx + 2, 1:ny + 1, 2:nz + 2]
    flux = np.zeros((nx, ny + 1, nz))               # flux at nodes 1..ny+1
    flux[:, 1:ny, :] = (f[:, 1:, :] - f[:, :-1, :]) / (yc[1:] - yc[:-1])[None, :, None]
    L = (flux[:, 1:, :] - flux[:, :-1, :]) / dyp[None, :, None]
    got = (p - p0)[2:nx + 2, 1:ny + 1, 2:nz + 2]
    assert np.abs(got - (f - b * L)).max() < 1e-12 * np.abs(f).max()
    o.close()


def test_wall_ghosts4(oracle_mod):
    O = oracle_mod
    for ft in (1, 2):
        o = O.Oracle4(dict(C4, flow_type=ft, uxbc=(0, 0, 0.2, -0.3, 0, 0),
                           uzbc=(0, 0, 0.1, 0.4, 0, 0), uybc=(0, 0, 0, 0.05, 0, 0)))
        nx, ny, nz = o.nxc, o.nyc, o.nzc
        rng = np.random.default_rng(1)
        f = [np.asfortranarray(rng.standard_normal((nx + 4, ny + 2, nz + 4))) for _ in range(3)]
        ym = np.asfortranarray(0.01 * rng.standard_normal((nx, nz)))
        o.call("SetBC_and_UpdateHalo", *f, ym)
        ux, uy, uz = f
        I, K = slice(2, nx + 2), slice(2, nz + 2)
        assert np.abs(0.5 * (ux[I, 0, K] + ux[I, 1, K]) - 0.2).max() < 1e-15
        assert np.abs(0.5 * (uz[I, 0, K] + uz[I, 1, K]) - 0.1).max() < 1e-15
        assert np.array_equal(uy[I, 1, K], ym)
        if ft == 1:
            assert np.abs(0.5 * (ux[I, ny + 1, K] + ux[I, ny, K]) + 0.3).max() < 1e-15
            assert np.all(uy[I, ny + 1, K] == 0.05)
        else:
            assert np.array_equal(ux[I, ny + 1, K], ux[I, ny, K])
            assert np.all(uy[I, ny + 1, K] == 0.0)
        for a in f:   # two periodic layers each side
            assert np.array_equal(a[0:2], a[nx:nx + 2]) and np.array_equal(a[nx + 2:], a[2:4])
            assert np.array_equal(a[:, :, 0:2], a[:, :, nz:nz + 2])
            assert np.array_equal(a[:, :, nz + 2:], a[:, :, 2:4])
        o.close()


def test_constant_flow_forcing4_is_mean_wall_shear(oracle_mod):
    """IsUxConst in the 4th-order scheme subtracts pmAlpha*dpmdxns from RhsX
    with dpmdxns = nu <Lap ux>, the dyp-weighted volume mean: the x and z
    second differences sum to zero over a period (their weights sum to
    zero), so only the wall-normal operator (pinned above) remains"""
    O = oracle_mod
    rng = np.random.default_rng(9)
    outs = {}
    for forced in (False, True):
        o = O.Oracle4(dict(C4, is_ux_const=forced, gamma=1.7))
        nx, ny, nz = o.nxc, o.nyc, o.nzc
        rng = np.random.default_rng(9)
        f = [np.asfortranarray(rng.standard_normal((nx + 4, ny + 2, nz + 4))) for _ in range(4)]
        o.call("SetBC_and_UpdateHalo", f[0], f[1], f[2], np.zeros((nx, nz), order="F"))
        o.call("PMcoeUpdate", 2)
        r = [np.zeros((nx, ny, nz), order="F"), np.zeros((nx, ny, nz), order="F"),
             np.zeros((nx + 4, ny + 2, nz + 4), order="F")]
        h = [np.zeros((nx, ny, nz), order="F") for _ in range(3)]
        o.call("clcRhsXYZ", f[0], f[1], f[2], *r, *h, f[3])
        outs[forced] = (r[0].copy(), o.scalar(O.S_PRGRADNOW), o.scalar(O.S_PMALPHA), o, f[0])
    base = outs[False][0]
    rhs, dpm, alpha, o, ux = outs[True]
    ny = o.nyc
    am, ap = o.metric(O.M_AM2C, 0, ny), o.metric(O.M_AP2C, 0, ny)
    dyp = o.metric(O.M_DYP, 0, ny + 1)
    u = ux[2:-2, :, 2:-2]
    j = np.arange(1, ny + 1)
    lap_y = (am[j][None, :, None] * (u[:, j - 1] - u[:, j]) +
             ap[j][None, :, None] * (u[:, j + 1] - u[:, j]))
    want = C4["xnu"] * (lap_y * dyp[j][None, :, None]).sum(axis=1).mean() / C4["yly"]
    assert abs(dpm - want) < 1e-10 * abs(want)
    assert np.abs((base - rhs) - alpha * dpm).max() < 1e-13 * max(1.0, np.abs(base).max())
