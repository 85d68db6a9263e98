"""Pins the CPU oracle against closed forms, library routines and brute force.

The reference ships no tests or golden vectors for this path, so each oracle
function is checked here against what the mathematics fixes: numpy's FFT for
FFTW's half-complex transforms, scipy's banded LU and dense solves for the
Thomas solvers, the circulant closed form for the periodic solver, the tanh
closed form for the grid, a sparse direct solve of the assembled 7-point
operator for the Poisson solver, and the discrete div(grad) identity for the
projection.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from tests import cases

SIZES = [4, 6, 8, 9, 12, 15, 16, 30, 32, 36, 40, 64, 72, 96, 160, 192, 384, 576, 768]


def hc_from_numpy(x):
    n = x.size
    X = np.fft.fft(x)
    out = np.zeros(n)
    out[: n // 2 + 1] = X.real[: n // 2 + 1]
    for k in range(1, (n + 1) // 2):
        out[n - k] = X.imag[k]
    return out


@pytest.mark.parametrize("n", SIZES)
def test_r2hc_matches_numpy(oracle_mod, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    ref = hc_from_numpy(x)
    got = oracle_mod.r2hc(x)
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("n", SIZES)
def test_hc2r_inverts_r2hc(oracle_mod, n):
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n)
    back = oracle_mod.hc2r(oracle_mod.r2hc(x))
    assert np.abs(back - n * x).max() <= 1e-13 * n


def test_hc2r_against_definition(oracle_mod):
    # FFTW HC2R: x_j = r_0 + 2 sum_k (r_k cos - i_k sin) + r_{n/2} (-1)^j
    n = 12
    rng = np.random.default_rng(7)
    hc = rng.standard_normal(n)
    j = np.arange(n)
    ref = hc[0] + hc[n // 2] * (-1.0) ** j
    for k in range(1, n // 2):
        ref += 2 * (hc[k] * np.cos(2 * np.pi * j * k / n) - hc[n - k] * np.sin(2 * np.pi * j * k / n))
    assert np.abs(oracle_mod.hc2r(hc) - ref).max() < 1e-13


def test_reproducible_sum(oracle_mod):
    import math
    rng = np.random.default_rng(11)
    # bulk-integral-like terms: O(1) values times cell widths, plus tiny and negative ones
    v = rng.uniform(-0.2, 1.6, 200000) * rng.uniform(1e-3, 5e-2, 200000)
    v[::97] = rng.standard_normal(v[::97].size) * 1e-18
    s = oracle_mod.xsum(v)
    exact = math.fsum(v)
    assert abs(s - exact) <= 4 * np.finfo(float).eps * abs(exact) + v.size * 2.0 ** -80
    # independent of the order of the terms, bit for bit
    for seed in range(3):
        assert oracle_mod.xsum(np.random.default_rng(seed).permutation(v)) == s
    # single terms round-trip exactly when on the 2^-80 grid
    for t in (1.0, -3.5, 2.0 ** -70, 1234.5678):
        assert oracle_mod.xsum(np.array([t])) == t


def test_inverse_tridiagonal_vs_banded_lu(oracle_mod):
    rng = np.random.default_rng(1)
    m, n = 7, 50
    a = rng.uniform(-1, 1, (m, n))
    c = rng.uniform(-1, 1, (m, n))
    b = 3.0 + rng.uniform(0, 1, (m, n))
    f = rng.standard_normal((m, n))
    x = oracle_mod.inverse_tridiagonal(a, b, c, f)
    for i in range(m):
        ab = np.zeros((3, n))
        ab[0, 1:] = c[i, :-1]
        ab[1] = b[i]
        ab[2, :-1] = a[i, 1:]
        ref = sla.solve_banded((1, 1), ab, f[i])
        assert np.abs(x[i] - ref).max() < 1e-13


def test_inverse_tridiagonal_dense_tiny(oracle_mod):
    rng = np.random.default_rng(2)
    n = 5
    a, b, c = rng.uniform(-1, 1, (1, n)), 4 + rng.uniform(0, 1, (1, n)), rng.uniform(-1, 1, (1, n))
    f = rng.standard_normal((1, n))
    A = np.diag(b[0]) + np.diag(a[0, 1:], -1) + np.diag(c[0, :-1], 1)
    assert np.abs(oracle_mod.inverse_tridiagonal(a, b, c, f)[0] - np.linalg.solve(A, f[0])).max() < 1e-14


@pytest.mark.parametrize("n,e", [(7, 1e-3), (16, 0.3), (64, 5.0), (768, 8e-4), (576, 2.0)])
def test_periodic_fixed_solver_vs_circulant(oracle_mod, n, e):
    # tridiag(a, b, a) periodic: eigenvalues b + 2a cos(2 pi k / n)
    a, b = -e, 1.0 + 2.0 * e
    rng = np.random.default_rng(n)
    f = rng.standard_normal((3, n))
    x = oracle_mod.inverse_ptri_fixed(a, b, a, f)
    lam = b + 2 * a * np.cos(2 * np.pi * np.arange(n) / n)
    ref = np.real(np.fft.ifft(np.fft.fft(f, axis=1) / lam, axis=1))
    assert np.abs(x - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def test_periodic_fixed_solver_dense_tiny(oracle_mod):
    n = 5
    a, b, c = -0.3, 1.9, -0.7  # non-symmetric fixed coefficients
    A = np.diag([b] * n) + np.diag([a] * (n - 1), -1) + np.diag([c] * (n - 1), 1)
    A[0, n - 1] = a
    A[n - 1, 0] = c
    f = np.random.default_rng(3).standard_normal((1, n))
    x = oracle_mod.inverse_ptri_fixed(a, b, c, f)
    assert np.abs(x[0] - np.linalg.solve(A, f[0])).max() < 1e-14


@pytest.mark.parametrize("gamma", [1.2, 1.5, 2.0])
def test_tanh_grid_closed_form(oracle_mod, gamma):
    nyc, yly = 64, 2.0
    yp = oracle_mod.channel_yp(nyc, yly, gamma)
    xi = np.arange(nyc + 1) / nyc
    ref = 1.0 + np.tanh(gamma * (2 * xi - 1)) / np.tanh(gamma)
    assert np.abs(yp[1:] - ref).max() < 1e-14


def _make(oracle_mod, case):
    return oracle_mod.Oracle(case)


def test_metric_identities(oracle_mod):
    O = oracle_mod
    o = _make(O, cases.SMALL)
    ny = o.nyc
    yp = o.metric(O.M_YP, 1, ny + 1)
    dyp = o.metric(O.M_DYP, 1, ny)
    assert np.allclose(dyp, np.diff(yp), rtol=0, atol=1e-15)
    am, ac, ap = (o.metric(m, 1, ny) for m in (O.M_AM2PR, O.M_AC2PR, O.M_AP2PR))
    assert np.array_equal(ac, -(am + ap))
    assert am[0] == 0.0 and ap[-1] == 0.0
    # Neumann operator in y is diag(1/dyp) times a symmetric matrix
    rdyc = o.metric(O.M_RDYC, 1, ny + 1)
    assert np.allclose(ap[:-1] * dyp[:-1], rdyc[1:-1], rtol=1e-14)
    assert np.allclose(am[1:] * dyp[1:], rdyc[1:-1], rtol=1e-14)
    wx = o.metric(O.M_WAVEX, 1, o.nxc)
    k = np.arange(o.nxc)
    rdx = o.nxc / cases.SMALL["xlx"]
    assert np.allclose(wx, -4 * rdx ** 2 * np.sin(np.pi * k / o.nxc) ** 2, rtol=0, atol=1e-10)
    # CN metrics fold the no-slip ghost: am2cCN(1) = 2 am2c(1)
    assert o.metric(O.M_AM2CCN, 1, 1)[0] == 2 * o.metric(O.M_AM2C, 1, 1)[0]


def assemble_laplacian(o, O):
    """7-point pressure operator of the reference (m_Poisson.f90 in physical
    space): periodic second differences in x and z, am2Pr/ac2Pr/ap2Pr in y."""
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    rdx2 = (nx / o.case["xlx"]) ** 2
    rdz2 = (nz / o.case["zlz"]) ** 2
    am, ac, ap = (o.metric(m, 1, ny) for m in (O.M_AM2PR, O.M_AC2PR, O.M_AP2PR))
    Dx = sp.diags([np.ones(nx - 1), -2 * np.ones(nx), np.ones(nx - 1)], [-1, 0, 1]).tolil()
    Dx[0, nx - 1] = 1
    Dx[nx - 1, 0] = 1
    Dz = sp.diags([np.ones(nz - 1), -2 * np.ones(nz), np.ones(nz - 1)], [-1, 0, 1]).tolil()
    Dz[0, nz - 1] = 1
    Dz[nz - 1, 0] = 1
    Dy = sp.diags([am[1:], ac, ap[:-1]], [-1, 0, 1])
    Ix, Iy, Iz = sp.identity(nx), sp.identity(ny), sp.identity(nz)
    # Fortran order: index = i + nx*(j + ny*k) -> kron(z, kron(y, x))
    return (rdx2 * sp.kron(Iz, sp.kron(Iy, Dx)) + sp.kron(Iz, sp.kron(Dy, Ix)) +
            rdz2 * sp.kron(Dz, sp.kron(Iy, Ix))).tocsr()


@pytest.mark.parametrize("shape", [(8, 6, 8), (12, 9, 10)])
@pytest.mark.parametrize("route", [0, 1], ids=["x1", "z1"])
def test_ppe_vs_sparse_direct(oracle_mod, shape, route):
    """both PPE routes (clcPPE_x1 and clcPPE_z1, m_Poisson.f90:567-688,
    1078-1197) against a direct solve of the assembled operator"""
    O = oracle_mod
    nx, ny, nz = shape
    case = dict(cases.SMALL, nxc=nx, nyc=ny, nzc=nz)
    o = O.Oracle(case)
    o.set_ppe_route(route)
    Lop = assemble_laplacian(o, O)
    dyp = o.metric(O.M_DYP, 1, ny)
    f = np.random.default_rng(5).standard_normal((nx, ny, nz))
    w = dyp[None, :, None] * np.ones_like(f)
    f = f - (f * w).sum() / w.sum()  # compatible with the Neumann walls
    src = np.asfortranarray(f.copy())
    phi_h = np.zeros((nx + 3, ny + 3, nz + 3), order="F")
    o.call("clcPPE", src, phi_h)
    phi = phi_h[1:nx + 1, 1:ny + 1, 1:nz + 1]
    # gauge of the reference: x-z mean of phi in the first cell is zero
    assert abs(phi[:, 0, :].mean()) < 1e-12 * np.abs(phi).max()
    # brute force: the singular operator plus the gauge row
    n = nx * ny * nz
    g = np.zeros(n)
    g[[i + nx * (0 + ny * k) for i in range(nx) for k in range(nz)]] = 1.0
    A = sp.vstack([sp.hstack([Lop, sp.csr_matrix(g[:, None])]),
                   sp.hstack([sp.csr_matrix(g[None, :]), sp.csr_matrix((1, 1))])]).tocsc()
    rhs = np.concatenate([f.reshape(-1, order="F"), [0.0]])
    sol = spla.spsolve(A, rhs)
    ref = sol[:n].reshape((nx, ny, nz), order="F")
    assert abs(sol[n]) < 1e-9  # multiplier vanishes for a compatible source
    assert np.abs(phi - ref).max() <= 1e-11 * np.abs(ref).max()


def test_ppe_manufactured_mode(oracle_mod):
    # phi = cos(2 pi m x / Lx) cos(2 pi q z / Lz) g(y): L phi = (lx + lz) phi + Dyy phi exactly
    O = oracle_mod
    case = dict(cases.SMALL)
    o = O.Oracle(case)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    am, ac, ap = (o.metric(m, 1, ny) for m in (O.M_AM2PR, O.M_AC2PR, O.M_AP2PR))
    wx = o.metric(O.M_WAVEX, 1, nx)
    wz = o.metric(O.M_WAVEZ, 1, nz)
    m, q = 3, 2
    gy = np.cos(np.linspace(0, 2.5, ny)) + 0.3
    cx = np.cos(2 * np.pi * m * np.arange(nx) / nx)
    cz = np.cos(2 * np.pi * q * np.arange(nz) / nz)
    Dyg = ac * gy
    Dyg[1:] += am[1:] * gy[:-1]
    Dyg[:-1] += ap[:-1] * gy[1:]
    f = cx[:, None, None] * cz[None, None, :] * ((wx[m] + wz[q]) * gy + Dyg)[None, :, None]
    src = np.asfortranarray(f)
    phi_h = np.zeros((nx + 3, ny + 3, nz + 3), order="F")
    o.call("clcPPE", src, phi_h)
    ref = cx[:, None, None] * cz[None, None, :] * gy[None, :, None]
    assert np.abs(phi_h[1:nx + 1, 1:ny + 1, 1:nz + 1] - ref).max() < 1e-12


def _laminar_state(O, case, seed=0):
    from channel3d_amd import synth
    o = O.Oracle(case)
    ux, uy, uz = synth.laminar_noise(o.nxc, o.nyc, o.nzc, o.yp, yly=case["yly"],
                                     ubulk=case.get("ubulk", 1.0), seed=seed)
    o.field(O.UX)[...] = ux
    o.field(O.UY)[...] = uy
    o.field(O.UZ)[...] = uz
    o.project()
    return o


def test_projection_is_divergence_free(oracle_mod):
    O = oracle_mod
    o = _laminar_state(O, cases.SMALL)
    d = o.call("CheckDivergence", o.field(O.UX), o.field(O.UY), o.field(O.UZ))
    assert d < 1e-12


def test_step_invariants(oracle_mod):
    O = oracle_mod
    o = _laminar_state(O, cases.SMALL, seed=3)
    dev0 = abs(o.call("CalcUxAver", o.field(O.UX)) - 1.0)
    for _ in range(6):
        o.iterate()
        assert o.scalar(O.S_DIVMAX2) < 1e-12            # projected velocity
    # the constant-flow-rate forcing pulls the bulk velocity to ubulk
    assert abs(o.scalar(O.S_UXAVER) - 1.0) < 1e-2 * dev0
    uy = o.field(O.UY)
    assert np.abs(uy[1:-2, 1, 1:-2]).max() == 0.0       # no-slip wall node
    assert np.abs(uy[1:-2, o.nyc + 1, 1:-2]).max() == 0.0


def test_convection_conserves_momentum(oracle_mod):
    # inviscid, pressure-free, periodic-in-x/z RHS: the conservative x-convection
    # sums to zero over the domain (weighted by dyp) when the walls carry no flux
    O = oracle_mod
    case = dict(cases.SMALL, xnu=0.0, is_ux_const=False)
    o = _laminar_state(O, case, seed=4)
    o.call("PMcoeUpdate", 1)
    rhs = np.zeros((o.nxc, o.nyc, o.nzc), order="F")
    hist = np.zeros_like(rhs)
    pr = np.zeros((o.nxc + 3, o.nyc + 3, o.nzc + 3), order="F")
    o.call("clcRhsX", o.field(O.UX), o.field(O.UY), o.field(O.UZ), rhs, hist, pr)
    dyp = o.metric(O.M_DYP, 1, o.nyc)
    total = (hist * dyp[None, :, None]).sum()
    scale = (np.abs(hist) * dyp[None, :, None]).sum()
    assert abs(total) < 1e-12 * scale
