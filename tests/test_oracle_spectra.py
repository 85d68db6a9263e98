"""Pins of the oracle's 1-D spectra (src/CFD_4th/m_FlowCase.f90:644-842,
clcEnergySpectraX/Z, clcCospectraX/Z, clcLTS_x/z, transform_uy) against
closed forms of the discrete Fourier transform:

* Parseval: the energy spectrum of a row sums to the plane mean of f^2;
* a single Fourier mode A cos(2 pi m i/n + phi) puts A^2/2 on mode m alone
  (A^2 at the Nyquist mode with phi = 0);
* the linear coherence spectra of rows carrying A_j cos(2 pi m i/n + phi_j)
  are A_j A_r cos(phi_j - phi_r)/4 (real part) and A_j A_r sin(phi_j -
  phi_r)/4 (imaginary part) against the reference row r (jForLCS in the
  lower half of a channel, its mirror in the upper half);
* the uv cospectra carry the 4th-order face-to-centre interpolation's
  transfer function T(k) = 9/8 cos(k dx/2) - 1/8 cos(3 k dx/2) (centred ux)
  and the half-cell phase shift k dx/2 (ux on its face);
* the vv spectrum takes uy on the node away from the nearer wall
  (transform_uy).
"""
import ctypes
import math

import numpy as np
import pytest

from tests import cases

NCOL = 18


def _case(**kw):
    c = dict(cases.SMALL4, nxc=16, nyc=8, nzc=12, u_crf=0.0)
    c.update(kw)
    return c


def _halo(o, fn):
    """4th-order halo array (nxc+4, nyc+2, nzc+4) from fn(ic, jc, kc) on the
    reference's 1-based indices, x and z halos periodic"""
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    ic = (np.arange(-1, nx + 3) - 1) % nx + 1
    jc = np.arange(0, ny + 2)
    kc = (np.arange(-1, nz + 3) - 1) % nz + 1
    a = fn(ic[:, None, None].astype(float), jc[None, :, None].astype(float),
           kc[None, None, :].astype(float))
    return np.asfortranarray(np.broadcast_to(a, (nx + 4, ny + 2, nz + 4)).astype(float))


def _spectra(O, o, ux, uy, uz, pr, jlcs=(2, 3)):
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    sx = np.zeros((nx // 2 + 1, ny, NCOL), order="F")
    sz = np.zeros((nz // 2 + 1, ny, NCOL), order="F")
    j = (ctypes.c_int * 2)(*jlcs)
    o.call("clcSpectra1D", ux, uy, uz, pr, j, sx, sz)
    return sx, sz


def _zero(o):
    return _halo(o, lambda i, j, k: 0.0 * i)


@pytest.mark.parametrize("ucrf", [0.0, 0.37])
def test_parseval(oracle_mod, ucrf):
    O = oracle_mod
    o = O.Oracle4(_case(u_crf=ucrf))
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    rng = np.random.default_rng(4)
    f = [rng.normal(size=(nx, ny + 2, nz)) for _ in range(4)]
    ux, uy, uz, pr = (_halo(o, lambda i, j, k, a=a: a[i.astype(int) - 1, j.astype(int),
                                                      k.astype(int) - 1]) for a in f)
    sx, sz = _spectra(O, o, ux, uy, uz, pr)
    for col, a, add in ((1, ux, ucrf), (3, uz, 0.0), (4, pr, 0.0)):
        mean_sq = ((a[2:nx + 2, 1:ny + 1, 2:nz + 2] + add) ** 2).mean(axis=(0, 2))
        assert np.allclose(sx[:, :, col - 1].sum(axis=0), mean_sq, rtol=1e-13, atol=0), col
        assert np.allclose(sz[:, :, col - 1].sum(axis=0), mean_sq, rtol=1e-13, atol=0), col
    o.close()


@pytest.mark.parametrize("axis", ["x", "z"])
def test_single_mode(oracle_mod, axis):
    O = oracle_mod
    o = O.Oracle4(_case())
    n = o.nxc if axis == "x" else o.nzc
    z = _zero(o)
    for m, phi, want in ((3, 0.7, None), (n // 2, 0.0, None), (1, -1.2, None)):
        A = 1.3
        if axis == "x":
            fn = lambda i, j, k: A * np.cos(2 * np.pi * m * (i - 1) / n + phi) + 0 * j * k
        else:
            fn = lambda i, j, k: A * np.cos(2 * np.pi * m * (k - 1) / n + phi) + 0 * i * j
        pr = _halo(o, fn)
        sx, sz = _spectra(O, o, z, z, z, pr)
        s = (sx if axis == "x" else sz)[:, :, 3]
        expect = np.zeros(n // 2 + 1)
        expect[m] = A * A if m == n // 2 else A * A / 2
        for j in range(o.nyc):
            assert np.allclose(s[:, j], expect, rtol=1e-13, atol=1e-14), (m, j, s[:, j])
    o.close()


@pytest.mark.parametrize("flow", [1, 2])
@pytest.mark.parametrize("axis", ["x", "z"])
def test_linear_coherence_closed_form(oracle_mod, flow, axis):
    O = oracle_mod
    o = O.Oracle4(_case(flow_type=flow))
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    n = nx if axis == "x" else nz
    m = 2
    rng = np.random.default_rng(7)
    Aj = rng.uniform(0.5, 1.5, ny + 2)
    Pj = rng.uniform(-np.pi, np.pi, ny + 2)

    def fn(i, j, k):
        t = (i - 1) if axis == "x" else (k - 1)
        jj = j.astype(int)
        return Aj[jj] * np.cos(2 * np.pi * m * t / n + Pj[jj])
    z = _zero(o)
    uz = _halo(o, fn)
    jl = (2, 3)
    sx, sz = _spectra(O, o, z, z, uz, z, jl)
    s = sx if axis == "x" else sz

    def ref(j, jr):
        if flow == 2:
            return jr
        return jr if j <= ny // 2 else ny + 1 - jr
    for j in range(1, ny + 1):
        for col, jr, part in ((13, jl[0], np.cos), (14, jl[0], np.sin), (15, jl[1], np.cos)):
            r = ref(j, jr)
            expect = np.zeros(n // 2 + 1)
            expect[m] = Aj[j] * Aj[r] * part(Pj[j] - Pj[r]) / 4
            assert np.allclose(s[:, j - 1, col - 1], expect, rtol=1e-12, atol=1e-14), (col, j)
        # the energy of the same rows: A_j^2 / 2 on mode m
        e = np.zeros(n // 2 + 1)
        e[m] = Aj[j] ** 2 / 2
        assert np.allclose(s[:, j - 1, 2], e, rtol=1e-13, atol=1e-14)
    o.close()


def test_uv_cospectra_interpolation_and_phase(oracle_mod):
    O = oracle_mod
    o = O.Oracle4(_case(u_crf=0.25))
    nx, ny = o.nxc, o.nyc
    dx = o.case["xlx"] / nx
    m, A, B, psi = 3, 0.8, 1.1, 0.6
    k = 2 * np.pi * m / o.case["xlx"]
    # ux on its face x=(i-1)dx; uy on node j equal to B_j cos(k x_c + psi) with
    # B_j linear in j, so the cell-centre uy (the mean of its two nodes) is
    # B_{j+1/2} cos(k x_c + psi)
    Bj = lambda j: B * (1.0 + 0.3 * j)
    ux = _halo(o, lambda i, j, kk: A * np.cos(k * (i - 1) * dx) + 0 * j * kk)
    uy = _halo(o, lambda i, j, kk: Bj(j) * np.cos(k * (i - 0.5) * dx + psi) + 0 * kk)
    z = _zero(o)
    sx, sz = _spectra(O, o, ux, uy, z, z)
    T = 9.0 / 8.0 * np.cos(k * dx / 2) - 1.0 / 8.0 * np.cos(3 * k * dx / 2)
    for j in range(ny):
        Bc = Bj(j + 1.5)                       # row jc = j+1, nodes jc and jc+1
        uv = np.zeros(nx // 2 + 1)
        uv[m] = T * A * Bc * np.cos(psi) / 2
        uv2 = np.zeros(nx // 2 + 1)
        uv2[m] = A * Bc * np.cos(psi + k * dx / 2) / 2
        assert np.allclose(sx[:, j, 4], uv, rtol=1e-12, atol=1e-14)
        assert np.allclose(sx[:, j, 5], uv2, rtol=1e-12, atol=1e-14)
        # uCRF only enters the mean mode of u
        e = np.zeros(nx // 2 + 1)
        e[0] = 0.25 ** 2
        e[m] = A * A / 2
        assert np.allclose(sx[:, j, 0], e, rtol=1e-13, atol=1e-14)
    o.close()


@pytest.mark.parametrize("flow", [1, 2])
def test_vv_takes_the_node_away_from_the_wall(oracle_mod, flow):
    O = oracle_mod
    o = O.Oracle4(_case(flow_type=flow))
    ny = o.nyc
    cj = np.arange(ny + 2) * 0.1 + 0.3     # a distinct constant on every node
    uy = _halo(o, lambda i, j, k: cj[j.astype(int)] + 0 * i * k)
    z = _zero(o)
    sx, sz = _spectra(O, o, z, uy, z, z)
    for j in range(1, ny + 1):
        if flow == 1:
            t = j + 1 if j <= ny // 2 else j
            want = cj[t] ** 2
        else:
            want = cj[j + 1] ** 2 if j < ny else 0.0
        assert sx[0, j - 1, 1] == pytest.approx(want, rel=1e-13, abs=1e-30)
        assert sz[0, j - 1, 1] == pytest.approx(want, rel=1e-13, abs=1e-30)
        assert np.all(np.abs(sx[1:, j - 1, 1]) < 1e-28)
    o.close()


# ------------------------------------------------------------ 2-D spectra --
def _spectra2d(O, o, ux, uy, uz, pr, jl=2):
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    nys = ny // 2 if o.params.flow_type == 1 else ny
    s = np.zeros((nx // 2 + 1, nys, nz // 2 + 1, 9), order="F")
    o.call("clcSpectra2D", ux, uy, uz, pr, jl, s)
    return s


def _ratio(nx, nz):
    """the weight of each half-plane mode in the full (kx, kz) plane: the
    modes kx = 0, nx/2 and kz = 0, nz/2 have no conjugate partner"""
    r = np.full((nx // 2 + 1, nz // 2 + 1), 4.0)
    r[0, :] = r[-1, :] = 2.0
    r[:, 0] = r[:, -1] = 2.0
    r[0, 0] = r[0, -1] = r[-1, 0] = r[-1, -1] = 1.0
    return r


@pytest.mark.parametrize("flow", [1, 2])
def test_parseval_2d(oracle_mod, flow):
    O = oracle_mod
    o = O.Oracle4(_case(flow_type=flow, u_crf=0.2))
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    rng = np.random.default_rng(9)
    f = [rng.normal(size=(nx, ny + 2, nz)) for _ in range(4)]
    ux, uy, uz, pr = (_halo(o, lambda i, j, k, a=a: a[i.astype(int) - 1, j.astype(int),
                                                      k.astype(int) - 1]) for a in f)
    jl = 2
    s = _spectra2d(O, o, ux, uy, uz, pr, jl)
    R = _ratio(nx, nz)[:, None, :]
    nys = s.shape[1]

    def plane(a, add=0.0):
        return a[2:nx + 2, 1:ny + 1, 2:nz + 2] + add
    for col, a, add in ((1, ux, 0.2), (3, uz, 0.0), (4, pr, 0.0)):
        p = plane(a, add)
        msq = (p ** 2).mean(axis=(0, 2))
        total = (R * s[:, :, :, col - 1]).sum(axis=(0, 2))
        want = 0.5 * (msq[:nys] + msq[::-1][:nys]) if flow == 1 else msq
        assert np.allclose(total, want, rtol=1e-12, atol=0), col
    # cross-Parseval of the coherence of w against row jl (and its mirror)
    p = plane(uz)
    cross = lambda j, r: (p[:, j - 1, :] * p[:, r - 1, :]).mean()
    total = (R * s[:, :, :, 7]).sum(axis=(0, 2))
    for j in range(1, nys + 1):
        want = (0.5 * (cross(j, jl) + cross(ny + 1 - j, ny + 1 - jl)) if flow == 1
                else cross(j, jl))
        assert total[j - 1] == pytest.approx(want, rel=1e-12, abs=1e-14)
    o.close()


def test_single_mode_2d(oracle_mod):
    O = oracle_mod
    o = O.Oracle4(_case(flow_type=2))
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    z = _zero(o)
    for mx, mz, sgn in ((3, 2, 1), (3, 2, -1), (0, 4, 1), (nx // 2, 1, 1)):
        A, phi = 0.9, 0.4
        pr = _halo(o, lambda i, j, k: A * np.cos(2 * np.pi * (mx * (i - 1) / nx
                                                             + sgn * mz * (k - 1) / nz) + phi)
                   + 0 * j)
        s = _spectra2d(O, o, z, z, z, pr)[:, :, :, 3]
        R = _ratio(nx, nz)
        for j in range(ny):
            e = s[:, j, :]
            mask = np.zeros_like(e, dtype=bool)
            mask[mx, mz] = True
            assert np.abs(e[~mask]).max() < 1e-14, (mx, mz, sgn)
            # every case varies along one direction at least: mean square A^2/2
            assert R[mx, mz] * e[mx, mz] == pytest.approx(A * A / 2, rel=1e-12)
    o.close()


def test_uv_2d_folds_antisymmetrically(oracle_mod):
    O = oracle_mod
    o = O.Oracle4(_case(flow_type=1, u_crf=0.3))
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    a = lambda j: 0.5 + 0.2 * j              # ux: a function of the row only
    b = lambda j: np.sin(0.7 * j) + 0.1      # uy on the nodes
    ux = _halo(o, lambda i, j, k: a(j) + 0 * i * k)
    uy = _halo(o, lambda i, j, k: b(j) + 0 * i * k)
    z = _zero(o)
    s = _spectra2d(O, o, ux, uy, z, z)[:, :, :, 4]
    # the centre-interpolated ux is a(j) + uCRF (the 4-point weights sum to
    # one), the centre uy the mean of its nodes; only the mean mode is set
    uv = lambda j: (a(j) + 0.3) * 0.5 * (b(j) + b(j + 1))
    for j in range(1, ny // 2 + 1):
        want = 0.5 * (uv(j) - uv(ny + 1 - j))
        assert s[0, j - 1, 0] == pytest.approx(want, rel=1e-12, abs=1e-15)
        assert np.abs(s[:, j - 1, :]).sum() - abs(s[0, j - 1, 0]) < 1e-14
    o.close()
