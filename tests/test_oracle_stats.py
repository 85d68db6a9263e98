"""Pins of the oracle's channel statistics (clcStat_CH, 35 columns,
src/CFD_2nd/m_FlowType_Channel.f90:175-318) against closed forms.

The reference ships no statistics files, so each column is pinned by fields
whose plane averages are known exactly on the discrete grid:

* constant fields give every product column as the product of constants and
  every derivative column as zero;
* fields linear in y (x, z independent) give the wall-normal derivatives,
  the node interpolation of the pressure and the vorticity columns exactly,
  because the staggered differences and the YinterpCoe interpolation are
  exact for linear data on any stretched grid;
* single Fourier modes in x (or z) give the streamwise (spanwise) derivative
  columns through the discrete identities
  cos(a+2t)-cos(a) = -2 sin t sin(a+t), sin(a)-sin(a-2t) = 2 sin t cos(a-t),
  and plane means of sin^2 = cos^2 = 1/2, cos^4 = 3/8, odd powers = 0.

Column numbers in the comments are the reference's SumVec indices (1-based).
"""
import math

import numpy as np
import pytest

from tests import cases

CASE = dict(cases.SMALL, is_ux_const=False)


def _grids(O, o):
    ny = o.nyc
    nx, nz = o.nxc, o.nzc
    dx, dz = CASE["xlx"] / nx, CASE["zlz"] / nz
    i = np.arange(nx + 3)
    k = np.arange(nz + 3)
    return dict(
        xf=(i - 1) * dx, xc=(i - 0.5) * dx, zf=(k - 1) * dz, zc=(k - 0.5) * dz,
        yp=o.metric(O.M_YP, 0, ny + 1), yc=o.metric(O.M_YC, 0, ny + 1), dx=dx, dz=dz)


def _set(o, O, ux, uy, uz, pr):
    for fid, v in ((O.UX, ux), (O.UY, uy), (O.UZ, uz), (O.PR, pr)):
        o.field(fid)[...] = v


def _stats(O, o, acc=None):
    s = np.zeros((35, o.nyc + 1), order="F") if acc is None else acc
    o.call("clcStat_CH", o.field(O.UX), o.field(O.UY), o.field(O.UZ), o.field(O.PR), s)
    return s


def _bcast(o, fx=None, fy=None, fz=None):
    """field on the halo array from separable factors (index-space vectors)"""
    a = np.ones((o.nxc + 3, o.nyc + 3, o.nzc + 3))
    if fx is not None:
        a = a * fx[:, None, None]
    if fy is not None:
        v = np.zeros(o.nyc + 3)
        v[:len(fy)] = fy
        a = a * v[None, :, None]
    if fz is not None:
        a = a * fz[None, None, :]
    return a


def test_stats_constant_fields(oracle_mod):
    O = oracle_mod
    o = O.Oracle(CASE)
    U, V, W, P = 2.0, 3.0, 5.0, 7.0
    one = _bcast(o)
    _set(o, O, U * one, V * one, W * one, P * one)
    s = _stats(O, o)
    want = [U, V, W, P, U * U, V * V, W * W, P * P, U * V, V * W, U * W, U * P, V * P, W * P,
            U * U * V, V ** 3, W * W * V, U * V * V, U ** 3, W ** 3, U ** 4, V ** 4, W ** 4]
    ny = o.nyc
    for m, w in enumerate(want):
        np.testing.assert_allclose(s[m, :ny], w, rtol=1e-14, atol=0, err_msg=f"column {m + 1}")
    # derivative columns vanish (the y Laplacian rows sum to zero up to rounding)
    scale = np.abs(o.metric(O.M_AC2C, 1, ny)).max() * max(U, V, W) ** 2
    assert np.abs(s[23:, :ny]).max() <= 1e-12 * scale
    # plane nyp: only columns 27, 33, 35 are accumulated, and they vanish here
    assert np.all(s[:26, ny] == 0) and np.all(s[27:32, ny] == 0) and s[33, ny] == 0
    assert np.abs(s[[26, 32, 34], ny]).max() <= 1e-12 * scale
    # a second call adds the same increments (SumStat accumulates)
    s2 = _stats(O, o, s.copy(order="F"))
    np.testing.assert_allclose(s2, 2 * s, rtol=1e-15, atol=1e-300)
    o.close()


@pytest.mark.parametrize("P0,PI", [(7.0, 0.0), (0.0, 1.3)])
def test_stats_linear_in_y(oracle_mod, P0, PI):
    """ux = a yc + u0, uy = b yp, uz = g yc, p = P0 + PI yc on a stretched grid"""
    O = oracle_mod
    o = O.Oracle(CASE)
    G = _grids(O, o)
    yc, yp = G["yc"], G["yp"]
    a, u0, b, g = 0.7, 0.4, -0.3, 1.9
    ny = o.nyc
    _set(o, O, _bcast(o, fy=a * yc + u0), _bcast(o, fy=b * yp), _bcast(o, fy=g * yc),
         _bcast(o, fy=P0 + PI * yc))
    s = _stats(O, o)
    j = np.arange(1, ny + 1)
    ux, uy, uz, p = a * yc[j] + u0, b * yp[j], g * yc[j], P0 + PI * yc[j]
    uyc = b * yc[j]                       # (uy(j) + uy(j+1))/2 at the cell centre
    pnode = P0 + PI * yp[j]               # YinterpCoe interpolation to the node
    want = {1: ux, 2: uy, 3: uz, 4: p, 5: ux ** 2, 6: uy ** 2, 7: uz ** 2, 8: p ** 2,
            9: ux * uyc, 10: uyc * uz, 11: ux * uz, 12: ux * p, 13: uyc * p, 14: uz * p,
            15: ux * ux * uyc, 16: uy ** 3, 17: uz * uz * uyc, 18: ux * uyc * uyc,
            19: ux ** 3, 20: uz ** 3, 21: ux ** 4, 22: uy ** 4, 23: uz ** 4,
            24: 0 * j, 25: p * b, 26: 0 * j, 27: pnode * a,
            31: a * b + 0 * j, 32: 0 * j, 33: g * g + 0 * j, 34: 0 * j, 35: a * a + 0 * j}
    for col, w in want.items():
        np.testing.assert_allclose(s[col - 1, :ny], w, rtol=1e-11, atol=1e-11,
                                   err_msg=f"column {col}")
    # Laplacian columns: zero for linear data away from the wall rows
    # (the wall rows see the linearly extended ghosts, also exact)
    assert np.abs(s[27:30, :ny]).max() < 1e-9
    # plane nyp: dudy = a, dwdy = g from the ghost row; p interpolated to yp(nyp)
    np.testing.assert_allclose(s[[32, 34], ny], [g * g, a * a], rtol=1e-11)
    if PI == 0.0:
        np.testing.assert_allclose(s[26, ny], P0 * a, rtol=1e-12)
    o.close()


def test_stats_x_modes(oracle_mod):
    """ux = A cos(k x_face), uy = B cos(k x_centre), p = C sin(k x_centre), uz = 0"""
    O = oracle_mod
    o = O.Oracle(CASE)
    G = _grids(O, o)
    m = 3
    k = 2 * math.pi * m / CASE["xlx"]
    th = k * G["dx"] / 2
    s_, c_ = math.sin(th), math.cos(th)
    dx = G["dx"]
    A, B, C = 1.3, 0.8, 0.6
    _set(o, O, _bcast(o, fx=A * np.cos(k * G["xf"])), _bcast(o, fx=B * np.cos(k * G["xc"])),
         0 * _bcast(o), _bcast(o, fx=C * np.sin(k * G["xc"])))
    s = _stats(O, o)
    ny = o.nyc
    want = {1: 0, 2: 0, 3: 0, 4: 0, 5: A * A / 2, 6: B * B / 2, 7: 0, 8: C * C / 2,
            9: A * B * c_ / 2, 12: 0, 13: 0, 15: 0, 16: 0, 18: 0, 19: 0,
            21: 3 * A ** 4 / 8, 22: 3 * B ** 4 / 8,
            24: -A * C * s_ / dx,                        # p dudx
            25: 0, 26: 0,
            27: -C * B * s_ * c_ / dx,                   # p@node (dvdx)
            28: -2 * A * A * s_ * s_ / dx ** 2,          # ux d2ux/dx2
            29: -2 * B * B * s_ * s_ / dx ** 2,          # uy d2uy/dx2
            30: 0,
            31: A * B * s_ * math.sin(2 * th) / dx ** 2,  # dudx(centred) dvdx
            32: 0, 33: 0, 34: 0,
            35: 2 * B * B * s_ * s_ / dx ** 2}           # (dvdx)^2
    scale = 10 * (1 + 1 / dx ** 2)
    for col, w in want.items():
        np.testing.assert_allclose(s[col - 1, :ny], w, rtol=1e-12, atol=1e-13 * scale,
                                   err_msg=f"column {col}")
    o.close()


def test_stats_z_modes(oracle_mod):
    """ux = A sin(k zc), uy = B sin(k zc), uz = D cos(k z_face), p = C sin(k zc)"""
    O = oracle_mod
    o = O.Oracle(CASE)
    G = _grids(O, o)
    m = 2
    k = 2 * math.pi * m / CASE["zlz"]
    th = k * G["dz"] / 2
    s_, c_ = math.sin(th), math.cos(th)
    dz = G["dz"]
    A, B, C, D = 1.1, 0.7, 0.5, 0.9
    zc, zf = G["zc"], G["zf"]
    _set(o, O, _bcast(o, fz=A * np.sin(k * zc)), _bcast(o, fz=B * np.sin(k * zc)),
         _bcast(o, fz=D * np.cos(k * zf)), _bcast(o, fz=C * np.sin(k * zc)))
    s = _stats(O, o)
    ny = o.nyc
    want = {1: 0, 3: 0, 5: A * A / 2, 7: D * D / 2, 9: A * B / 2, 10: 0, 11: 0, 14: 0,
            17: 0, 20: 0, 23: 3 * D ** 4 / 8,
            26: -C * D * s_ / dz,                            # p dwdz
            30: -2 * D * D * s_ * s_ / dz ** 2,              # uz d2uz/dz2
            32: A * B * math.sin(2 * th) ** 2 / (2 * dz ** 2),  # dudz dvdz at the edge
            33: 2 * B * B * s_ * s_ / dz ** 2,               # (dvdz)^2
            34: 2 * A * A * s_ * s_ / dz ** 2,               # (dudz)^2
            35: 0}
    scale = 10 * (1 + 1 / dz ** 2)
    for col, w in want.items():
        np.testing.assert_allclose(s[col - 1, :ny], w, rtol=1e-12, atol=1e-13 * scale,
                                   err_msg=f"column {col}")
    # uz is cos(k z_face): its cell value (uz(k)+uz(k+1))/2 = D cos(t) cos(k zc)
    np.testing.assert_allclose(s[9, :ny], 0, atol=1e-13)
    o.close()


def test_stats_cross_terms(oracle_mod):
    """vorticity components whose two terms are correlated, and the cell
    interpolation inside the triple products:
    ux = A yc cos(k x_f) + E sin(kx x_f + kz zc), uy = B sin(k xc),
    uz = D yc cos(kz z_f) + F sin(kx xc + kz z_f), p = C sin(k xc)"""
    O = oracle_mod
    o = O.Oracle(CASE)
    G = _grids(O, o)
    kx = 2 * math.pi * 2 / CASE["xlx"]
    kz = 2 * math.pi * 1 / CASE["zlz"]
    dx, dz = G["dx"], G["dz"]
    tx, tz = kx * dx / 2, kz * dz / 2
    sx, sz = math.sin(tx), math.sin(tz)
    yc = G["yc"]
    A, B, C = 0.9, 0.6, 0.4
    ny = o.nyc
    x_f, xc, z_f, zc = G["xf"], G["xc"], G["zf"], G["zc"]

    def y(v):
        return _bcast(o, fy=v)

    # (1) vor_z = dvdx - dudy = (2 B sx/dx - A) cos(kx x_f); p at the node
    # is C cos(tx) sin(kx x_f), uncorrelated with vor_z terms
    _set(o, O, y(A * yc) * _bcast(o, fx=np.cos(kx * x_f)), _bcast(o, fx=B * np.sin(kx * xc)),
         0 * _bcast(o), _bcast(o, fx=C * np.sin(kx * xc)))
    s = _stats(O, o)
    np.testing.assert_allclose(s[34, :ny], (2 * B * sx / dx - A) ** 2 / 2, rtol=1e-11)
    np.testing.assert_allclose(s[26, :ny], 0, atol=1e-12 * (A + B / dx))
    o.close()

    # (2) vor_x = dwdy - dvdz = (D - 2 B sz/dz) cos(kz z_f)
    o = O.Oracle(CASE)
    D = 1.2
    _set(o, O, 0 * _bcast(o), _bcast(o, fz=B * np.sin(kz * zc)),
         y(D * yc) * _bcast(o, fz=np.cos(kz * z_f)), 0 * _bcast(o))
    s = _stats(O, o)
    np.testing.assert_allclose(s[32, :ny], (D - 2 * B * sz / dz) ** 2 / 2, rtol=1e-11)
    o.close()

    # (3) vor_y = dudz - dwdx = (2 E sz/dz - 2 F sx/dx) cos(kx x_f + kz z_f)
    o = O.Oracle(CASE)
    E, F = 0.8, 0.5
    ux = E * np.sin(kx * x_f[:, None] + kz * zc[None, :])
    uz = F * np.sin(kx * xc[:, None] + kz * z_f[None, :])
    _set(o, O, ux[:, None, :] * np.ones((1, ny + 3, 1)), 0 * _bcast(o),
         uz[:, None, :] * np.ones((1, ny + 3, 1)), 0 * _bcast(o))
    s = _stats(O, o)
    np.testing.assert_allclose(s[33, :ny], (2 * E * sz / dz - 2 * F * sx / dx) ** 2 / 2,
                               rtol=1e-11)
    o.close()

    # (4) triple products with the cell values: ux = A cos(k x_f), uy = B cos(2k xc)
    # uxCell = A cos(t) cos(k xc): mean(uxCell^2 uyCell) = A^2 B cos(t)^2 / 4
    o = O.Oracle(CASE)
    _set(o, O, _bcast(o, fx=A * np.cos(kx * x_f)), _bcast(o, fx=B * np.cos(2 * kx * xc)),
         _bcast(o, fz=A * np.cos(kz * z_f)), 0 * _bcast(o))
    s = _stats(O, o)
    np.testing.assert_allclose(s[14, :ny], A * A * B * math.cos(tx) ** 2 / 4, rtol=1e-11)
    # uzCell = A cos(tz) cos(kz zc), uyCell independent of z: mean(uzCell^2 uyCell) = 0
    # and mean(uzCell^2) enters only column 17 through uyCell's zero mean
    np.testing.assert_allclose(s[16, :ny], 0, atol=1e-13)
    np.testing.assert_allclose(s[10, :ny], 0, atol=1e-13)
    o.close()
