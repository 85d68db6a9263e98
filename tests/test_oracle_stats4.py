"""Pins of the 4th-order oracle's statistics (clcStat of src/CFD_4th/
m_FlowCase.f90:306-633: the 35 columns and the 98 HighOrderGradStat
moments) against closed forms on the discrete grid.

* constant fields: products of constants (ux shifted by uCRF), derivative
  columns and moments zero;
* fields linear in y: the staggered y differences, the YinterpCoe node
  interpolation and the vorticities are exact on the stretched grid, and
  the x/z stencils (coefficients summing to 0 for derivatives, 1 for the
  interpolation) drop out;
* single Fourier modes: the 4th-order stencils act on exp(ikx) through
  their symbols, with t = k dx / 2,
    interpolation (-1, 9, 9, -1)/16       ->  (9 cos t - cos 3t) / 8
    derivative (1, -27, 27, -1)/(24 dx)   ->  i (27 sin t - sin 3t) / (12 dx)
    second derivative (-1,16,-30,16,-1)/(12 dx^2)
                                          ->  (-cos 4t + 16 cos 2t - 15) / (6 dx^2)
    centred (-1, 8, 0, -8, 1)/(12 dx)     ->  i (8 sin 2t - sin 4t) / (6 dx)
  and plane means of sin^2 = cos^2 = 1/2, sin^4 = 3/8, odd powers 0.
"""
import math

import numpy as np
import pytest

CASE = dict(nxc=32, nyc=24, nzc=24, xlx=4 * math.pi, yly=2.0, zlz=4 * math.pi / 3,
            xnu=3.4969e-4, dt=0.005, gamma=1.5, is_ux_const=False)


def _run(O, o, grad=True, acc=None, gacc=None):
    ny = o.nyc
    s = np.zeros((35, ny + 1), order="F") if acc is None else acc
    q = (np.zeros((98, ny + 1), order="F") if gacc is None else gacc) if grad else None
    o.call("clcStat", o.field(O.UX), o.field(O.UY), o.field(O.UZ), o.field(O.PR), s, q)
    return s, q


def _field(o, fx=None, fy=None, fz=None):
    """a 4th-order halo array (nx+4, ny+2, nz+4) from separable factors"""
    a = np.ones((o.nxc + 4, o.nyc + 2, o.nzc + 4))
    if fx is not None:
        a = a * fx[:, None, None]
    if fy is not None:
        a = a * fy[None, :o.nyc + 2, None]
    if fz is not None:
        a = a * fz[None, None, :]
    return np.asfortranarray(a)


def _grid(O, o):
    nx, nz = o.nxc, o.nzc
    dx, dz = CASE["xlx"] / nx, CASE["zlz"] / nz
    i = np.arange(-1, nx + 3)            # halo index -1 .. nx+2
    k = np.arange(-1, nz + 3)
    return dict(xf=(i - 1) * dx, xc=(i - 0.5) * dx, zf=(k - 1) * dz, zc=(k - 0.5) * dz,
                yp=o.metric(O.M_YP, 0, o.nyc + 1), yc=o.metric(O.M_YC, 0, o.nyc + 1),
                dx=dx, dz=dz)


def _set(o, O, ux, uy, uz, p):
    for fid, a in ((O.UX, ux), (O.UY, uy), (O.UZ, uz), (O.PR, p)):
        o.field(fid)[...] = a


def test_stats4_constant_fields(oracle_mod):
    O = oracle_mod
    crf = 0.25
    o = O.Oracle4(dict(CASE, u_crf=crf))
    U, V, W, P = 2.0, 3.0, 5.0, 7.0
    one = _field(o)
    _set(o, O, U * one, V * one, W * one, P * one)
    s, q = _run(O, o)
    Ut = U + crf        # statistics see ux + uCRF (:352, :356)
    want = [Ut, V, W, P, Ut * Ut, V * V, W * W, P * P, Ut * V, V * W, Ut * W, Ut * P, V * P,
            W * P, Ut * Ut * V, V ** 3, W * W * V, Ut * V * V, Ut ** 3, W ** 3, Ut ** 4, V ** 4,
            W ** 4]
    ny = o.nyc
    for m, w in enumerate(want):
        np.testing.assert_allclose(s[m, :ny], w, rtol=1e-14, err_msg=f"column {m + 1}")
    scale = 1e-12 * np.abs(o.metric(O.M_AC2C, 1, ny)).max() * 100.0
    assert np.abs(s[23:, :ny]).max() <= scale
    np.testing.assert_allclose(q[0, :ny], P ** 3, rtol=1e-14)
    np.testing.assert_allclose(q[1, :ny], P ** 4, rtol=1e-14)
    assert np.abs(q[2:, :ny]).max() <= scale ** 4 + scale
    o.close()


@pytest.mark.parametrize("P0,PI", [(7.0, 0.0), (0.0, 1.3)])
def test_stats4_linear_in_y(oracle_mod, P0, PI):
    O = oracle_mod
    o = O.Oracle4(CASE)
    G = _grid(O, o)
    yc, yp = G["yc"], G["yp"]
    a, u0, b, g = 0.7, 0.4, -0.3, 1.9
    ny = o.nyc
    _set(o, O, _field(o, fy=a * yc + u0), _field(o, fy=b * yp), _field(o, fy=g * yc),
         _field(o, fy=P0 + PI * yc))
    s, q = _run(O, o)
    j = np.arange(1, ny + 1)
    ux, uy, uz, p = a * yc[j] + u0, b * yp[j], g * yc[j], P0 + PI * yc[j]
    uyc = b * yc[j]
    pnode = P0 + PI * yp[j]
    z = 0 * j
    want = {1: ux, 2: uy, 3: uz, 4: p, 5: ux ** 2, 6: uy ** 2, 7: uz ** 2, 8: p ** 2,
            9: ux * uyc, 10: uyc * uz, 11: ux * uz, 12: ux * p, 13: uyc * p, 14: uz * p,
            15: ux * ux * uyc, 16: uy ** 3, 17: uz * uz * uyc, 18: ux * uyc * uyc,
            19: ux ** 3, 20: uz ** 3, 21: ux ** 4, 22: uy ** 4, 23: uz ** 4,
            24: z, 25: p * b, 26: z, 27: pnode * a,
            31: z, 32: z, 33: g * g + z, 34: z, 35: a * a + z}
    # column 31 for 4th order: (dudxC (dvdxU + dvdx) + (dudyU + dudy)(uy_xp - uy_xc) caj)/2,
    # uy_xp - uy_xc = b dyp, caj = 1/dyp: (2a)(b)/2 = a b
    want[31] = a * b + z
    for col, w in want.items():
        np.testing.assert_allclose(s[col - 1, :ny], w, rtol=1e-11, atol=1e-11,
                                   err_msg=f"column {col}")
    # gradient moments: dudy = a, dvdy = b, dwdy = g, dpdy = PI everywhere
    for base, v in ((7, a), (19, b), (31, g), (43, PI)):
        for m in range(4):
            np.testing.assert_allclose(q[base - 1 + m, :ny], v ** (m + 1), rtol=1e-10,
                                       atol=1e-11, err_msg=f"moment {base + m}")
    for base in (3, 11, 15, 23, 27, 35, 39, 47):          # x and z derivatives
        assert np.abs(q[base - 1:base + 3, :ny]).max() < 1e-11
    # plane nyp: columns 27, 33, 35 and the dudy, dwdy moments
    np.testing.assert_allclose(s[[32, 34], ny], [g * g, a * a], rtol=1e-11)
    np.testing.assert_allclose(q[6:10, ny], [a, a ** 2, a ** 3, a ** 4], rtol=1e-10)
    np.testing.assert_allclose(q[30:34, ny], [g, g ** 2, g ** 3, g ** 4], rtol=1e-10)
    if PI == 0.0:
        np.testing.assert_allclose(s[26, ny], P0 * a, rtol=1e-12)
    o.close()


def test_stats4_x_modes(oracle_mod):
    """ux = A cos(k x_f) + U0, uy = B sin(k xc), p = C sin(k xc), uz = 0"""
    O = oracle_mod
    o = O.Oracle4(CASE)
    G = _grid(O, o)
    k = 2 * math.pi * 3 / CASE["xlx"]
    dx = G["dx"]
    t = k * dx / 2
    Si = (9 * math.cos(t) - math.cos(3 * t)) / 8                 # interpolation
    Sd = (27 * math.sin(t) - math.sin(3 * t)) / (12 * dx)       # staggered derivative
    S2 = (-math.cos(4 * t) + 16 * math.cos(2 * t) - 15) / (6 * dx * dx)
    A, B, C, U0 = 1.3, 0.8, 0.6, 0.5
    ny = o.nyc
    _set(o, O, _field(o, fx=A * np.cos(k * G["xf"]) + U0), _field(o, fx=B * np.sin(k * G["xc"])),
         0 * _field(o), _field(o, fx=C * np.sin(k * G["xc"])))
    s, q = _run(O, o)
    # dudx = -A Sd sin(k xc); dvdx = B Sd cos(k xf); uxCell = U0 + A Si cos(k xc)
    # p at the node (x_f) = C Si sin(k xf); dudxx = A S2 cos(k xf); dudxC = -A Sc sin(k xf)
    # dvdxU = dvdx (y-independent); dudy = 0; uy_xp = uy_xc
    # column 9 uxCell*uyCell: U0 B <sin> + A Si B <cos sin> = 0; column 27:
    # p_node dvdx ~ <sin cos> = 0; column 31: dudxC (2 dvdx) ~ <sin cos> = 0
    want = {5: U0 * U0 + A * A / 2, 9: 0.0,
            21: U0 ** 4 + 3 * U0 ** 2 * A ** 2 + 3 * A ** 4 / 8,
            24: -A * Sd * C / 2, 27: 0.0, 28: A * A * S2 / 2, 31: 0.0,
            35: B * B * Sd * Sd / 2}
    scale = 10 * (1 + 1 / dx ** 2)
    for col, w in want.items():
        np.testing.assert_allclose(s[col - 1, :ny], w, rtol=1e-12, atol=1e-13 * scale,
                                   err_msg=f"column {col}")
    # gradient moments of dudx, dvdx, dudxx, dvdxx, dpdx (amplitude a: 0, a^2/2, 0, 3a^4/8)
    for base, amp in ((3, A * Sd), (15, B * Sd), (51, A * S2), (63, B * S2), (39, C * Sd)):
        np.testing.assert_allclose(q[base - 1:base + 3, :ny].T,
                                   np.tile([0, amp ** 2 / 2, 0, 3 * amp ** 4 / 8], (ny, 1)),
                                   rtol=1e-11, atol=1e-11 * (1 + amp ** 4), err_msg=f"moment {base}")
    o.close()


def test_stats4_z_modes(oracle_mod):
    """ux = A sin(k zc), uy = B sin(k zc), uz = D cos(k z_f)"""
    O = oracle_mod
    o = O.Oracle4(CASE)
    G = _grid(O, o)
    k = 2 * math.pi * 2 / CASE["zlz"]
    dz = G["dz"]
    t = k * dz / 2
    Si = (9 * math.cos(t) - math.cos(3 * t)) / 8
    Sd = (27 * math.sin(t) - math.sin(3 * t)) / (12 * dz)
    S2 = (-math.cos(4 * t) + 16 * math.cos(2 * t) - 15) / (6 * dz * dz)
    Sc = (8 * math.sin(2 * t) - math.sin(4 * t)) / (6 * dz)
    A, B, D = 1.1, 0.7, 0.9
    ny = o.nyc
    zc, zf = G["zc"], G["zf"]
    _set(o, O, _field(o, fz=A * np.sin(k * zc)), _field(o, fz=B * np.sin(k * zc)),
         _field(o, fz=D * np.cos(k * zf)), 0 * _field(o))
    s, q = _run(O, o)
    # dudz = A Sd cos(k zf), dvdz = B Sd cos(k zf), dwdz = -D Sd sin(k zc)
    # uzCell = D Si cos(k zc); dwdzz = D S2 cos(k zf)
    # dudzC = A Sc cos(k zc), dvdzC = B Sc cos(k zc) (x-independent interpolation weights sum to 1)
    want = {7: D * D / 2, 9: A * B / 2, 11: 0, 30: D * D * S2 / 2,
            32: A * B * Sc * Sc / 2, 33: B * B * Sd * Sd / 2, 34: A * A * Sd * Sd / 2}
    scale = 10 * (1 + 1 / dz ** 2)
    for col, w in want.items():
        np.testing.assert_allclose(s[col - 1, :ny], w, rtol=1e-12, atol=1e-13 * scale,
                                   err_msg=f"column {col}")
    for base, amp in ((11, A * Sd), (23, B * Sd), (35, D * Sd), (83, D * S2), (59, A * S2)):
        np.testing.assert_allclose(q[base - 1:base + 3, :ny].T,
                                   np.tile([0, amp ** 2 / 2, 0, 3 * amp ** 4 / 8], (ny, 1)),
                                   rtol=1e-11, atol=1e-11 * (1 + amp ** 4), err_msg=f"moment {base}")
    o.close()


def test_stats4_vorticity_cross_terms(oracle_mod):
    """vorticity components whose two terms are correlated"""
    O = oracle_mod
    ny = CASE["nyc"]
    kx = 2 * math.pi * 2 / CASE["xlx"]
    kz = 2 * math.pi * 1 / CASE["zlz"]

    def symbols(k, d):
        t = k * d / 2
        return (27 * math.sin(t) - math.sin(3 * t)) / (12 * d)

    # vor_z = dvdx - dudy: ux = A yc cos(kx x_f), uy = B sin(kx xc)
    o = O.Oracle4(CASE)
    G = _grid(O, o)
    Sdx, Sdz = symbols(kx, G["dx"]), symbols(kz, G["dz"])
    A, B, D, E, F = 0.9, 0.6, 1.2, 0.8, 0.5
    _set(o, O, _field(o, fx=np.cos(kx * G["xf"]), fy=A * G["yc"]),
         _field(o, fx=B * np.sin(kx * G["xc"])), 0 * _field(o), 0 * _field(o))
    s, _ = _run(O, o, grad=False)
    np.testing.assert_allclose(s[34, :ny], (B * Sdx - A) ** 2 / 2, rtol=1e-11)
    o.close()
    # vor_x = dwdy - dvdz: uz = D yc cos(kz z_f), uy = B sin(kz zc)
    o = O.Oracle4(CASE)
    _set(o, O, 0 * _field(o), _field(o, fz=B * np.sin(kz * G["zc"])),
         _field(o, fz=np.cos(kz * G["zf"]), fy=D * G["yc"]), 0 * _field(o))
    s, _ = _run(O, o, grad=False)
    np.testing.assert_allclose(s[32, :ny], (D - B * Sdz) ** 2 / 2, rtol=1e-11)
    o.close()
    # vor_y = dudz - dwdx: ux = E sin(kx x_f + kz zc), uz = F sin(kx xc + kz z_f)
    o = O.Oracle4(CASE)
    ux = E * np.sin(kx * G["xf"][:, None] + kz * G["zc"][None, :])
    uz = F * np.sin(kx * G["xc"][:, None] + kz * G["zf"][None, :])
    one_y = np.ones((1, ny + 2, 1))
    _set(o, O, np.asfortranarray(ux[:, None, :] * one_y), 0 * _field(o),
         np.asfortranarray(uz[:, None, :] * one_y), 0 * _field(o))
    s, _ = _run(O, o, grad=False)
    np.testing.assert_allclose(s[33, :ny], (E * Sdz - F * Sdx) ** 2 / 2, rtol=1e-11)
    o.close()
