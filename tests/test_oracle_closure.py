"""Pins of the closure oracle (oracle/closure.py): box filter, SGS stress,
LES-grid gradients, network forward/backward, Adam and the dynamic
Smagorinsky comparator, each against a closed form, a brute-force loop or
finite differences.  No reference counterpart exists for any of it.
"""
import numpy as np
import pytest

from oracle import closure as C


def _yp(ny, gamma=1.5, yly=2.0):
    from channel3d_amd.channel import channel_yp
    return channel_yp(ny, yly, gamma)


def test_box_filter_closed_forms():
    yp = _yp(8)
    dy = np.diff(yp[1:])
    yc = 0.5 * (yp[1:-1] + yp[2:])
    nx, ny, nz = 8, 8, 4
    f = np.full((nx, ny, nz), 3.25)
    assert np.allclose(C.box(f, dy, 4, 2, 2), 3.25, rtol=0, atol=1e-15)
    # a linear function of y: the dy-weighted block mean is its value at the
    # block's geometric centre (midpoint rule exact for linear functions)
    f = np.broadcast_to(2.0 * yc[None, :, None] - 0.5, (nx, ny, nz)).copy()
    dY, Yc = C.les_grid(yp, 2)
    assert np.allclose(C.box(f, dy, 4, 2, 2), (2.0 * Yc - 0.5)[None, :, None], atol=1e-14)
    assert np.allclose(dY.sum(), 2.0) and np.all(dY > 0)


def test_sgs_stress_is_block_covariance():
    """tau_ab = bar(u_a u_b) - bar(u_a) bar(u_b) equals the dy-weighted
    covariance within each block, computed by brute force"""
    rng = np.random.default_rng(3)
    nx, ny, nz, fx, fy, fz = 8, 8, 8, 4, 2, 4
    yp = _yp(ny)
    ux, uy, uz = (rng.standard_normal((nx + 3, ny + 3, nz + 3)) for _ in range(3))
    ub, tau = C.filtered(ux, uy, uz, yp, fx, fy, fz)
    u = C.cell_centred(ux, uy, uz)
    dy = np.diff(yp[1:])
    for I in range(nx // fx):
        for J in range(ny // fy):
            for K in range(nz // fz):
                sl = (slice(I * fx, I * fx + fx), slice(J * fy, J * fy + fy),
                      slice(K * fz, K * fz + fz))
                w = np.broadcast_to(dy[None, J * fy:J * fy + fy, None], (fx, fy, fz))
                w = w / w.sum()
                m = [(w * c[sl]).sum() for c in u]
                for q, (a, b) in enumerate(C.PAIRS):
                    cov = (w * (u[a][sl] - m[a]) * (u[b][sl] - m[b])).sum()
                    assert abs(tau[q][I, J, K] - cov) < 1e-13
                    assert abs(ub[a][I, J, K] - m[a]) < 1e-14
    assert all((tau[q] >= -1e-15).all() for q in range(3))
    # cell centring: face averages of the staggered fields
    assert np.allclose(u[0][2, 3, 1], 0.5 * (ux[3, 4, 2] + ux[4, 4, 2]))
    assert np.allclose(u[1][2, 3, 1], 0.5 * (uy[3, 4, 2] + uy[3, 5, 2]))
    assert np.allclose(u[2][2, 3, nz - 1], 0.5 * (uz[3, 4, nz] + uz[3, 4, 1]))


def test_gradients_and_strain():
    nI, nJ, nK = 16, 6, 8
    dX, dZ = 0.3, 0.2
    yc = np.cumsum(np.linspace(0.1, 0.4, nJ))
    kx = 2 * np.pi * 3 / (nI * dX)
    x = (np.arange(nI) + 0.5) * dX
    f0 = np.sin(kx * x)[:, None, None] + 0 * yc[None, :, None] + np.zeros((1, 1, nK))
    f1 = np.broadcast_to(1.5 * yc[None, :, None] + 0.2, (nI, nJ, nK)).copy()
    f2 = np.zeros((nI, nJ, nK))
    g = C.gradients([f0, f1, f2], dX, dZ, yc)
    # central difference of a Fourier mode: sin(k dX)/dX cos(k x)
    assert np.allclose(g[0][0], (np.sin(kx * dX) / dX * np.cos(kx * x))[:, None, None], atol=1e-12)
    # linear in y: exact, one-sided rows included, on the non-uniform centres
    assert np.allclose(g[1][1], 1.5, atol=1e-12)
    assert np.allclose(g[0][1], 0.0, atol=1e-12) and np.allclose(g[2][2], 0.0)
    # pure shear s: |S| = |s|
    z = np.zeros((2, 2, 2))
    gs = [[z, z + 0.7, z], [z, z, z], [z, z, z]]
    assert np.allclose(C.strain_norm(gs), 0.7)


def test_samples_layout():
    rng = np.random.default_rng(5)
    nx, ny, nz = 16, 8, 8
    yp = _yp(ny)
    ux, uy, uz = (rng.standard_normal((nx + 3, ny + 3, nz + 3)) for _ in range(3))
    X, T, scale = C.samples(ux, uy, uz, yp, 0.1, 0.1, 2.0)
    n = (nx // 4) * (ny // 2) * (nz // 4)
    assert X.shape == (n, 16) and T.shape == (n, 6) and scale.shape == (n,)
    assert np.all(X[:, 10:] == 0) and np.all(np.abs(X[:, :9]) <= C.CLIP)
    assert np.all((X[:, 9] > 0) & (X[:, 9] <= 1))
    assert np.all(np.abs(T) <= C.CLIP) and np.all(scale > 0)
    # the first LES row is nearest the lower wall; x runs fastest
    assert X[0, 9] == X[1, 9] and X[0, 9] < X[nx // 4, 9]


def test_bf16_rounding():
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, -2.5, 1 + 2 ** -8 + 2 ** -20],
                 dtype=np.float32)
    r = C.bf16(v)
    assert list(r) == [1.0, 1.0, 1 + 2 ** -7, -2.5, 1 + 2 ** -7]


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = {k: v.astype(np.float64) for k, v in C.init_params(4).items()}
    X = np.zeros((7, 16))
    X[:, :10] = rng.standard_normal((7, 10))
    T = rng.standard_normal((7, 6))
    g = C.backward(p, X, T, emulate=False)
    h = 1e-6
    for k in ("W1", "W2", "W3", "W4", "b1", "b2", "b4"):
        for _ in range(6):
            idx = tuple(rng.integers(0, s) for s in p[k].shape)
            if k == "W1" and idx[0] >= 10:
                continue
            p[k][idx] += h
            lp = C.loss(p, X, T, emulate=False)
            p[k][idx] -= 2 * h
            lm = C.loss(p, X, T, emulate=False)
            p[k][idx] += h
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[k][idx]) <= 1e-7 * max(1.0, abs(fd)), (k, idx, fd, g[k][idx])


def test_emulated_forward_is_bf16_close():
    rng = np.random.default_rng(2)
    p = C.init_params(0)
    X = np.zeros((50, 16), dtype=np.float32)
    X[:, :10] = rng.standard_normal((50, 10))
    y32, _ = C.forward(p, X, emulate=True)
    y64, _ = C.forward({k: v.astype(np.float64) for k, v in p.items()}, X, emulate=False)
    err = np.abs(y32 - y64).max() / np.abs(y64).max()
    assert 1e-6 < err < 3e-2


def test_adam_first_step_is_sign_step():
    p = C.init_params(0)
    q = {k: v.copy() for k, v in p.items()}
    grads = {k: np.full_like(v, -0.3) for k, v in p.items()}
    C.adam(q, grads, {}, lr=1e-3)
    for k in p:
        d = q[k] - p[k]
        if k == "W1":
            assert np.all(q[k][10:] == 0)
            d = d[:10]
        assert np.allclose(d, 1e-3, rtol=1e-4)


def test_dynamic_smagorinsky_least_squares():
    """Cs^2(J) minimises the plane sum of (L^d - Cs^2 M)^2 (weights 1 on the
    diagonal, 2 off it); a field without x-z variation gives L = 0, Cs^2 = 0"""
    rng = np.random.default_rng(9)
    nI, nJ, nK = 12, 6, 10
    dX, dZ = 0.2, 0.15
    yc = np.cumsum(np.linspace(0.05, 0.3, nJ))
    dY = np.full(nJ, 0.1)
    ub = [rng.standard_normal((nI, nJ, nK)) for _ in range(3)]
    tau, cs2 = C.dynamic_smagorinsky(ub, dX, dY, dZ, yc)
    # recompute L and M as in the model, then check optimality per plane
    g = C.gradients(ub, dX, dZ, yc)
    S = {(a, b): 0.5 * (g[a][b] + g[b][a]) for a in range(3) for b in range(3)}
    sn = C.strain_norm(g)
    tub = [C.test_filter(u) for u in ub]
    tg = C.gradients(tub, dX, dZ, yc)
    tS = {(a, b): 0.5 * (tg[a][b] + tg[b][a]) for a in range(3) for b in range(3)}
    tsn = C.strain_norm(tg)
    d2 = ((dX * dY * dZ) ** (2 / 3))[None, :, None]
    Lm = {ab: C.test_filter(ub[ab[0]] * ub[ab[1]]) - tub[ab[0]] * tub[ab[1]] for ab in C.PAIRS}
    tr = (Lm[0, 0] + Lm[1, 1] + Lm[2, 2]) / 3
    for a in range(3):
        Lm[a, a] = Lm[a, a] - tr
    Mm = {ab: -2 * d2 * (4 * tsn * tS[ab] - C.test_filter(sn * S[ab])) for ab in C.PAIRS}

    def resid(J, c):
        return sum((1 if a == b else 2) * ((Lm[a, b] - c * Mm[a, b])[:, J, :] ** 2).sum()
                   for a, b in C.PAIRS)
    for J in range(nJ):
        r0 = resid(J, cs2[J])
        assert r0 <= resid(J, cs2[J] * 1.01 + 1e-12) and r0 <= resid(J, cs2[J] * 0.99 - 1e-12)
    # the model stress is -2 Cs^2 Delta^2 |S| S
    assert np.allclose(tau[3], -2 * cs2[None, :, None] * d2 * sn * S[0, 1])
    # no x-z variation: no resolved test-scale stress
    flat = [np.broadcast_to(rng.standard_normal(nJ)[None, :, None], (nI, nJ, nK)).copy()
            for _ in range(3)]
    _, c0 = C.dynamic_smagorinsky(flat, dX, dY, dZ, yc)
    assert np.allclose(c0, 0.0, atol=1e-20)


def test_correlation():
    a = [np.arange(10.0)]
    assert C.correlation(a, [2 * a[0] + 1]) == pytest.approx(1.0)
    assert C.correlation(a, [-a[0]]) == pytest.approx(-1.0)
