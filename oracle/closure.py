"""CPU restatement of the learned subgrid-stress closure's data path and
network, and of the dynamic Smagorinsky comparator.  Test infrastructure
only: the tests and the benchmark's comparator use it, the library never
does.

None of this has a counterpart in the reference (it has WALE and MTS only,
src/CFD_2nd/m_LESModels.f90:53-58, no Smagorinsky, no dynamic model, no
network).  Everything below is written from textbook definitions and pinned
by closed forms and finite differences in tests/test_oracle_closure.py:

* DNS -> LES: cell-centred velocity (face averages of the staggered fields),
  top-hat box filter over blocks of fx x fy x fz cells, volume-weighted in
  the stretched y direction; SGS stress tau_ab = bar(u_a u_b) - bar(u_a)
  bar(u_b) (Leonard/Germano definition).
* LES-grid gradients: central differences in the periodic x and z, central
  differences over the block centres in y, one-sided in the first and last
  block rows.
* Network inputs: the 9 gradient components over |S| = sqrt(2 S:S), the
  wall distance over the half height (y+/Re_tau); padded to 16.  Targets:
  tau_ab / (Delta^2 |S|^2), Delta = (dX dY dZ)^(1/3).  Inputs and targets are
  clipped to +-CLIP.
* MLP 10 -> 64 -> 64 -> 64 -> 6, SiLU hidden layers, linear output, loss the
  mean square error over samples and the 6 outputs.  The GPU multiplies
  bf16 operands with FP32 accumulation; `forward`/`backward` round the same
  operands to bf16 (round to nearest even) so the two agree to FP32
  summation-order rounding.
* Dynamic Smagorinsky (Germano, Piomelli, Moin & Cabot 1991, with Lilly's
  1992 least squares, averaged over x-z planes): test filter (1/4, 1/2,
  1/4) in x and z, width ratio alpha = 2.
"""
import numpy as np

NIN, NH, NOUT, NPAD = 10, 64, 6, 16
CLIP = 10.0
PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))   # tau_11 22 33 12 13 23


# ------------------------------------------------------------ DNS -> LES --
def cell_centred(ux, uy, uz):
    """halo arrays (nx+3, ny+3, nz+3), reference indexing (interior 1..n)
    -> cell-centred (nx, ny, nz) velocities"""
    u = 0.5 * (ux[1:-2, 1:-2, 1:-2] + np.roll(ux[1:-2, 1:-2, 1:-2], -1, axis=0))
    v = 0.5 * (uy[1:-2, 1:-2, 1:-2] + uy[1:-2, 2:-1, 1:-2])
    w = 0.5 * (uz[1:-2, 1:-2, 1:-2] + np.roll(uz[1:-2, 1:-2, 1:-2], -1, axis=2))
    return u, v, w


def box(f, dy, fx, fy, fz):
    """top-hat filter of a cell field (nx, ny, nz) over fx x fy x fz blocks,
    weighted by the cell heights dy (ny,)"""
    nx, ny, nz = f.shape
    g = (f * dy[None, :, None]).reshape(nx // fx, fx, ny // fy, fy, nz // fz, fz)
    h = dy.reshape(ny // fy, fy).sum(axis=1)
    return g.sum(axis=(1, 3, 5)) / (fx * fz) / h[None, :, None]


def les_grid(yp, fy):
    """block heights and centres of the LES rows from the DNS nodes yp
    (1-based nodes yp[1..ny+1] as the reference stores them)"""
    nodes = np.asarray(yp[1:], dtype=np.float64)            # y of nodes 1..ny+1
    edges = nodes[::fy]
    return np.diff(edges), 0.5 * (edges[1:] + edges[:-1])


def filtered(ux, uy, uz, yp, fx, fy, fz):
    """(ubar[3], tau[6]) on the LES grid"""
    u = cell_centred(ux, uy, uz)
    dy = np.diff(np.asarray(yp[1:], dtype=np.float64))
    ub = [box(c, dy, fx, fy, fz) for c in u]
    tau = [box(u[a] * u[b], dy, fx, fy, fz) - ub[a] * ub[b] for a, b in PAIRS]
    return ub, tau


def gradients(ub, dX, dZ, yc):
    """g[a][b] = d ubar_a / d x_b on the LES grid (x, z periodic)"""
    g = [[None] * 3 for _ in range(3)]
    for a in range(3):
        f = ub[a]
        g[a][0] = (np.roll(f, -1, 0) - np.roll(f, 1, 0)) / (2.0 * dX)
        g[a][2] = (np.roll(f, -1, 2) - np.roll(f, 1, 2)) / (2.0 * dZ)
        d = np.empty_like(f)
        d[:, 1:-1, :] = (f[:, 2:, :] - f[:, :-2, :]) / (yc[2:] - yc[:-2])[None, :, None]
        d[:, 0, :] = (f[:, 1, :] - f[:, 0, :]) / (yc[1] - yc[0])
        d[:, -1, :] = (f[:, -1, :] - f[:, -2, :]) / (yc[-1] - yc[-2])
        g[a][1] = d
    return g


def strain_norm(g):
    s2 = 0.0
    for a in range(3):
        for b in range(3):
            s = 0.5 * (g[a][b] + g[b][a])
            s2 = s2 + s * s
    return np.sqrt(2.0 * s2)


def samples(ux, uy, uz, yp, dx, dz, yly, fx=4, fy=2, fz=4):
    """network inputs X (n, 16), targets T (n, 6) and the stress scale
    Delta^2 |S|^2 (n,), LES cells in x-fastest order (I, J, K)"""
    ub, tau = filtered(ux, uy, uz, yp, fx, fy, fz)
    dY, yc = les_grid(yp, fy)
    dX, dZ = fx * dx, fz * dz
    g = gradients(ub, dX, dZ, yc)
    sn = strain_norm(g)
    inv = 1.0 / np.maximum(sn, 1e-30)
    nI, nJ, nK = ub[0].shape
    n = nI * nJ * nK
    X = np.zeros((n, NPAD))
    q = 0
    for a in range(3):
        for b in range(3):
            X[:, q] = np.clip(g[a][b] * inv, -CLIP, CLIP).ravel(order="F")
            q += 1
    wall = np.minimum(yc, yly - yc) / (0.5 * yly)
    X[:, 9] = np.broadcast_to(wall[None, :, None], (nI, nJ, nK)).ravel(order="F")
    delta2 = (dX * dY * dZ) ** (2.0 / 3.0)
    scale = (delta2[None, :, None] * sn * sn).ravel(order="F")
    sinv = 1.0 / np.maximum(scale, 1e-30)
    T = np.stack([np.clip(t.ravel(order="F") * sinv, -CLIP, CLIP) for t in tau], axis=1)
    return X, T, scale


# ----------------------------------------------------------------- MLP --
def bf16(a):
    """round FP32 values to bf16 (nearest even), returned as FP32"""
    b = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def init_params(seed=0):
    """W1 (16, 64) (rows >= 10 zero), W2, W3 (64, 64), W4 (64, 6), biases;
    uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)), FP32 masters"""
    rng = np.random.default_rng(seed)
    shapes = [(NPAD, NH), (NH, NH), (NH, NH), (NH, NOUT)]
    fans = [NIN, NH, NH, NH]
    p = {}
    for l, (s, fan) in enumerate(zip(shapes, fans), 1):
        lim = 1.0 / np.sqrt(fan)
        w = rng.uniform(-lim, lim, size=s).astype(np.float32)
        if l == 1:
            w[NIN:, :] = 0.0
        p[f"W{l}"] = w
        p[f"b{l}"] = rng.uniform(-lim, lim, size=s[1]).astype(np.float32)
    return p


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, emulate=True):
    """returns (output, cache); emulate: bf16 operands, FP32 accumulation"""
    r = bf16 if emulate else (lambda a: a)
    dt = np.float32 if emulate else np.float64
    h = np.asarray(X, dtype=dt)
    cache = []
    for l in range(1, 5):
        a_in = r(h).astype(dt)
        z = a_in @ r(p[f"W{l}"]).astype(dt) + p[f"b{l}"].astype(dt)
        cache.append((a_in, z))
        h = silu(z) if l < 4 else z
    return h, cache


def loss(p, X, T, emulate=True):
    y, _ = forward(p, X, emulate)
    return float(np.mean((y - T) ** 2))


def backward(p, X, T, emulate=True):
    """gradients of the mean square error; the GPU forms the same products
    from bf16 operands: dH = dZ W^T with bf16 dZ and W, dW = A^T dZ with
    bf16 A (the layer input as the forward used it) and bf16 dZ"""
    r = bf16 if emulate else (lambda a: a)
    dt = np.float32 if emulate else np.float64
    y, cache = forward(p, X, emulate)
    n = X.shape[0]
    dy = (2.0 / (n * NOUT)) * (y - T.astype(dt))
    grads = {}
    dz = dy.astype(dt)
    for l in range(4, 0, -1):
        a_in, z = cache[l - 1]
        if l < 4:
            dz = (dz * silu_grad(z)).astype(dt)
        dzr = r(dz).astype(dt)
        grads[f"W{l}"] = a_in.T @ dzr
        grads[f"b{l}"] = dz.sum(axis=0)
        if l > 1:
            dz = dzr @ r(p[f"W{l}"]).astype(dt).T
    return grads


def adam(p, grads, state, lr, b1=0.9, b2=0.999, eps=1e-8):
    """one Adam step (Kingma & Ba 2015) on the FP32 masters, in place"""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    for k in p:
        m = state.setdefault("m" + k, np.zeros_like(p[k]))
        v = state.setdefault("v" + k, np.zeros_like(p[k]))
        g = grads[k].astype(np.float32)
        m[...] = b1 * m + (1 - b1) * g
        v[...] = b2 * v + (1 - b2) * g * g
        mh = m / (1 - b1 ** t)
        vh = v / (1 - b2 ** t)
        p[k] -= (lr * mh / (np.sqrt(vh) + eps)).astype(np.float32)
        if k == "W1":
            p[k][NIN:, :] = 0.0


# ------------------------------------------------- dynamic Smagorinsky --
def test_filter(f):
    """(1/4, 1/2, 1/4) in x and z (periodic)"""
    f = 0.25 * np.roll(f, 1, 0) + 0.5 * f + 0.25 * np.roll(f, -1, 0)
    return 0.25 * np.roll(f, 1, 2) + 0.5 * f + 0.25 * np.roll(f, -1, 2)


def dynamic_smagorinsky(ub, dX, dY, dZ, yc, alpha=2.0):
    """deviatoric model stress [6] and the plane coefficient Cs^2(J)"""
    g = gradients(ub, dX, dZ, yc)
    S = {(a, b): 0.5 * (g[a][b] + g[b][a]) for a in range(3) for b in range(3)}
    sn = strain_norm(g)
    tub = [test_filter(u) for u in ub]
    tg = gradients(tub, dX, dZ, yc)
    tS = {(a, b): 0.5 * (tg[a][b] + tg[b][a]) for a in range(3) for b in range(3)}
    tsn = strain_norm(tg)
    d2 = ((dX * dY * dZ) ** (2.0 / 3.0))[None, :, None]
    Lm = {}
    for a, b in PAIRS:
        Lm[a, b] = test_filter(ub[a] * ub[b]) - tub[a] * tub[b]
    tr = (Lm[0, 0] + Lm[1, 1] + Lm[2, 2]) / 3.0
    for a in range(3):
        Lm[a, a] = Lm[a, a] - tr
    num = 0.0
    den = 0.0
    for a, b in PAIRS:
        w = 1.0 if a == b else 2.0
        M = -2.0 * d2 * (alpha * alpha * tsn * tS[a, b] - test_filter(sn * S[a, b]))
        num = num + w * Lm[a, b] * M
        den = den + w * M * M
    cs2 = num.mean(axis=(0, 2)) / np.maximum(den.mean(axis=(0, 2)), 1e-300)
    tau = [-2.0 * cs2[None, :, None] * d2 * sn * S[a, b] for a, b in PAIRS]
    return tau, cs2


def deviatoric(t):
    tr = (t[0] + t[1] + t[2]) / 3.0
    return [t[0] - tr, t[1] - tr, t[2] - tr, t[3], t[4], t[5]]


def correlation(a, b):
    a = np.concatenate([np.ravel(x) for x in a])
    b = np.concatenate([np.ravel(x) for x in b])
    a = a - a.mean()
    b = b - b.mean()
    return float((a * b).sum() / np.sqrt((a * a).sum() * (b * b).sum()))
