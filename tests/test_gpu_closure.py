"""The learned subgrid closure on the GPU (c3d_closure_*) against its
oracle (oracle/closure.py): samples filtered from a DNS state, the bf16
MFMA network's outputs and gradients, the Adam step, and a training run
whose network beats the dynamic Smagorinsky model a priori on held-out
filtered DNS.  The network computes with bf16 operands: a 1-ulp bf16
difference of an activation (0.4 %) moves later values, so the network
bars are a small maximum relative error plus a tight bound on the 99th
percentile, both written in each test.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DNS = dict(nxc=64, nyc=32, nzc=48, xlx=2 * np.pi, yly=2.0, zlz=np.pi, xnu=1.0 / 2000.0,
           dt=0.004, gamma=1.5, ubulk=1.0)


def _dns(case, seed, amp=0.3):
    from channel3d_amd import Channel, synth, lib as L
    g = Channel(case)
    ux, uy, uz = synth.laminar_noise(g.nxc, g.nyc, g.nzc, g.yp, yly=case["yly"], ubulk=1.0,
                                     amp=amp, seed=seed)
    for f, v in ((L.UX, ux), (L.UY, uy), (L.UZ, uz)):
        g.upload(f, v)
    g.project()
    return g


def _smooth_dns(case, seed):
    """laminar flow plus random smooth modes, projected: resolved structure
    at and above the filter scale, whose subgrid stress the filtered
    gradients determine (white noise would leave nothing to learn)"""
    from channel3d_amd import Channel, synth, lib as L
    g = Channel(case)
    ux, uy, uz = synth.smooth_modes(g.nxc, g.nyc, g.nzc, g.yp, case["xlx"], case["yly"],
                                    case["zlz"], seed=seed)
    for f, v in ((L.UX, ux), (L.UY, uy), (L.UZ, uz)):
        g.upload(f, v)
    g.project()
    return g


def _oracle_samples(g, case):
    from oracle import closure as C
    from channel3d_amd import lib as L
    ux, uy, uz = (g.download(f) for f in (L.UX, L.UY, L.UZ))
    X, T, S = C.samples(ux, uy, uz, g.yp, case["xlx"] / case["nxc"], case["zlz"] / case["nzc"],
                        case["yly"])
    return X, T, S, (ux, uy, uz)


def test_samples_match_oracle():
    from channel3d_amd.closure import Closure
    g = _dns(DNS, seed=4)
    for _ in range(2):
        g.step()
    cl = Closure(1 << 16)
    first, n = cl.add_samples(g)
    assert first == 0 and n == (64 // 4) * (32 // 2) * (48 // 4)
    X, T, S = cl.samples(first, n)
    Xo, To, So, _ = _oracle_samples(g, DNS)
    assert np.allclose(X, Xo, rtol=1e-6, atol=1e-6 * np.abs(Xo).max())
    assert np.allclose(T, To, rtol=1e-6, atol=1e-6 * np.abs(To).max())
    assert np.allclose(S, So, rtol=1e-6, atol=0)
    # a second state appends after the first
    g.step()
    first2, n2 = cl.add_samples(g)
    assert first2 == n and n2 == n and cl.size() == 2 * n
    cl.close()
    g.close()


def _net_data(seed=4):
    from channel3d_amd.closure import Closure
    g = _dns(DNS, seed=seed)
    g.step()
    cl = Closure(1 << 16)
    first, n = cl.add_samples(g)
    g.close()
    return cl, first, n


def test_forward_matches_oracle():
    from oracle import closure as C
    cl, first, n = _net_data()
    p = C.init_params(3)
    cl.set_params(p)
    X, _, _ = cl.samples(first, n)
    Y = cl.predict(first, n)
    Yo, _ = C.forward(p, X, emulate=True)
    scale = np.abs(Yo).max()
    err = np.abs(Y - Yo) / scale
    # bars: maximum 1e-2, 99th percentile 1e-5 of the largest output
    assert err.max() < 1e-2 and np.quantile(err, 0.99) < 1e-5, (err.max(), np.quantile(err, 0.99))
    # and the bf16 network is the FP64 network to bf16 accuracy
    Y64, _ = C.forward({k: v.astype(np.float64) for k, v in p.items()}, X.astype(np.float64),
                       emulate=False)
    assert np.abs(Y - Y64).max() / np.abs(Y64).max() < 3e-2
    assert cl.get_params()["W2"].tobytes() == p["W2"].tobytes()
    cl.close()


def test_gradient_matches_oracle():
    from oracle import closure as C
    cl, first, n = _net_data(seed=5)
    p = C.init_params(8)
    cl.set_params(p)
    X, T, _ = cl.samples(first, n)
    g, loss = cl.grad(first, n)
    go = C.backward(p, X, T, emulate=True)
    lo = C.loss(p, X, T, emulate=True)
    assert abs(loss - lo) <= 1e-5 * lo
    # Each weight-gradient element is a sum over the samples of products of
    # bf16 operands, accumulated in FP32.  Its rounding scale is the sum of
    # the terms' magnitudes, s = |A|^T |dZ|: FP32 accumulation in another
    # order moves it by ~2^-24 s per term, and an operand that rounds to the
    # neighbouring bf16 value (its FP32 input differing in the last bits)
    # moves one term by 2^-8 of itself.  Bars on |g - go| / s: 99th
    # percentile 2^-12, maximum 2^-7; the biases, sums of one operand,
    # likewise with s = sum |dZ|.
    a_in = _layer_inputs(C, p, X)
    dzs = _layer_dz(C, p, X, T)
    worst = {}
    for l in range(1, 5):
        sw = np.abs(a_in[l - 1]).T @ np.abs(C.bf16(dzs[l - 1]).astype(np.float64))
        sb = np.abs(dzs[l - 1]).sum(axis=0)
        for k, sc in ((f"W{l}", sw), (f"b{l}", sb)):
            live = sc > 0
            r = np.abs(g[k] - go[k])[live] / sc[live]
            worst[k] = (float(np.quantile(r, 0.99)), float(r.max()))
            assert worst[k][0] < 2.0 ** -12 and worst[k][1] < 2.0 ** -7, (k, worst[k])
    print("gradient error / rounding scale (p99, max):",
          {k: f"{v[0]:.1e} {v[1]:.1e}" for k, v in worst.items()})
    assert not g["W1"][10:].any()
    # a well-conditioned case: output bias +100, so every output error
    # y - T (targets clipped to +-10) is positive and the b4 and W4 sums have
    # no cancellation: per element relative to the value, p99 1e-4, max 1e-3
    # (a systematic error of 0.1 % in these blocks fails)
    q = {k: v.copy() for k, v in p.items()}
    q["b4"] = q["b4"] + np.float32(100.0)
    cl.set_params(q)
    g2, _ = cl.grad(first, n)
    go2 = C.backward(q, X, T, emulate=True)
    for k in ("b4", "W4"):
        big = np.abs(go2[k]) > 1e-3 * np.abs(go2[k]).max()
        r = np.abs(g2[k] - go2[k])[big] / np.abs(go2[k])[big]
        print(f"{k} one-signed: rel err p99 {np.quantile(r, 0.99):.1e} max {r.max():.1e}")
        assert big.mean() > 0.9 and np.quantile(r, 0.99) < 1e-4 and r.max() < 1e-3, (k, r.max())
    cl.close()


def _layer_inputs(C, p, X):
    """the bf16-rounded input of each layer, as the forward pass uses it"""
    _, cache = C.forward(p, X, emulate=True)
    return [a.astype(np.float64) for a, _ in cache]


def _layer_dz(C, p, X, T):
    """dL/dZ of each layer (FP32, before bf16 rounding), the backward pass's"""
    y, cache = C.forward(p, X, emulate=True)
    n = X.shape[0]
    dz = ((2.0 / (n * C.NOUT)) * (y - T.astype(np.float32))).astype(np.float32)
    out = [None] * 4
    for l in range(4, 0, -1):
        _, z = cache[l - 1]
        if l < 4:
            dz = (dz * C.silu_grad(z)).astype(np.float32)
        out[l - 1] = dz
        if l > 1:
            dz = C.bf16(dz).astype(np.float32) @ C.bf16(p[f"W{l}"]).astype(np.float32).T
    return out


def test_adam_step_matches_oracle():
    from oracle import closure as C
    cl, first, n = _net_data(seed=6)
    p = C.init_params(1)
    cl.set_params(p)
    rng = np.random.default_rng(2)
    state = {}
    q = {k: v.copy() for k, v in p.items()}
    for _ in range(3):
        grads = {k: rng.standard_normal(v.shape).astype(np.float32) * 1e-2 for k, v in p.items()}
        grads["W1"][10:] = 0.0
        cl.adam_step(grads, 1e-3)
        C.adam(q, grads, state, lr=1e-3)
    got = cl.get_params()
    for k in q:
        assert np.allclose(got[k], q[k], rtol=1e-5, atol=1e-7), k
    cl.close()


def test_dsm_matches_oracle():
    """the GPU dynamic Smagorinsky model against the oracle's on the same
    filtered state: Cs^2 per row, model and filtered-DNS stresses"""
    from oracle import closure as C
    from channel3d_amd.closure import Closure
    from channel3d_amd import lib as L
    g = _smooth_dns(DNS, seed=2)
    for _ in range(5):
        g.step()
    cl = Closure(1 << 12)
    tm, tt, cs2 = cl.dsm(g)
    ux, uy, uz = (g.download(f) for f in (L.UX, L.UY, L.UZ))
    g.close()
    ub, tau = C.filtered(ux, uy, uz, _yp(), 4, 2, 4)
    dY, yc = C.les_grid(_yp(), 2)
    dx, dz = DNS["xlx"] / DNS["nxc"], DNS["zlz"] / DNS["nzc"]
    to, co = C.dynamic_smagorinsky(ub, 4 * dx, dY, 4 * dz, yc)
    assert np.allclose(cs2, co, rtol=1e-10, atol=1e-14 * np.abs(co).max())
    for q in range(6):
        assert np.allclose(tm[q], to[q], rtol=1e-9, atol=1e-12 * np.abs(to[q]).max()), q
        assert np.allclose(tt[q], tau[q], rtol=1e-9, atol=1e-12 * np.abs(tau[q]).max()), q
    cl.close()


def test_training_beats_dynamic_smagorinsky_a_priori():
    """train on four DNS snapshots, evaluate on two later ones: correlation of
    the predicted with the filtered-DNS stresses (tau_12 and the deviatoric
    tensor) against the dynamic Smagorinsky model's"""
    from oracle import closure as C
    from channel3d_amd.closure import Closure
    from channel3d_amd import lib as L
    g = _smooth_dns(DNS, seed=9)
    cl = Closure(1 << 18)
    snaps = []
    for s in range(6):
        for _ in range(20):
            g.step()
        first, n = cl.add_samples(g)
        dsm = cl.dsm(g) if s >= 4 else None
        snaps.append((first, n, [g.download(f) for f in (L.UX, L.UY, L.UZ)], dsm))
    g.close()
    cl.set_params(C.init_params(0))
    ntrain = snaps[4][0]
    losses = cl.train(0, ntrain, batch=2048, steps=2000, lr=2e-3, seed=1)
    print(f"loss {losses[:10].mean():.4g} -> {losses[-50:].mean():.4g}")
    assert losses[-50:].mean() < 0.5 * losses[:10].mean()
    dx, dz = DNS["xlx"] / DNS["nxc"], DNS["zlz"] / DNS["nzc"]
    nI, nJ, nK = DNS["nxc"] // 4, DNS["nyc"] // 2, DNS["nzc"] // 4
    dY, yc = C.les_grid(_yp(), 2)
    true, mlp, dsm = [], [], []
    for first, n, (ux, uy, uz), (tm, tt, _) in snaps[4:]:
        ub, tau = C.filtered(ux, uy, uz, _yp(), 4, 2, 4)
        _, _, S = cl.samples(first, n)
        Y = cl.predict(first, n)
        that = [(Y[:, q] * S).reshape((nI, nJ, nK), order="F") for q in range(6)]
        true.append(C.deviatoric(tau))
        mlp.append(C.deviatoric(that))
        dsm.append(list(tm))     # the GPU model (its parity: test_dsm_matches_oracle)
    r12 = [C.correlation([t[3] for t in m], [t[3] for t in true]) for m in (mlp, dsm)]
    rdev = [C.correlation([x for t in m for x in t], [x for t in true for x in t])
            for m in (mlp, dsm)]
    print(f"a priori: tau_12 corr mlp {r12[0]:.3f} dsm {r12[1]:.3f}; "
          f"deviatoric corr mlp {rdev[0]:.3f} dsm {rdev[1]:.3f}")
    assert r12[0] > r12[1] and rdev[0] > rdev[1]
    cl.close()


def _yp():
    from channel3d_amd.channel import channel_yp
    return channel_yp(DNS["nyc"], DNS["yly"], DNS["gamma"])


def _nut_oracle(g, case, p, dsm_only=False):
    """the oracle chain of the closure's LES viscosity on the context's own
    grid: (nut in the interior cells, fallback mask); dsm_only: the dynamic
    Smagorinsky value in every cell"""
    from oracle import closure as C
    from channel3d_amd import lib as L
    ux, uy, uz = (g.download(f) for f in (L.UX, L.UY, L.UZ))
    dx, dz = case["xlx"] / case["nxc"], case["zlz"] / case["nzc"]
    yp = g.yp
    X, _, S = C.samples(ux, uy, uz, yp, dx, dz, case["yly"], fx=1, fy=1, fz=1)
    Y, _ = C.forward(p, X.astype(np.float32), emulate=True)
    nI, nJ, nK = case["nxc"], case["nyc"], case["nzc"]
    t = [(Y[:, q].astype(np.float64) * S.astype(np.float32).astype(np.float64))
         .reshape((nI, nJ, nK), order="F") for q in range(6)]
    t = C.deviatoric(t)
    ub, _ = C.filtered(ux, uy, uz, yp, 1, 1, 1)
    dY, yc = C.les_grid(yp, 1)
    g3 = C.gradients(ub, dx, dz, yc)
    Sij = [0.5 * (g3[a][b] + g3[b][a]) for a, b in C.PAIRS]
    sn = C.strain_norm(g3)
    _, cs2 = C.dynamic_smagorinsky(ub, dx, dY, dz, yc)
    wt = [1, 1, 1, 2, 2, 2]
    ts = sum(w * a * b for w, a, b in zip(wt, t, Sij))
    ss = sum(w * b * b for w, b in zip(wt, Sij))
    nu = -ts / (2 * ss)
    fb = ~(nu >= 0) | ~np.isfinite(nu)
    if dsm_only:
        fb = np.ones_like(fb)
    d2 = ((dx * dY * dz) ** (2 / 3))[None, :, None]
    nu = np.where(fb, np.maximum(cs2, 0)[None, :, None] * d2 * sn, nu)
    return case["xnu"] + nu, fb


def _check_nut(g, case, p, nut, nfb):
    """bf16 operands: bar on the 99th percentile of the nut error, and the
    fallback counts agree to within the cells a bf16 flip can move across
    nu = 0"""
    ref, fb = _nut_oracle(g, case, p)
    got = nut[1:-2, 1:-2, 1:-2]
    err = np.abs(got - ref) / np.abs(ref).max()
    print(f"nut {case['nxc']}x{case['nyc']}x{case['nzc']}: fallback {nfb} "
          f"(oracle {int(fb.sum())}), err max {err.max():.2e}, p99 {np.quantile(err, 0.99):.2e}")
    assert np.quantile(err, 0.99) < 1e-4
    assert abs(nfb - int(fb.sum())) <= max(2, fb.size // 1000)
    assert (got >= case["xnu"] * (1 - 1e-12)).all()


LESC = dict(nxc=32, nyc=32, nzc=24, xlx=2 * np.pi, yly=2.0, zlz=np.pi, xnu=1.0 / 2000.0,
            dt=0.004, gamma=1.5, ubulk=1.0, les_type=2)


def test_les_viscosity_from_closure():
    """the closure as an LES context's SGS model: nut per cell against the
    oracle chain (samples at the LES grid, network, least-squares eddy
    viscosity, dynamic Smagorinsky where it backscatters), then LES steps
    with it stay finite and solenoidal"""
    from oracle import closure as C
    from channel3d_amd.closure import Closure
    from channel3d_amd import lib as L
    # a briefly trained network (any weights would do for the parity part)
    cl, first, n = _net_data(seed=7)
    cl.set_params(C.init_params(5))
    cl.train(first, n, batch=1024, steps=100, lr=2e-3, seed=3)
    p = cl.get_params()
    g = _smooth_dns(LESC, seed=11)
    g.set_les_closure(cl)
    g.Clc_SGS_Vis()
    g.sync()
    nut = g.download(L.NUT)
    nfb = cl.fallbacks()
    _check_nut(g, LESC, p, nut, nfb)
    # LES steps with the learned model
    for _ in range(5):
        g.step()
    d = g.diagnostics()
    assert np.isfinite(d["vmax_x"]) and d["divmax2"] < 1e-10
    g.set_les_closure(None)
    g.close()
    cl.close()


def test_les_viscosity_dynamic_smagorinsky_mode():
    """mode 1 (dynamic Smagorinsky in every cell): nut against the oracle's
    FP64 dynamic Smagorinsky chain to 1e-9 (no bf16 in this path), every cell
    counted as a fallback; mode 0 restores the network's viscosity"""
    from oracle import closure as C
    from channel3d_amd.closure import Closure
    from channel3d_amd import lib as L
    cl = Closure(1 << 12)
    p = C.init_params(4)
    cl.set_params(p)
    g = _smooth_dns(LESC, seed=13)
    g.set_les_closure(cl)
    cl.set_mode(1)
    g.Clc_SGS_Vis()
    g.sync()
    nut = g.download(L.NUT)[1:-2, 1:-2, 1:-2]
    ref, _ = _nut_oracle(g, LESC, p, dsm_only=True)
    err = np.abs(nut - ref).max() / np.abs(ref).max()
    print(f"dynamic Smagorinsky mode: nut max rel err {err:.2e}", flush=True)
    assert err < 1e-9
    assert cl.fallbacks() == LESC["nxc"] * LESC["nyc"] * LESC["nzc"]
    for _ in range(3):
        g.step()
    d = g.diagnostics()
    assert np.isfinite(d["vmax_x"]) and d["divmax2"] < 1e-10
    cl.set_mode(0)
    g.Clc_SGS_Vis()
    g.sync()
    _check_nut(g, LESC, p, g.download(L.NUT), cl.fallbacks())
    g.set_les_closure(None)
    g.close()
    cl.close()


def test_samples_and_nut_at_the_les96_shape():
    """the learned-closure configuration's own shapes: samples of a
    384x192x288 DNS state filtered 4x2x4 onto 96x96x72 against the oracle
    (FP32 storage, 1e-6), and the closure's viscosity on a 96x96x72 LES
    context against the oracle chain (the bars of _check_nut)"""
    from oracle import closure as C
    from channel3d_amd.closure import Closure
    from tests import cases
    dns = dict(cases.LES384)
    dns.pop("les_type")
    dns.pop("filter_type")
    g = _smooth_dns(dns, seed=31)
    g.step()
    cl = Closure(96 * 96 * 72)
    first, n = cl.add_samples(g)
    assert n == 96 * 96 * 72
    X, T, S = cl.samples(first, n)
    Xo, To, So, _ = _oracle_samples(g, dns)
    g.close()
    for a, b, nm in ((X, Xo, "X"), (T, To, "T")):
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"96x96x72 samples {nm}: max err {err:.2e} of the largest", flush=True)
        assert np.allclose(a, b, rtol=1e-6, atol=1e-6 * np.abs(b).max()), nm
    assert np.allclose(S, So, rtol=1e-6, atol=0)
    # a network trained briefly on these samples as the LES model
    p = C.init_params(9)
    cl.set_params(p)
    cl.train(first, n, batch=4096, steps=50, lr=2e-3, seed=5)
    p = cl.get_params()
    les = dict(cases.LES96)
    h = _smooth_dns(les, seed=32)
    h.set_les_closure(cl)
    h.Clc_SGS_Vis()
    h.sync()
    from channel3d_amd import lib as L
    _check_nut(h, les, p, h.download(L.NUT), cl.fallbacks())
    h.set_les_closure(None)
    h.close()
    cl.close()


def test_les_closure_steps_replay_from_graphs():
    """LES steps with the closure as SGS model, launched and replayed from
    HIP graphs, agree bit for bit (the closure's scratch is allocated when
    it is attached, so a captured step allocates nothing)"""
    from oracle import closure as C
    from channel3d_amd.closure import Closure
    from channel3d_amd import lib as L
    cl = Closure(1 << 12)
    cl.set_params(C.init_params(2))
    a = _smooth_dns(LESC, seed=12)
    b = _smooth_dns(LESC, seed=12)
    for g in (a, b):
        g.set_les_closure(cl)
    b.set_graph(True)
    for n in range(4):
        a.step()
        b.step()
        a.sync()
        b.sync()
        for f in (L.UX, L.UY, L.UZ, L.PR, L.NUT):
            assert np.array_equal(a.download(f), b.download(f)), (n, f)
        assert a.diagnostics() == b.diagnostics()
    for g in (a, b):
        g.set_les_closure(None)
        g.close()
    cl.close()


def test_one_closure_two_contexts():
    """one network attached to two LES contexts of different grids (ny and
    stretching differ): each context's viscosity follows its own y tables
    and scratch, before and after the other evaluates, against the oracle"""
    from oracle import closure as C
    from channel3d_amd.closure import Closure
    from channel3d_amd import lib as L
    cl = Closure(1 << 12)
    p = C.init_params(6)
    cl.set_params(p)
    case_b = dict(LESC, nyc=40, gamma=2.0)
    a = _smooth_dns(LESC, seed=21)
    b = _smooth_dns(case_b, seed=22)
    a.set_les_closure(cl)
    b.set_les_closure(cl)
    a.Clc_SGS_Vis()
    b.Clc_SGS_Vis()
    a.sync()
    b.sync()
    nfb_b = cl.fallbacks()            # b evaluated last
    nut_a, nut_b = a.download(L.NUT), b.download(L.NUT)
    _check_nut(b, case_b, p, nut_b, nfb_b)
    a.Clc_SGS_Vis()                   # a again, after b's evaluation
    a.sync()
    assert np.array_equal(a.download(L.NUT), nut_a)
    _check_nut(a, LESC, p, nut_a, cl.fallbacks())
    for g in (a, b):
        g.set_les_closure(None)
        g.close()
    cl.close()


def test_two_contexts_share_closure_concurrently():
    """two LES contexts with different states share one network and step
    interleaved on their own streams, without synchronising between them;
    each ends bit for bit where it ends when it runs alone"""
    from oracle import closure as C
    from channel3d_amd.closure import Closure
    from channel3d_amd import lib as L
    cl = Closure(1 << 12)
    cl.set_params(C.init_params(2))
    ref = {}
    for seed in (31, 32):
        g = _smooth_dns(LESC, seed=seed)
        g.set_les_closure(cl)
        for _ in range(3):
            g.step()
        g.sync()
        ref[seed] = [g.download(f) for f in (L.UX, L.UY, L.UZ, L.PR, L.NUT)]
        g.set_les_closure(None)
        g.close()
    a = _smooth_dns(LESC, seed=31)
    b = _smooth_dns(LESC, seed=32)
    for g in (a, b):
        g.set_les_closure(cl)
    for _ in range(3):
        a.step()
        b.step()
    a.sync()
    b.sync()
    for seed, g in ((31, a), (32, b)):
        got = [g.download(f) for f in (L.UX, L.UY, L.UZ, L.PR, L.NUT)]
        for x, y in zip(got, ref[seed]):
            assert np.array_equal(x, y), seed
    # destroying the closure detaches it: the contexts step on with WALE
    cl.close()
    a.step()
    a.sync()
    assert np.isfinite(a.diagnostics()["vmax_x"])
    for g in (a, b):
        g.close()
