"""The production size, 768x384x576 FP64 (170M cells, the benchmark case),
checked through properties that do not depend on an oracle run at that
size: the projection leaves round-off divergence, the pressure solve is
linear, the FP32 + refinement solve matches the FP64 solve, whole steps
keep the divergence at round-off, and graph replay equals launched steps
bit for bit.  The smaller cases of the other GPU test files carry the
oracle parity; this file shows the same kernels hold at the full shape
(largest grids, workgroup counts and transform lengths 384 and 576).
"""
import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


def _log(msg):
    print(f"[fullsize] {msg}", flush=True)


def _init(seed=0):
    from channel3d_amd import Channel, synth, lib as L
    case = cases.C5
    g = Channel(case)
    ux, uy, uz = synth.laminar_noise(g.nxc, g.nyc, g.nzc, g.yp, yly=case["yly"],
                                     ubulk=1.0, seed=seed)
    for f, v in ((L.UX, ux), (L.UY, uy), (L.UZ, uz)):
        g.upload(f, v)
    del ux, uy, uz
    return g


def _first_substep_source(g, L, dst):
    """div(u^)/(alpha dt) of a first substep from the context state, into
    dst, by the per-routine entries; the velocity is restored after"""
    saved = {f: g.download(f) for f in (L.UX, L.UY, L.UZ)}
    g.PMcoeUpdate(1)
    g.Update_uy_ym()
    g.clcRhsX(L.UX, L.UY, L.UZ, L.ARR1, L.HISTX, L.PR)
    g.clcRhsY(L.UX, L.UY, L.UZ, L.ARR2, L.HISTY, L.PR)
    g.clcRhsZ(L.UX, L.UY, L.UZ, L.HALOARR, L.HISTZ, L.PR)
    g.clcU1Hat(L.UX, L.ARR1)
    g.clcU2Hat(L.UY, L.ARR2, L.DUY_YM)
    g.clcU3Hat(L.UZ, L.HALOARR)
    g.SetBC_and_UpdateHaloForPrSrc(L.UX, L.UY, L.UZ, L.UY_YM)
    g.clcPrSrc(L.UX, L.UY, L.UZ, dst)
    for f, a in saved.items():
        g.upload(f, a)


def _interior(a):
    return a[1:-2, 1:-2, 1:-2]


def test_fullsize_projection_and_solve():
    from channel3d_amd import lib as L
    _log("create + upload")
    g = _init(seed=1)
    g.project()
    div = g.CheckDivergence()
    _log(f"divergence after projection {div:.3e}")
    assert div < 1e-12
    # two sources: a first substep's from the projected state, and one from
    # a second noise field; the solve is linear in the source
    _first_substep_source(g, L, L.ARR1)
    f1 = g.download(L.ARR1)
    g.ppe(L.ARR1, L.HALOARR)
    p1 = _interior(g.download(L.HALOARR)).copy()
    # FP32 + 2 refinement sweeps against the FP64 solve on the same source
    g.set_ppe_refine(2)
    g.ppe(L.ARR1, L.HALOARR)
    pr = _interior(g.download(L.HALOARR))
    err = float(np.abs(pr - p1).max() / np.abs(p1).max())
    _log(f"fp32 + 2 sweeps vs fp64: {err:.3e}")
    assert err < 1e-10
    g.set_ppe_refine(-1)
    rng = np.random.default_rng(7)
    f2 = rng.standard_normal(f1.shape)
    f2 -= f2.mean()
    f2 = np.asfortranarray(f2)
    g.upload(L.ARR1, f2)
    g.ppe(L.ARR1, L.HALOARR)
    p2 = _interior(g.download(L.HALOARR)).copy()
    g.upload(L.ARR1, np.asfortranarray(2.0 * f1 + f2))
    del f1, f2
    g.ppe(L.ARR1, L.HALOARR)
    p3 = _interior(g.download(L.HALOARR))
    lin = float(np.abs(p3 - (2.0 * p1 + p2)).max() / np.abs(p3).max())
    _log(f"linearity {lin:.3e}")
    assert lin < 1e-12
    g.close()


def test_fullsize_steps_and_graph():
    from channel3d_amd import lib as L
    _log("create + upload (two contexts)")
    a = _init(seed=2)
    a.project()
    b = type(a)(cases.C5)
    for f in (L.UX, L.UY, L.UZ, L.PR, L.HISTX, L.HISTY, L.HISTZ):
        b.upload(f, a.download(f))
    b.set_graph(True)
    for n in range(3):
        a.step()
        b.step()
        da, db = a.diagnostics(), b.diagnostics()
        _log(f"step {n}: divmax {da['divmax2']:.3e}, cfl {da['cfl']:.3f}")
        assert da == db
        assert da["divmax2"] < 1e-12 and np.isfinite(da["vmax_x"]) and 0 < da["cfl"] < 10
    for f in (L.UX, L.UY, L.UZ, L.PR):
        assert np.array_equal(a.download(f), b.download(f)), f
    a.close()
    b.close()
