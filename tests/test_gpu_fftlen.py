"""FFT lengths outside the compiled list (any 2^a 3^b 5^c): the transforms
then run from a radix plan made at run time.

* On compiled lengths the run-time plan (forced with C3D_FFT_PLAN=runtime)
  runs the same butterflies on the same operands: pressure solves and whole
  steps equal the compiled kernels' bit for bit.
* The reference's own 4th-order input Input/TurbulentChannel_4th_1.prm:
  180x132x120 (nxc/2 = 90 = 2 3^2 5, nzc = 120 = 2^3 3 5), channel,
  ubulk = 2/3, uCRF = 0.6666666, xnu = 2.331e-4, Lx = 4 pi, Lz = 2 pi
  (iCFL = 1 there; here a fixed dt), per-step parity with the 4th-order
  oracle at 1e-12.
* A 2nd-order channel with the same odd-factor lengths: per substep and over
  steps at 1e-12 against the oracle.
"""
import numpy as np
import pytest

from tests import cases
from tests.test_gpu_parity import compare_state, make_pair, interior, relerr
from tests import test_gpu_parity4 as P4

pytestmark = pytest.mark.gpu

TOL = 1e-12
# Input/TurbulentChannel_4th_1.prm:13-28
PRM4 = dict(nxc=180, nyc=132, nzc=120, xlx=12.56637061, yly=2.0, zlz=6.283185307,
            xnu=2.3310e-4, dt=0.01, gamma=1.5, ubulk=0.666666666666666666666666666667,
            u_crf=0.6666666, order=4)
# 2nd order with the same x and z lengths
ODD2 = dict(nxc=180, nyc=40, nzc=120, xlx=4 * np.pi, yly=2.0, zlz=2 * np.pi, xnu=2.0 / 5600.0,
            dt=0.005, gamma=1.5, ubulk=1.0)
# lengths 2 * 75 = 150 (3 5^2) and 50 (2 5^2): radix-5 passes back to back
ODD5 = dict(nxc=150, nyc=24, nzc=50, xlx=3.0, yly=2.0, zlz=1.0, xnu=1.0 / 1500.0, dt=0.004,
            gamma=1.3, ubulk=1.0)


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


@pytest.mark.parametrize("case", [cases.SMALL, cases.ODD, cases.C1], ids=["small", "odd", "c1"])
def test_runtime_plan_equals_compiled(O, case, monkeypatch):
    from channel3d_amd import Channel, lib as L
    o, g0 = make_pair(O, case, seed=2)
    monkeypatch.setenv("C3D_FFT_PLAN", "runtime")
    g1 = Channel(case, yp=o.yp)
    monkeypatch.delenv("C3D_FFT_PLAN")
    for gid in (L.UX, L.UY, L.UZ, L.PR):
        g1.upload(gid, g0.download(gid))
    dyp = o.metric(O.M_DYP, 1, o.nyc)
    from channel3d_amd import synth
    f = synth.spectrum_field(o.nxc, o.nyc, o.nzc, seed=5, dyp=dyp)
    for s in (-1, 2):
        res = []
        for g in (g0, g1):
            g.set_ppe_refine(s)
            g.upload(L.ARR1, f)
            g.ppe(L.ARR1, L.HALOARR)
            res.append(interior(g.download(L.HALOARR)))
            g.set_ppe_refine(-1)
        assert np.array_equal(res[0], res[1]), s
    for _ in range(2):
        g0.step()
        g1.step()
    for gid in (L.UX, L.UY, L.UZ, L.PR, L.HISTX, L.HISTY, L.HISTZ):
        assert np.array_equal(g0.download(gid), g1.download(gid)), gid
    g0.close()
    g1.close()


@pytest.mark.parametrize("case", [ODD2, ODD5], ids=["180x40x120", "150x24x50"])
def test_general_lengths_2nd_order(O, case):
    from channel3d_amd import lib as L
    o, g = make_pair(O, case, seed=3)
    # the pressure solve alone
    dyp = o.metric(O.M_DYP, 1, o.nyc)
    from channel3d_amd import synth
    f = synth.spectrum_field(o.nxc, o.nyc, o.nzc, seed=9, dyp=dyp)
    ph = np.zeros((o.nxc + 3, o.nyc + 3, o.nzc + 3), order="F")
    o.call("clcPPE", np.asfortranarray(f.copy()), ph)
    g.upload(L.ARR1, f)
    g.clcPPE(L.ARR1, L.HALOARR)
    assert relerr(interior(g.download(L.HALOARR)), interior(ph)) < TOL
    # fused substeps, then free-running steps
    for ns in (1, 2, 3):
        o.substep(ns)
        g.substep(ns)
        compare_state(O, o, g, TOL)
    for _ in range(3):
        o.iterate()
        g.step()
    compare_state(O, o, g, TOL)
    assert g.diagnostics()["divmax2"] < 1e-12
    g.close()


def test_reference_input_4th_order(O):
    """Input/TurbulentChannel_4th_1.prm's grid: per-step parity (each step
    from the oracle's state) for two steps"""
    o, g = P4.make_pair(O, PRM4, seed=11)
    print(f"4th order {o.nxc}x{o.nyc}x{o.nzc} (reference input), per-step parity", flush=True)
    for _ in range(2):
        P4.upload_state4(O, o, g)
        o.iterate()
        g.step()
        P4.compare_state4(O, o, g, TOL)
    assert g.diagnostics()["divmax2"] < 1e-12
    g.close()
    o.close()
