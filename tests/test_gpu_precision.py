"""The mixed-precision pressure solve (c3d_set_ppe_refine) against the FP64
solve and the oracle.

FP32 transforms and wall-normal solves (the kx=kz=0 mode in FP64), then
refinement sweeps with the FP64 residual of the discrete operator.  Bars:
0 sweeps at FP32 accuracy (<= 1e-4 relative max norm on these grids), each
sweep at least 100x better than the last until the FP64 floor, and 2 sweeps
within 1e-10 of the FP64 oracle, the pressure bar of the precision form.
A full step with the refined solve keeps the divergence at FP64 round-off.
"""
import numpy as np
import pytest

from tests import cases
from tests.test_gpu_parity import make_pair, relerr

pytestmark = pytest.mark.gpu

CASES = {"small": cases.SMALL, "c1": cases.C1, "odd": cases.ODD, "freeslip": cases.FREESLIP}


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def compatible_rhs(O, o, seed):
    """N(0,1) source with zero dyp-weighted mean (the Neumann operator's
    solvability condition: dyp is its left null vector)"""
    rng = np.random.default_rng(seed)
    f = np.asfortranarray(rng.standard_normal((o.nxc, o.nyc, o.nzc)))
    dyp = o.metric(O.M_DYP, 1, o.nyc)
    m = (f.mean(axis=(0, 2)) * dyp).sum() / dyp.sum()
    return np.asfortranarray(f - m)


def interior_of(a):
    nx, ny, nz = a.shape[0] - 3, a.shape[1] - 3, a.shape[2] - 3
    return a[1:nx + 1, 1:ny + 1, 1:nz + 1]


@pytest.mark.parametrize("name", list(CASES))
def test_refined_ppe_converges_to_fp64(O, name):
    from channel3d_amd import lib as L
    o, g = make_pair(O, CASES[name], laminar=False)
    f = compatible_rhs(O, o, seed=11)
    # oracle FP64
    ph_o = np.zeros((o.nxc + 3, o.nyc + 3, o.nzc + 3), order="F")
    fo = f.copy(order="F")
    o.call("clcPPE", fo, ph_o)
    ref = interior_of(ph_o)
    errs = []
    for s in (-1, 0, 1, 2):
        g.set_ppe_refine(s)
        g.upload(L.ARR1, f)
        g.ppe(L.ARR1, L.HALOARR)
        phi = interior_of(g.download(L.HALOARR))
        errs.append(relerr(phi, ref))
    print(name, "PPE rel err vs oracle: fp64 %.2e  fp32 %.2e  +1 %.2e  +2 %.2e" % tuple(errs),
          flush=True)
    assert errs[0] < 1e-12
    assert errs[1] < 1e-4
    assert errs[2] < max(errs[1] * 1e-2, 1e-12)
    assert errs[3] < 1e-10
    g.close()
    o.close()


def test_refined_ppe_in_the_step(O):
    """two steps with the refined solve (2 sweeps) against the FP64 path"""
    from channel3d_amd import lib as L
    case = cases.C1
    o, g = make_pair(O, case, seed=6)
    g2 = type(g)(case, yp=o.yp)
    for fid in (L.UX, L.UY, L.UZ, L.PR):
        g2.upload(fid, g.download(fid))
    g2.set_ppe_refine(2)
    for _ in range(2):
        g.step()
        g2.step()
    d1, d2 = g.diagnostics(), g2.diagnostics()
    print("divmax fp64 %.2e refined %.2e" % (d1["divmax2"], d2["divmax2"]), flush=True)
    assert d2["divmax2"] < 1e-10
    for fid in (L.UX, L.UY, L.UZ):
        assert relerr(g2.download(fid), g.download(fid)) < 1e-10
    assert relerr(interior_of(g2.download(L.PR)), interior_of(g.download(L.PR))) < 1e-9
    g.close()
    g2.close()
    o.close()


def test_refine_rejects_other_contexts():
    from channel3d_amd import Channel
    from channel3d_amd import lib as L
    g = Channel(cases.SMALL4)
    with pytest.raises(Exception, match="2nd-order"):
        g.set_ppe_refine(1)
    g.close()
    g = Channel(cases.SMALL)
    with pytest.raises(Exception):
        g.set_ppe_refine(-2)
    g.set_ppe_refine(1)
    g.set_ppe_refine(-1)
    g.close()
