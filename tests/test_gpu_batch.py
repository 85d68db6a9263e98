"""Batched pressure and Helmholtz solves (c3d_ppe_batch, c3d_helmholtz_batch)
in their FP64 and FP32 + refinement forms, against the oracle's clcPPE and
clcU1/2/3Hat (oracle/c3d_oracle.c) on the same sources.

The sources follow the precision-sweep workload: random phases with the
spectrum E(k) ~ k^4 exp(-2 (k/8)^2), compatible (dyp-weighted zero mean),
eight in one batch.  Bars, relative max norm per field against the oracle:
FP64 1e-12 (and bit for bit the single-field solve, whose kernels the batch
runs); FP32 without refinement at FP32 accuracy (1e-4 pressure, 1e-5
Helmholtz: the Helmholtz operator is close to the identity); one sweep
1e-9 (pressure) / 1e-12 (Helmholtz); two sweeps 1e-10 (pressure, the bar of
the precision form) / 1e-12 (Helmholtz).  At 512x256x512 x 8 the oracle
would take minutes per field: there the checks are linearity of the FP64
solves and the refined forms against the FP64 solve.
"""
import numpy as np
import pytest

from tests import cases
from tests.test_gpu_parity import interior, relerr

pytestmark = pytest.mark.gpu

NB = 8
# reduced precision-sweep shape, with the production step's nu and dt
C3S = dict(nxc=64, nyc=48, nzc=64, xlx=2 * np.pi, yly=2.0, zlz=np.pi, xnu=9.9369e-5, dt=0.002,
           gamma=1.5, ubulk=1.0)
# the precision-sweep shape itself
C3 = dict(C3S, nxc=512, nyc=256, nzc=512)
BATCH_CASES = {"c3s": C3S, "odd": cases.ODD, "small": cases.SMALL}


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def sources(g, nb, seed0, dyp=None):
    from channel3d_amd import synth
    return [synth.spectrum_field(g.nxc, g.nyc, g.nzc, seed=seed0 + q, dyp=dyp) for q in range(nb)]


def _ctx(O, case):
    from channel3d_amd import Channel
    o = O.Oracle(case)
    g = Channel(case, yp=o.yp)
    return o, g


@pytest.mark.parametrize("name", list(BATCH_CASES))
def test_ppe_batch_forms(O, name):
    from channel3d_amd import lib as L
    case = BATCH_CASES[name]
    o, g = _ctx(O, case)
    dyp = o.metric(O.M_DYP, 1, o.nyc)
    fs = sources(g, NB, 100, dyp)
    src = g.alloc_batch(L.INTERIOR, NB)
    phi = g.alloc_batch(L.HALO, NB)
    for q, f in enumerate(fs):
        g.upload(src[q], f, L.INTERIOR)
    refs = []
    for f in fs:
        ph = np.zeros((o.nxc + 3, o.nyc + 3, o.nzc + 3), order="F")
        o.call("clcPPE", np.asfortranarray(f.copy()), ph)
        refs.append(interior(ph))
    bars = {-1: 1e-12, 0: 1e-4, 1: 1e-9, 2: 1e-10}
    for s, bar in bars.items():
        g.ppe_batch(src, phi, s)
        errs = [relerr(interior(g.download(phi[q], L.HALO)), refs[q]) for q in range(NB)]
        print(f"{name} ppe batch sweeps {s}: max rel err vs oracle {max(errs):.2e}", flush=True)
        assert max(errs) < bar, (s, errs)
        # the batch runs the single solve's kernels: equal bit for bit
        g.set_ppe_refine(s)
        for q in (0, NB - 1):
            g.upload(L.ARR1, fs[q])
            g.ppe(L.ARR1, L.HALOARR)
            assert np.array_equal(interior(g.download(L.HALOARR)),
                                  interior(g.download(phi[q], L.HALO))), (s, q)
        g.set_ppe_refine(-1)
    g.close()


@pytest.mark.parametrize("name", list(BATCH_CASES))
@pytest.mark.parametrize("ns", [1, 3])
def test_helmholtz_batch_forms(O, name, ns):
    from channel3d_amd import lib as L
    case = BATCH_CASES[name]
    o, g = _ctx(O, case)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    o.call("PMcoeUpdate", ns)
    fs = [0.01 * f for f in sources(g, NB, 200)]
    rhs = g.alloc_batch(L.INTERIOR, NB)
    du = g.alloc_batch(L.INTERIOR, NB)
    for q, f in enumerate(fs):
        g.upload(rhs[q], f, L.INTERIOR)
    bars = {-1: 1e-12, 0: 1e-5, 1: 1e-12, 2: 1e-12}
    zero_wall = np.zeros((nx, nz), order="F")
    for comp, (routine, fid) in enumerate((("clcU1Hat", O.UX), ("clcU2Hat", O.UY),
                                           ("clcU3Hat", O.UZ))):
        refs = []
        for f in fs:
            u = o.field(fid)
            u[...] = 0.0
            r = np.asfortranarray(f.copy())
            if comp == 2:   # clcU3Hat takes the halo-shaped RhsZ
                rh = np.zeros((nx + 3, ny + 3, nz + 3), order="F")
                rh[1:nx + 1, 1:ny + 1, 1:nz + 1] = r
                o.call(routine, u, rh)
            elif comp == 1:
                o.call(routine, u, r, zero_wall)
            else:
                o.call(routine, u, r)
            refs.append(interior(u).copy())
        for s, bar in bars.items():
            g.helmholtz_batch(comp, ns, rhs, du, s)
            errs = [relerr(g.download(du[q], L.INTERIOR), refs[q]) for q in range(NB)]
            print(f"{name} ns {ns} comp {comp} helmholtz batch sweeps {s}: "
                  f"max rel err vs oracle {max(errs):.2e}", flush=True)
            assert max(errs) < bar, (comp, s, errs)
    g.close()


def test_batch_full_size_properties():
    """512x256x512 x 8: the FP64 solves are linear in the source (a batch
    holding f0, f1 and f0 + 2 f1), and the refined forms land on the FP64
    solution: pressure within 1e-10 after two sweeps, Helmholtz within 1e-12
    after one"""
    from channel3d_amd import Channel, lib as L
    g = Channel(C3)
    dyp = g.metric(L.M_DYP, 1, g.nyc)
    fs = sources(g, NB - 1, 300, dyp)
    fs.append(np.asfortranarray(fs[0] + 2.0 * fs[1]))
    src = g.alloc_batch(L.INTERIOR, NB)
    out = g.alloc_batch(L.HALO, NB)
    for q, f in enumerate(fs):
        g.upload(src[q], f, L.INTERIOR)
    del fs
    g.ppe_batch(src, out, -1)
    ph = [interior(g.download(out[q], L.HALO)) for q in (0, 1, NB - 1)]
    lin = relerr(ph[2], ph[0] + 2.0 * ph[1])
    print(f"ppe batch 512x256x512: linearity {lin:.2e}", flush=True)
    assert lin < 1e-12
    ref = [interior(g.download(out[q], L.HALO)) for q in range(NB)]
    for s, bar in ((2, 1e-10),):
        g.ppe_batch(src, out, s)
        errs = [relerr(interior(g.download(out[q], L.HALO)), ref[q]) for q in range(NB)]
        print(f"ppe batch 512x256x512 sweeps {s}: max rel err vs fp64 {max(errs):.2e}",
              flush=True)
        assert max(errs) < bar
    del ref
    du = g.alloc_batch(L.INTERIOR, NB)
    g.helmholtz_batch(0, 1, src, du, -1)
    h = [g.download(du[q], L.INTERIOR) for q in range(NB)]
    lin = relerr(h[NB - 1], h[0] + 2.0 * h[1])
    print(f"helmholtz batch 512x256x512: linearity {lin:.2e}", flush=True)
    assert lin < 1e-12
    g.helmholtz_batch(0, 1, src, du, 1)
    errs = [relerr(g.download(du[q], L.INTERIOR), h[q]) for q in range(NB)]
    print(f"helmholtz batch 512x256x512 sweeps 1: max rel err vs fp64 {max(errs):.2e}",
          flush=True)
    assert max(errs) < 1e-12
    g.close()
