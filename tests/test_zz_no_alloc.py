"""c3d_ppe_batch_noalloc: the batched pressure solve on the library's fixed
device work space, against c3d_ppe_batch (the same kernels on the context's
own batch buffers) and, for one field, the oracle's clcPPE
(m_Poisson.f90:567-688 restated in oracle/c3d_oracle.c).

Bar: equal to c3d_ppe_batch bit for bit (FP64 solve and FP32 + 1 refinement
sweep), over two calls in a row on one context, the second with a smaller
batch of other sources (stale work-space contents would show), then a call
from a second context on a smaller grid; the FP64 result within 1e-12 of the
oracle.  A batch whose work space exceeds the fixed size is rejected.
"""
import numpy as np
import pytest

from tests import cases
from tests.test_gpu_parity import interior, relerr

pytestmark = pytest.mark.gpu

C3S = dict(nxc=64, nyc=48, nzc=64, xlx=2 * np.pi, yly=2.0, zlz=np.pi, xnu=9.9369e-5, dt=0.002,
           gamma=1.5, ubulk=1.0)
LIMIT = 128 * 1024 * 1024


def _work_bytes(case, nb, sweeps):
    a = lambda n: (n + 255) // 256 * 256  # noqa: E731
    nx, ny, nz = case["nxc"], case["nyc"], case["nzc"]
    P = ((nx // 2 + 1) + 7) // 8 * 8
    nspec = P * ny * nz
    nfield = nx * (ny + 3) * nz
    mean = 3 * (ny + 1) + ((nx + 63) // 64) * nz
    return (a(nb * nspec * 16) + a(nb * nspec * 8) + a(nb * mean * 8) +
            (a(nb * nfield * 4) if sweeps >= 0 else 0))


def _compare(g, fs, sweeps, tag):
    """c3d_ppe_batch, then c3d_ppe_batch_noalloc, on the same sources"""
    from channel3d_amd import lib as L
    nb = len(fs)
    src = g.alloc_batch(L.INTERIOR, nb)
    phi_a = g.alloc_batch(L.HALO, nb)
    phi_b = g.alloc_batch(L.HALO, nb)
    for q, f in enumerate(fs):
        g.upload(src[q], f, L.INTERIOR)
    g.ppe_batch(src, phi_a, sweeps)
    g.ppe_batch_noalloc(src, phi_b, sweeps)
    out = []
    for q in range(nb):
        a = interior(g.download(phi_a[q], L.HALO))
        b = interior(g.download(phi_b[q], L.HALO))
        assert np.array_equal(a, b), (tag, sweeps, q, float(np.abs(a - b).max()))
        out.append(b)
    print(f"{tag}: nbatch {nb} sweeps {sweeps}: no-alloc == batch bit for bit", flush=True)
    for h in (src, phi_a, phi_b):
        g.free(h)
    return out


def test_noalloc_matches_batch_two_calls(oracle_mod):
    O = oracle_mod
    from channel3d_amd import Channel
    o = O.Oracle(C3S)
    dyp = o.metric(O.M_DYP, 1, o.nyc)
    from channel3d_amd import synth
    with Channel(C3S, yp=o.yp) as g:
        for sweeps in (-1, 1):
            assert _work_bytes(C3S, 8, sweeps) <= LIMIT
            f8 = [synth.spectrum_field(64, 48, 64, seed=300 + q, dyp=dyp) for q in range(8)]
            got = _compare(g, f8, sweeps, "first call")
            if sweeps == -1:
                ph = np.zeros((o.nxc + 3, o.nyc + 3, o.nzc + 3), order="F")
                o.call("clcPPE", np.asfortranarray(f8[5].copy()), ph)
                err = relerr(got[5], interior(ph))
                print(f"no-alloc FP64 field 5 vs oracle {err:.2e}", flush=True)
                assert err < 1e-12
            # second call: a smaller batch of other sources
            f3 = [synth.spectrum_field(64, 48, 64, seed=900 + q, dyp=dyp) for q in range(3)]
            _compare(g, f3, sweeps, "second call")
    # a second context on a smaller grid shares the same work space
    oo = O.Oracle(cases.ODD)
    with Channel(cases.ODD, yp=oo.yp) as g2:
        dy2 = oo.metric(O.M_DYP, 1, oo.nyc)
        fs = [synth.spectrum_field(g2.nxc, g2.nyc, g2.nzc, seed=40 + q, dyp=dy2) for q in range(2)]
        _compare(g2, fs, -1, "smaller grid")
        _compare(g2, fs, 1, "smaller grid")


def test_noalloc_rejects_oversized_batch(oracle_mod):
    O = oracle_mod
    from channel3d_amd import Channel, lib as L
    o = O.Oracle(C3S)
    nb = 40
    assert _work_bytes(C3S, nb, 1) > LIMIT
    with Channel(C3S, yp=o.yp) as g:
        src = g.alloc_batch(L.INTERIOR, nb)
        phi = g.alloc_batch(L.HALO, nb)
        rc = g.lib.c3d_ppe_batch_noalloc(g._h, nb, src.base, phi.base, 1)
        assert rc == L.C3D_ERR_ARG
        msg = g.lib.c3d_last_error().decode()
        assert "work space" in msg and str(LIMIT) in msg, msg
        for h in (src, phi):
            g.free(h)
