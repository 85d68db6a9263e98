"""c3d_step replayed from HIP graphs (c3d_set_graph) against the launched
step: after every step both contexts hold the same state bit for bit and
report the same diagnostics, through both CFL-slot graphs, a change of dt
(which drops the graphs), switching graphs off, and the 4th-order, LES and
mixed-precision pressure forms.  Each replay is followed by host work and a
synchronisation before the next one, and every replay is checked.
"""
import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu


def _pair(case, seed):
    from channel3d_amd import Channel, synth, lib as L
    a = Channel(case)
    nx, ny, nz = a.nxc, a.nyc, a.nzc
    ux, uy, uz = synth.laminar_noise(nx, ny, nz, a.yp, yly=case["yly"],
                                     ubulk=case.get("ubulk", 1.0), seed=seed)
    if a.order == 4:
        from tests.test_gpu_parity4 import to4
        ux, uy, uz = (to4(v, nx, ny, nz) for v in (ux, uy, uz))
    for f, v in ((L.UX, ux), (L.UY, uy), (L.UZ, uz)):
        a.upload(f, v)
    a.project()
    b = Channel(case)
    for f in _fields(L):
        b.upload(f, a.download(f))
    return a, b


def _fields(L):
    return (L.UX, L.UY, L.UZ, L.PR, L.HISTX, L.HISTY, L.HISTZ)


def _same(a, b, L, what):
    a.sync()
    b.sync()
    for f in _fields(L):
        assert np.array_equal(a.download(f), b.download(f)), (what, f)
    assert a.diagnostics() == b.diagnostics(), what


@pytest.mark.parametrize("name", ["small", "small4", "les96", "small_refined"])
def test_graph_step_matches_launched_step(name):
    from channel3d_amd import lib as L
    case = {"small": cases.SMALL, "small4": cases.SMALL4, "les96": cases.LES96,
            "small_refined": cases.SMALL}[name]
    a, b = _pair(case, seed=11)
    if name == "small_refined":
        a.set_ppe_refine(2)
        b.set_ppe_refine(2)
    b.set_graph(True)
    for n in range(5):
        a.step()
        b.step()
        _same(a, b, L, f"step {n}")
    # a new dt drops the graphs; the next steps capture new ones
    dt = 0.8 * case["dt"]
    a.set_dt(dt)
    b.set_dt(dt)
    for n in range(3):
        a.step()
        b.step()
        _same(a, b, L, f"step {n} after set_dt")
    # a state changed from the host: the next step recomputes the CFL eagerly
    ux = a.download(L.UX)
    ux *= 1.01
    a.upload(L.UX, ux)
    b.upload(L.UX, ux)
    for n in range(2):
        a.step()
        b.step()
        _same(a, b, L, f"step {n} after upload")
    b.set_graph(False)
    a.step()
    b.step()
    _same(a, b, L, "graphs off")
    a.close()
    b.close()
