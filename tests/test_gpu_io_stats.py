"""Restart files and channel statistics of the HIP library, through the C ABI.

* statistics: the 35 plane averages of clcStat_CH against the oracle on the
  same fields, the running sums over several samples, PrGradsum, and the
  text of the stats and PrGrad files.  The GPU sums a plane in a different
  order than the reference's sequential loop, so the two agree to the
  rounding error of an n-term sum: the bar is 2 n eps times the plane
  average of |term| (the oracle's clcStat_CH_mag), n = nxc*nzc.  Columns
  whose exact value is zero (the mean of v in a divergence-free channel)
  are held to that bar rather than to their own round-off-sized value;
* restart: the file holds exactly the interior points of ux, uy, uz, p (and
  the histories) as real(8) in Fortran order, a read restores them bit for
  bit, a run continued from a restart matches the uninterrupted run bit for
  bit, a 4th-order context stores ux + uCRF, and wrong sizes or missing files
  fail with the reference's messages.
"""

import numpy as np
import pytest

from tests import cases
from tests.test_gpu_parity import make_pair, random_state

pytestmark = pytest.mark.gpu

STAT_CASES = {"small": cases.SMALL, "odd": cases.ODD, "freeslip": cases.FREESLIP}


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def _oracle_stats(O, o, acc=None):
    s = np.zeros((35, o.nyc + 1), order="F") if acc is None else acc
    o.call("clcStat_CH", o.field(O.UX), o.field(O.UY), o.field(O.UZ), o.field(O.PR), s)
    return s


def _oracle_mag(O, o, acc=None):
    s = np.zeros((35, o.nyc + 1), order="F") if acc is None else acc
    o.call("clcStat_CH_mag", o.field(O.UX), o.field(O.UY), o.field(O.UZ), o.field(O.PR), s)
    return s


def _sum_err(a, b, mag, n):
    """largest |a - b| in units of the rounding bound 2 n eps mag"""
    bound = 2.0 * n * np.finfo(np.float64).eps * mag
    d = np.abs(a - b)
    return float((d / np.where(bound > 0, bound, np.inf)).max()), bool(np.all(d[bound == 0] == 0))


@pytest.mark.parametrize("name", list(STAT_CASES))
def test_stats_random_fields(O, name):
    o, g = make_pair(O, STAT_CASES[name], laminar=False)
    random_state(O, o, g, seed=5)
    g.stat_reset()
    g.clcStat()
    s_g, n, _ = g.stat_get()
    s_o = _oracle_stats(O, o)
    mag = _oracle_mag(O, o)
    r, zeros_ok = _sum_err(s_g, s_o, mag, o.nxc * o.nzc)
    print(name, "stats error / rounding bound", r, flush=True)
    assert n == 1
    assert r <= 1.0 and zeros_ok
    # plane nyp holds only the three wall-shear columns
    assert np.all(s_g[[q for q in range(35) if q not in (26, 32, 34)], -1] == 0)
    g.close()
    o.close()


def test_stats_accumulate_over_steps(O):
    """three samples taken after steps of a developing flow; SumStat, nfstime
    and PrGradsum follow the reference's running sums"""
    case = cases.SMALL
    o, g = make_pair(O, case, seed=3)
    from channel3d_amd import lib as L
    g.SetBC_and_UpdateHalo(L.UX, L.UY, L.UZ, L.UY_YM)
    g.SetBC_and_UpdateHalo_pr(L.PR)
    g.stat_reset()
    acc = np.zeros((35, o.nyc + 1), order="F")
    mag = np.zeros((35, o.nyc + 1), order="F")
    prg = 0.0
    for it in range(3):
        o.iterate()
        g.step()
        g.clcStat()
        _oracle_stats(O, o, acc)
        _oracle_mag(O, o, mag)
        prg += g.diagnostics()["PrGradAver"]
    s_g, n, pg = g.stat_get()
    assert n == 3
    # the fields agree to ~1e-13 after the steps (the step parity tests):
    # allow that on top of the summation bound
    d = np.abs(s_g - acc)
    bound = 2.0 * o.nxc * o.nzc * np.finfo(np.float64).eps * mag + 1e-12 * mag
    print("accumulated stats error / bound", float((d / np.maximum(bound, 1e-300)).max()),
          flush=True)
    assert np.all(d <= bound)
    assert pg == prg            # PrGradsum = sum of PrGradAver, in sample order
    g.close()
    o.close()


def test_stat_files(tmp_path, O):
    case = dict(cases.SMALL)
    o, g = make_pair(O, case, seed=2)
    g.stat_reset()
    for it in range(2):
        g.step()
        g.clcStat()
    s_g, n, pg = g.stat_get()
    path = str(tmp_path / "stats0000000020")
    g.write_stat(path, itime=20, ivstats=10)
    lines = open(path).read().split("\n")
    assert lines[0] == "  The time step range for this fluid statistics is %7d:%7d:%7d" % (10, 10, 20)
    assert lines[1] == "  "
    assert lines[2] == "  Constant velocity in x-dir by adding a pressure gradient."
    assert lines[3].startswith("    time averaged pressure gradient is: ")
    assert float(lines[3].split(":")[1]) == float("%.15E" % (pg / 2))
    assert lines[4] == "  "
    rows = [ln for ln in lines[5:] if ln]
    assert len(rows) == o.nyc + 1
    for j, ln in enumerate(rows):
        assert len(ln) == 35 * 24
        vals = [float(ln[24 * q:24 * q + 24]) for q in range(35)]
        want = [float("%.15E" % v) for v in s_g[:, j] * 0.5]
        assert vals == want, j
        for q in range(35):
            f = ln[24 * q:24 * q + 24]
            assert f[-4] == "E" and f[-3] in "+-" and f[-20] == "."
    # writing resets the sums
    _, n2, pg2 = g.stat_get()
    assert n2 == 0 and pg2 == 0.0
    with pytest.raises(Exception):
        g.write_stat(path, itime=30, ivstats=10)   # no samples
    # PrGrad file: I7, 2ES24.15, appended
    pgp = str(tmp_path / "PrGrad0000000100")
    open(pgp, "w").close()
    g.append_prgrad(pgp, 21)
    g.append_prgrad(pgp, 22)
    d = g.diagnostics()
    got = open(pgp).read().split("\n")
    assert got[0] == "%7d%24.15E%24.15E" % (21, d["SimTime"], d["PrGradAver"])
    assert got[1].startswith("     22")
    g.close()
    o.close()


def test_stats_entry_points_check_the_order():
    """clcStat_CH (35 columns) belongs to the 2nd-order solver and the
    4th-order clcStat to the 4th-order one; each rejects the other context."""
    from channel3d_amd import Channel
    from channel3d_amd import lib as L
    g4 = Channel(cases.SMALL4)
    with pytest.raises(Exception, match="2nd-order"):
        g4._call("clcStat_CH", L.UX, L.UY, L.UZ, L.PR)
    g4.close()
    g2 = Channel(cases.SMALL)
    with pytest.raises(Exception, match="order-4"):
        g2._call("clcStat", L.UX, L.UY, L.UZ, L.PR)
    g2.close()


# ------------------------------------------------------------------ restart --
def _interiors(g, O=None):
    from channel3d_amd import lib as L
    out = []
    for fid in (L.UX, L.UY, L.UZ, L.PR):
        a = g.download(fid)
        if g.order == 4:
            out.append(np.asfortranarray(a[2:g.nxc + 2, 1:g.nyc + 1, 2:g.nzc + 2]))
        else:
            out.append(np.asfortranarray(a[1:g.nxc + 1, 1:g.nyc + 1, 1:g.nzc + 1]))
    return out


def _file_fields(path, nx, ny, nz, nf):
    raw = np.fromfile(path, dtype="<f8")
    assert raw.size == nf * nx * ny * nz
    return [raw[q * nx * ny * nz:(q + 1) * nx * ny * nz].reshape((nx, ny, nz), order="F")
            for q in range(nf)]


def test_restart_file_layout_and_round_trip(tmp_path, O):
    from channel3d_amd import lib as L
    case = cases.ODD
    o, g = make_pair(O, case, laminar=False)
    random_state(O, o, g, seed=9)
    want = _interiors(g)
    hist = [g.download(f) for f in (L.HISTX, L.HISTY, L.HISTZ)]
    p4 = str(tmp_path / "r4")
    p7 = str(tmp_path / "r7")
    g.write_restart(p4)
    g.write_restart(p7, hist=(L.HISTX, L.HISTY, L.HISTZ))
    got = _file_fields(p4, o.nxc, o.nyc, o.nzc, 4)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    got7 = _file_fields(p7, o.nxc, o.nyc, o.nzc, 7)
    for a, b in zip(got7, want + hist):
        assert np.array_equal(a, b)
    # wipe, read back: interior restored bit for bit, histories too
    z = np.zeros(g.shape(L.HALO), order="F")
    for fid in (L.UX, L.UY, L.UZ, L.PR):
        g.upload(fid, z)
    zi = np.zeros(g.shape(L.INTERIOR), order="F")
    for fid in (L.HISTX, L.HISTY, L.HISTZ):
        g.upload(fid, zi)
    g.read_restart(p7, hist=(L.HISTX, L.HISTY, L.HISTZ))
    for a, b in zip(_interiors(g), want):
        assert np.array_equal(a, b)
    for f, h in zip((L.HISTX, L.HISTY, L.HISTZ), hist):
        assert np.array_equal(g.download(f), h)
    # a 4-array file restores the fields when no histories are asked for
    g.read_restart(p4)
    # size checks (m_IOAndVisu.f90:455-467) and a missing file
    with pytest.raises(Exception, match="file byte wrong1"):
        g.read_restart(p4, hist=(L.HISTX, L.HISTY, L.HISTZ))
    bad = str(tmp_path / "bad")
    np.zeros(5 * o.nxc * o.nyc * o.nzc).tofile(bad)
    with pytest.raises(Exception, match="file byte wrong2"):
        g.read_restart(bad)
    with pytest.raises(Exception, match="cannot open"):
        g.read_restart(str(tmp_path / "missing"))
    with pytest.raises(Exception, match="history"):
        g.write_restart(p4, hist=(L.HISTX, L.HISTY, None))
    g.close()
    o.close()


def test_restart_continues_the_run(tmp_path, O):
    """2 steps, restart file, 2 more steps in a fresh context == 4 steps"""
    from channel3d_amd import Channel
    from channel3d_amd import lib as L
    case = cases.SMALL
    o, g = make_pair(O, case, seed=4)
    g.SetBC_and_UpdateHalo(L.UX, L.UY, L.UZ, L.UY_YM)
    g.SetBC_and_UpdateHalo_pr(L.PR)
    for _ in range(2):
        g.step()
    path = str(tmp_path / "RestartForTest0000000002")
    g.write_restart(path)
    for _ in range(2):
        g.step()
    ref = _interiors(g)
    g2 = Channel(case, yp=o.yp)
    g2.read_restart(path)
    # ChannelInitialize after read_restart (m_ChannelSystem.f90:54-62)
    g2.Update_uy_ym()
    g2.SetBC_and_UpdateHalo(L.UX, L.UY, L.UZ, L.UY_YM)
    g2.SetBC_and_UpdateHalo_pr(L.PR)
    for _ in range(2):
        g2.step()
    for a, b in zip(_interiors(g2), ref):
        assert np.array_equal(a, b)
    g.close()
    g2.close()
    o.close()


def test_restart_fourth_order_crf(tmp_path):
    """the 4th-order file stores ux + uCRF and reading subtracts it again"""
    from channel3d_amd import Channel, synth
    from channel3d_amd import lib as L
    case = dict(cases.SMALL4, u_crf=0.75)
    g = Channel(case)
    nx, ny, nz = g.nxc, g.nyc, g.nzc
    rng = np.random.default_rng(1)
    fields = []
    for fid in (L.UX, L.UY, L.UZ, L.PR):
        a = np.asfortranarray(rng.standard_normal(g.shape(L.HALO)))
        g.upload(fid, a)
        fields.append(a[2:nx + 2, 1:ny + 1, 2:nz + 2])
    path = str(tmp_path / "r4th")
    g.write_restart(path)
    got = _file_fields(path, nx, ny, nz, 4)
    assert np.array_equal(got[0], fields[0] + 0.75)
    for a, b in zip(got[1:], fields[1:]):
        assert np.array_equal(a, b)
    g.read_restart(path)
    back = _interiors(g)
    assert np.array_equal(back[0], (fields[0] + 0.75) - 0.75)
    for a, b in zip(back[1:], fields[1:]):
        assert np.array_equal(a, b)
    g.close()


# ------------------------------------------------------- 4th-order statistics --
def _periodic4(a):
    """2-deep periodic x and z halo layers of a 4th-order halo array"""
    nx, nz = a.shape[0] - 4, a.shape[2] - 4
    a[0:2] = a[nx:nx + 2]
    a[nx + 2:nx + 4] = a[2:4]
    a[:, :, 0:2] = a[:, :, nz:nz + 2]
    a[:, :, nz + 2:nz + 4] = a[:, :, 2:4]
    return a


@pytest.mark.parametrize("crf", [0.0, 0.4])
def test_stats4_random_fields(O, crf):
    from channel3d_amd import Channel
    from channel3d_amd import lib as L
    case = dict(cases.SMALL4, u_crf=crf)
    o = O.Oracle4(case)
    g = Channel(case, yp=o.yp)
    rng = np.random.default_rng(8)
    for fid, gid in ((O.UX, L.UX), (O.UY, L.UY), (O.UZ, L.UZ), (O.PR, L.PR)):
        a = _periodic4(np.asfortranarray(rng.standard_normal(g.shape(L.HALO))))
        o.field(fid)[...] = a
        g.upload(gid, a)
    g.stat_reset()
    g.stat_set_grad(True)
    g.clcStat()
    g.clcStat()
    s_g, n, _ = g.stat_get()
    q_g = g.stat_get_grad()
    ny = o.nyc
    s_o = np.zeros((35, ny + 1), order="F")
    q_o = np.zeros((98, ny + 1), order="F")
    s_m = np.zeros((35, ny + 1), order="F")
    q_m = np.zeros((98, ny + 1), order="F")
    for _ in range(2):
        o.call("clcStat", o.field(O.UX), o.field(O.UY), o.field(O.UZ), o.field(O.PR), s_o, q_o)
        o.call("clcStat_mag", o.field(O.UX), o.field(O.UY), o.field(O.UZ), o.field(O.PR), s_m,
               q_m)
    assert n == 2
    nt = o.nxc * o.nzc
    r1, z1 = _sum_err(s_g, s_o, s_m, nt)
    r2, z2 = _sum_err(q_g, q_o, q_m, nt)
    print("4th-order stats error / rounding bound", r1, "grad moments", r2, flush=True)
    assert r1 <= 1.0 and z1 and r2 <= 1.0 and z2
    g.close()
    o.close()


def test_stat4_files(tmp_path, O):
    from channel3d_amd import Channel, synth
    from channel3d_amd import lib as L
    case = dict(cases.SMALL4, is_ux_const=False)
    g = Channel(case)
    rng = np.random.default_rng(3)
    for gid in (L.UX, L.UY, L.UZ, L.PR):
        g.upload(gid, _periodic4(np.asfortranarray(rng.standard_normal(g.shape(L.HALO)))))
    g.stat_reset()
    g.stat_set_grad(True)
    g.clcStat()
    s_g, n, _ = g.stat_get()
    q_g = g.stat_get_grad()
    vp = str(tmp_path / "vGrad0000000010")
    g.write_vgrad(vp)
    sp = str(tmp_path / "stats0000000010")
    g.write_stat(sp, itime=10, ivstats=10)
    rows = [ln for ln in open(vp).read().split("\n") if ln]
    assert len(rows) == g.nyc + 1 and all(len(ln) == 98 * 24 for ln in rows)
    got = np.array([[float(ln[24 * q:24 * q + 24]) for q in range(98)] for ln in rows]).T
    assert np.array_equal(got, np.array([[float("%.15E" % v) for v in col] for col in q_g]))
    lines = open(sp).read().split("\n")
    assert lines[2].startswith("  Variable velocity in x-dir while adding a constant body force.")
    rdyc = g.metric(L.M_RDYC, 0, g.nyc + 1)
    dudy = abs(s_g[0, 0]) * 2.0 * rdyc[1] + abs(s_g[0, g.nyc - 1]) * 2.0 * rdyc[g.nyc + 1]
    assert float(lines[2][len("  Variable velocity in x-dir while adding a constant body force."):]) \
        == float("%.15E" % (case["xnu"] * dudy / case["yly"]))
    g.close()
