"""Long channel runs: the reference's time loop with its statistics, and the
reduction of those statistics to wall units.

``start`` puts the reference's turbulent-channel initial field
(synth.channel_init) into a context and projects it.  ``run`` advances it
as ChannelIterate does (m_ChannelSystem.f90:73-154: one fused RK3 step per
time step), samples clcStat every ``ivstats`` steps after ``stat_from``
(4th m_ChannelSystem.f90:144-158; readStat_CH.m's iTimeSet drops the
statistics files before ``stat_from``) and records the step diagnostics.  ``reduce``
turns the accumulated SumStat into the quantities Tool/readStat_CH.m
prints: u_tau from the mean pressure gradient and from the wall shear
(readStat_CH.m:163-193), Re_tau, and the mean-velocity and Reynolds-stress
profiles in wall units, averaged over the two halves of the channel.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import lib as L
from . import synth


def _place(g, arr):
    """an interior array (nxc, nyc, nzc) inside a zero halo array of the
    context's layout (channel.py)"""
    nx, ny, nz = arr.shape
    h = np.zeros(g.shape(L.HALO), order="F")
    if g.order == 4:
        h[2:nx + 2, 1:ny + 1, 2:nz + 2] = arr
    else:
        h[1:nx + 1, 1:ny + 1, 1:nz + 1] = arr
    return h


def start(g, seed=0):
    """the reference's initial field for this context's case, projected"""
    c = g.case
    ux, uy, uz = synth.channel_init(g.nxc, g.nyc, g.nzc, g.yp, c["xlx"], c["zlz"], c["xnu"],
                                    ubulk=c.get("ubulk", 1.0), yly=c.get("yly", 2.0), seed=seed,
                                    order=g.order, u_crf=c.get("u_crf", 0.0))
    for fid, a in ((L.UX, ux), (L.UY, uy), (L.UZ, uz)):
        g.upload(fid, _place(g, a))
    g.upload(L.PR, np.zeros(g.shape(L.HALO), order="F"))
    g.project()
    g.sync()


def run(g, nsteps, stat_from, ivstats=10, report=1000, vel_limit=4.0, log=print, t0=None):
    """nsteps fused steps; clcStat after every ivstats-th step past
    stat_from; every `report` steps the step diagnostics are read and logged.
    Stops with RuntimeError if the velocity leaves vel_limit or turns NaN
    (the reference's vel_limit check, 4th m_ChannelSystem.f90:146-148)."""
    hist = []
    t0 = time.time() if t0 is None else t0
    for itime in range(1, nsteps + 1):
        g.step()
        if itime > stat_from and itime % ivstats == 0:
            g.clcStat()
        if itime % report == 0 or itime == nsteps:
            d = g.diagnostics()
            vmax = max(d["vmax_x"], d["vmax_y"], d["vmax_z"])
            rec = dict(itime=itime, SimTime=d["SimTime"], PrGradAver=d["PrGradAver"],
                       cfl=d["cfl"], divmax=d["divmax2"], vmax=vmax,
                       wall_s=round(time.time() - t0, 1))
            hist.append(rec)
            log(f"step {itime:6d} t {d['SimTime']:9.3f} PrGradAver {d['PrGradAver']:.6e} "
                f"cfl {d['cfl']:.3f} div {d['divmax2']:.2e} vmax {vmax:.3f} "
                f"({rec['wall_s']} s)")
            if not (vmax < vel_limit) or not math.isfinite(d["divmax2"]):
                raise RuntimeError(f"run left the velocity limit at step {itime}: {rec}")
    return hist


def reduce(g):
    """readStat_CH.m's reduction of this context's accumulated statistics.

    Column numbers are clcStat's SumVec indices (1-based): 1 ux, 2 uy (at
    the yp nodes), 3 uz, 5 ux ux, 6 uy uy (nodes), 7 uz uz, 9 uxCell uyCell.
    Returns a dict: u_tau from the pressure gradient (utau1) and the wall
    shear (utau2) and their mean, Re_tau, Re_b of the mean profile, and the
    lower/upper-averaged profiles over y+ at the cell centres (U+, u'+, w'+,
    -u'v'+) and the nodes (v'+)."""
    s, n, pgsum = g.stat_get()
    if n < 1:
        raise RuntimeError("no statistics samples")
    d = np.asarray(s).T / n                      # (nyp, 35), row j-1 = jc
    xnu = g.case["xnu"]
    yly = g.case.get("yly", 2.0)
    height = 0.5 * yly
    yp = np.asarray(g.yp[1:g.nyc + 2])            # yp(1..nyp)
    yc = 0.5 * (yp[:-1] + yp[1:])
    nyc = g.nyc
    dyp = np.diff(yp)
    dyc = np.empty(nyc + 1)
    dyc[1:nyc] = yc[1:] - yc[:-1]
    dyc[0], dyc[nyc] = dyp[0], dyp[nyc - 1]
    U = d[:nyc, 0]
    # readStat_CH.m:160-170: wall gradients from the first cell centre
    dudy_bot = 2.0 * U[0] / dyc[0]
    dudy_top = -2.0 * U[nyc - 1] / dyc[nyc]
    prgrad = abs(pgsum) / n
    utau1 = math.sqrt(prgrad * height)
    utau2 = 0.5 * (math.sqrt(xnu * dudy_bot) + math.sqrt(-xnu * dudy_top))
    utau = 0.5 * (utau1 + utau2)
    ubulk = float((U * dyp).sum() / yly)
    uu = d[:nyc, 4] - U ** 2
    ww = d[:nyc, 6] - d[:nyc, 2] ** 2
    V = d[:, 1]
    vv = d[:, 5] - V ** 2
    Vc = 0.5 * (V[:nyc] + V[1:nyc + 1])
    uv = d[:nyc, 8] - U * Vc
    h = nyc // 2
    lo, hi = slice(0, h), slice(nyc - 1, nyc - 1 - h, -1)
    yplus = yc[lo] * utau / xnu
    prof = dict(
        yplus=yplus.tolist(),
        Uplus=(0.5 * (U[lo] + U[hi]) / utau).tolist(),
        urms=(np.sqrt(np.maximum(0.5 * (uu[lo] + uu[hi]), 0.0)) / utau).tolist(),
        wrms=(np.sqrt(np.maximum(0.5 * (ww[lo] + ww[hi]), 0.0)) / utau).tolist(),
        uv=(-0.5 * (uv[lo] - uv[hi]) / utau ** 2).tolist(),
    )
    hn = (nyc + 1) // 2
    vnode = 0.5 * (vv[:hn] + vv[nyc:nyc - hn:-1])
    prof["yplus_node"] = (yp[:hn] * utau / xnu).tolist()
    prof["vrms"] = (np.sqrt(np.maximum(vnode, 0.0)) / utau).tolist()
    urms = np.asarray(prof["urms"])
    uvp = np.asarray(prof["uv"])
    return dict(nfstime=n, prgrad=prgrad, utau1=utau1, utau2=utau2, utau=utau,
                Re_tau=utau * height / xnu, Re_b=ubulk * height / xnu, ubulk=ubulk,
                Uplus_centre=prof["Uplus"][-1], urms_peak=float(urms.max()),
                urms_peak_yplus=float(yplus[int(urms.argmax())]),
                vrms_max=float(np.max(prof["vrms"])), wrms_max=float(np.max(prof["wrms"])),
                uv_peak=float(uvp.max()), uv_peak_yplus=float(yplus[int(uvp.argmax())]),
                profile=prof)


# Named long-run cases.  cha180_4th is Input/TurbulentChannel_4th_direct.prm
# (4th order, 384x192x256, two sine-stretched sections, uCRF, fixed dt 0.02,
# 27000 steps, statistics every 10 steps; readStat_CH.m takes iTimeSet 6000).
# re395_2nd is the learned closure's DNS: 2nd order, 384x192x288, Lx = 2 pi,
# Lz = pi, Re_tau 395 by the reference's own regression (nu = 1.4428e-4),
# two tanh sections with gamma 1.5.
CASES = {
    "cha180_4th": dict(
        # IsUseCRF: the wall values of ux become uxBcValue - uCRF
        # (4th m_Parameters.f90:152-153)
        case=dict(nxc=384, nyc=192, nzc=256, xlx=12.56637061, yly=2.0, zlz=6.283185307,
                  xnu=2.3310e-4, dt=0.02, order=4, ubulk=2.0 / 3.0, u_crf=0.6666666,
                  uxbc=(-0.6666666,) * 6),
        mesh=dict(section_length=[1.0, 1.0], nyc_section=[96, 96], stret_type=[2, 2],
                  stret_option=[0, 1], cstret=[1.0, 1.0]),
        nsteps=27000, stat_from=6000, ivstats=10),
    "re395_2nd": dict(
        case=dict(nxc=384, nyc=192, nzc=288, xlx=2.0 * math.pi, yly=2.0, zlz=math.pi,
                  xnu=1.4428e-4, dt=0.005, gamma=1.5, ubulk=1.0),
        mesh=None, nsteps=40000, stat_from=20000, ivstats=10),
}


def make(name, **over):
    """a Channel for a named case (keyword overrides go into its dict)"""
    from .channel import Channel, channel_yp, mesh_coords
    spec = CASES[name]
    case = dict(spec["case"], **over)
    if spec["mesh"]:
        m = spec["mesh"]
        yp = mesh_coords(case["nyc"], case["yly"], m["section_length"], m["nyc_section"],
                         m["stret_type"], m["stret_option"], m["cstret"])
    else:
        yp = channel_yp(case["nyc"], case["yly"], case.get("gamma", 1.5))
    return Channel(case, yp=yp)


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="long channel run with statistics")
    ap.add_argument("case", choices=sorted(CASES))
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--stat-from", type=int, default=None)
    ap.add_argument("--report", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None, help="JSON file for the reduced statistics")
    a = ap.parse_args(argv)
    spec = CASES[a.case]
    nsteps = a.steps or spec["nsteps"]
    stat_from = spec["stat_from"] if a.stat_from is None else a.stat_from
    t0 = time.time()
    g = make(a.case)
    start(g, seed=a.seed)
    print(f"{a.case}: {g.nxc}x{g.nyc}x{g.nzc} order {g.order}, {nsteps} steps of dt "
          f"{g.case['dt']}, statistics every {spec['ivstats']} after step {stat_from}",
          flush=True)
    hist = run(g, nsteps, stat_from, spec["ivstats"], a.report, t0=t0,
               log=lambda s: print(s, flush=True))
    res = reduce(g)
    res.update(case=a.case, steps=nsteps, stat_from=stat_from, history=hist,
               wall_s=round(time.time() - t0, 1),
               ms_per_step=round(1e3 * (hist[-1]["wall_s"] - 0.0) / nsteps, 3))
    print(json.dumps({k: v for k, v in res.items() if k not in ("profile", "history")}),
          flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)
    g.close()


if __name__ == "__main__":
    main()
