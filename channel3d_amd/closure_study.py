"""The learned subgrid closure on developed channel turbulence, a priori and
a posteriori.

1. DNS: the long-run case re395_2nd (2nd order, 384x192x288, Re_tau 395)
   from the reference's streak start, ``transient`` steps to developed
   turbulence, then ``nsnap`` snapshots ``every`` steps apart.  Each adds its
   LES samples (box filter 4x2x4 onto 96x96x72) to the closure's data set;
   every ``fstat``-th is also downloaded and box-filtered on the host for the
   filtered-DNS statistics, and clcStat_CH samples the DNS every 10 steps.
2. Training on the first 80 % of the snapshots (bf16 MFMA, Adam), and the
   a-priori correlation of the predicted with the filtered-DNS subgrid stress
   on the last 20 %, against the dynamic Smagorinsky model on the same states.
3. A posteriori: LES on the 96x96x72 grid (every second DNS y node) from a
   filtered DNS state, with the network (dynamic Smagorinsky where it
   backscatters), with dynamic Smagorinsky alone and with WALE (the
   reference's model, m_LESModels.f90:72-135); after a transient, their mean
   velocity, resolved Reynolds stresses and Re_tau against the filtered DNS.

No counterpart in the reference; its LES host is m_LESModels.f90:47-62 and
ChannelIterate's Clc_SGS_Vis call (m_ChannelSystem.f90:89).
"""
from __future__ import annotations

import json
import math
import time

import numpy as np

from . import lib as L
from . import longrun

FX, FY, FZ = 4, 2, 4


def centred(g, ux, uy, uz):
    """cell-centred velocity (nxc, nyc, nzc) from a 2nd-order context's halo
    arrays: ux(ic) sits on the face left of cell ic, uy(jc) on the node below,
    uz(kc) on the face behind (m_Variables.f90 staggering)"""
    nx, ny, nz = g.nxc, g.nyc, g.nzc
    u = 0.5 * (ux[1:nx + 1, 1:ny + 1, 1:nz + 1] + ux[2:nx + 2, 1:ny + 1, 1:nz + 1])
    v = 0.5 * (uy[1:nx + 1, 1:ny + 1, 1:nz + 1] + uy[1:nx + 1, 2:ny + 2, 1:nz + 1])
    w = 0.5 * (uz[1:nx + 1, 1:ny + 1, 1:nz + 1] + uz[1:nx + 1, 1:ny + 1, 2:nz + 2])
    return u, v, w


def box_filter(a, dyp, fx=FX, fy=FY, fz=FZ):
    """means over fx x fy x fz cell blocks, y weighted by the cell heights"""
    nx, ny, nz = a.shape
    w = np.asarray(dyp, dtype=np.float64)[None, :, None]
    s = (a * w).reshape(nx // fx, fx, ny // fy, fy, nz // fz, fz).sum(axis=(1, 3, 5))
    ws = np.asarray(dyp).reshape(ny // fy, fy).sum(axis=1)
    return s / (fx * fz * ws[None, :, None])


class RowStats:
    """per-row means and second moments of cell-centred velocity samples"""

    def __init__(self, ny):
        self.n = 0
        self.m = np.zeros((3, ny))
        self.s = np.zeros((4, ny))            # uu vv ww uv

    def add(self, u, v, w):
        self.n += 1
        for q, a in enumerate((u, v, w)):
            self.m[q] += a.mean(axis=(0, 2))
        self.s[0] += (u * u).mean(axis=(0, 2))
        self.s[1] += (v * v).mean(axis=(0, 2))
        self.s[2] += (w * w).mean(axis=(0, 2))
        self.s[3] += (u * v).mean(axis=(0, 2))

    def result(self):
        m = self.m / self.n
        s = self.s / self.n
        return dict(U=m[0], uu=s[0] - m[0] ** 2, vv=s[1] - m[1] ** 2, ww=s[2] - m[2] ** 2,
                    uv=s[3] - m[0] * m[1])


def fold(a, sign=1.0):
    """average of the lower half and the mirrored upper half"""
    n = a.size // 2
    return 0.5 * (a[:n] + sign * a[::-1][:n])


def les_case(dns_case, dt):
    return dict(nxc=dns_case["nxc"] // FX, nyc=dns_case["nyc"] // FY, nzc=dns_case["nzc"] // FZ,
                xlx=dns_case["xlx"], yly=dns_case["yly"], zlz=dns_case["zlz"],
                xnu=dns_case["xnu"], dt=dt, ubulk=dns_case.get("ubulk", 1.0))


def les_start(h, fil):
    """an LES state from filtered DNS cell averages: face values as the mean
    of the two neighbouring cells (periodic in x and z), v on the nodes as the
    mean of the cells below and above (0 on the walls), then projected"""
    u, v, w = fil
    nx, ny, nz = u.shape
    shp = h.shape(L.HALO)
    ux = np.zeros(shp, order="F")
    uy = np.zeros(shp, order="F")
    uz = np.zeros(shp, order="F")
    ux[1:nx + 1, 1:ny + 1, 1:nz + 1] = 0.5 * (u + np.roll(u, 1, axis=0))
    vv = np.zeros((nx, ny + 1, nz))
    vv[:, 1:ny, :] = 0.5 * (v[:, 1:, :] + v[:, :-1, :])
    uy[1:nx + 1, 1:ny + 2, 1:nz + 1] = vv
    uz[1:nx + 1, 1:ny + 1, 1:nz + 1] = 0.5 * (w + np.roll(w, 1, axis=2))
    for f, a in ((L.UX, ux), (L.UY, uy), (L.UZ, uz)):
        h.upload(f, a)
    h.upload(L.PR, np.zeros(shp, order="F"))
    h.project()


def run_les(h, steps, stat_from, every, log):
    st = RowStats(h.nyc)
    pg = []
    t0 = time.time()
    for n in range(1, steps + 1):
        h.step()
        if n > stat_from and n % every == 0:
            u, v, w = centred(h, h.download(L.UX), h.download(L.UY), h.download(L.UZ))
            st.add(u, v, w)
            pg.append(h.diagnostics()["PrGradAver"])
        if n % 2000 == 0:
            d = h.diagnostics()
            log(f"  LES step {n}: PrGradAver {d['PrGradAver']:.4e} cfl {d['cfl']:.3f} "
                f"div {d['divmax2']:.1e} vmax {max(d['vmax_x'], d['vmax_y'], d['vmax_z']):.3f}")
            if not math.isfinite(d["vmax_x"]) or d["vmax_x"] > 4.0:
                raise RuntimeError(f"LES left the velocity limit at step {n}")
    h.sync()
    return st.result(), float(np.mean(pg)), (time.time() - t0) / steps


def compare(les, ref):
    """relative L2 errors over the half channel of the folded profiles"""
    out = {}
    for k, sgn in (("U", 1.0), ("uu", 1.0), ("vv", 1.0), ("ww", 1.0), ("uv", -1.0)):
        a, b = fold(les[k], sgn), fold(ref[k], sgn)
        out[k] = float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
    return out


def _corr(a, b):
    a = np.concatenate([np.ravel(x) for x in a])
    b = np.concatenate([np.ravel(x) for x in b])
    a = a - a.mean()
    b = b - b.mean()
    return float((a * b).sum() / np.sqrt((a * a).sum() * (b * b).sum()))


def _dev(t):
    tr = (t[0] + t[1] + t[2]) / 3.0
    return [t[0] - tr, t[1] - tr, t[2] - tr, t[3], t[4], t[5]]


def main(argv=None):
    import argparse
    from .channel import Channel
    from .closure import Closure, unflatten
    ap = argparse.ArgumentParser(description="learned closure on channel turbulence")
    ap.add_argument("--transient", type=int, default=24000)
    ap.add_argument("--nsnap", type=int, default=200)
    ap.add_argument("--every", type=int, default=40)
    ap.add_argument("--fstat", type=int, default=10)
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--les-dt", type=float, default=0.01)
    ap.add_argument("--les-steps", type=int, default=9000)
    ap.add_argument("--les-stat-from", type=int, default=3000)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    t0 = time.time()
    res = dict(workload="learned SGS closure on Re_tau=395 channel turbulence: DNS 384x192x288 "
                        "(2nd order, FP64) box-filtered 4x2x4 to LES 96x96x72")
    # 1. DNS ----------------------------------------------------------------
    g = longrun.make("re395_2nd")
    c = g.case
    longrun.start(g, seed=0)
    log(f"DNS {g.nxc}x{g.nyc}x{g.nzc}: {a.transient} transient steps of dt {c['dt']}")
    hist = longrun.run(g, a.transient, a.transient, report=2000, t0=t0, log=log)
    nI, nJ, nK = g.nxc // FX, g.nyc // FY, g.nzc // FZ
    nl = nI * nJ * nK
    cl = Closure(a.nsnap * nl)
    dyp = np.diff(np.asarray(g.yp[1:g.nyc + 2]))
    fstats = RowStats(nJ)
    ntrain_snap = int(round(0.8 * a.nsnap))
    held = []
    fil_last = None
    log(f"sampling {a.nsnap} snapshots every {a.every} steps")
    itime = 0
    for s in range(a.nsnap):
        for _ in range(a.every):
            g.step()
            itime += 1
            if itime % 10 == 0:
                g.clcStat()
        first, cnt = cl.add_samples(g, FX, FY, FZ)
        if s >= ntrain_snap:
            tm, tt, _ = cl.dsm(g, FX, FY, FZ)
            held.append((first, cnt, tm.astype(np.float32), tt.astype(np.float32)))
        if s % a.fstat == 0 or s == a.nsnap - 1:
            u, v, w = centred(g, g.download(L.UX), g.download(L.UY), g.download(L.UZ))
            fil_last = tuple(box_filter(x, dyp) for x in (u, v, w))
            fstats.add(*fil_last)
        if s % 20 == 19:
            d = g.diagnostics()
            log(f"  snapshot {s + 1}: PrGradAver {d['PrGradAver']:.4e} "
                f"({time.time() - t0:.0f} s)")
    dns = longrun.reduce(g)
    res["dns"] = {k: v for k, v in dns.items() if k != "profile"}
    res["dns"]["history"] = hist
    res["dns"]["profile"] = dns["profile"]
    utau = dns["utau"]
    xnu = c["xnu"]
    g.close()
    log(f"DNS: Re_tau {dns['Re_tau']:.1f} (u_tau {utau:.5f}), {time.time() - t0:.0f} s")
    fref = fstats.result()
    # 2. training and a priori ----------------------------------------------
    ntrain = held[0][0]
    rng = np.random.default_rng(0)
    v = []
    for fan, shape in ((10, (16, 64)), (10, (64,)), (64, (64, 64)), (64, (64,)),
                       (64, (64, 64)), (64, (64,)), (64, (64, 6)), (64, (6,))):
        v.append(rng.uniform(-1 / np.sqrt(fan), 1 / np.sqrt(fan), size=shape).astype(np.float32))
    v[0][10:] = 0.0
    cl.set_params(unflatten(np.concatenate([x.ravel() for x in v])))
    tt0 = time.time()
    losses = []
    for frac, lr in ((0.5, 1e-3), (0.25, 3e-4), (0.25, 1e-4)):
        k = max(1, int(a.train_steps * frac))
        losses.append(cl.train(0, ntrain, batch=a.batch, steps=k, lr=lr, seed=len(losses)))
    t_train = time.time() - tt0
    losses = np.concatenate(losses)
    pred = {"mlp": [], "dsm": [], "true": []}
    for first, cnt, tm, tt in held:
        _, _, S = cl.samples(first, cnt)
        Y = cl.predict(first, cnt)
        pred["mlp"].append(_dev([(Y[:, q] * S).reshape((nI, nJ, nK), order="F")
                                 for q in range(6)]))
        pred["dsm"].append(list(tm))
        pred["true"].append(_dev(list(tt)))
    apriori = {}
    for key in ("mlp", "dsm"):
        apriori[key] = dict(
            tau12=round(_corr([d[3] for d in pred[key]], [d[3] for d in pred["true"]]), 4),
            deviatoric=round(_corr([x for d in pred[key] for x in d],
                                   [x for d in pred["true"] for x in d]), 4))
    del pred, held
    res["training"] = dict(samples=int(ntrain), snapshots_train=ntrain_snap,
                           snapshots_held_out=a.nsnap - ntrain_snap, batch=a.batch,
                           steps=int(losses.size), seconds=round(t_train, 2),
                           samples_per_s=round(losses.size * a.batch / t_train, 1),
                           loss_first=float(losses[:50].mean()),
                           loss_last=float(losses[-50:].mean()))
    res["a_priori_corr_held_out"] = apriori
    log(f"training: {res['training']}; a priori {apriori}")
    # 3. a posteriori ---------------------------------------------------------
    lc = les_case(c, a.les_dt)
    yp_les = np.zeros(nJ + 2)
    yp_les[1:] = np.asarray(g.yp)[1::FY][:nJ + 1]
    post = {}
    for model in ("network", "dynamic_smagorinsky", "wale"):
        h = Channel(lc, yp=yp_les)
        h.set_les(2)
        if model != "wale":
            cl.set_mode(0 if model == "network" else 1)
            h.set_les_closure(cl)
        les_start(h, fil_last)
        log(f"LES {model}: {a.les_steps} steps of dt {a.les_dt}")
        st, pg, spt = run_les(h, a.les_steps, a.les_stat_from, 50, log)
        nfb = cl.fallbacks() if model == "network" else None
        if model != "wale":
            h.set_les_closure(None)
        h.close()
        ut = math.sqrt(abs(pg) * 0.5 * lc["yly"])
        err = compare(st, fref)
        post[model] = dict(Re_tau=ut * 0.5 * lc["yly"] / xnu, u_tau=ut,
                           rel_l2_err_vs_filtered_dns=err, ms_per_step=round(1e3 * spt, 3),
                           fallback_fraction=(round(nfb / nl, 4) if nfb is not None else None),
                           profile={k: fold(v_, -1.0 if k == "uv" else 1.0).tolist()
                                    for k, v_ in st.items()})
        log(f"  {model}: Re_tau {post[model]['Re_tau']:.1f}, errors {err}")
    cl.close()
    yc_les = 0.5 * (yp_les[1:nJ + 1] + yp_les[2:nJ + 2])
    res["filtered_dns"] = dict(Re_tau=dns["Re_tau"], u_tau=utau,
                               yplus=(fold(yc_les) * utau / xnu).tolist(),
                               profile={k: fold(v_, -1.0 if k == "uv" else 1.0).tolist()
                                        for k, v_ in fref.items()}, samples=fstats.n)
    res["a_posteriori"] = post
    res["wall_s"] = round(time.time() - t0, 1)
    print(json.dumps(dict(dns_Re_tau=dns["Re_tau"], a_priori=apriori,
                          a_posteriori={k: dict(Re_tau=round(v_["Re_tau"], 1),
                                                err=v_["rel_l2_err_vs_filtered_dns"])
                                        for k, v_ in post.items()})), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
