"""Host-side mirror of the reference's channel time-step interface.

``Channel`` owns one device context of libchannel3d.so.  Its methods carry
the reference's routine names and argument order (m_TScheme.f90:28-30,
m_Poisson.f90:16, m_ChannelSystem.f90:91-130); arrays move in and out in
the reference's Fortran layouts:

  halo fields      (nxc+3, nyc+3, nzc+3), view[ic, jc, kc] == ux(ic, jc, kc)
                   4th order (case["order"] == 4, src/CFD_4th):
                   (nxc+4, nyc+2, nzc+4), view[ic+1, jc, kc+1] == ux(ic, jc, kc)
  interior fields  (nxc, nyc, nzc),       view[ic-1, jc-1, kc-1]
  wall planes      (nxc, nzc)

Device fields are referred to by the state ids of ``lib`` (UX, UY, ...) or by
handles returned from ``alloc``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import lib as L


def mesh_coords(nyc, yly, section_length, nyc_section, stret_type, stret_option, cstret):
    """yp(1..nyc+1) of the y grid (m_MeshAndMetries.f90:144-160, 292-360);
    returns nyc+2 values indexed like the reference (entry 0 unused)."""
    lib = L.load()
    ns = len(section_length)
    sl = np.ascontiguousarray(section_length, dtype=np.float64)
    cs = np.ascontiguousarray(cstret, dtype=np.float64)
    ints = [np.ascontiguousarray(v, dtype=np.int32) for v in (nyc_section, stret_type,
                                                             stret_option)]
    out = np.zeros(nyc + 2)
    L.check(lib.c3d_mesh_coords(int(nyc), float(yly), ns, sl.ctypes.data_as(L._DP),
                                *(v.ctypes.data_as(L._IP) for v in ints),
                                cs.ctypes.data_as(L._DP), out.ctypes.data_as(L._DP)),
            "c3d_mesh_coords")
    return out


def restart_name(restart_dir, run_name, ntime):
    """RestartDir//"RestartFor"//RunName//I10.10 (m_IOAndVisu.f90:363)"""
    lib = L.load()
    buf = ctypes.create_string_buffer(4096)
    L.check(lib.c3d_restart_name(os.fsencode(restart_dir), os.fsencode(run_name), int(ntime),
                                 buf, len(buf)), "c3d_restart_name")
    return buf.value.decode()


def channel_yp(nyc, yly, gamma):
    """Two tanh sections (bottom, top), StretType 1, as the channel inputs use
    (Input/TurbulentChannel_2nd.prm MeshOptions)."""
    h = nyc // 2
    return mesh_coords(nyc, yly, [1.0, 1.0], [h, nyc - h], [1, 1], [0, 1], [gamma, gamma])


def make_params(case: dict, precision: int = L.FP64) -> L.Params:
    p = L.Params()
    p.nxc, p.nyc, p.nzc = int(case["nxc"]), int(case["nyc"]), int(case["nzc"])
    p.xlx, p.yly, p.zlz = float(case["xlx"]), float(case["yly"]), float(case["zlz"])
    p.xnu, p.dt = float(case["xnu"]), float(case["dt"])
    bc = case.get("bc", (0, 0, L.BC_NOSLIP, L.BC_NOSLIP, 0, 0))
    for d in range(6):
        p.bc[d] = int(bc[d])
    for name in ("uxbc", "uybc", "uzbc"):
        vals = case.get(name, (0.0,) * 6)
        arr = getattr(p, name)
        for d in range(6):
            arr[d] = float(vals[d])
    for d, g in enumerate(case.get("gravity", (0.0, 0.0, 0.0))):
        p.gravity[d] = float(g)
    p.is_ux_const = int(case.get("is_ux_const", True))
    p.ubulk = float(case.get("ubulk", 1.0))
    p.ybulk1 = float(case.get("ybulk1", 0.0))
    p.ybulk2 = float(case.get("ybulk2", case["yly"]))
    p.ischeme = int(case.get("ischeme", 3))
    p.precision = int(precision)
    p.order = int(case.get("order", 2))
    p.u_crf = float(case.get("u_crf", 0.0))
    if p.order == 4 and "bc" not in case and case.get("flow_type", 1) == 2:
        p.bc[3] = L.BC_FREESLIP        # half channel (FT_HC)
    return p


class Batch:
    """a batch of device fields in one allocation: base address and the
    handle of each field"""

    def __init__(self, base, fields):
        self.base = base
        self.fields = fields

    def __len__(self):
        return len(self.fields)

    def __getitem__(self, q):
        return self.fields[q]


class Channel:
    def __init__(self, case: dict, yp=None, precision: int = L.FP64):
        self.lib = L.load()
        self.case = dict(case)
        self.params = make_params(case, precision)
        if yp is None:
            yp = channel_yp(self.params.nyc, self.params.yly, case.get("gamma", 1.5))
        self.yp = np.ascontiguousarray(yp, dtype=np.float64)
        h = ctypes.c_void_p()
        L.check(self.lib.c3d_create(ctypes.byref(self.params), self.yp.ctypes.data_as(L._DP),
                                    ctypes.byref(h)), "c3d_create")
        self._h = h
        self.nxc, self.nyc, self.nzc = self.params.nxc, self.params.nyc, self.params.nzc
        self.order = self.params.order
        self._state = {}
        for fid in range(12):
            v = ctypes.c_void_p()
            L.check(self.lib.c3d_state_field(self._h, fid, ctypes.byref(v)), "c3d_state_field")
            self._state[fid] = v
        if case.get("les_type", 0):
            self.set_les(case["les_type"], case.get("filter_type", 1))

    # lifetime ----------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self.lib.c3d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # fields ------------------------------------------------------------------
    def shape(self, layout):
        if layout == L.HALO:
            if self.order == 4:
                return (self.nxc + 4, self.nyc + 2, self.nzc + 4)
            return (self.nxc + 3, self.nyc + 3, self.nzc + 3)
        if layout == L.INTERIOR:
            return (self.nxc, self.nyc, self.nzc)
        return (self.nxc, self.nzc)

    def ptr(self, f):
        """device pointer of a state id or an alloc() handle"""
        if isinstance(f, int):
            return self._state[f]
        return f

    def alloc(self, layout):
        v = ctypes.c_void_p()
        L.check(self.lib.c3d_field_alloc(self._h, layout, ctypes.byref(v)), "c3d_field_alloc")
        return v

    def alloc_batch(self, layout, nbatch):
        """one device allocation of nbatch fields of a layout, back to back;
        returns the handles of its fields (field(q) addresses)"""
        v = ctypes.c_void_p()
        L.check(self.lib.c3d_field_alloc_batch(self._h, layout, int(nbatch), ctypes.byref(v)),
                "c3d_field_alloc_batch")
        fb = int(self.lib.c3d_field_bytes(self._h, layout))
        return Batch(v, [ctypes.c_void_p(v.value + q * fb) for q in range(nbatch)])

    def ppe_batch(self, prsrc, phi, sweeps=-1):
        """nbatch pressure solves (Batch handles: INTERIOR sources, HALO phi)"""
        assert len(prsrc) == len(phi)
        L.check(self.lib.c3d_ppe_batch(self._h, len(prsrc), prsrc.base, phi.base, int(sweeps)),
                "c3d_ppe_batch")

    def ppe_batch_noalloc(self, prsrc, phi, sweeps=-1):
        """ppe_batch on the library's fixed work space (no device allocation;
        calls that overlap on one GPU race)"""
        assert len(prsrc) == len(phi)
        L.check(self.lib.c3d_ppe_batch_noalloc(self._h, len(prsrc), prsrc.base, phi.base,
                                               int(sweeps)), "c3d_ppe_batch_noalloc")

    def helmholtz_batch(self, comp, ns, rhs, du, sweeps=-1):
        """(1 - b Dxx)(1 - b Dzz)(1 - b Dyy) du = rhs of component comp (0..2)
        at substep ns, for every field of the batch (INTERIOR layouts)"""
        assert len(rhs) == len(du)
        L.check(self.lib.c3d_helmholtz_batch(self._h, int(comp), int(ns), len(rhs), rhs.base,
                                             du.base, int(sweeps)), "c3d_helmholtz_batch")

    def free(self, handle):
        if isinstance(handle, Batch):
            handle = handle.base
        L.check(self.lib.c3d_field_free(self._h, handle), "c3d_field_free")

    def upload(self, f, host, layout=None):
        layout = L.STATE_LAYOUT[f] if layout is None else layout
        a = np.asfortranarray(host, dtype=np.float64)
        if a.shape != self.shape(layout):
            raise ValueError(f"shape {a.shape} != {self.shape(layout)}")
        L.check(self.lib.c3d_upload(self._h, layout, self.ptr(f), L.dptr(a)), "c3d_upload")

    def download(self, f, layout=None):
        layout = L.STATE_LAYOUT[f] if layout is None else layout
        a = np.zeros(self.shape(layout), dtype=np.float64, order="F")
        L.check(self.lib.c3d_download(self._h, layout, self.ptr(f), L.dptr(a)), "c3d_download")
        return a

    def copy(self, dst, src, layout=L.HALO):
        """device-to-device copy of a whole field"""
        L.check(self.lib.c3d_field_copy(self._h, layout, self.ptr(dst), self.ptr(src)),
                "c3d_field_copy")

    def metric(self, mid, lo, hi):
        out = np.zeros(hi - lo + 1)
        L.check(self.lib.c3d_get_metric(self._h, mid, lo, hi, out.ctypes.data_as(L._DP)),
                "c3d_get_metric")
        return out

    def sync(self):
        L.check(self.lib.c3d_sync(self._h), "c3d_sync")

    def set_dt(self, dt):
        L.check(self.lib.c3d_set_dt(self._h, float(dt)), "c3d_set_dt")

    def set_ppe_refine(self, sweeps):
        """-1: FP64 pressure solve; s >= 0: FP32 solve + s FP64-residual sweeps"""
        L.check(self.lib.c3d_set_ppe_refine(self._h, int(sweeps)), "c3d_set_ppe_refine")

    def set_graph(self, on=True):
        """step() replays HIP graphs of the whole step (captured on first use)"""
        L.check(self.lib.c3d_set_graph(self._h, int(bool(on))), "c3d_set_graph")

    def clcSpectra1D(self, jlcs=(1, 2), ux=L.UX, uy=L.UY, uz=L.UZ, pressure=L.PR):
        """one sample of the 4th-order 1-D spectra (clcStat with clcSpectra1D)"""
        L.check(self.lib.c3d_clcSpectra1D(self._h, *[self.ptr(f) for f in (ux, uy, uz, pressure)],
                                          int(jlcs[0]), int(jlcs[1])), "c3d_clcSpectra1D")

    def spectra_get(self):
        """(EnergySpecX (nxc/2+1, nyc, 18), EnergySpecZ (nzc/2+1, nyc, 18), nSpectime)"""
        sx = np.zeros((self.nxc // 2 + 1, self.nyc, L.NSPEC1D), order="F")
        sz = np.zeros((self.nzc // 2 + 1, self.nyc, L.NSPEC1D), order="F")
        n = ctypes.c_int()
        L.check(self.lib.c3d_spectra_get(self._h, L.dptr(sx), L.dptr(sz), ctypes.byref(n)),
                "c3d_spectra_get")
        return sx, sz, n.value

    def spectra_reset(self):
        L.check(self.lib.c3d_spectra_reset(self._h), "c3d_spectra_reset")

    def write_spectra(self, path_x, path_z):
        L.check(self.lib.c3d_write_spectra(self._h, os.fsencode(path_x), os.fsencode(path_z)),
                "c3d_write_spectra")

    def clcSpectra2D(self, jlcs=1, ux=L.UX, uy=L.UY, uz=L.UZ, pressure=L.PR):
        """one sample of the 4th-order 2-D spectra (clcStat with clcSpectra2D)"""
        L.check(self.lib.c3d_clcSpectra2D(self._h, *[self.ptr(f) for f in (ux, uy, uz, pressure)],
                                          int(jlcs)), "c3d_clcSpectra2D")

    def spectra2d_get(self):
        """(EnergySpec2D (nxc/2+1, nySpec2D, nzc/2+1, 9), nSpectime)"""
        nys = self.nyc // 2 if self.params.bc[3] == L.BC_NOSLIP else self.nyc
        s = np.zeros((self.nxc // 2 + 1, nys, self.nzc // 2 + 1, 9), order="F")
        n = ctypes.c_int()
        L.check(self.lib.c3d_spectra2d_get(self._h, L.dptr(s), ctypes.byref(n)),
                "c3d_spectra2d_get")
        return s, n.value

    def write_spectra2d(self, path):
        L.check(self.lib.c3d_write_spectra2d(self._h, os.fsencode(path)), "c3d_write_spectra2d")

    def set_les(self, les_type, filter_type=1):
        """LES_type 0 (off), 2 (WALE) or 3 (MTS); FilterType 0 trapezoidal, 1 Simpson
        (InitFilterCoe and the LES branch of InitTimeScheme)"""
        L.check(self.lib.c3d_set_les(self._h, int(les_type), int(filter_type)), "c3d_set_les")
        if les_type:
            v = ctypes.c_void_p()
            L.check(self.lib.c3d_state_field(self._h, L.NUT, ctypes.byref(v)), "c3d_state_field")
            self._state[L.NUT] = v

    def set_les_closure(self, closure):
        """the learned closure (channel3d_amd.closure.Closure) as this LES
        context's SGS model; None restores WALE / MTS"""
        L.check(self.lib.c3d_set_les_closure(self._h, closure._h if closure else None),
                "c3d_set_les_closure")
        self._les_closure = closure   # kept alive while in use

    def Clc_SGS_Vis(self, ux=L.UX, uy=L.UY, uz=L.UZ, nut=L.NUT):
        self._call("Clc_SGS_Vis", ux, uy, uz, nut)

    # reference routines --------------------------------------------------------
    def _call(self, name, *fields):
        fn = getattr(self.lib, "c3d_" + name)
        L.check(fn(self._h, *[self.ptr(f) for f in fields]), "c3d_" + name)

    def PMcoeUpdate(self, ns):
        L.check(self.lib.c3d_PMcoeUpdate(self._h, int(ns)), "c3d_PMcoeUpdate")

    def Update_uy_ym(self, uy_ym=L.UY_YM, duy_ym=L.DUY_YM):
        self._call("Update_uy_ym", uy_ym, duy_ym)

    def clcRhsX(self, ux, uy, uz, RhsX, HistXOld, pressure):
        self._call("clcRhsX", ux, uy, uz, RhsX, HistXOld, pressure)

    def clcRhsY(self, ux, uy, uz, RhsY, HistYOld, pressure):
        self._call("clcRhsY", ux, uy, uz, RhsY, HistYOld, pressure)

    def clcRhsZ(self, ux, uy, uz, RhsZ, HistZOld, pressure):
        self._call("clcRhsZ", ux, uy, uz, RhsZ, HistZOld, pressure)

    def clcRhsXYZ(self, ux, uy, uz, RhsX, RhsY, RhsZ, HistXOld, HistYOld, HistZOld, pressure):
        """4th-order contexts (src/CFD_4th/m_TScheme.f90:189-438)"""
        self._call("clcRhsXYZ", ux, uy, uz, RhsX, RhsY, RhsZ, HistXOld, HistYOld, HistZOld,
                   pressure)

    def clcU1Hat(self, ux, RhsX):
        self._call("clcU1Hat", ux, RhsX)

    def clcU2Hat(self, uy, RhsY, duy_ym):
        self._call("clcU2Hat", uy, RhsY, duy_ym)

    def clcU3Hat(self, uz, RhsZ):
        self._call("clcU3Hat", uz, RhsZ)

    def SetBC_and_UpdateHaloForPrSrc(self, ux, uy, uz, uy_ym):
        self._call("SetBC_and_UpdateHaloForPrSrc", ux, uy, uz, uy_ym)

    def SetBC_and_UpdateHalo_pr(self, pressure):
        self._call("SetBC_and_UpdateHalo_pr", pressure)

    def SetBC_and_UpdateHalo(self, ux, uy, uz, uy_ym):
        self._call("SetBC_and_UpdateHalo", ux, uy, uz, uy_ym)

    def clcPrSrc(self, ux, uy, uz, prsrc, pressure=None):
        """pressure is corrected by -nu/2 div on the 2nd-order path; the
        4th-order routine has no pressure argument (pass None)"""
        out = ctypes.c_double()
        pr = None if pressure is None else self.ptr(pressure)
        L.check(self.lib.c3d_clcPrSrc(self._h, *[self.ptr(f) for f in (ux, uy, uz, prsrc)], pr,
                                      ctypes.byref(out)), "c3d_clcPrSrc")
        return out.value

    def clcPPE(self, prsrc, prphiHalo):
        self._call("clcPPE", prsrc, prphiHalo)

    def FluidVelUpdate(self, prphiHalo, ux, uy, uz):
        self._call("FluidVelUpdate", prphiHalo, ux, uy, uz)

    def PressureUpdate(self, pressure, prphiHalo):
        self._call("PressureUpdate", pressure, prphiHalo)

    def _diag(self, name, fields, n):
        out = (ctypes.c_double * n)()
        fn = getattr(self.lib, "c3d_" + name)
        L.check(fn(self._h, *[self.ptr(f) for f in fields], out), "c3d_" + name)
        return list(out) if n > 1 else out[0]

    def CalcMaxCFL(self, ux=L.UX, uy=L.UY, uz=L.UZ):
        return self._diag("CalcMaxCFL", (ux, uy, uz), 1)

    def CheckDivergence(self, ux=L.UX, uy=L.UY, uz=L.UZ):
        return self._diag("CheckDivergence", (ux, uy, uz), 1)

    def CalcVmax(self, ux=L.UX, uy=L.UY, uz=L.UZ):
        return self._diag("CalcVmax", (ux, uy, uz), 3)

    def CalcUxAver(self, ux=L.UX):
        return self._diag("CalcUxAver", (ux,), 1)

    # fused, device-resident --------------------------------------------------
    def substep(self, ns):
        L.check(self.lib.c3d_substep(self._h, int(ns)), "c3d_substep")

    def step(self):
        L.check(self.lib.c3d_step(self._h), "c3d_step")

    def project(self):
        L.check(self.lib.c3d_project(self._h), "c3d_project")

    def ppe(self, prsrc, phi):
        L.check(self.lib.c3d_ppe(self._h, self.ptr(prsrc), self.ptr(phi)), "c3d_ppe")

    def diagnostics(self):
        out = (ctypes.c_double * 9)()
        L.check(self.lib.c3d_step_diagnostics(self._h, out), "c3d_step_diagnostics")
        keys = ("cfl", "divmax1", "divmax2", "vmax_x", "vmax_y", "vmax_z", "PrGradNow",
                "PrGradAver", "SimTime")
        return dict(zip(keys, list(out)))

    # restart files (m_IOAndVisu.f90:352-558) ----------------------------------
    def write_restart(self, path, ux=L.UX, uy=L.UY, uz=L.UZ, pressure=L.PR, hist=None):
        """hist: None or the three history fields (HistXOld, HistYOld, HistZOld)"""
        h = [None] * 3 if hist is None else [self.ptr(f) for f in hist]
        L.check(self.lib.c3d_write_restart(self._h, os.fsencode(path), self.ptr(ux),
                                           self.ptr(uy), self.ptr(uz), self.ptr(pressure), *h),
                "c3d_write_restart")

    def read_restart(self, path, ux=L.UX, uy=L.UY, uz=L.UZ, pressure=L.PR, hist=None):
        h = [None] * 3 if hist is None else [self.ptr(f) for f in hist]
        L.check(self.lib.c3d_read_restart(self._h, os.fsencode(path), self.ptr(ux),
                                          self.ptr(uy), self.ptr(uz), self.ptr(pressure), *h),
                "c3d_read_restart")

    # channel statistics (m_FlowType_Channel.f90:135-363) ----------------------
    def clcStat(self, ux=L.UX, uy=L.UY, uz=L.UZ, pressure=L.PR):
        """clcStat_CH (2nd order) or the 4th-order code's clcStat"""
        self._call("clcStat" if self.order == 4 else "clcStat_CH", ux, uy, uz, pressure)

    def stat_set_grad(self, on=True):
        L.check(self.lib.c3d_stat_set_grad(self._h, int(on)), "c3d_stat_set_grad")

    def stat_get_grad(self):
        q = np.zeros((L.NGRADSTAT, self.nyc + 1), order="F")
        L.check(self.lib.c3d_stat_get_grad(self._h, L.dptr(q)), "c3d_stat_get_grad")
        return q

    def write_vgrad(self, path):
        L.check(self.lib.c3d_write_vgrad(self._h, os.fsencode(path)), "c3d_write_vgrad")

    def stat_get(self):
        """(SumStat as a (35, nyc+1) array, nfstime, PrGradsum)"""
        s = np.zeros((L.NCHASTAT, self.nyc + 1), order="F")
        n = ctypes.c_int()
        pg = ctypes.c_double()
        L.check(self.lib.c3d_stat_get(self._h, L.dptr(s), ctypes.byref(n), ctypes.byref(pg)),
                "c3d_stat_get")
        return s, n.value, pg.value

    def stat_reset(self):
        L.check(self.lib.c3d_stat_reset(self._h), "c3d_stat_reset")

    def write_stat(self, path, itime, ivstats):
        L.check(self.lib.c3d_write_stat(self._h, os.fsencode(path), int(itime), int(ivstats)),
                "c3d_write_stat")

    def append_prgrad(self, path, itime):
        L.check(self.lib.c3d_append_prgrad(self._h, os.fsencode(path), int(itime)),
                "c3d_append_prgrad")

    # profiling ----------------------------------------------------------------
    def profile(self, on=True):
        L.check(self.lib.c3d_profile_enable(self._h, int(on)), "c3d_profile_enable")

    def profile_reset(self):
        L.check(self.lib.c3d_profile_reset(self._h), "c3d_profile_reset")

    def profile_stats(self):
        n = self.lib.c3d_profile_count(self._h)
        out = {}
        for i in range(n):
            name = ctypes.c_char_p()
            calls = ctypes.c_long()
            ms = ctypes.c_double()
            b = ctypes.c_double()
            L.check(self.lib.c3d_profile_get(self._h, i, ctypes.byref(name), ctypes.byref(calls),
                                             ctypes.byref(ms), ctypes.byref(b)),
                    "c3d_profile_get")
            if calls.value:
                out[name.value.decode()] = dict(calls=calls.value, total_ms=ms.value,
                                                bytes_per_call=b.value)
        return out

    def device_bytes(self):
        return int(self.lib.c3d_device_bytes(self._h))
