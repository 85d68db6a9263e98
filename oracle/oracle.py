"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a C restatement of Channel3d's 2nd-order full-implicit channel
step (src/CFD_2nd/m_TSchemeFIMP_inc.f90, m_TScheme.f90, m_Poisson.f90,
m_Tools.f90, m_BC_and_Halo.f90, m_MeshAndMetries.f90, m_ChannelSystem.f90).
Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
may import this module; the product library never does.

Arrays are exposed as numpy views in the reference's own Fortran layout:
halo fields have shape (nxc+3, nyc+3, nzc+3) indexed [i, j, k] with the
reference's 0-based halo index (i.e. view[ic, jc, kc] == ux(ic, jc, kc)),
interior fields have shape (nxc, nyc, nzc) with view[ic-1, jc-1, kc-1].
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libc3d_oracle.so")

BC_PERIOD, BC_NOSLIP, BC_FREESLIP = 0, -1, -2

# field ids (c3d_oracle.h)
UX, UY, UZ, PR, HISTX, HISTY, HISTZ, ARR1, ARR2, HALO, UY_YM, DUY_YM, NUT = range(13)
HALO_FIELDS = (UX, UY, UZ, PR, HALO, NUT)
INTERIOR_FIELDS = (HISTX, HISTY, HISTZ, ARR1, ARR2)
WALL_FIELDS = (UY_YM, DUY_YM)

(M_YP, M_YC, M_DYP, M_DYC, M_RDYP, M_RDYC, M_YINTERP, M_AM2PR, M_AC2PR, M_AP2PR,
 M_AM2C, M_AC2C, M_AP2C, M_AM2P, M_AC2P, M_AP2P, M_AM2CCN, M_AP2CCN,
 M_WAVEX, M_WAVEZ, M_DELTA) = range(21)

(S_PMGAMMA, S_PMTHETA, S_PMALPHA, S_PMALPHAC, S_PMBETA, S_SIMTIME, S_PRGRADNOW,
 S_PRGRADAVER, S_NORMFFT, S_DIVMAX1, S_DIVMAX2, S_CFL, S_UXAVER, S_VMAX1, S_VMAX2,
 S_VMAX3, S_DT, S_PMBETAT) = range(18)

# 4th-order oracle (c3d_oracle4.h)
M4_AC2CCN = 21
FT_CH, FT_HC = 1, 2


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nxc", ctypes.c_int), ("nyc", ctypes.c_int), ("nzc", ctypes.c_int),
        ("xlx", ctypes.c_double), ("yly", ctypes.c_double), ("zlz", ctypes.c_double),
        ("xnu", ctypes.c_double), ("dt", ctypes.c_double),
        ("bc", ctypes.c_int * 6),
        ("uxbc", ctypes.c_double * 6), ("uybc", ctypes.c_double * 6),
        ("uzbc", ctypes.c_double * 6),
        ("gravity", ctypes.c_double * 3),
        ("is_ux_const", ctypes.c_int),
        ("ubulk", ctypes.c_double), ("ybulk1", ctypes.c_double), ("ybulk2", ctypes.c_double),
        ("ischeme", ctypes.c_int),
        ("les_type", ctypes.c_int), ("filter_type", ctypes.c_int),
    ]


class Orc4Params(ctypes.Structure):
    _fields_ = [
        ("nxc", ctypes.c_int), ("nyc", ctypes.c_int), ("nzc", ctypes.c_int),
        ("xlx", ctypes.c_double), ("yly", ctypes.c_double), ("zlz", ctypes.c_double),
        ("xnu", ctypes.c_double), ("dt", ctypes.c_double),
        ("flow_type", ctypes.c_int),
        ("uxbc", ctypes.c_double * 6), ("uybc", ctypes.c_double * 6),
        ("uzbc", ctypes.c_double * 6),
        ("gravity", ctypes.c_double * 3),
        ("is_ux_const", ctypes.c_int),
        ("u_crf", ctypes.c_double),
        ("ischeme", ctypes.c_int),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


_lib = None
_DP = ctypes.POINTER(ctypes.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams), _DP]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_field.restype = _DP
        L.orc_field.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_metric.restype = _DP
        L.orc_metric.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_scalar.restype = ctypes.c_double
        L.orc_scalar.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_scalar.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_mesh_coords.restype = ctypes.c_int
        L.orc_mesh_coords.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, _DP,
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_int), _DP, _DP]
        for name in ("orc_substep",):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int]
        for name in ("orc_iterate", "orc_project"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.orc_PMcoeUpdate.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_Update_uy_ym.argtypes = [ctypes.c_void_p, _DP, _DP]
        for name in ("orc_clcRhsX", "orc_clcRhsY", "orc_clcRhsZ"):
            getattr(L, name).argtypes = [ctypes.c_void_p] + [_DP] * 6
        L.orc_clcU1Hat.argtypes = [ctypes.c_void_p, _DP, _DP]
        L.orc_clcU2Hat.argtypes = [ctypes.c_void_p, _DP, _DP, _DP]
        L.orc_clcU3Hat.argtypes = [ctypes.c_void_p, _DP, _DP]
        L.orc_SetBC_and_UpdateHaloForPrSrc.argtypes = [ctypes.c_void_p] + [_DP] * 4
        L.orc_SetBC_and_UpdateHalo.argtypes = [ctypes.c_void_p] + [_DP] * 4
        L.orc_SetBC_and_UpdateHalo_pr.argtypes = [ctypes.c_void_p, _DP]
        L.orc_clcPrSrc.argtypes = [ctypes.c_void_p] + [_DP] * 6
        L.orc_clcPPE.argtypes = [ctypes.c_void_p, _DP, _DP]
        L.orc_set_ppe_route.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_sum_order.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_FluidVelUpdate.argtypes = [ctypes.c_void_p] + [_DP] * 4
        L.orc_PressureUpdate.argtypes = [ctypes.c_void_p, _DP, _DP]
        for name in ("orc_CalcMaxCFL", "orc_CheckDivergence"):
            getattr(L, name).restype = ctypes.c_double
            getattr(L, name).argtypes = [ctypes.c_void_p] + [_DP] * 3
        L.orc_CalcVmax.argtypes = [ctypes.c_void_p] + [_DP] * 4
        L.orc_CalcUxAver.restype = ctypes.c_double
        L.orc_CalcUxAver.argtypes = [ctypes.c_void_p, _DP]
        L.orc_clcStat_CH.argtypes = [ctypes.c_void_p] + [_DP] * 5
        L.orc_clcStat_CH_mag.argtypes = [ctypes.c_void_p] + [_DP] * 5
        L.orc_InverseTridiagonal.argtypes = [_DP, _DP, _DP, _DP, ctypes.c_int, ctypes.c_int]
        L.orc_InversePTriFixedCoe.argtypes = [ctypes.c_double] * 3 + [_DP, ctypes.c_int,
                                                                       ctypes.c_int]
        for name in ("orc_Clc_SGS_Vis", "orc_Clc_SGS_WALE", "orc_Clc_SGS_MTS",
                     "orc_ClcVelStrainRate"):
            getattr(L, name).argtypes = [ctypes.c_void_p] + [_DP] * 4
        L.orc_SetBC_and_UpdateHalo_SGS.argtypes = [ctypes.c_void_p, _DP]
        for name in ("orc_clcRhsX_les", "orc_clcRhsY_les", "orc_clcRhsZ_les"):
            getattr(L, name).argtypes = [ctypes.c_void_p] + [_DP] * 7
        L.orc_clcU1Hat_les.argtypes = [ctypes.c_void_p, _DP, _DP, _DP]
        L.orc_clcU2Hat_les.argtypes = [ctypes.c_void_p, _DP, _DP, _DP, _DP]
        L.orc_clcU3Hat_les.argtypes = [ctypes.c_void_p, _DP, _DP, _DP]
        L.orc_clcPrSrcOther.argtypes = [ctypes.c_void_p] + [_DP] * 5
        L.orc_PressureUpdate_les.argtypes = [ctypes.c_void_p, _DP, _DP, _DP]
        L.orc_xsum.restype = ctypes.c_double
        L.orc_xsum.argtypes = [ctypes.c_int, _DP]
        L.orc_r2hc.argtypes = [ctypes.c_int, _DP, _DP]
        L.orc_hc2r.argtypes = [ctypes.c_int, _DP, _DP]
        # 4th-order restatement
        L.orc4_create.restype = ctypes.c_void_p
        L.orc4_create.argtypes = [ctypes.POINTER(Orc4Params), _DP]
        L.orc4_destroy.argtypes = [ctypes.c_void_p]
        L.orc4_field.restype = _DP
        L.orc4_field.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc4_metric.restype = _DP
        L.orc4_metric.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc4_scalar.restype = ctypes.c_double
        L.orc4_scalar.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc4_set_scalar.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.orc4_substep.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc4_iterate.argtypes = [ctypes.c_void_p]
        L.orc4_project.argtypes = [ctypes.c_void_p]
        L.orc4_PMcoeUpdate.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc4_Update_uy_ym.argtypes = [ctypes.c_void_p, _DP, _DP]
        L.orc4_clcRhsXYZ.argtypes = [ctypes.c_void_p] + [_DP] * 10
        L.orc4_clcU1Hat.argtypes = [ctypes.c_void_p, _DP, _DP]
        L.orc4_clcU2Hat.argtypes = [ctypes.c_void_p, _DP, _DP, _DP]
        L.orc4_clcU3Hat.argtypes = [ctypes.c_void_p, _DP, _DP]
        L.orc4_SetBC_and_UpdateHalo.argtypes = [ctypes.c_void_p] + [_DP] * 4
        L.orc4_SetBC_and_UpdateHaloForPrSrc.argtypes = [ctypes.c_void_p] + [_DP] * 4
        L.orc4_SetBC_and_UpdateHalo_pr.argtypes = [ctypes.c_void_p, _DP]
        L.orc4_clcPrSrc.argtypes = [ctypes.c_void_p] + [_DP] * 5
        L.orc4_clcPPE.argtypes = [ctypes.c_void_p, _DP, _DP]
        L.orc4_set_ppe_route.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc4_FluidVelUpdate.argtypes = [ctypes.c_void_p] + [_DP] * 4
        L.orc4_PressureUpdate.argtypes = [ctypes.c_void_p, _DP, _DP]
        for name in ("orc4_CalcMaxCFL", "orc4_CheckDivergence"):
            getattr(L, name).restype = ctypes.c_double
            getattr(L, name).argtypes = [ctypes.c_void_p] + [_DP] * 3
        L.orc4_CalcVmax.argtypes = [ctypes.c_void_p] + [_DP] * 4
        L.orc4_CalcUxAver.restype = ctypes.c_double
        L.orc4_CalcUxAver.argtypes = [ctypes.c_void_p, _DP]
        L.orc4_clcStat.argtypes = [ctypes.c_void_p] + [_DP] * 6
        L.orc4_clcStat_mag.argtypes = [ctypes.c_void_p] + [_DP] * 6
        L.orc4_clcSpectra1D.argtypes = [ctypes.c_void_p] + [_DP] * 4 + [
            ctypes.POINTER(ctypes.c_int), _DP, _DP]
        L.orc4_clcSpectra2D.argtypes = [ctypes.c_void_p] + [_DP] * 4 + [ctypes.c_int, _DP]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(_DP)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


# ---------------------------------------------------------------- line solvers
def inverse_tridiagonal(a, b, c, f):
    """m_Tools.f90:218-245 on (m, n) Fortran arrays (line index fastest)."""
    a, b, c = (np.asfortranarray(x, dtype=np.float64) for x in (a, b, c))
    f = np.array(f, dtype=np.float64, order="F", copy=True)
    m, n = f.shape
    lib().orc_InverseTridiagonal(_ptr(a), _ptr(b), _ptr(c), _ptr(f), m, n)
    return f


def inverse_ptri_fixed(a, b, c, f):
    """m_Tools.f90:161-213 on an (m, n) Fortran array (line index fastest)."""
    f = np.array(f, dtype=np.float64, order="F", copy=True)
    m, n = f.shape
    lib().orc_InversePTriFixedCoe(float(a), float(b), float(c), _ptr(f), m, n)
    return f


def xsum(v):
    """order-independent sum used for the bulk-velocity integral"""
    v = np.ascontiguousarray(v, dtype=np.float64)
    return float(lib().orc_xsum(v.size, _ptr(v)))


def r2hc(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_r2hc(x.size, _ptr(x), _ptr(out))
    return out


def hc2r(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_hc2r(x.size, _ptr(x), _ptr(out))
    return out


def mesh_coords(nyc, yly, section_length, nyc_section, stret_type, stret_option, cstret):
    """y node coordinates yp(1..nyc+1) (m_MeshAndMetries.f90:144-160, 292-360).
    Returns an array of length nyc+2 indexed like the reference (entry 0 unused)."""
    ns = len(section_length)
    sl = np.ascontiguousarray(section_length, dtype=np.float64)
    cs = np.ascontiguousarray(cstret, dtype=np.float64)
    ints = [np.ascontiguousarray(v, dtype=np.int32) for v in (nyc_section, stret_type,
                                                             stret_option)]
    out = np.zeros(nyc + 2)
    ip = ctypes.POINTER(ctypes.c_int)
    rc = lib().orc_mesh_coords(nyc, float(yly), ns, _ptr(sl), *(v.ctypes.data_as(ip) for v in ints),
                               _ptr(cs), _ptr(out))
    if rc != 0:
        raise ValueError(f"orc_mesh_coords failed ({rc})")
    return out


def channel_yp(nyc, yly, gamma):
    """Two tanh sections (bottom/top), StretType 1, as in the channel inputs."""
    h = nyc // 2
    return mesh_coords(nyc, yly, [1.0, 1.0], [h, nyc - h], [1, 1], [0, 1], [gamma, gamma])


# ---------------------------------------------------------------- full context
class Oracle:
    def __init__(self, case: dict, yp=None):
        self.case = dict(case)
        p = OrcParams()
        p.nxc, p.nyc, p.nzc = case["nxc"], case["nyc"], case["nzc"]
        p.xlx, p.yly, p.zlz = case["xlx"], case["yly"], case["zlz"]
        p.xnu, p.dt = case["xnu"], case["dt"]
        bc = case.get("bc", (0, 0, BC_NOSLIP, BC_NOSLIP, 0, 0))
        for d in range(6):
            p.bc[d] = bc[d]
        for name in ("uxbc", "uybc", "uzbc"):
            vals = case.get(name, (0.0,) * 6)
            arr = getattr(p, name)
            for d in range(6):
                arr[d] = vals[d]
        for d, g in enumerate(case.get("gravity", (0.0, 0.0, 0.0))):
            p.gravity[d] = g
        p.is_ux_const = int(case.get("is_ux_const", True))
        p.ubulk = case.get("ubulk", 1.0)
        p.ybulk1 = case.get("ybulk1", 0.0)
        p.ybulk2 = case.get("ybulk2", case["yly"])
        p.ischeme = case.get("ischeme", 3)
        p.les_type = case.get("les_type", 0)
        p.filter_type = case.get("filter_type", 1)
        self.params = p
        if yp is None:
            yp = channel_yp(p.nyc, p.yly, case.get("gamma", 1.5))
        self.yp = np.ascontiguousarray(yp, dtype=np.float64)
        h = lib().orc_create(ctypes.byref(p), _ptr(self.yp))
        if not h:
            raise ValueError("orc_create rejected the parameters")
        self._h = ctypes.c_void_p(h)
        self.nxc, self.nyc, self.nzc = p.nxc, p.nyc, p.nzc

    def close(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # views -----------------------------------------------------------------
    def shape(self, fid):
        if fid in HALO_FIELDS:
            return (self.nxc + 3, self.nyc + 3, self.nzc + 3)
        if fid in INTERIOR_FIELDS:
            return (self.nxc, self.nyc, self.nzc)
        return (self.nxc, self.nzc)

    def field(self, fid) -> np.ndarray:
        shp = self.shape(fid)
        ptr = lib().orc_field(self._h, fid)
        n = int(np.prod(shp))
        buf = np.ctypeslib.as_array(ptr, shape=(n,))
        return buf.reshape(shp, order="F")

    def metric(self, mid, lo, hi) -> np.ndarray:
        ptr = lib().orc_metric(self._h, mid)
        base = ctypes.cast(ctypes.addressof(ptr.contents) + 8 * lo, _DP)
        return np.ctypeslib.as_array(base, shape=(hi - lo + 1,)).copy()

    def scalar(self, sid) -> float:
        return float(lib().orc_scalar(self._h, sid))

    def set_scalar(self, sid, v):
        lib().orc_set_scalar(self._h, sid, float(v))

    # routines --------------------------------------------------------------
    def substep(self, ns):
        lib().orc_substep(self._h, int(ns))

    def iterate(self):
        lib().orc_iterate(self._h)

    def project(self):
        lib().orc_project(self._h)

    def set_ppe_route(self, route):
        """0: clcPPE_x1 (x first), 1: clcPPE_z1 (z first)"""
        lib().orc_set_ppe_route(self._h, int(route))

    def set_sum_order(self, seq):
        """1: s1tot and CalcUxAver summed in the reference's sequential order;
        0: the exact fixed-point sum (default, what the library computes)"""
        lib().orc_set_sum_order(self._h, int(seq))

    def call(self, name, *arrays):
        fn = getattr(lib(), "orc_" + name)
        args = [self._h] + [_ptr(a) if isinstance(a, np.ndarray) else a for a in arrays]
        return fn(*args)

    def f(self, fid):
        """raw pointer argument for a context field (for .call)"""
        return self.field(fid)


# ------------------------------------------------------- 4th-order context
class Oracle4:
    """The 4th-order channel step (src/CFD_4th).  Halo fields have shape
    (nxc+4, nyc+2, nzc+4) and view[ic+1, jc, kc+1] == ux(ic, jc, kc) for the
    reference's halo index range (-1:nxc+2, 0:nyc+1, -1:nzc+2)."""

    def __init__(self, case: dict, yp=None):
        self.case = dict(case)
        p = Orc4Params()
        p.nxc, p.nyc, p.nzc = case["nxc"], case["nyc"], case["nzc"]
        p.xlx, p.yly, p.zlz = case["xlx"], case["yly"], case["zlz"]
        p.xnu, p.dt = case["xnu"], case["dt"]
        p.flow_type = case.get("flow_type", FT_CH)
        for name in ("uxbc", "uybc", "uzbc"):
            vals = case.get(name, (0.0,) * 6)
            arr = getattr(p, name)
            for d in range(6):
                arr[d] = vals[d]
        for d, g in enumerate(case.get("gravity", (0.0, 0.0, 0.0))):
            p.gravity[d] = g
        p.is_ux_const = int(case.get("is_ux_const", True))
        p.u_crf = case.get("u_crf", 0.0)
        p.ischeme = case.get("ischeme", 3)
        self.params = p
        if yp is None:
            yp = channel_yp(p.nyc, p.yly, case.get("gamma", 1.5))
        self.yp = np.ascontiguousarray(yp, dtype=np.float64)
        h = lib().orc4_create(ctypes.byref(p), _ptr(self.yp))
        if not h:
            raise ValueError("orc4_create rejected the parameters")
        self._h = ctypes.c_void_p(h)
        self.nxc, self.nyc, self.nzc = p.nxc, p.nyc, p.nzc

    def close(self):
        if getattr(self, "_h", None):
            lib().orc4_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def shape(self, fid):
        if fid in HALO_FIELDS:
            return (self.nxc + 4, self.nyc + 2, self.nzc + 4)
        if fid in INTERIOR_FIELDS:
            return (self.nxc, self.nyc, self.nzc)
        return (self.nxc, self.nzc)

    def field(self, fid) -> np.ndarray:
        shp = self.shape(fid)
        ptr = lib().orc4_field(self._h, fid)
        buf = np.ctypeslib.as_array(ptr, shape=(int(np.prod(shp)),))
        return buf.reshape(shp, order="F")

    def metric(self, mid, lo, hi) -> np.ndarray:
        ptr = lib().orc4_metric(self._h, mid)
        base = ctypes.cast(ctypes.addressof(ptr.contents) + 8 * lo, _DP)
        return np.ctypeslib.as_array(base, shape=(hi - lo + 1,)).copy()

    def scalar(self, sid) -> float:
        return float(lib().orc4_scalar(self._h, sid))

    def set_scalar(self, sid, v):
        lib().orc4_set_scalar(self._h, sid, float(v))

    def substep(self, ns):
        lib().orc4_substep(self._h, int(ns))

    def iterate(self):
        lib().orc4_iterate(self._h)

    def project(self):
        lib().orc4_project(self._h)

    def set_ppe_route(self, route):
        """0: clcPPE_x1 (x first), 1: clcPPE_z1 (z first)"""
        lib().orc4_set_ppe_route(self._h, int(route))

    def call(self, name, *arrays):
        fn = getattr(lib(), "orc4_" + name)
        args = [self._h] + [_ptr(a) if isinstance(a, np.ndarray) else a for a in arrays]
        return fn(*args)
