"""Seeded synthetic inputs for the channel workloads.

Every random value is a pure function of (seed, field id, linear cell index)
through the SplitMix64 finaliser, so any process (oracle, GPU run, bench)
regenerates bit-identical fields without sharing state.  Arrays come out in
the reference's Fortran layouts (see channel.py).
"""
from __future__ import annotations

import numpy as np

_K1 = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + _K1
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, field: int, n: int, start: int = 0) -> np.ndarray:
    """n values in [0, 1) for linear indices start .. start+n-1"""
    key = _mix(np.array([np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(field)],
                        dtype=np.uint64))[0]
    idx = np.arange(start, start + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = _mix(idx * np.uint64(0xD1B54A32D192ED03) ^ key)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform_pm1(seed, field, shape):
    n = int(np.prod(shape))
    return (2.0 * uniform01(seed, field, n) - 1.0).reshape(shape, order="F")


def normal(seed, field, shape):
    n = int(np.prod(shape))
    u1 = uniform01(seed, field, n)
    u2 = uniform01(seed, field + 1000003, n)
    r = np.sqrt(-2.0 * np.log1p(-u1))
    return (r * np.cos(2.0 * np.pi * u2)).reshape(shape, order="F")


def yc_from_yp(yp: np.ndarray) -> np.ndarray:
    """cell-centre y, yc(j) for j = 1..nyc (m_MeshAndMetries.f90:172-174)"""
    nyc = yp.size - 2
    return 0.5 * (yp[1:nyc + 1] + yp[2:nyc + 2])


def laminar_noise(nxc, nyc, nzc, yp, yly=2.0, ubulk=1.0, amp=0.1, seed=0):
    """u = 1.5 ub (1 - (y/h-1)^2) + amp ub U(-1,1), v = amp ub U(-1,1) off the
    wall node, w = amp ub U(-1,1); returns halo arrays (halo layers zero)."""
    shp = (nxc + 3, nyc + 3, nzc + 3)
    ux = np.zeros(shp, order="F")
    uy = np.zeros(shp, order="F")
    uz = np.zeros(shp, order="F")
    h = 0.5 * yly
    yc = yc_from_yp(yp)
    prof = 1.5 * ubulk * (1.0 - (yc / h - 1.0) ** 2)
    ish = (nxc, nyc, nzc)
    ux[1:nxc + 1, 1:nyc + 1, 1:nzc + 1] = prof[None, :, None] + amp * ubulk * uniform_pm1(seed, 1, ish)
    noise_v = amp * ubulk * uniform_pm1(seed, 2, ish)
    noise_v[:, 0, :] = 0.0
    uy[1:nxc + 1, 1:nyc + 1, 1:nzc + 1] = noise_v
    uz[1:nxc + 1, 1:nyc + 1, 1:nzc + 1] = amp * ubulk * uniform_pm1(seed, 3, ish)
    return ux, uy, uz


def ppe_rhs(nxc, nyc, nzc, seed=0):
    """N(0,1) interior field with its mean removed (a compatible PPE source)"""
    f = normal(seed, 7, (nxc, nyc, nzc))
    return np.asfortranarray(f - f.mean())


def random_halo(nxc, nyc, nzc, seed, field, amp=1.0):
    """uniform(-amp, amp) everywhere in the halo array, halo layers included"""
    return amp * uniform_pm1(seed, field, (nxc + 3, nyc + 3, nzc + 3))


def random_interior(nxc, nyc, nzc, seed, field, amp=1.0):
    return amp * uniform_pm1(seed, field, (nxc, nyc, nzc))


def smooth_modes(nxc, nyc, nzc, yp, xlx, yly, zlz, seed=0, nmodes=48, kmax=6, amp=0.25,
                 ubulk=1.0):
    """laminar profile plus random smooth modes on the staggered grid: x and
    z wavenumbers 1..kmax, wall-normal sines sin(m pi y / yly), m = 1..4,
    vanishing at the walls, amplitudes amp N(0,1)/sqrt(nmodes); halo arrays
    (halo layers zero), to be projected.  A field with resolved structure at
    and above a few cells' scale (the learned closure's training data)."""
    rng = np.random.default_rng(seed)
    xf = np.arange(nxc) * xlx / nxc
    xc = xf + 0.5 * xlx / nxc
    zf = np.arange(nzc) * zlz / nzc
    zc = zf + 0.5 * zlz / nzc
    yn = np.asarray(yp[1:nyc + 1])
    yc = 0.5 * (np.asarray(yp[1:nyc + 1]) + np.asarray(yp[2:nyc + 2]))
    ux, uy, uz = laminar_noise(nxc, nyc, nzc, yp, yly=yly, ubulk=ubulk, amp=0.0, seed=seed)
    for comp, (x, y, z) in zip((ux, uy, uz), ((xf, yc, zc), (xc, yn, zc), (xc, yc, zf))):
        f = np.zeros((nxc, nyc, nzc))
        for _ in range(nmodes):
            kx, kz, m = rng.integers(1, kmax + 1), rng.integers(1, kmax + 1), rng.integers(1, 5)
            a = amp * ubulk * rng.standard_normal() / np.sqrt(nmodes)
            px, pz = rng.uniform(0, 2 * np.pi, 2)
            f += a * (np.sin(2 * np.pi * kx * x / xlx + px)[:, None, None]
                      * np.sin(m * np.pi * y / yly)[None, :, None]
                      * np.sin(2 * np.pi * kz * z / zlz + pz)[None, None, :])
        comp[1:nxc + 1, 1:nyc + 1, 1:nzc + 1] += f
    return ux, uy, uz


def spectrum_field(nxc, nyc, nzc, seed=0, k0=8.0, kcut=40, dyp=None):
    """a random-phase field whose energy spectrum over the index wavenumber
    |k| = |(kx, ky, kz)| is E(k) ~ k^4 exp(-2 (k/k0)^2): modal amplitudes
    k exp(-(k/k0)^2) (E(k) over the shell area 4 pi k^2, square-rooted),
    uniform phases from numpy's default_rng(seed), modes with |k| > kcut left
    out (their amplitude is below 1e-9 of the peak for k0 = 8), unit RMS.
    With dyp (nyc cell heights) the dyp-weighted mean is removed, so the field
    is a compatible source of the Neumann pressure equation (dyp is the left
    null vector of its wall-normal operator)."""
    import scipy.fft as sfft
    rng = np.random.default_rng(seed)
    K = int(kcut)
    kx = np.arange(-K, K + 1)
    ky = np.arange(-K, K + 1)
    kz = np.arange(0, K + 1)
    kk = np.sqrt(kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2)
    amp = kk * np.exp(-(kk / k0) ** 2)
    amp[kk > K] = 0.0
    ph = rng.uniform(0.0, 2.0 * np.pi, size=amp.shape)
    spec = np.zeros((nxc, nyc, nzc // 2 + 1), dtype=np.complex128)
    ix = kx % nxc
    iy = ky % nyc
    iz = kz[kz <= nzc // 2]
    band = amp * np.exp(1j * ph)
    # modes beyond the grid's own Nyquist limits fold onto lower indices only
    # on grids smaller than 2 kcut; they are added there
    for a, i in enumerate(ix):
        for b, jy in enumerate(iy):
            spec[i, jy, iz] += band[a, b, :iz.size]
    f = sfft.irfftn(spec, s=(nxc, nyc, nzc), workers=-1)
    f = np.asfortranarray(f / f.std())
    if dyp is not None:
        w = np.asarray(dyp, dtype=np.float64)
        m = (f.mean(axis=(0, 2)) * w).sum() / w.sum()
        f = np.asfortranarray(f - m)
    return f


def channel_init(nxc, nyc, nzc, yp, xlx, zlz, xnu, ubulk=1.0, yly=2.0, seed=0, order=2,
                 u_crf=0.0):
    """The reference's turbulent-channel start: a laminar profile
    3 ub (ybar - ybar^2/2) plus streamwise streaks and spanwise rolls of
    amplitude ub ybar exp(-4.5 ybar^2), modulated by cos(wz z+) and
    sin(wx x+) with wavelengths near 200 and 500 wall units and a random
    factor in [0.8, 1.2]; the streamwise velocity rescaled to the bulk velocity
    and the plane mean of the spanwise one removed (2nd order
    m_FlowType_Channel.f90:37-115 for ybulk = [0, yly]; 4th order
    m_FlowCase.f90:99-157, which rescales by CalcUxAver and subtracts uCRF).
    The random factor comes from this module's seeded generator instead of
    the clock-seeded Fortran random_number.  Returns interior arrays
    (nxc, nyc, nzc) of ux, uy (zero), uz; 2nd order: v is zero on the walls
    already."""
    height = 0.5 * yly
    rem = ubulk * height / xnu
    retau = 0.1538 * rem ** 0.887741
    utau = retau * xnu / height
    dev = 0.2 * uniform01(seed, 41, nxc * nyc * nzc).reshape((nxc, nyc, nzc), order="F") + 0.9
    wx, wz = 2.0 * np.pi / 500.0, 2.0 * np.pi / 200.0
    xlxp, zlzp = xlx * utau / xnu, zlz * utau / xnu
    m1 = np.floor(xlxp * wx / (2.0 * np.pi)) + 1
    wx = m1 * 2.0 * np.pi / xlxp
    m2 = np.floor(zlzp * wz / (2.0 * np.pi)) + 1
    wz = m2 * 2.0 * np.pi / zlzp
    yc = yc_from_yp(np.asarray(yp, dtype=np.float64))
    yct = height - np.abs(height - yc)
    ybar = yct / height
    dx, dz = xlx / nxc, zlz / nzc
    xplus = utau * ((np.arange(nxc) + 0.5) * dx) / xnu
    zplus = utau * ((np.arange(nzc) + 0.5) * dz) / xnu
    env = (ubulk * ybar * np.exp(-4.5 * ybar * ybar))[None, :, None]
    ux = env * np.cos(wz * zplus)[None, None, :] * dev
    uz = env * np.sin(wx * xplus)[:, None, None] * dev
    ux = ux + (3.0 * ubulk * (ybar - 0.5 * ybar * ybar))[None, :, None]
    if order == 4:
        # CalcUxAver of the 4th-order code: the dyp-weighted volume mean
        dyp = np.diff(np.asarray(yp[1:nyc + 2], dtype=np.float64))
        uxav = (ux.mean(axis=(0, 2)) * dyp).sum() / yly
        ux = ux * (ubulk / uxav) - u_crf
    else:
        dyp = np.diff(np.asarray(yp[1:nyc + 2], dtype=np.float64))
        ratio = (ux.mean(axis=(0, 2)) * dyp).sum() / yly / ubulk
        ux = ux * ratio
    uz = uz - uz.mean(axis=(0, 2))[None, :, None]
    uy = np.zeros((nxc, nyc, nzc))
    return np.asfortranarray(ux), np.asfortranarray(uy), np.asfortranarray(uz)
