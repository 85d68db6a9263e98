"""Parity of the 4th-order 1-D spectra on the GPU (c3d_clcSpectra1D,
src/CFD_4th/m_FlowCase.f90:644-842) with the CPU oracle, and the SpecX /
SpecZ files.  The GPU transforms with its own Stockham FFTs (along z two x
points per complex line), the oracle with its restated FFTW R2HC, so the
columns agree to rounding: bar 1e-12 relative to each column's largest
value.
"""
import ctypes
import os

import numpy as np
import pytest

from tests import cases
from tests.test_gpu_parity4 import make_pair

pytestmark = pytest.mark.gpu

NCOL = 18


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def _oracle_spectra(O, o, jl):
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    sx = np.zeros((nx // 2 + 1, ny, NCOL), order="F")
    sz = np.zeros((nz // 2 + 1, ny, NCOL), order="F")
    o.call("clcSpectra1D", o.field(O.UX), o.field(O.UY), o.field(O.UZ), o.field(O.PR),
           (ctypes.c_int * 2)(*jl), sx, sz)
    return sx, sz


def _same_state(O, o, g):
    """the oracle's fields in the device context, so that both compute the
    spectra of identical inputs"""
    from channel3d_amd import lib as L
    for fid, gid in ((O.UX, L.UX), (O.UY, L.UY), (O.UZ, L.UZ), (O.PR, L.PR)):
        g.upload(gid, o.field(fid))


def _colerr(got, ref):
    errs = []
    for c in range(NCOL):
        s = np.abs(ref[:, :, c]).max()
        errs.append(np.abs(got[:, :, c] - ref[:, :, c]).max() / (s if s > 0 else 1.0))
    return np.array(errs)


@pytest.mark.parametrize("name,ucrf", [("small4", 0.0), ("small4", 0.4), ("half4", 0.0)])
def test_spectra_match_oracle(O, name, ucrf):
    case = dict({"small4": cases.SMALL4, "half4": cases.HALF4}[name], u_crf=ucrf)
    o, g = make_pair(O, case, seed=6)
    for _ in range(2):
        o.iterate()
    _same_state(O, o, g)
    jl = (2, 5)
    sx, sz = _oracle_spectra(O, o, jl)
    g.clcSpectra1D(jl)
    gx, gz, n = g.spectra_get()
    assert n == 1
    ex, ez = _colerr(gx, sx), _colerr(gz, sz)
    assert ex.max() < 1e-12 and ez.max() < 1e-12, (ex, ez)
    # the energy columns are positive; the uu spectrum sums to <(ux+uCRF)^2>
    assert (gx[:, :, :4] >= 0).all() and (gz[:, :, :4] >= 0).all()
    g.close()
    o.close()


def test_spectra_accumulate_and_file(O, tmp_path):
    case = cases.SMALL4
    o, g = make_pair(O, case, seed=3)
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    jl = (1, 3)
    ax = np.zeros((nx // 2 + 1, ny, NCOL), order="F")
    az = np.zeros((nz // 2 + 1, ny, NCOL), order="F")
    for _ in range(2):
        o.iterate()
        _same_state(O, o, g)
        sx, sz = _oracle_spectra(O, o, jl)
        ax += sx
        az += sz
        g.clcSpectra1D(jl)
    gx, gz, n = g.spectra_get()
    assert n == 2
    assert _colerr(gx, ax).max() < 1e-12 and _colerr(gz, az).max() < 1e-12
    px, pz = tmp_path / "SpecX0000000002", tmp_path / "SpecZ0000000002"
    g.write_spectra(str(px), str(pz))
    fx = np.fromfile(px, dtype="<f8")
    fz = np.fromfile(pz, dtype="<f8")
    assert fx.size == NCOL * (nx // 2 + 1) * ny and fz.size == NCOL * (nz // 2 + 1) * ny
    # column-major (n/2+1, nyc) blocks, one per column, times 1/nSpectime
    assert np.array_equal(fx, (gx * 0.5).ravel(order="F"))
    assert np.array_equal(fz, (gz * 0.5).ravel(order="F"))
    gx, gz, n = g.spectra_get()
    assert n == 0 and not gx.any() and not gz.any()
    g.close()
    o.close()


def test_spectra_rejects(O):
    from channel3d_amd import Channel, lib as L
    g2 = Channel(cases.SMALL)
    with pytest.raises(L.C3DError, match="4th-order"):
        g2.clcSpectra1D((1, 2))
    g2.close()
    g4 = Channel(cases.SMALL4)
    with pytest.raises(L.C3DError, match="jForLCS"):
        g4.clcSpectra1D((0, 2))
    with pytest.raises(L.C3DError, match="no spectra samples"):
        g4.write_spectra(os.devnull, os.devnull)
    g4.close()


# ------------------------------------------------------------ 2-D spectra --
def _oracle_spectra2d(O, o, jl):
    nx, ny, nz = o.nxc, o.nyc, o.nzc
    nys = ny // 2 if o.params.flow_type == 1 else ny
    s = np.zeros((nx // 2 + 1, nys, nz // 2 + 1, 9), order="F")
    o.call("clcSpectra2D", o.field(O.UX), o.field(O.UY), o.field(O.UZ), o.field(O.PR), jl, s)
    return s


@pytest.mark.parametrize("name,ucrf", [("small4", 0.3), ("half4", 0.0)])
def test_spectra2d_match_oracle(O, name, ucrf, tmp_path):
    case = dict({"small4": cases.SMALL4, "half4": cases.HALF4}[name], u_crf=ucrf)
    o, g = make_pair(O, case, seed=8)
    o.iterate()
    _same_state(O, o, g)
    jl = 3
    ref = _oracle_spectra2d(O, o, jl)
    g.clcSpectra2D(jl)
    got, n = g.spectra2d_get()
    assert n == 1 and got.shape == ref.shape
    for c in range(9):
        s = np.abs(ref[..., c]).max()
        err = np.abs(got[..., c] - ref[..., c]).max() / s
        assert err < 1e-12, (c, err)
    # the Spec2D file: weights 4 / 2 on the kx or kz edges / 1 at the corners
    p = tmp_path / "Spec2D0000000001"
    g.write_spectra2d(str(p))
    f = np.fromfile(p, dtype="<f8").reshape(got.shape, order="F")
    nxh, nzh = got.shape[0] - 1, got.shape[2] - 1
    w = np.full((nxh + 1, nzh + 1), 4.0)
    w[[0, nxh], :] = 2.0
    w[:, [0, nzh]] = 2.0
    w[0, 0] = w[0, nzh] = w[nxh, 0] = w[nxh, nzh] = 1.0
    assert np.allclose(f, got * w[:, None, :, None], rtol=1e-15, atol=0)
    got2, n2 = g.spectra2d_get()
    assert n2 == 0 and not got2.any()
    g.close()
    o.close()
