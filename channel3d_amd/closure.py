"""Python side of the learned subgrid-stress closure (c3d_closure_* in
include/channel3d.h): a data set of LES samples filtered from DNS states
of a Channel and a bf16-MFMA network trained on it.  No counterpart in the
reference, whose LES models are WALE and MTS (m_LESModels.f90:53-58)."""
import ctypes

import numpy as np

from . import lib as L

NPAD, NOUT = 16, 6
SHAPES = (("W1", (16, 64)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)),
          ("W3", (64, 64)), ("b3", (64,)), ("W4", (64, 6)), ("b4", (6,)))


def flatten(p):
    return np.concatenate([np.asarray(p[k], dtype=np.float32).ravel() for k, _ in SHAPES])


def unflatten(v):
    out, o = {}, 0
    for k, s in SHAPES:
        n = int(np.prod(s))
        out[k] = v[o:o + n].reshape(s).copy()
        o += n
    return out


def _fp(a):
    return a.ctypes.data_as(L._FP)


class Closure:
    def __init__(self, max_samples):
        self.lib = L.load()
        h = ctypes.c_void_p()
        L.check(self.lib.c3d_closure_create(int(max_samples), ctypes.byref(h)),
                "c3d_closure_create")
        self._h = h
        self.nparam = self.lib.c3d_closure_nparam()

    def close(self):
        if self._h:
            self.lib.c3d_closure_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        v = np.ascontiguousarray(flatten(p))
        assert v.size == self.nparam
        L.check(self.lib.c3d_closure_set_params(self._h, _fp(v)), "c3d_closure_set_params")

    def get_params(self):
        v = np.zeros(self.nparam, dtype=np.float32)
        L.check(self.lib.c3d_closure_get_params(self._h, _fp(v)), "c3d_closure_get_params")
        return unflatten(v)

    def size(self):
        n = ctypes.c_longlong()
        L.check(self.lib.c3d_closure_size(self._h, ctypes.byref(n)), "c3d_closure_size")
        return n.value

    def add_samples(self, channel, fx=4, fy=2, fz=4):
        """samples of the channel's current state; returns (first, count)"""
        first = ctypes.c_longlong()
        n0 = self.size()
        L.check(self.lib.c3d_closure_add_samples(self._h, channel._h, int(fx), int(fy), int(fz),
                                                 ctypes.byref(first)), "c3d_closure_add_samples")
        return first.value, self.size() - n0

    def samples(self, first, count):
        X = np.zeros((count, NPAD), dtype=np.float32)
        T = np.zeros((count, NOUT), dtype=np.float32)
        S = np.zeros(count, dtype=np.float32)
        L.check(self.lib.c3d_closure_get_samples(self._h, int(first), int(count), _fp(X), _fp(T),
                                                 _fp(S)), "c3d_closure_get_samples")
        return X, T, S

    def predict(self, first, count):
        Y = np.zeros((count, NOUT), dtype=np.float32)
        L.check(self.lib.c3d_closure_predict(self._h, int(first), int(count), _fp(Y)),
                "c3d_closure_predict")
        return Y

    def grad(self, first, count):
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        L.check(self.lib.c3d_closure_grad(self._h, int(first), int(count), _fp(g), _fp(loss)),
                "c3d_closure_grad")
        return unflatten(g), float(loss[0])

    def train(self, first, count, batch, steps, lr, seed=0):
        losses = np.zeros(steps, dtype=np.float32)
        L.check(self.lib.c3d_closure_train(self._h, int(first), int(count), int(batch),
                                           int(steps), float(lr), int(seed), _fp(losses)),
                "c3d_closure_train")
        return losses

    def adam_step(self, grads, lr):
        v = np.ascontiguousarray(flatten(grads))
        L.check(self.lib.c3d_closure_adam_step(self._h, _fp(v), float(lr)),
                "c3d_closure_adam_step")

    def set_mode(self, mode):
        """0: network viscosity with the dynamic Smagorinsky fallback;
        1: dynamic Smagorinsky in every cell"""
        L.check(self.lib.c3d_closure_set_mode(self._h, int(mode)), "c3d_closure_set_mode")

    def fallbacks(self):
        """cells of the last LES viscosity evaluation that fell back to
        dynamic Smagorinsky"""
        n = ctypes.c_longlong()
        L.check(self.lib.c3d_closure_fallbacks(self._h, ctypes.byref(n)), "c3d_closure_fallbacks")
        return n.value

    def dsm(self, channel, fx=4, fy=2, fz=4):
        """dynamic Smagorinsky on the LES grid of the channel's state:
        (model stress (6, nI, nJ, nK), filtered-DNS stress, Cs^2 per row)"""
        nI, nJ, nK = channel.nxc // fx, channel.nyc // fy, channel.nzc // fz
        n = nI * nJ * nK
        tm = np.zeros(6 * n)
        tt = np.zeros(6 * n)
        cs2 = np.zeros(nJ)
        L.check(self.lib.c3d_closure_dsm(self._h, channel._h, int(fx), int(fy), int(fz),
                                         tm.ctypes.data_as(L._DP), tt.ctypes.data_as(L._DP),
                                         cs2.ctypes.data_as(L._DP)), "c3d_closure_dsm")
        def fields(a):
            return np.stack([a[q * n:(q + 1) * n].reshape((nI, nJ, nK), order="F")
                             for q in range(6)])
        return fields(tm), fields(tt), cs2
