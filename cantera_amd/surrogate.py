"""Learned one-step surrogate for constant-pressure reactors (no counterpart in
the reference), trained on pairs produced by the FP64 GPU integrator and used
behind an FP64 guard with the BDF integrator as fallback.

    sur = Surrogate(sol, dt=1e-6)
    y0, P, y1 = generate_pairs(sol, T0, P, Y0, dt=1e-6, nsteps=200)
    sur.fit(y0, P, y1, epochs=20)
    counts = batch.advance_surrogate(t_out, sur)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib


class DeviceArray:
    """A float64 (or other) array copied into device memory, freed with the object."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        self.shape, self.dtype, self.nbytes = host.shape, host.dtype, host.nbytes
        self.ptr = lib().cg_dev_alloc(max(1, host.nbytes))
        if not self.ptr:
            check(-1, "cg_dev_alloc")
        check(lib().cg_memcpy_h2d(self.ptr, host.ctypes.data_as(C.c_void_p), host.nbytes), "cg_memcpy_h2d")

    @classmethod
    def empty(cls, shape, dtype=np.float64):
        return cls(np.zeros(shape, dtype=dtype))

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().cg_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes), "cg_memcpy_d2h")
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().cg_dev_free(self.ptr)
            self.ptr = None


class Surrogate:
    def __init__(self, sol, dt=1e-6, hidden=256, depth=4, seed=0):
        self.sol = sol
        self.dt = float(dt)
        self.K = sol.n_species
        self._h = check(lib().cg_surrogate_create(sol.handle, self.dt, int(hidden), int(depth), int(seed)),
                        "cg_surrogate_create")
        self.n_params = int(lib().cg_surrogate_nparams(self._h))

    @property
    def handle(self) -> int:
        return self._h

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            lib().cg_surrogate_free(self._h)
            self._h = None

    @staticmethod
    def _dev(y0, P, y1=None):
        out = [DeviceArray(np.asarray(y0, np.float64)), DeviceArray(np.asarray(P, np.float64))]
        if y1 is not None:
            out.append(DeviceArray(np.asarray(y1, np.float64)))
        return out

    def fit_normalization(self, y0, P, y1):
        d0, dP, d1 = self._dev(y0, P, y1)
        check(lib().cg_surrogate_fit_normalization(self._h, len(y0), d0.ptr, dP.ptr, d1.ptr),
              "cg_surrogate_fit_normalization")

    def fit(self, y0, P, y1, epochs=10, batch=4096, lr=1e-3, normalize=True, log=None):
        """Train on host pairs; returns the mean loss of each epoch."""
        d0, dP, d1 = self._dev(y0, P, y1)
        if normalize:
            check(lib().cg_surrogate_fit_normalization(self._h, len(y0), d0.ptr, dP.ptr, d1.ptr),
                  "cg_surrogate_fit_normalization")
        losses = []
        for e in range(epochs):
            loss = C.c_double()
            check(lib().cg_surrogate_train_epoch(self._h, len(y0), d0.ptr, dP.ptr, d1.ptr, int(batch), float(lr),
                                                 C.byref(loss)), "cg_surrogate_train_epoch")
            losses.append(loss.value)
            if log:
                log(e, loss.value)
        return losses

    def loss_grad(self, y0, P, y1):
        d0, dP, d1 = self._dev(y0, P, y1)
        loss = C.c_double()
        g = np.zeros(self.n_params, dtype=np.float32)
        check(lib().cg_surrogate_loss_grad(self._h, len(y0), d0.ptr, dP.ptr, d1.ptr, C.byref(loss),
                                           g.ctypes.data_as(C.POINTER(C.c_float))), "cg_surrogate_loss_grad")
        return loss.value, g

    def predict(self, y0, P):
        d0, dP = self._dev(y0, P)
        d1 = DeviceArray.empty(np.shape(y0))
        check(lib().cg_surrogate_predict(self._h, len(y0), d0.ptr, dP.ptr, d1.ptr), "cg_surrogate_predict")
        return d1.get()

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        check(lib().cg_surrogate_get_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))), "get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        check(lib().cg_surrogate_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))), "set_params")

    def _norm_len(self):
        pad = lambda x: (x + 63) // 64 * 64
        return 2 * pad(self.K + 2) + 3 * pad(self.K + 1)

    @property
    def normalization(self) -> np.ndarray:
        v = np.zeros(self._norm_len())
        check(lib().cg_surrogate_get_normalization(self._h, v.ctypes.data_as(C.POINTER(C.c_double))), "get_norm")
        return v

    @normalization.setter
    def normalization(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.size == self._norm_len()
        check(lib().cg_surrogate_set_normalization(self._h, v.ctypes.data_as(C.POINTER(C.c_double))), "set_norm")


def regime_gate(sol, y, P, gate_T=1500.0):
    """The guard's regime gate for states y (rows [m, T, Y]): T > gate_T and
    the H + O + OH pool being destroyed (sum of their net production rates
    < 0, the sign of sum_k f_k / W_k that the device guard tests)."""
    y = np.asarray(y, np.float64)
    rad = [sol.species_index(nm) for nm in ("H", "O", "OH")]
    w = sol.net_production_rates(y[:, 1], P, np.ascontiguousarray(y[:, 2:].T))
    return (y[:, 1] > gate_T) & (w[rad].sum(axis=0) < 0.0)


def generate_pairs(sol, T0, P, Y0, dt=1e-6, nsteps=100, rtol=1e-8, atol=1e-15, stride=1, gate_T=None):
    """State pairs (y(t), y(t + dt)) along FP64 GPU trajectories of
    constant-pressure reactors started at (T0, P, Y0); returns y0, P, y1 with
    rows [m, T, Y].  Every `stride`-th step is kept; with gate_T, only pairs
    whose first state passes the guard's regime gate (regime_gate)."""
    from . import ReactorBatch
    rb = ReactorBatch(sol, T0, P, Y0, rtol=rtol, atol=atol)
    n = len(np.atleast_1d(T0))
    Pn = np.ascontiguousarray(np.broadcast_to(np.asarray(P, np.float64), (n,)))
    prev = rb.state.copy()
    a, b, p = [], [], []
    for k in range(1, nsteps + 1):
        rb.advance(k * dt)
        cur = rb.state.copy()
        if k % stride == 0:
            keep = np.ones(n, bool) if gate_T is None else regime_gate(sol, prev, Pn, gate_T)
            a.append(prev[keep])
            b.append(cur[keep])
            p.append(Pn[keep])
        prev = cur
    return np.concatenate(a), np.concatenate(p), np.concatenate(b)
