"""cg_surrogate_predict_noalloc against cg_surrogate_predict, through the C ABI.

The allocation-free entry point runs the same kernels on the same rows, only
in passes through its fixed device workspace, so its predictions must equal
those of cg_surrogate_predict bit for bit.  Two calls in a row, the second on
fewer rows (the workspace keeps the first call's contents), both spanning
several passes; an argument error is reported without touching the device.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import mech_file

pytestmark = pytest.mark.gpu


def _pairs(info, n, seed):
    g = np.random.default_rng(seed)
    K = info.n_species
    X = 10.0 ** g.uniform(-8, 0, (n, K))
    X /= X.sum(axis=1, keepdims=True)
    Y = info.mass_fractions_from_moles(X)
    y0 = np.concatenate([np.ones((n, 1)), g.uniform(900, 2200, (n, 1)), Y], axis=1)
    P = 101325.0 * g.uniform(1, 20, n)
    y1 = y0.copy()
    y1[:, 1] += g.normal(0, 5.0, n)
    y1[:, 2:] += 1e-4 * g.normal(0, 1, (n, K)) * Y
    return y0, P, y1


def test_predict_noalloc_matches_predict():
    import cantera_amd as ca
    from cantera_amd._lib import check, last_error, lib
    from cantera_amd.surrogate import DeviceArray, Surrogate
    from cantera_amd.workloads import MechInfo
    L = lib()
    info = MechInfo(mech_file("gri30"))
    sol = ca.Solution("gri30")
    sur = Surrogate(sol, dt=1e-5, hidden=256, depth=4, seed=3)
    y0, P, y1 = _pairs(info, 4096, 1)
    sur.fit(y0, P, y1, epochs=2, batch=1024, lr=1e-3)
    for n, seed in ((30000, 2), (5000, 3)):
        a0, aP, _ = _pairs(info, n, seed)
        d0, dP = DeviceArray(a0), DeviceArray(aP)
        ref = DeviceArray.empty(a0.shape)
        out = DeviceArray.empty(a0.shape)
        print(f"noalloc: {n} rows", flush=True)
        check(L.cg_surrogate_predict(sur.handle, n, d0.ptr, dP.ptr, ref.ptr), "cg_surrogate_predict")
        check(L.cg_surrogate_predict_noalloc(sur.handle, n, d0.ptr, dP.ptr, out.ptr, None),
              "cg_surrogate_predict_noalloc")
        check(L.cg_device_synchronize())
        r, o = ref.get(), out.get()
        assert np.all(np.isfinite(r))
        assert np.array_equal(r, o), f"{n} rows: max difference {np.abs(r - o).max():.3e}"
    assert L.cg_surrogate_predict_noalloc(sur.handle, -1, d0.ptr, dP.ptr, out.ptr, None) == -999
    assert "negative row count" in last_error()
    assert L.cg_surrogate_predict_noalloc(sur.handle, 10, None, dP.ptr, out.ptr, None) == -999
    assert C.c_int32(0).value == 0
