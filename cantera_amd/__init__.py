"""cantera_amd: MI355X-native batched gas kinetics and constant-pressure reactors.

Host-side mirror of the reference interface for this path:

* ``Solution(mech)`` -- an ideal-gas phase with bulk kinetics loaded from a
  compiled mechanism (reference ``newSolution``, ``src/base/Solution.cpp:242-282``).
* ``Solution.net_production_rates(T, P, Y)`` -- batched
  ``Kinetics::getNetProductionRates`` (``src/kinetics/Kinetics.cpp:447-457``).
* ``ReactorBatch`` -- n independent ``IdealGasConstPressureReactor`` +
  ``ReactorNet`` pairs with ``advance(t)`` (``src/zeroD/ReactorNet.cpp:351-357``).

All compute runs in libcantera_amd.so on the GPU; nothing here falls back to
the CPU.
"""
from __future__ import annotations

import os

import ctypes as C

import numpy as np

from . import _lib
from ._lib import CG_FP32, CG_FP64, CG_MIXED, CanteraGpuError, check, lib

MECH_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mechanisms")

__all__ = ["Solution", "ReactorBatch", "MoleReactorBatch", "CanteraGpuError", "CG_FP64", "CG_MIXED", "CG_FP32", "mech_path", "lib"]


def mech_path(name: str) -> str:
    """Resolve a mechanism name ('gri30', 'gri30.yaml', or a path) to a .mech file."""
    if os.path.exists(name) and name.endswith(".mech"):
        return name
    base = os.path.splitext(os.path.basename(name))[0]
    alias = {"nDodecane_Reitz": "nDodecane"}.get(base, base)
    p = os.path.join(MECH_DIR, alias + ".mech")
    if os.path.exists(p):
        return p
    if os.path.exists(name) and name.endswith((".yaml", ".yml")):
        from .mechc import compile_mechanism, write_mech
        out = os.path.join(MECH_DIR, base + ".mech")
        write_mech(compile_mechanism(name), out)
        return out
    if os.path.exists(name) and name.endswith((".inp", ".ck", ".dat")):
        # Chemkin format with the thermo data in the same file
        from .ckc import compile_chemkin
        from .mechc import write_mech
        out = os.path.join(MECH_DIR, base + ".mech")
        write_mech(compile_chemkin(name), out)
        return out
    raise FileNotFoundError(f"no compiled mechanism for {name!r}")


class Solution:
    """Ideal-gas phase + bulk kinetics on the GPU (one mechanism handle)."""

    def __init__(self, mechanism: str = "gri30"):
        self.path = mech_path(mechanism)
        L = lib()
        self._h = check(L.cg_mech_load(self.path.encode()), "cg_mech_load")
        self.n_species = check(L.cg_mech_nspecies(self._h))
        self.n_reactions = check(L.cg_mech_nreactions(self._h))
        self.n_elements = check(L.cg_mech_nelements(self._h))
        self.molecular_weights = np.zeros(self.n_species)
        check(L.cg_mech_molecular_weights(self._h, self.n_species, self.molecular_weights))
        buf = _lib.C.create_string_buffer(256)
        names = []
        for k in range(self.n_species):
            check(L.cg_mech_species_name(self._h, k, buf, 256))
            names.append(buf.value.decode())
        self.species_names = names

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.cg_mech_free(h)
            self._h = None

    @property
    def handle(self) -> int:
        return self._h

    def species_index(self, name: str) -> int:
        k = lib().cg_mech_species_index(self._h, name.encode())
        if k < 0:
            raise KeyError(name)
        return k

    def mass_fractions_from_moles(self, X) -> np.ndarray:
        X = np.asarray(X, dtype=np.float64)
        Y = X * self.molecular_weights
        return Y / Y.sum(axis=-1, keepdims=True)

    def net_production_rates(self, T, P, Y, precision=CG_FP64) -> np.ndarray:
        """wdot [kmol/m^3/s] for n states.  Y is species-major (K, n); returns (K, n).
        precision CG_FP32 evaluates the rate constants' transcendentals in FP32."""
        T = np.ascontiguousarray(np.atleast_1d(T), dtype=np.float64)
        n = T.shape[0]
        P = np.ascontiguousarray(np.broadcast_to(np.asarray(P, dtype=np.float64), (n,)))
        Y = np.ascontiguousarray(np.asarray(Y, dtype=np.float64).reshape(self.n_species, n))
        wdot = np.zeros((self.n_species, n))
        check(lib().cg_rates_batch(self._h, n, T, P, Y, wdot, int(precision)), "cg_rates_batch")
        return wdot

    def last_rates_kernel_ms(self) -> float:
        return float(lib().cg_rates_last_kernel_ms(self._h))

    def reactor_jacobians(self, T, P, Y, precision=CG_FP64) -> np.ndarray:
        """d(ydot)/dy of a constant-pressure reactor at n states, shape (n, K+2, K+2),
        state order [m, T, Y_0..Y_{K-1}] (as ``Reactor.jacobian`` orders it).
        Y is species-major (K, n).  precision=CG_FP32 returns float32 (the FP64
        assembly rounded at the store)."""
        T = np.ascontiguousarray(np.atleast_1d(T), dtype=np.float64)
        n = T.shape[0]
        P = np.ascontiguousarray(np.broadcast_to(np.asarray(P, dtype=np.float64), (n,)))
        Y = np.ascontiguousarray(np.asarray(Y, dtype=np.float64).reshape(self.n_species, n))
        N = self.n_species + 2
        J = np.zeros((n, N, N), dtype=np.float32 if precision == CG_FP32 else np.float64)
        check(lib().cg_jac_batch(self._h, n, T, P, Y, J.ctypes.data_as(C.c_void_p), int(precision)), "cg_jac_batch")
        return J


    def equilibrate(self, T, P, Y, mode="HP"):
        """Equilibrium states of n mixtures (ThermoPhase::equilibrate): mode
        "HP" (fixed specific enthalpy and pressure) or "TP".  Y species-major
        (K, n).  Returns (T_eq (n,), Y_eq (K, n), iterations (n,))."""
        T = np.ascontiguousarray(np.atleast_1d(T), dtype=np.float64)
        n = T.shape[0]
        P = np.ascontiguousarray(np.broadcast_to(np.asarray(P, dtype=np.float64), (n,)))
        Y = np.ascontiguousarray(np.asarray(Y, dtype=np.float64).reshape(self.n_species, n))
        To = np.zeros(n)
        Yo = np.zeros((self.n_species, n))
        info = np.zeros(n, dtype=np.int32)
        check(lib().cg_equilibrate_batch(self._h, n, mode.encode(), T, P, Y, To, Yo, info), "cg_equilibrate_batch")
        return To, Yo, info


class ReactorBatch:
    """n independent constant-pressure ideal-gas reactors, each with its own
    ReactorNet; ``advance`` mirrors ``ReactorNet::advance``.  ``const_volume=True``
    makes them closed rigid IdealGasReactors (V = 1 m^3) started at (T0, P, Y0)."""

    def __init__(self, sol: Solution, T0, P, Y0, rtol=1e-9, atol=1e-15, precision=CG_FP64, const_volume=False):
        self.sol = sol
        T0 = np.ascontiguousarray(np.atleast_1d(T0), dtype=np.float64)
        n = T0.shape[0]
        P = np.ascontiguousarray(np.broadcast_to(np.asarray(P, dtype=np.float64), (n,)))
        Y0 = np.ascontiguousarray(np.asarray(Y0, dtype=np.float64).reshape(n, sol.n_species))
        self.n = n
        self.neq = sol.n_species + 2
        create = lib().cg_reactors_create_const_volume if const_volume else lib().cg_reactors_create
        self._h = check(create(sol.handle, n, T0, P, Y0, float(rtol), float(atol), int(precision)),
                        "cg_reactors_create")
        self.time = np.zeros(n)
        self.sensitivity_reactions = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib._lib is not None:
            self._free_tbuf()
            _lib._lib.cg_reactors_free(h)
            self._h = None

    @property
    def handle(self) -> int:
        return self._h

    def add_sensitivity_reaction(self, rxn) -> int:
        """Reactor::addSensitivityReaction for every reactor of the batch (a
        multiplier on reaction ``rxn``'s rate, nominal 1); before the first
        advance.  Returns the parameter index."""
        p = check(lib().cg_reactors_add_sensitivity_reaction(self._h, int(rxn)), "add_sensitivity_reaction")
        self.sensitivity_reactions.append(int(rxn))
        return p

    def set_sensitivity_tolerances(self, rtol=1e-4, atol=1e-6):
        """ReactorNet.rtol_sensitivity / atol_sensitivity."""
        check(lib().cg_reactors_set_sensitivity_tolerances(self._h, float(rtol), float(atol)))

    def sensitivities(self, normalized=True) -> np.ndarray:
        """(n, K+2, n_params): ReactorNet.sensitivities() of each reactor,
        (dy_i/dp_j) / y_i; ``normalized=False`` gives dy_i/dp_j."""
        npar = len(self.sensitivity_reactions)
        S = np.zeros((self.n, self.neq, npar))
        if npar:
            check(lib().cg_reactors_get_sensitivities(self._h, 1 if normalized else 0, S))
        return S

    def set_options(self, max_steps=0, waves_per_cu=-1):
        check(lib().cg_reactors_set_options(self._h, int(max_steps), int(waves_per_cu)))

    # guard defaults: the end-to-end budget (1 K, 1e-4 in Y over 100 steps)
    # divided by the number of steps; gate at 1500 K
    GUARD = {"tol_T": 0.01, "tol_Y": 1e-6, "gate_T": 1500.0}

    def advance_surrogate(self, t, surrogate, tol_T=GUARD["tol_T"], tol_Y=GUARD["tol_Y"],
                          gate_T=GUARD["gate_T"]) -> np.ndarray:
        """advance() with one guarded surrogate step for every reactor whose
        output time is one surrogate step ahead; the rest (and every rejected
        step) run the integrator.  Returns [eligible, kept, integrated]."""
        t_out = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (self.n,)))
        counts = (C.c_int64 * 3)()
        check(lib().cg_reactors_advance_surrogate(self._h, surrogate.handle, t_out, float(tol_T), float(tol_Y),
                                                  float(gate_T), counts), "cg_reactors_advance_surrogate")
        self.time = t_out.copy()
        return np.array(counts[:], dtype=np.int64)

    def advance_surrogate_device(self, t_dev, surrogate, row=0, tol_T=GUARD["tol_T"], tol_Y=GUARD["tol_Y"],
                                 gate_T=GUARD["gate_T"]):
        """Asynchronous advance_surrogate to the output times at t_dev (row `row`
        of an n_steps x n device array); counts accumulate on the device."""
        ptr = (t_dev if isinstance(t_dev, int) else t_dev.ptr) + row * self.n * 8
        check(lib().cg_reactors_advance_surrogate_device(self._h, surrogate.handle, C.c_void_p(ptr), float(tol_T),
                                                         float(tol_Y), float(gate_T)),
              "cg_reactors_advance_surrogate_device")

    def surrogate_counts(self, reset=False) -> np.ndarray:
        """[eligible reactor-steps, kept, integrated] accumulated on the device."""
        counts = (C.c_int64 * 3)()
        check(lib().cg_reactors_surrogate_counts(self._h, counts, 1 if reset else 0), "cg_reactors_surrogate_counts")
        return np.array(counts[:], dtype=np.int64)

    def rates_to_device(self, w_dev, row=0):
        """Queue the net production rates at the current states (species-major
        K x n) into device memory at row `row` of an n_rows x K x n array."""
        ptr = (w_dev if isinstance(w_dev, int) else w_dev.ptr) + row * self.n * self.sol.n_species * 8
        check(lib().cg_reactors_net_production_rates_device(self._h, C.c_void_p(ptr)),
              "cg_reactors_net_production_rates_device")

    @property
    def net_production_rates(self) -> np.ndarray:
        """Net production rates (K, n) at the current reactor states."""
        w = np.zeros((self.sol.n_species, self.n))
        L = lib()
        p = L.cg_dev_alloc(w.nbytes)
        try:
            self.rates_to_device(p)
            check(L.cg_device_synchronize())
            check(L.cg_memcpy_d2h(w.ctypes.data_as(C.c_void_p), p, w.nbytes))
        finally:
            L.cg_dev_free(p)
        return w

    def state_to_device(self, y_dev, row=0):
        """Queue a copy of the states (n x (K+2), [m, T, Y]) into device memory
        at row `row` of an n_rows x n x (K+2) array."""
        ptr = (y_dev if isinstance(y_dev, int) else y_dev.ptr) + row * self.n * self.neq * 8
        check(lib().cg_reactors_get_state_device(self._h, C.c_void_p(ptr)), "cg_reactors_get_state_device")

    def set_matrix_storage(self, storage):
        """'lds' or 'global' (or 'default') iteration-matrix storage for 32 < N <= 64."""
        check(lib().cg_reactors_set_matrix_storage(self._h, {"default": 0, "lds": 1, "global": 2}[storage]))

    def set_schedule(self, longest_first=True):
        check(lib().cg_reactors_set_schedule(self._h, 1 if longest_first else 0))

    def advance(self, t):
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(t, dtype=np.float64), (self.n,)))
        check(lib().cg_reactors_advance(self._h, t), "cg_reactors_advance")
        self.time = t.copy()

    def advance_device(self, t_dev, row=0):
        """Asynchronous advance to the output times held in device memory at
        t_dev (row `row` of an n_steps x n array); check .status afterwards."""
        ptr = (t_dev if isinstance(t_dev, int) else t_dev.ptr) + row * self.n * 8
        check(lib().cg_reactors_advance_device(self._h, C.c_void_p(ptr)), "cg_reactors_advance_device")

    def advance_steps(self, t_steps):
        """Advance through the rows of t_steps (n_steps x n output times, or
        n_steps scalars) with one replayed HIP graph of n_steps integrator
        launches (cg_reactors_advance_graph).  The device copy of the times is
        kept, so later calls with the same number of steps replay the graph."""
        t = np.asarray(t_steps, dtype=np.float64)
        if t.ndim == 1:
            t = np.repeat(t[:, None], self.n, axis=1)
        t = np.ascontiguousarray(t.reshape(-1, self.n))
        L = lib()
        if getattr(self, "_tbuf", None) is None or self._tbuf[1] != t.nbytes:
            self._free_tbuf()
            p = L.cg_dev_alloc(t.nbytes)
            if not p:
                check(-1, "cg_dev_alloc")
            self._tbuf = (p, t.nbytes)
        check(L.cg_memcpy_h2d(self._tbuf[0], t.ctypes.data_as(C.c_void_p), t.nbytes), "cg_memcpy_h2d")
        check(L.cg_reactors_advance_graph(self._h, t.shape[0], self._tbuf[0]), "cg_reactors_advance_graph")
        st = np.zeros(self.n, dtype=np.int32)
        bad = check(L.cg_reactors_status(self._h, st), "cg_reactors_status")
        if bad:
            raise CanteraGpuError(f"{bad} reactor(s) failed to integrate; see ReactorBatch.status")
        self.time = t[-1].copy()

    def autotune(self, t_probe):
        """Time each integrator launch candidate on one advance to t_probe
        (scalar or per reactor) from the current state, restore the state and
        keep the fastest (cg_reactors_autotune).  Returns (choice, times_ms):
        choice = (matrix storage, waves per CU)."""
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(t_probe, dtype=np.float64), (self.n,)))
        L = lib()
        p = L.cg_dev_alloc(t.nbytes)
        if not p:
            check(-1, "cg_dev_alloc")
        try:
            check(L.cg_memcpy_h2d(p, t.ctypes.data_as(C.c_void_p), t.nbytes), "cg_memcpy_h2d")
            choice = np.zeros(2, dtype=np.int32)
            times = np.full(8, -1.0)
            nc = check(L.cg_reactors_autotune(self._h, p, choice.ctypes.data_as(C.c_void_p),
                                              times.ctypes.data_as(C.c_void_p), 8), "cg_reactors_autotune")
        finally:
            L.cg_dev_free(p)
        return (int(choice[0]), int(choice[1])), times[:nc]

    def _free_tbuf(self):
        if getattr(self, "_tbuf", None) is not None:
            lib().cg_dev_free(self._tbuf[0])
            self._tbuf = None

    @property
    def state(self) -> np.ndarray:
        y = np.zeros((self.n, self.neq))
        check(lib().cg_reactors_get_state(self._h, y), "cg_reactors_get_state")
        return y

    @property
    def T(self):
        return self.state[:, 1]

    @property
    def Y(self):
        return self.state[:, 2:]

    @property
    def status(self) -> np.ndarray:
        s = np.zeros(self.n, dtype=np.int32)
        check(lib().cg_reactors_status(self._h, s))
        return s

    @property
    def stats(self) -> dict:
        s = np.zeros((self.n, 9), dtype=np.int64)
        check(lib().cg_reactors_stats(self._h, s))
        keys = ("steps", "rhs", "jac", "lu", "newton", "conv_fail", "err_fail", "order", "rhs_jac")
        return {k: s[:, i] for i, k in enumerate(keys)}

    def last_kernel_ms(self) -> float:
        return float(lib().cg_reactors_last_kernel_ms(self._h))


class MoleReactorBatch(ReactorBatch):
    """n IdealGasConstPressureMoleReactors, each in a ReactorNet with the
    AdaptivePreconditioner and GMRES (``net.preconditioner =
    ct.AdaptivePreconditioner()``): state [T, n_1..n_K] in kmol, V = 1 m^3 at
    (T0, P, Y0).  ``skip_third_bodies`` / ``skip_falloff`` are the derivative
    settings of the preconditioner Jacobian; ``user_atol`` applies atol to every
    component, as ``net.atol = ...`` does (otherwise the species' atol is atol
    times the initial total moles).  ``const_volume=True``: IdealGasMoleReactors
    (closed, rigid), state [T, V, n_1..n_K]."""

    def __init__(self, sol: Solution, T0, P, Y0, rtol=1e-9, atol=1e-15, skip_third_bodies=False,
                 skip_falloff=False, user_atol=False, const_volume=False):
        self.sol = sol
        T0 = np.ascontiguousarray(np.atleast_1d(T0), dtype=np.float64)
        n = T0.shape[0]
        P = np.ascontiguousarray(np.broadcast_to(np.asarray(P, dtype=np.float64), (n,)))
        Y0 = np.ascontiguousarray(np.asarray(Y0, dtype=np.float64).reshape(n, sol.n_species))
        self.n = n
        self.neq = sol.n_species + (2 if const_volume else 1)
        self._s0 = 2 if const_volume else 1
        flags = ((1 if skip_third_bodies else 0) | (2 if skip_falloff else 0) | (4 if user_atol else 0) |
                 (8 if const_volume else 0))
        self._h = check(lib().cg_reactors_create_mole(sol.handle, n, T0, P, Y0, float(rtol), float(atol), flags),
                        "cg_reactors_create_mole")
        self.time = np.zeros(n)
        self.sensitivity_reactions = []

    @property
    def T(self):
        return self.state[:, 0]

    @property
    def moles(self):
        return self.state[:, self._s0:]

    @property
    def Y(self):
        m = self.moles * self.sol.molecular_weights
        return m / m.sum(axis=1, keepdims=True)

    @property
    def linear_stats(self) -> dict:
        s = np.zeros((self.n, 6), dtype=np.int64)
        check(lib().cg_reactors_linear_stats(self._h, s))
        keys = ("lin_iters", "prec_evals", "prec_solves", "jt_vec_prod_evals", "lin_rhs_evals", "lin_conv_fails")
        return {k: s[:, i] for i, k in enumerate(keys)}
