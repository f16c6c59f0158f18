"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py``; the product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run `make -C oracle`")
        L = C.CDLL(path)
        L.or_mech_load.restype = C.c_void_p
        L.or_mech_load.argtypes = [C.c_char_p]
        L.or_mech_free.argtypes = [C.c_void_p]
        for f in ("or_mech_nspecies", "or_mech_nreactions", "or_mech_nelements"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_int
        L.or_mech_mw.argtypes = [C.c_void_p, _dp]
        L.or_mech_composition.argtypes = [C.c_void_p, _dp]
        L.or_mech_species_index.argtypes = [C.c_void_p, C.c_char_p]
        L.or_mech_species_index.restype = C.c_int
        L.or_last_error.restype = C.c_char_p
        L.or_thermo.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp]
        opt = C.c_void_p
        L.or_rates.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp, opt, opt, opt, opt, opt]
        L.or_equilibrium_constants.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp, _dp]
        L.or_rates_batch.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, C.c_int]
        L.or_jacobian_batch.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, C.c_int]
        L.or_jacobian_batch.restype = C.c_int
        L.or_equilibrate.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _dp, C.POINTER(C.c_double), _dp,
                                     C.c_int, C.POINTER(C.c_int)]
        L.or_equilibrate.restype = C.c_int
        L.or_igcpr_rhs.argtypes = [C.c_void_p, C.c_double, _dp, _dp]
        L.or_igr_rhs.argtypes = [C.c_void_p, C.c_double, _dp, _dp]
        L.or_igcpr_rhs.restype = C.c_int
        L.or_igcpr_jacobian.argtypes = [C.c_void_p, C.c_double, _dp, _dp, C.c_int]
        L.or_reactor_new.restype = C.c_void_p
        L.or_reactor_new.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp, C.c_double, C.c_double, C.c_int]
        L.or_reactor_free.argtypes = [C.c_void_p]
        L.or_reactor_advance.argtypes = [C.c_void_p, C.c_double]
        L.or_reactor_advance.restype = C.c_int
        L.or_reactor_step.argtypes = [C.c_void_p]
        L.or_reactor_step.restype = C.c_int
        L.or_reactor_time.argtypes = [C.c_void_p]
        L.or_reactor_time.restype = C.c_double
        L.or_reactor_state.argtypes = [C.c_void_p, _dp]
        L.or_reactor_stats.argtypes = [C.c_void_p, _lp]
        L.or_reactor_add_sensitivity.argtypes = [C.c_void_p, C.c_int]
        L.or_reactor_add_sensitivity.restype = C.c_int
        L.or_reactor_set_sensitivity_tolerances.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.or_reactor_sensitivities.argtypes = [C.c_void_p, _dp]
        L.or_reactor_set_multipliers.argtypes = [C.c_void_p, _dp]
        L.or_mole_reactor_new.restype = C.c_void_p
        L.or_mole_reactor_new.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp, C.c_double, C.c_double, C.c_int]
        L.or_reactor_lin_stats.argtypes = [C.c_void_p, _lp]
        L.or_mole_rhs.argtypes = [C.c_void_p, C.c_double, _dp, _dp, C.c_int]
        L.or_mole_rhs.restype = C.c_int
        L.or_mole_jacobian.argtypes = [C.c_void_p, C.c_double, _dp, _dp, C.c_int]
        L.or_mole_jacobian.restype = C.c_int
        L.or_mole_reactors_run.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_double,
                                           C.c_int, opt, C.c_int]
        L.or_mole_reactors_run.restype = C.c_int
        L.or_reactors_run.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_double,
                                      C.c_double, C.c_int, opt, opt, C.c_int]
        L.or_reactors_run.restype = C.c_int
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Mech:
    def __init__(self, path: str):
        L = lib()
        self._h = L.or_mech_load(path.encode())
        if not self._h:
            raise RuntimeError(f"oracle: cannot load {path}: {L.or_last_error().decode()}")
        self.K = L.or_mech_nspecies(self._h)
        self.I = L.or_mech_nreactions(self._h)
        self.E = L.or_mech_nelements(self._h)
        self.mw = np.zeros(self.K)
        L.or_mech_mw(self._h, self.mw)
        comp = np.zeros(self.K * max(self.E, 1))
        L.or_mech_composition(self._h, comp)
        self.composition = comp.reshape(self.K, max(self.E, 1))[:, : self.E]

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.or_mech_free(self._h)
            self._h = None

    def index(self, name: str) -> int:
        k = lib().or_mech_species_index(self._h, name.encode())
        if k < 0:
            raise KeyError(name)
        return k

    def thermo(self, T):
        cp, h, s = (np.zeros(self.K) for _ in range(3))
        lib().or_thermo(self._h, float(T), cp, h, s)
        return cp, h, s

    def rates(self, T, P, Y):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        out = {k: np.zeros(n) for k, n in (("wdot", self.K), ("ropf", self.I), ("ropr", self.I),
                                          ("kf", self.I), ("rkcn", self.I))}
        lib().or_rates(self._h, float(T), float(P), Y, *(_ptr(out[k]) for k in ("wdot", "ropf", "ropr", "kf", "rkcn")))
        return out

    def equilibrium_constants(self, T, P, Y):
        Kc = np.zeros(self.I)
        lib().or_equilibrium_constants(self._h, float(T), float(P), np.ascontiguousarray(Y, dtype=np.float64), Kc)
        return Kc

    def rates_batch(self, T, P, Y_soa, nthreads=0):
        n = len(T)
        wdot = np.zeros((self.K, n))
        lib().or_rates_batch(self._h, n, np.ascontiguousarray(T, np.float64), np.ascontiguousarray(P, np.float64),
                             np.ascontiguousarray(Y_soa, np.float64), wdot, nthreads)
        return wdot

    def run_mole_reactors(self, T0, P, Y0, dt, nout, rtol, atol, skip_third_bodies=False, skip_falloff=False,
                          user_atol=False, nthreads=0, const_volume=False):
        """Mole reactors with the adaptive preconditioner and GMRES to
        k dt, k = 1..nout; returns the final states [T, (V,) n_1..n_K]."""
        n = len(T0)
        y = np.zeros((n, self.K + (2 if const_volume else 1)))
        flags = ((1 if skip_third_bodies else 0) | (2 if skip_falloff else 0) | (4 if user_atol else 0) |
                 (8 if const_volume else 0))
        r = lib().or_mole_reactors_run(self._h, n, np.ascontiguousarray(T0, np.float64),
                                       np.ascontiguousarray(P, np.float64), np.ascontiguousarray(Y0, np.float64),
                                       np.ascontiguousarray(dt, np.float64), nout, rtol, atol, flags, _ptr(y), nthreads)
        if r:
            raise RuntimeError(f"oracle integration failed for reactor {-r - 1}")
        return y

    def mole_rhs(self, P, y, const_volume=False):
        """[T, n] (or, constant volume, [T, V, n]; P unused) -> ydot"""
        f = np.zeros(len(y))
        lib().or_mole_rhs(self._h, float(P), np.ascontiguousarray(y, np.float64), f, 1 if const_volume else 0)
        return f

    def mole_jacobian(self, P, y, skip_third_bodies=False, skip_falloff=False, const_volume=False):
        N = len(y)
        J = np.zeros((N, N))
        lib().or_mole_jacobian(self._h, float(P), np.ascontiguousarray(y, np.float64), J,
                               (1 if skip_third_bodies else 0) | (2 if skip_falloff else 0) |
                               (8 if const_volume else 0))
        return J

    def jacobian_batch(self, T, P, Y_soa, nthreads=0):
        """Analytic reactor Jacobians (n, N, N) at n states (Y species-major)."""
        n = len(T)
        N = self.K + 2
        J = np.zeros((n, N, N))
        if lib().or_jacobian_batch(self._h, n, np.ascontiguousarray(T, np.float64), np.ascontiguousarray(P, np.float64),
                                   np.ascontiguousarray(Y_soa, np.float64), J.reshape(-1), nthreads):
            raise FloatingPointError("non-finite Jacobian")
        return J

    def equilibrate(self, T, P, Y, mode="HP", max_iter=500):
        """(T_eq, Y_eq, iterations) at fixed enthalpy and pressure ("HP") or
        fixed temperature and pressure ("TP")."""
        Yo = np.zeros(self.K)
        To = C.c_double()
        it = C.c_int()
        r = lib().or_equilibrate(self._h, 1 if mode == "HP" else 0, float(T), float(P),
                                 np.ascontiguousarray(Y, np.float64), C.byref(To), Yo, int(max_iter), C.byref(it))
        if r:
            raise RuntimeError(f"oracle equilibrium did not converge ({r})")
        return To.value, Yo, it.value

    def rhs(self, P, y, const_volume=False):
        """ydot of the constant-pressure reactor at P, or of the constant-volume
        reactor at volume P."""
        ydot = np.zeros(self.K + 2)
        f = lib().or_igr_rhs if const_volume else lib().or_igcpr_rhs
        r = f(self._h, float(P), np.ascontiguousarray(y, np.float64), ydot)
        if r:
            raise FloatingPointError("non-finite RHS")
        return ydot

    def jacobian(self, P, y, analytic=True, const_volume=False):
        N = self.K + 2
        J = np.zeros(N * N)
        lib().or_igcpr_jacobian(self._h, float(P), np.ascontiguousarray(y, np.float64), J,
                                (1 if analytic else 0) | (2 if const_volume else 0))
        return J.reshape(N, N)

    def run_reactors(self, T0, P, Y0, dt, nout, rtol=1e-9, atol=1e-15, analytic=False, nthreads=0, keep_states=True,
                     const_volume=False):
        n = len(T0)
        N = self.K + 2
        y_out = np.zeros((n, nout + 1, N)) if keep_states else None
        stats = np.zeros((n, 9), dtype=np.int64)
        r = lib().or_reactors_run(self._h, n, np.ascontiguousarray(T0, np.float64), np.ascontiguousarray(P, np.float64),
                                  np.ascontiguousarray(Y0, np.float64), np.ascontiguousarray(dt, np.float64), nout,
                                  rtol, atol, (1 if analytic else 0) | (2 if const_volume else 0), _ptr(y_out),
                                  _ptr(stats), nthreads)
        if r:
            raise RuntimeError(f"oracle integration failed for reactor {-r - 1}")
        return y_out, stats


class Reactor:
    """One IdealGasConstPressureReactor + ReactorNet (oracle)."""

    neq_extra = 2

    def __init__(self, mech: Mech, T0, P, Y0, rtol=1e-9, atol=1e-15, analytic=False, const_volume=False):
        self.mech = mech
        self._h = lib().or_reactor_new(mech._h, float(T0), float(P), np.ascontiguousarray(Y0, np.float64),
                                       rtol, atol, (1 if analytic else 0) | (2 if const_volume else 0))

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.or_reactor_free(self._h)
            self._h = None

    def advance(self, t):
        r = lib().or_reactor_advance(self._h, float(t))
        if r:
            raise RuntimeError(f"oracle advance failed ({r})")

    def step(self):
        r = lib().or_reactor_step(self._h)
        if r:
            raise RuntimeError(f"oracle step failed ({r})")
        return self.time

    @property
    def time(self):
        return lib().or_reactor_time(self._h)

    @property
    def state(self):
        y = np.zeros(self.mech.K + self.neq_extra)
        lib().or_reactor_state(self._h, y)
        return y

    def add_sensitivity_reaction(self, rxn):
        """Reactor::addSensitivityReaction (src/zeroD/Reactor.cpp:278-289)."""
        r = lib().or_reactor_add_sensitivity(self._h, int(rxn))
        if r:
            raise RuntimeError(f"oracle add_sensitivity failed ({r})")
        self._ns = getattr(self, "_ns", 0) + 1

    def set_sensitivity_tolerances(self, rtol, atol):
        lib().or_reactor_set_sensitivity_tolerances(self._h, float(rtol), float(atol))

    def set_multipliers(self, mult):
        lib().or_reactor_set_multipliers(self._h, np.ascontiguousarray(mult, np.float64))

    @property
    def sensitivities(self):
        """dy_i/dp_j (neq x ns): the raw CVODES sensitivities."""
        ns = getattr(self, "_ns", 0)
        S = np.zeros((self.mech.K + self.neq_extra, ns))
        if ns:
            lib().or_reactor_sensitivities(self._h, S)
        return S

    @property
    def stats(self):
        s = np.zeros(9, dtype=np.int64)
        lib().or_reactor_stats(self._h, s)
        return dict(zip(("steps", "rhs", "jac", "lu", "newton", "conv_fail", "err_fail", "order", "rhs_jac"), s.tolist()))


class MoleReactor(Reactor):
    """IdealGasConstPressureMoleReactor (or, const_volume, IdealGasMoleReactor)
    + ReactorNet with the adaptive preconditioner and GMRES (oracle); state
    [T, n_1..n_K] (constant volume [T, V, n_1..n_K]) in kmol."""

    def __init__(self, mech: Mech, T0, P, Y0, rtol=1e-9, atol=1e-15, skip_third_bodies=False, skip_falloff=False,
                 user_atol=False, const_volume=False):
        self.mech = mech
        self.neq_extra = 2 if const_volume else 1
        flags = ((1 if skip_third_bodies else 0) | (2 if skip_falloff else 0) | (4 if user_atol else 0) |
                 (8 if const_volume else 0))
        self._h = lib().or_mole_reactor_new(mech._h, float(T0), float(P), np.ascontiguousarray(Y0, np.float64),
                                            rtol, atol, flags)

    @property
    def lin_stats(self):
        s = np.zeros(6, dtype=np.int64)
        lib().or_reactor_lin_stats(self._h, s)
        return dict(zip(("lin_iters", "prec_evals", "prec_solves", "jt_vec_prod_evals", "lin_rhs_evals",
                         "lin_conv_fails"), s.tolist()))

    def mass_fractions(self, y=None):
        y = self.state if y is None else y
        m = y[self.neq_extra:] * self.mech.mw
        return m / m.sum()
