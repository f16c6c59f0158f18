"""Benchmark: GRI-3.0 constant-pressure reactor integration throughput.

Headline workload (default, ``--workload gri30-reactors``): 65,536 independent
gri30 IdealGasConstPressureReactors, T0 ~ U[900, 2000] K, CH4/air at
phi ~ U[0.5, 3.0], P ~ 1 atm x exp(U[ln 1, ln 40]), each advanced 100 times by
its own dt ~ exp(U[ln 1e-6, ln 1e-4]) s, rtol 1e-8, atol 1e-15, FP64.  A "step"
is one ReactorNet::advance of every reactor by its dt (65,536 reactor-steps).

Warm-up steps run on a separate batch with the same inputs, so the timed
region covers steps 1..K of a fresh batch from t = 0, with all inputs already
resident in HBM.  One JSON line is printed.

Other workloads of the same path, one JSON line each:
  gri30-rates        1,048,576 gri30 states, net production rates (states/s);
                     --precision fp32 adds the FP32 rate-constant option and its
                     error against the FP64 kernel
  gri30-jacobian     1,048,576 gri30 states, reactor Jacobians (states/s)
  ndodecane-reactors 262,144 n-dodecane reactors x 100 steps of 1e-6 s
  h2o2-reactors      1,024 h2o2 reactors x 100 outputs every 1e-5 s
  gri30-surrogate    BF16 MLP surrogate: training on integrator pairs,
                     inference on the matrix cores, guarded steps vs plain BDF
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0
METRIC = "GRI30 reactor-steps/s: plain FP64 BDF vs mixed+surrogate at ≤1 K T error"
PIPE = "FP64 (MI355X vector FMA and MFMA share the 78.6 TFLOP/s FP64 peak)"

# Algorithmic cost convention: every FP64 add/sub/mul/div counts 1, an FMA 2,
# and each transcendental (exp, log, log10, exp10, pow) counts TRANS_FLOPS.
TRANS_FLOPS = 20.0


def rhs_cost(mech_path):
    """(flops, transcendentals, K) of one reactor RHS evaluation, from the tables."""
    K = 0
    fl = 0.0
    tr = 0.0
    nnz = 0
    with open(mech_path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "species":
                K = int(t[1])
            elif t[0] == "R":
                typ, rev = int(t[2]), int(t[3])
                p = 15
                nr = int(t[p]); p += 1 + 3 * nr
                np_ = int(t[p]); p += 1 + 2 * np_
                ntb = int(t[p + 1])
                fl += 4 + 1                       # Arrhenius argument + A*exp
                tr += 1
                if rev:
                    fl += nr + np_ + 4            # delta mu, scaling, clip
                    tr += 1
                if typ != 0:
                    fl += 2 * ntb + 2             # third-body sum
                if typ >= 2:
                    fl += 10                      # low-pressure rate, Pr, blend
                    tr += 1
                    if typ == 3:
                        fl += 20                  # Fcent, f1, lgf
                        tr += 5                   # exps of Fcent, 2 log10, exp10
                fl += nr + (np_ if rev else 0) + 3  # concentration products, net
                nnz += nr + np_
    fl += 2 * nnz                                 # wdot gather
    fl += K * (18 + 6 + 4)                        # NASA7, mu, concentrations, energy sums
    tr += 4                                       # log T, log C0, log(P/Pref), 1/T
    return fl, tr, K


def flop_model(mech_path, stats):
    """Algorithmic FP64 flops of the integration, from the integrator counters.

    W = N_rhs F_rhs + N_jac F_jac + N_lu (2/3) N^3 + N_newton 2 N^2, with
    F_jac = 2 F_rhs + the analytic derivative scatter and chain rule (~F_rhs + 8 K^2).
    """
    fl, tr, K = rhs_cost(mech_path)
    N = K + 2
    F_rhs = fl + TRANS_FLOPS * tr
    F_jac = 3 * F_rhs + 8 * K * K
    W = (stats["rhs"].sum() * F_rhs + stats["jac"].sum() * F_jac
         + stats["lu"].sum() * (2.0 / 3.0) * N ** 3 + stats["newton"].sum() * 2.0 * N * N)
    return float(W), F_rhs


def measured_traffic(name, n):
    """HBM bytes per launch of the dominant kernel from a committed rocprofv3
    --pmc summary (profiles/<name>_*.md holding a `traffic_n: <n>` line).

    The summary holds FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the
    kernel and of a plain copy kernel of known size; each counter is scaled by
    (copy bytes / copy counter).  None when no summary was taken at this batch
    size.
    """
    pdir = os.path.join(ROOT, "profiles")
    try:
        files = sorted(f for f in os.listdir(pdir) if f.startswith(name + "_") and f.endswith(".md"))
    except OSError:
        return None
    for fn in files:
        try:
            kv = {}
            with open(os.path.join(pdir, fn)) as f:
                for line in f:
                    if line.startswith("traffic_") and ":" in line:
                        k, v = line.split(":", 1)
                        kv[k.strip()] = float(v.split()[0])
            if int(kv["traffic_n"]) != n:
                continue
            return kv["traffic_bytes_per_launch"]
        except (OSError, KeyError, ValueError):
            continue
    return None


def cpu_threads():
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return max(1, min(threads, 16))


def cpu_reactors(mech_path, wl, steps, sample, threads, const_volume=False):
    from oracle.oracle import Mech
    om = Mech(mech_path)
    S = min(sample, len(wl.T0))
    t0 = time.perf_counter()
    om.run_reactors(wl.T0[:S], wl.P[:S], wl.Y0[:S], wl.dt[:S], steps, wl.rtol, wl.atol,
                    analytic=False, nthreads=threads, keep_states=False, const_volume=const_volume)
    el = time.perf_counter() - t0
    return {"value": S * steps / el, "unit": "reactor-steps/s", "cores": threads, "kind": "port",
            "sample": f"{S} of the {len(wl.T0)} reactors x {steps} steps, CVODES-method BDF with "
                      f"difference-quotient Jacobian (oracle), OpenMP schedule(static,1), {el:.1f} s"}


def reactors(a):
    import cantera_amd as ca
    from cantera_amd._lib import check, lib
    from cantera_amd.workloads import config1, config3, config5

    L = lib()
    mech = {"gri30-reactors": "gri30", "ndodecane-reactors": "nDodecane", "h2o2-reactors": "h2o2"}[a.workload]
    sol = ca.Solution(mech)
    steps = a.steps or 100
    if mech == "gri30":
        n = a.n or 65536
        wl = config3(sol, n=n, nsteps=steps)
        metric = METRIC
        desc = (f"gri30 ({sol.n_species} species, {sol.n_reactions} reactions) IdealGasConstPressureReactor, {n} "
                f"reactors x {steps} advance "
                f"steps, per-reactor dt log-uniform in [1e-6, 1e-4] s, T0 U[900,2000] K, CH4/air phi U[0.5,3], "
                f"P 1-40 atm, rtol 1e-8 atol 1e-15")
    elif mech == "nDodecane":
        n = a.n or 262144
        wl = config5(sol, n=n, nsteps=steps)
        metric = "nDodecane reactor-steps/s (FP64 BDF)"
        desc = (f"nDodecane ({sol.n_species} species, {sol.n_reactions} reactions) IdealGasConstPressureReactor, "
                f"{n} reactors x {steps} "
                f"advance steps of 1e-6 s, T0 U[700,2200] K, c12h26/air phi U[0.3,3], P 20-60 atm, rtol 1e-8 "
                f"atol 1e-15")
    else:
        n = a.n or 1024
        wl = config1(sol, n=n)
        metric = "h2o2 reactor-steps/s (FP64 BDF)"
        desc = (f"h2o2 ({sol.n_species} species, {sol.n_reactions} reactions) IdealGasConstPressureReactor, {n} "
                f"reactors x {steps} outputs "
                f"every 1e-5 s, T0 U[1000,1800] K, H2/air phi U[0.5,2], 1 atm, rtol 1e-8 atol 1e-15")
    if a.precision == "fp32":
        raise SystemExit("--precision fp32 applies to gri30-rates only (reactors: fp64 or mixed)")
    prec = ca.CG_FP64 if a.precision == "fp64" else ca.CG_MIXED
    if a.const_volume:
        # closed rigid reactors started at the same (T0, P0, Y0): a different
        # workload from the constant-pressure headline, named as such
        desc = desc.replace("IdealGasConstPressureReactor", "IdealGasReactor (constant volume, V = 1 m^3)")
        metric = f"{mech} IdealGasReactor (constant volume) reactor-steps/s (FP64 BDF)"

    # all output times resident in HBM before the timed region
    touts = np.ascontiguousarray(np.outer(np.arange(1, steps + 1, dtype=np.float64), wl.dt))
    d_tout = L.cg_dev_alloc(touts.nbytes)
    check(L.cg_memcpy_h2d(d_tout, touts.ctypes.data_as(C.c_void_p), touts.nbytes))

    tuned = None
    if a.warmup > 0 or a.autotune:
        wb = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol, precision=prec,
                         const_volume=a.const_volume)
        wb.set_options(0, a.waves_per_cu)
        done = min(max(a.warmup, 1 if a.autotune else 0), steps)
        for k in range(done):
            check(L.cg_reactors_advance_device(wb.handle, C.c_void_p(d_tout + k * n * 8)))
        check(L.cg_device_synchronize())
        if a.autotune:
            # measured launch choice on the warm-up batch's next step (the
            # first step from t = 0 is a start-up transient, not typical)
            probe = touts[done] if done < steps else touts[-1] + wl.dt
            choice, times = wb.autotune(probe)
            tuned = {"matrix_storage": ["default", "lds", "global"][choice[0]], "waves_per_cu": choice[1],
                     "candidate_ms": [round(float(x), 3) for x in times]}
        del wb

    rb = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol, precision=prec,
                         const_volume=a.const_volume)
    rb.set_options(0, a.waves_per_cu)
    rb.set_schedule(a.schedule == "longest-first")
    rb.set_matrix_storage(a.matrix_storage)
    if tuned:
        rb.set_matrix_storage(tuned["matrix_storage"])
        rb.set_options(0, tuned["waves_per_cu"])
    check(L.cg_device_synchronize())
    kernel_ms = []
    t0 = time.perf_counter()
    if a.graph:
        # all K advances captured as one HIP graph (capture included in the timing)
        check(L.cg_reactors_advance_graph(rb.handle, steps, C.c_void_p(d_tout)))
        kernel_ms = [rb.last_kernel_ms() / steps]
        print(f"{steps} steps in one graph: {kernel_ms[0] * steps:.2f} ms", file=sys.stderr, flush=True)
    else:
        for k in range(steps):
            check(L.cg_reactors_advance_device(rb.handle, C.c_void_p(d_tout + k * n * 8)))
            kernel_ms.append(rb.last_kernel_ms())
            print(f"step {k + 1}/{steps}: kernel {kernel_ms[-1]:.2f} ms", file=sys.stderr, flush=True)
    check(L.cg_device_synchronize())
    elapsed = time.perf_counter() - t0
    bad = int((rb.status != 0).sum())
    if bad:
        print(f"{bad} reactors failed to integrate", file=sys.stderr)
        return 1
    stats = rb.stats
    W, F_rhs = flop_model(sol.path, stats)
    avg_launch_s = float(np.mean(kernel_ms)) / 1e3
    achieved = W / steps / avg_launch_s / 1e12
    out = {
        "metric": metric,
        "value": n * steps / elapsed,
        "unit": "reactor-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f64 state/residual, f32 iteration matrix",
        "data": "synthetic",
        "config": {
            "workload": desc,
            "precision": a.precision,
            "parallelism": "1 GPU, one wavefront per reactor, persistent work-queue kernel",
            "schedule": a.schedule,
            "matrix_storage": a.matrix_storage,
            "launch": "one HIP graph of all steps" if a.graph else "one advance call per step",
            "autotune": tuned,
        },
        "roofline": {
            "bound": "mfma",
            "pipe": PIPE,
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": measured_traffic(f"bdf_advance_{mech}" + ("_cv" if a.const_volume else ""), n),
            "kernel": "bdf_advance",
            "avg_launch_ms": avg_launch_s * 1e3,
            "flops_per_launch": W / steps,
        },
        "integrator": {k: float(np.mean(v)) for k, v in stats.items() if k != "order"},
    }
    if a.workload in ("gri30-reactors", "ndodecane-reactors") and a.precision == "fp64" and not a.no_mixed:
        # the mixed form (FP32 iteration matrix / LU / solves, FP64 residual)
        # on a fresh batch, same inputs and steps, after the timed FP64 region;
        # its error is stated against the FP64 result
        y64 = rb.state
        del rb
        # warm the mixed kernels on a few reactors, then time all K steps
        wm = ca.ReactorBatch(sol, wl.T0[:64], wl.P[:64], wl.Y0[:64], rtol=wl.rtol, atol=wl.atol, precision=ca.CG_MIXED,
                             const_volume=a.const_volume)
        wm.advance(wl.dt[:64])
        del wm
        rm = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol, precision=ca.CG_MIXED,
                             const_volume=a.const_volume)
        check(L.cg_device_synchronize())
        t1 = time.perf_counter()
        for k in range(steps):
            rm.advance_device(d_tout, k)
        check(L.cg_device_synchronize())
        el_m = time.perf_counter() - t1
        if int((rm.status != 0).sum()) == 0:
            ym = rm.state
            out["mixed_precision"] = {
                "value": n * steps / el_m, "unit": "reactor-steps/s", "steps": steps,
                "speedup_vs_fp64": (n * steps / el_m) / (n * steps / elapsed),
                "max_abs_dT_vs_fp64_K": float(np.max(np.abs(ym[:, 1] - y64[:, 1]))),
                "max_abs_dY_vs_fp64": float(np.max(np.abs(ym[:, 2:] - y64[:, 2:]))),
            }
    if not a.no_cpu:
        sample = a.cpu_sample or {"gri30": 16384, "nDodecane": 512, "h2o2": 1024}[mech]
        out["cpu_baseline"] = cpu_reactors(sol.path, wl, steps, sample, cpu_threads(), a.const_volume)
    L.cg_dev_free(d_tout)
    copy_calib(a, L, check)
    print(json.dumps(out), flush=True)
    return 0


def batch_kernels(a):
    """Batched rates or Jacobians over gri30 states resident in HBM."""
    import cantera_amd as ca
    from cantera_amd._lib import check, lib
    from cantera_amd.workloads import config2_states

    L = lib()
    sol = ca.Solution("gri30")
    n = a.n or 1_048_576
    steps = a.steps or 20
    K = sol.n_species
    N = K + 2
    T, P, Y = config2_states(sol, n=n)
    jac = a.workload == "gri30-jacobian"
    out_doubles = n * N * N if jac else n * K
    d_in = L.cg_dev_alloc((2 + K) * n * 8)
    d_out = L.cg_dev_alloc(out_doubles * 8)
    for off, arr in ((0, T), (n, P), (2 * n, Y)):
        arr = np.ascontiguousarray(arr)
        check(L.cg_memcpy_h2d(d_in + off * 8, arr.ctypes.data_as(C.c_void_p), arr.nbytes))
    dT, dP, dY = d_in, d_in + n * 8, d_in + 2 * n * 8
    call = L.cg_jac_batch_device if jac else L.cg_rates_batch_device
    last = L.cg_jac_last_kernel_ms if jac else L.cg_rates_last_kernel_ms
    f32 = a.precision == "fp32"
    if f32 and jac:
        raise SystemExit("--precision fp32 applies to gri30-rates only")
    prec = ca.CG_FP32 if f32 else ca.CG_FP64
    for _ in range(a.warmup):
        check(call(sol.handle, n, dT, dP, dY, d_out, prec, None))
    check(L.cg_device_synchronize())
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        check(call(sol.handle, n, dT, dP, dY, d_out, prec, None))
        kernel_ms.append(last(sol.handle))
    check(L.cg_device_synchronize())
    elapsed = time.perf_counter() - t0
    fp32_error = None
    if f32:
        # error of the FP32 option against the FP64 kernel on the same states
        # (untimed): per state, max_k |wdot32 - wdot64| / max_k |wdot64|
        w32 = np.empty((K, n))
        check(L.cg_memcpy_d2h(w32.ctypes.data_as(C.c_void_p), d_out, w32.nbytes))
        check(call(sol.handle, n, dT, dP, dY, d_out, ca.CG_FP64, None))
        check(L.cg_device_synchronize())
        w64 = np.empty((K, n))
        check(L.cg_memcpy_d2h(w64.ctypes.data_as(C.c_void_p), d_out, w64.nbytes))
        e = np.abs(w32 - w64).max(axis=0) / (np.abs(w64).max(axis=0) + 1e-300)
        fp32_error = {"per_state_max_rel_to_largest_wdot": {"max": float(e.max()), "p99": float(np.percentile(e, 99)),
                                                             "median": float(np.median(e))}}
        del w32, w64
    fl, tr, _ = rhs_cost(sol.path)
    F_rates = fl + TRANS_FLOPS * tr
    avg_s = float(np.mean(kernel_ms)) / 1e3
    name = "jac_batch_gri30" if jac else "rates_gri30"
    if jac:
        # HBM-bound: T, P, Y in; the N x N Jacobian out
        alg_bytes = n * (8 * (2 + K) + 8 * N * N)
        roof = {"bound": "hbm", "achieved": alg_bytes / avg_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": alg_bytes / avg_s / 1e9 / HBM_PEAK_GBS, "traffic": measured_traffic(name, n),
                "kernel": "jac_batch_kernel", "avg_launch_ms": avg_s * 1e3, "bytes_per_launch": alg_bytes}
        metric = "gri30 reactor Jacobians/s (FP64, 55x55 dense per state)"
    else:
        flops = n * F_rates
        roof = {"bound": "mfma", "pipe": PIPE, "achieved": flops / avg_s / 1e12, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": flops / avg_s / 1e12 / FP64_PEAK_TFLOPS,
                "traffic": measured_traffic(name, n), "kernel": "rates_kernel", "avg_launch_ms": avg_s * 1e3,
                "flops_per_launch": flops, "hbm_bytes_algorithmic": n * 8 * (2 + 2 * K)}
        metric = "gri30 net production rates, states/s (" + ("FP32 rate constants" if f32 else "FP64") + ")"
    out = {
        "metric": metric, "value": n * steps / elapsed, "unit": "states/s", "n_gpus": 1, "steps": steps,
        "warmup": a.warmup, "ms_per_step": elapsed / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"gri30 states, {n} per launch: T U[300,2500] K, P 0.1-50 atm, CH4/O2/N2 majors at "
                               f"phi U[0.3,3], minors X 1e-12..1e-2; species-major SoA in HBM"},
        "roofline": roof,
    }
    if f32:
        out["dtype"] = "f32 rate-constant transcendentals, f64 otherwise"
        out["fp32_error_vs_fp64"] = fp32_error
    if not a.no_cpu:
        from oracle.oracle import Mech
        om = Mech(sol.path)
        th = cpu_threads()
        S = min(n, a.cpu_sample or (4096 if jac else 262144))
        Ys = np.ascontiguousarray(Y[:, :S])
        t1 = time.perf_counter()
        if jac:
            for s in range(S):
                om.jacobian(P[s], np.concatenate(([1.0, T[s]], Ys[:, s])), analytic=True)
            th = 1
        else:
            om.rates_batch(T[:S], P[:S], Ys, nthreads=th)
        el = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": S / el, "unit": "states/s", "cores": th, "kind": "port",
                               "sample": f"first {S} of the {n} states, oracle "
                                         f"{'analytic Jacobian, one thread' if jac else 'rates, OpenMP'}, {el:.1f} s"}
    L.cg_dev_free(d_in)
    L.cg_dev_free(d_out)
    copy_calib(a, L, check)
    print(json.dumps(out), flush=True)
    return 0


def surrogate(a):
    """Learned surrogate on gri30: train on integrator pairs, time inference on
    the matrix cores, and run guarded surrogate steps against the plain BDF
    integrator on the same reactors."""
    import cantera_amd as ca
    from cantera_amd._lib import check, lib
    from cantera_amd.surrogate import DeviceArray, Surrogate, generate_pairs
    from cantera_amd.workloads import config3

    L = lib()
    sol = ca.Solution("gri30")
    K = sol.n_species
    n = a.n or 65536
    steps = a.steps or 10
    dt = 1e-6
    hidden, depth = 256, 4
    # training pairs along FP64 integrator trajectories from the evaluation
    # distribution (another seed), dt = 1e-6 s
    ntr = max(1024, a.pairs // 16)
    tw = config3(sol, n=ntr, nsteps=16, seed=0x5EED0104)
    t0 = time.perf_counter()
    y0, P, y1 = generate_pairs(sol, tw.T0, tw.P, tw.Y0, dt=dt, nsteps=16, rtol=tw.rtol, atol=tw.atol)
    gen_s = time.perf_counter() - t0
    sur = Surrogate(sol, dt=dt, hidden=hidden, depth=depth, seed=1)
    t0 = time.perf_counter()
    losses = sur.fit(y0, P, y1, epochs=a.epochs, batch=4096, lr=1e-3)
    check(L.cg_device_synchronize())
    train_s = time.perf_counter() - t0
    n_in, n_out = K + 2, K + 1
    flop_fwd = 2.0 * (n_in * hidden + (depth - 1) * hidden * hidden + hidden * n_out)
    train_samples = len(y0) * a.epochs
    # inference on the evaluation states (device arrays, GEMMs on MFMA)
    wl = config3(sol, n=n, nsteps=steps)
    st0 = np.concatenate([np.ones((n, 1)), wl.T0[:, None], wl.Y0], axis=1)
    d0, dP, d1 = DeviceArray(st0), DeviceArray(np.asarray(wl.P, np.float64)), DeviceArray.empty(st0.shape)
    for _ in range(2):
        check(L.cg_surrogate_predict(sur.handle, n, d0.ptr, dP.ptr, d1.ptr))
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        check(L.cg_surrogate_predict(sur.handle, n, d0.ptr, dP.ptr, d1.ptr))
    inf_s = (time.perf_counter() - t0) / reps
    # guarded advance vs the plain integrator, same reactors and output times
    plain = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
    guarded = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
    t0 = time.perf_counter()
    for k in range(1, steps + 1):
        plain.advance(k * dt)
    plain_s = time.perf_counter() - t0
    counts = np.zeros(3, dtype=np.int64)
    t0 = time.perf_counter()
    for k in range(1, steps + 1):
        counts += guarded.advance_surrogate(k * dt, sur, dT_tol=a.dT_tol, rtol_g=a.rtol_g, atol_g=a.atol_g)
    guarded_s = time.perf_counter() - t0
    yp, yg = plain.state, guarded.state
    dT = np.abs(yp[:, 1] - yg[:, 1])
    dY = np.abs(yp[:, 2:] - yg[:, 2:]).max()
    achieved = n * flop_fwd / inf_s / 1e12
    out = {
        "metric": "gri30 reactor-steps/s at dt = 1e-6 s: guarded surrogate with BDF fallback",
        "value": n * steps / guarded_s, "unit": "reactor-steps/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": guarded_s / steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 network (fp32 accumulate), f64 guard and fallback", "data": "synthetic",
        "config": {"workload": f"gri30 MLP {n_in} -> {depth} x {hidden} GELU -> {n_out}, dt 1e-6 s; training on "
                               f"{len(y0)} integrator pairs ({ntr} reactors x 16 steps, T0 U[900,2000] K, CH4/air "
                               f"phi U[0.5,3], 1-40 atm), {a.epochs} epochs, batch 4096, Adam lr 1e-3; evaluation "
                               f"{n} reactors x {steps} steps of the same distribution (another seed)",
                   "guard": {"dT_tol_K": a.dT_tol, "rtol_g": a.rtol_g, "atol_g": a.atol_g}},
        "roofline": {"bound": "mfma", "pipe": "BF16 MFMA (dense 2.5 PFLOP/s)", "achieved": achieved, "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": achieved / 2500.0, "traffic": None,
                     "kernel": "gemm_bf16 (surrogate inference: features, 5 GEMMs, decode)",
                     "avg_launch_ms": inf_s * 1e3, "flops_per_launch": n * flop_fwd},
        "surrogate": {
            "plain_fp64_reactor_steps_per_s": n * steps / plain_s,
            "speedup_vs_plain": plain_s / guarded_s,
            "kept_surrogate_steps": int(counts[0]), "reactor_steps_by_surrogate_alone": int(counts[1]),
            "reactor_steps_by_bdf": int(counts[2]), "fallback_fraction": float(counts[2] / max(1, n * steps)),
            "max_abs_dT_vs_plain_K": float(dT.max()), "max_abs_dY_vs_plain": float(dY),
            "training_pairs": len(y0), "pair_generation_s": gen_s, "training_s": train_s,
            "training_samples_per_s": train_samples / train_s,
            "training_tflops": train_samples * 3 * flop_fwd / train_s / 1e12,
            "loss_first_last_epoch": [losses[0], losses[-1]],
            "inference_samples_per_s": n / inf_s,
        },
    }
    print(json.dumps(out), flush=True)
    return 0


def copy_calib(a, L, check):
    if a.copy_calib > 0:
        nb = a.copy_calib << 20
        src, dst = L.cg_dev_alloc(nb), L.cg_dev_alloc(nb)
        check(L.cg_memcpy_d2d(dst, src, nb))
        check(L.cg_device_synchronize())
        L.cg_dev_free(src)
        L.cg_dev_free(dst)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="gri30-reactors",
                    choices=["gri30-reactors", "gri30-rates", "gri30-jacobian", "ndodecane-reactors", "h2o2-reactors",
                             "gri30-surrogate"],
                    help="gri30-reactors is the headline line; the others measure the other workloads of the path")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--precision", choices=["fp64", "mixed", "fp32"], default="fp64",
                    help="reactors: fp64 or mixed (FP32 LU); gri30-rates: fp64 or fp32 (FP32 rate constants)")
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--schedule", choices=["longest-first", "index"], default="longest-first")
    ap.add_argument("--matrix-storage", choices=["default", "lds", "global"], default="default")
    ap.add_argument("--graph", action="store_true",
                    help="reactor workloads: run the K steps as one captured HIP graph")
    ap.add_argument("--autotune", action="store_true",
                    help="reactor workloads: pick the integrator launch by timing the candidates first")
    ap.add_argument("--pairs", type=int, default=1 << 20, help="gri30-surrogate: training pairs")
    ap.add_argument("--epochs", type=int, default=4, help="gri30-surrogate: training epochs")
    ap.add_argument("--dT-tol", type=float, default=0.05, help="gri30-surrogate: guard, K")
    ap.add_argument("--rtol-g", type=float, default=1e-3, help="gri30-surrogate: guard, relative")
    ap.add_argument("--atol-g", type=float, default=1e-9, help="gri30-surrogate: guard, absolute Y")
    ap.add_argument("--const-volume", action="store_true",
                    help="reactor workloads: closed rigid IdealGasReactor instead of IdealGasConstPressureReactor")
    ap.add_argument("--no-mixed", action="store_true",
                    help="gri30-/ndodecane-reactors: skip the mixed-precision run after the FP64 line")
    ap.add_argument("--cpu-sample", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="at the end, run one device-to-device copy (blit kernel) of MIB MiB "
                         "(known HBM bytes, to scale FETCH_SIZE/WRITE_SIZE counters)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the reactors are independent, "
              "so N GPUs would each run their own shard).", file=sys.stderr)
        return 2
    if a.workload in ("gri30-rates", "gri30-jacobian"):
        return batch_kernels(a)
    if a.workload == "gri30-surrogate":
        return surrogate(a)
    return reactors(a)


if __name__ == "__main__":
    sys.exit(main())
