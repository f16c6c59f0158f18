"""Benchmark: GRI-3.0 constant-pressure reactor integration throughput.

Headline workload (default): 65,536 independent gri30
IdealGasConstPressureReactors, T0 ~ U[900, 2000] K, CH4/air at
phi ~ U[0.5, 3.0], P ~ 1 atm x exp(U[ln 1, ln 40]), each advanced K times by
its own dt ~ exp(U[ln 1e-6, ln 1e-4]) s, rtol 1e-8, atol 1e-15, FP64.  A
"step" is one ReactorNet::advance of every reactor by its dt (65,536
reactor-steps).  Warm-up steps run on a separate batch with the same inputs,
so the timed region covers steps 1..K of a fresh batch from t = 0, with all
inputs already resident in HBM.

The default run prints ONE JSON line holding the headline plus:
  mixed_precision  the same workload in the mixed form (FP32 iteration matrix,
                   FP64 residual), error against the FP64 result;
  surrogate        the target case of the learned part: the same 65,536
                   reactors at a fixed dt = 1e-5 s, plain FP64 vs the mixed
                   form with guarded surrogate steps (network trained first on
                   gated pairs of FP64 GPU trajectories of other reactors from
                   the same distribution), kept fraction, speedup and the
                   largest T and Y differences over all outputs;
  workloads        the other single-GPU workloads of the path, each with its
                   roofline and CPU baseline: gri30 rates over 1,048,576 states
                   (FP64 and the FP32 rate-constant option with its error),
                   gri30 reactor Jacobians (1,048,576 states), n-dodecane
                   262,144 reactors (a few steps, FP64 and mixed), h2o2 1,024
                   reactors x 100 outputs with net production rates at every
                   recorded state.
--workload X runs one workload alone (one JSON line).

CPU baselines: the oracle (the reference's algorithms restated in C, built
-O3 -march=native) on the host of the GPU, OpenMP schedule(static,1) over
reactors/states like the reference's own batching sample, on
cpu_threads() threads (the host's CPU share for one GPU) and on 1 thread,
each on a bounded sample stated in the line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import gc
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 (vector and matrix)
BF16_PEAK_TFLOPS = 2500.0    # MI355X dense BF16 MFMA
HBM_PEAK_GBS = 8000.0
METRIC = "GRI30 reactor-steps/s: plain FP64 BDF vs mixed+surrogate at ≤1 K T error"
VALU_PIPE = "FP64 VALU (no MFMA issued; the MI355X FP64 vector peak is 78.6 TFLOP/s)"
# N > 32 with the iteration matrix in LDS (gri30, n-dodecane): the dense LU's
# rank-4 trailing updates run on the matrix cores
LDS_LU_PIPE = ("FP64 VALU, the dense LU's trailing updates on FP64 MFMA (v_mfma_f64_16x16x4f64); "
               "the MI355X FP64 peak is 78.6 TFLOP/s for both")


def reactor_pipe(n_species, storage):
    """Pipes of the integrator kernel: systems of N = K + 2 > 32 unknowns with
    the matrix in LDS factor on the matrix cores; N <= 32 (register LU) and the
    global-matrix policy factor on the VALU alone."""
    return LDS_LU_PIPE if n_species + 2 > 32 and storage != "global" else VALU_PIPE
SURROGATE_DT = 1e-5

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


def measured_traffic(name, n, steps=None):
    """HBM bytes per launch of the dominant kernel from a committed rocprofv3
    --pmc summary: profiles/<name>_<n>[x<steps>].md holding `traffic_n: <n>`
    (and `traffic_steps: <steps>` when given) lines.

    The summary holds FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the
    kernel and of a plain copy kernel of known size; each counter is scaled by
    (copy bytes / copy counter).  None when no summary was taken at this
    configuration.
    """
    pdir = os.path.join(ROOT, "profiles")
    try:
        files = sorted(os.listdir(pdir))
    except OSError:
        return None
    for fn in files:
        stem, ext = os.path.splitext(fn)
        if ext != ".md" or not stem.startswith(name + "_"):
            continue
        tail = stem[len(name) + 1:]
        if not tail.replace("x", "").isdigit():
            continue                      # another workload whose name extends this one
        try:
            kv = {}
            with open(os.path.join(pdir, fn)) as f:
                for line in f:
                    if line.startswith("traffic_") and ":" in line:
                        k, v = line.split(":", 1)
                        kv[k.strip()] = float(v.split()[0])
            if int(kv["traffic_n"]) != n:
                continue
            if steps is not None and int(kv.get("traffic_steps", -1)) != steps:
                continue
            return kv["traffic_bytes_per_launch"]
        except (OSError, KeyError, ValueError):
            continue
    return None


def cpu_threads():
    """Threads for the all-cores CPU baseline: the host's CPU share for one
    GPU (OMP_NUM_THREADS, 16 on the MI355X hosts), at most the cores present."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return max(1, min(threads, os.cpu_count() or 1, 16))


def physical_cores():
    """Physical cores of the host from lscpu (sockets x cores per socket)."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        return int(kv["Socket(s)"]) * int(kv["Core(s) per socket"])
    except (OSError, KeyError, ValueError, subprocess.SubprocessError):
        return None


def cpu_baseline(run, unit, sample_all, sample_one, describe):
    """Time run(S, threads) -> work units on the all-cores share and on one
    thread; returns the cpu_baseline object (value = the all-cores rate)."""
    th = cpu_threads()
    t0 = time.perf_counter()
    w_all = run(sample_all, th)
    el_all = time.perf_counter() - t0
    t0 = time.perf_counter()
    w_one = run(sample_one, 1)
    el_one = time.perf_counter() - t0
    return {"value": w_all / el_all, "unit": unit, "cores": th, "kind": "port",
            "host_physical_cores": physical_cores(),
            "sample": f"{describe(sample_all)}, {th} threads (the host's CPU share for one GPU), {el_all:.1f} s",
            "one_thread": {"value": w_one / el_one, "unit": unit, "sample": f"{describe(sample_one)}, {el_one:.1f} s"}}


def cpu_reactors(mech_path, wl, steps, sample_all, sample_one, const_volume=False):
    from oracle.oracle import Mech
    om = Mech(mech_path)

    def run(S, th):
        S = min(S, len(wl.T0))
        om.run_reactors(wl.T0[:S], wl.P[:S], wl.Y0[:S], wl.dt[:S], steps, wl.rtol, wl.atol,
                        analytic=False, nthreads=th, keep_states=False, const_volume=const_volume)
        return S * steps

    return cpu_baseline(run, "reactor-steps/s", sample_all, sample_one,
                        lambda S: f"{min(S, len(wl.T0))} of the {len(wl.T0)} reactors x {steps} steps, CVODES-method "
                                  f"BDF with difference-quotient Jacobian (oracle), OpenMP schedule(static,1)")


def _dev_array(L, check, host):
    host = np.ascontiguousarray(host)
    p = L.cg_dev_alloc(max(1, host.nbytes))
    if not p:
        raise MemoryError("cg_dev_alloc")
    check(L.cg_memcpy_h2d(p, host.ctypes.data_as(C.c_void_p), host.nbytes))
    return p


def _timed_advances(L, check, rb, d_tout, steps, n, log=None, history=None):
    """K advance calls on the batch (output times at row k of d_tout), timed
    between device synchronisations; returns (seconds, kernel ms per call)."""
    check(L.cg_device_synchronize())
    kernel_ms = []
    t0 = time.perf_counter()
    for k in range(steps):
        check(L.cg_reactors_advance_device(rb.handle, C.c_void_p(d_tout + k * n * 8)))
        if history is not None:
            rb.state_to_device(history, k)
        kernel_ms.append(rb.last_kernel_ms())
        if log:
            print(f"{log} step {k + 1}/{steps}: kernel {kernel_ms[-1]:.2f} ms", file=sys.stderr, flush=True)
    check(L.cg_device_synchronize())
    return time.perf_counter() - t0, kernel_ms


def reactors(a):
    """The reactor workloads: the headline (gri30) with its sub-results, or
    n-dodecane / h2o2 alone."""
    import cantera_amd as ca
    from cantera_amd._lib import check, lib
    from cantera_amd.workloads import config1, config3, config5

    L = lib()
    mech = {"gri30-reactors": "gri30", "ndodecane-reactors": "nDodecane", "h2o2-reactors": "h2o2"}[a.workload]
    if mech == "h2o2":
        out = h2o2_result(a)
        print(json.dumps(out), flush=True)
        return 0
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
    else:
        n = a.n or 262144
        wl = config5(sol, n=n, nsteps=steps)
        metric = "nDodecane reactor-steps/s (FP64 BDF)"
        desc = (f"nDodecane ({sol.n_species} species, {sol.n_reactions} reactions) IdealGasConstPressureReactor, "
                f"{n} reactors x {steps} "
                f"advance steps of 1e-6 s, T0 U[700,2200] K, c12h26/air phi U[0.3,3], P 20-60 atm, rtol 1e-8 "
                f"atol 1e-15")
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
    d_tout = _dev_array(L, check, touts)

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
        gc.collect()

    rb = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol, precision=prec,
                         const_volume=a.const_volume)
    rb.set_options(0, a.waves_per_cu)
    rb.set_schedule(a.schedule == "longest-first")
    rb.set_matrix_storage(a.matrix_storage)
    if tuned:
        rb.set_matrix_storage(tuned["matrix_storage"])
        rb.set_options(0, tuned["waves_per_cu"])
    check(L.cg_device_synchronize())
    if a.graph:
        # all K advances captured as one HIP graph (capture included in the timing)
        t0 = time.perf_counter()
        check(L.cg_reactors_advance_graph(rb.handle, steps, C.c_void_p(d_tout)))
        check(L.cg_device_synchronize())
        elapsed = time.perf_counter() - t0
        kernel_ms = [rb.last_kernel_ms() / steps]
        print(f"{steps} steps in one graph: {kernel_ms[0] * steps:.2f} ms", file=sys.stderr, flush=True)
    else:
        elapsed, kernel_ms = _timed_advances(L, check, rb, d_tout, steps, n, log="fp64")
    bad = int((rb.status != 0).sum())
    if bad:
        print(f"{bad} reactors failed to integrate", file=sys.stderr)
        return 1
    stats = rb.stats
    W, F_rhs = flop_model(sol.path, stats)
    avg_launch_s = float(np.mean(kernel_ms)) / 1e3
    achieved = W / steps / avg_launch_s / 1e12
    name = f"bdf_advance_{mech}" + ("_cv" if a.const_volume else "")
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
            "bound": "valu",
            "pipe": reactor_pipe(sol.n_species, tuned["matrix_storage"] if tuned else a.matrix_storage),
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": measured_traffic(name, n, steps),
            "kernel": "bdf_advance",
            "avg_launch_ms": avg_launch_s * 1e3,
            "flops_per_launch": W / steps,
            "flop_model": "W = N_rhs F_rhs + N_jac F_jac + N_lu (2/3) N^3 + N_newton 2 N^2 from the integrator's "
                          "counters; exp/log/pow counted 20 flops",
        },
        "integrator": {k: float(np.mean(v)) for k, v in stats.items() if k != "order"},
    }
    if a.const_volume:
        out["roofline"]["flop_model"] += " (the constant-pressure model: a lower bound for constant volume)"
    if a.precision == "fp64" and not a.no_mixed:
        # the mixed form (FP32 iteration matrix / LU / solves, FP64 residual)
        # on a fresh batch, same inputs and steps, after the timed FP64 region;
        # its error is stated against the FP64 result
        y64 = rb.state
        del rb
        gc.collect()
        out["mixed_precision"] = mixed_result(ca, L, check, sol, wl, d_tout, steps, n, elapsed, y64,
                                              a.const_volume)
    else:
        del rb
        gc.collect()
    L.cg_dev_free(d_tout)
    if not a.no_cpu:
        sample = a.cpu_sample or {"gri30": 16384, "nDodecane": 256}[mech]
        out["cpu_baseline"] = cpu_reactors(sol.path, wl, steps, sample, max(64, sample // 32), a.const_volume)
    if mech == "gri30" and not a.const_volume and not a.no_extras:
        out["surrogate"] = surrogate_block(a, sol, wl, steps)
        gc.collect()
        out["workloads"] = {
            "gri30_rates_fp64": batch_kernel_result(a, "rates", "fp64"),
            "gri30_rates_fp32": batch_kernel_result(a, "rates", "fp32"),
            "gri30_jacobian_fp64": batch_kernel_result(a, "jacobian", "fp64"),
            "ndodecane_reactors": ndodecane_result(a),
            "h2o2_reactors": h2o2_result(a),
        }
    copy_calib(a, L, check)
    print(json.dumps(out), flush=True)
    return 0


def mixed_result(ca, L, check, sol, wl, d_tout, steps, n, elapsed64, y64, const_volume=False):
    # warm the mixed kernels on a few reactors, then time all K steps
    wm = ca.ReactorBatch(sol, wl.T0[:64], wl.P[:64], wl.Y0[:64], rtol=wl.rtol, atol=wl.atol, precision=ca.CG_MIXED,
                         const_volume=const_volume)
    wm.advance(wl.dt[:64])
    del wm
    rm = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol, precision=ca.CG_MIXED,
                         const_volume=const_volume)
    el_m, _ = _timed_advances(L, check, rm, d_tout, steps, n)
    res = None
    if int((rm.status != 0).sum()) == 0:
        ym = rm.state
        res = {
            "value": n * steps / el_m, "unit": "reactor-steps/s", "steps": steps,
            "speedup_vs_fp64": elapsed64 / el_m,
            "max_abs_dT_vs_fp64_K": float(np.max(np.abs(ym[:, 1] - y64[:, 1]))),
            "max_abs_dY_vs_fp64": float(np.max(np.abs(ym[:, 2:] - y64[:, 2:]))),
        }
    del rm
    gc.collect()
    return res


def surrogate_block(a, sol, wl, steps):
    """The learned part on its target case: the headline's reactors at a fixed
    dt = 1e-5 s.  Trains a network on gated pairs of FP64 GPU trajectories of
    other reactors from the same distribution, then times K steps of (1) plain
    FP64, (2) the mixed form alone, (3) the mixed form with guarded surrogate
    steps, and compares (2) and (3) with (1) at every output."""
    import cantera_amd as ca
    from cantera_amd._lib import check, lib
    from cantera_amd.surrogate import Surrogate, generate_pairs
    from cantera_amd.workloads import config3

    L = lib()
    K = sol.n_species
    N = K + 2
    n = len(wl.T0)
    dt = SURROGATE_DT
    g = ca.ReactorBatch.GUARD
    hidden, depth = 256, 4
    # training pairs: 16,384 reactors of another seed, 100 steps of dt, FP64,
    # kept where the first state passes the guard's regime gate
    t0 = time.perf_counter()
    tr = config3(sol, n=a.train_reactors, nsteps=100, seed=0x5EED0104)
    y0, P, y1 = generate_pairs(sol, tr.T0, tr.P, tr.Y0, dt=dt, nsteps=100, rtol=tr.rtol, atol=tr.atol,
                               gate_T=g["gate_T"])
    gen_s = time.perf_counter() - t0
    sur = Surrogate(sol, dt=dt, hidden=hidden, depth=depth, seed=1)
    t0 = time.perf_counter()
    losses = sur.fit(y0, P, y1, epochs=a.epochs, batch=4096, lr=1e-3)
    check(L.cg_device_synchronize())
    train_s = time.perf_counter() - t0
    n_pairs = len(y0)
    del y0, P, y1
    gc.collect()
    n_in, n_out = K + 2, K + 1
    flop_fwd = 2.0 * (n_in * hidden + (depth - 1) * hidden * hidden + hidden * n_out)

    touts = np.ascontiguousarray(np.outer(np.arange(1, steps + 1, dtype=np.float64), np.full(n, dt)))
    d_tout = _dev_array(L, check, touts)
    hist_bytes = steps * n * N * 8
    h64 = L.cg_dev_alloc(hist_bytes)
    hx = L.cg_dev_alloc(hist_bytes)
    if not h64 or not hx:
        raise MemoryError("cg_dev_alloc(history)")

    def host_hist(p):
        h = np.empty((steps, n, N))
        check(L.cg_memcpy_d2h(h.ctypes.data_as(C.c_void_p), p, hist_bytes))
        return h

    # (1) plain FP64 at dt
    rb = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
    el64, _ = _timed_advances(L, check, rb, d_tout, steps, n, log="surrogate-case fp64", history=h64)
    ok64 = int((rb.status != 0).sum()) == 0
    stats64 = rb.stats
    del rb
    gc.collect()
    y64 = host_hist(h64)
    L.cg_dev_free(h64)

    def diff(h):
        return (float(np.max(np.abs(h[:, :, 1] - y64[:, :, 1]))), float(np.max(np.abs(h[:, :, 2:] - y64[:, :, 2:]))))

    # (2) mixed alone at dt
    rm = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol, precision=ca.CG_MIXED)
    el_m, _ = _timed_advances(L, check, rm, d_tout, steps, n, history=hx)
    okm = int((rm.status != 0).sum()) == 0
    del rm
    gc.collect()
    dTm, dYm = diff(host_hist(hx))

    # (3) mixed + guarded surrogate at dt (one device-resident call per step)
    rs = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol, precision=ca.CG_MIXED)
    check(L.cg_device_synchronize())
    t0 = time.perf_counter()
    for k in range(steps):
        rs.advance_surrogate_device(d_tout, sur, row=k, **g)
        rs.state_to_device(hx, k)
    check(L.cg_device_synchronize())
    el_s = time.perf_counter() - t0
    counts = rs.surrogate_counts()
    oks = int((rs.status != 0).sum()) == 0
    del rs
    gc.collect()
    dTs, dYs = diff(host_hist(hx))
    L.cg_dev_free(hx)

    # inference alone on the matrix cores (the evaluation states), for its roofline
    d0 = _dev_array(L, check, np.concatenate([np.ones((n, 1)), wl.T0[:, None], wl.Y0], axis=1))
    dP = _dev_array(L, check, np.asarray(wl.P, np.float64))
    d1 = L.cg_dev_alloc(n * N * 8)
    for _ in range(2):
        check(L.cg_surrogate_predict(sur.handle, n, d0, dP, d1))
    check(L.cg_device_synchronize())
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        check(L.cg_surrogate_predict(sur.handle, n, d0, dP, d1))
    check(L.cg_device_synchronize())
    inf_s = (time.perf_counter() - t0) / reps
    for p in (d0, dP, d1, d_tout):
        L.cg_dev_free(p)
    achieved = n * flop_fwd / inf_s / 1e12
    plain = n * steps / el64
    return {
        "workload": f"the headline's {n} gri30 reactors (same T0, P, Y0) at a fixed dt = 1e-5 s, {steps} steps; "
                    f"FP64 tolerances rtol 1e-8, atol 1e-15",
        "plain_fp64_reactor_steps_per_s": plain,
        "mixed_reactor_steps_per_s": n * steps / el_m,
        "mixed_plus_surrogate_reactor_steps_per_s": n * steps / el_s,
        "speedup_mixed_vs_plain": el64 / el_m,
        "speedup_mixed_plus_surrogate_vs_plain": el64 / el_s,
        "target_speedup": 3.0,
        "eligible_reactor_steps": int(counts[0]),
        "kept_surrogate_steps": int(counts[1]),
        "kept_fraction": float(counts[1] / max(1, n * steps)),
        "fallback_fraction": float(counts[2] / max(1, n * steps)),
        "max_abs_dT_vs_fp64_K": dTs, "max_abs_dY_vs_fp64": dYs,
        "mixed_alone_max_abs_dT_vs_fp64_K": dTm, "mixed_alone_max_abs_dY_vs_fp64": dYm,
        "error_measure": "largest difference over all reactors and all outputs against the plain FP64 trajectories",
        "budget": {"dT_K": 1.0, "dY": 1e-4},
        "within_budget": bool(dTs <= 1.0 and dYs <= 1e-4),
        "all_reactors_integrated": bool(ok64 and okm and oks),
        "plain_internal_steps_per_reactor": float(np.mean(stats64["steps"])),
        "guard": {"tol_T_K_per_step": g["tol_T"], "tol_Y_per_step": g["tol_Y"], "gate_T_K": g["gate_T"],
                  "derivation": "budget 1 K and 1e-4 over 100 steps, divided by the steps"},
        "network": f"MLP {n_in} -> {depth} x {hidden} GELU -> {n_out}, BF16 weights/activations, FP32 accumulation",
        "training": {"pairs": n_pairs, "source": f"{a.train_reactors} reactors (seed 0x5EED0104) x 100 FP64 GPU "
                                                 f"steps of 1e-5 s, pairs whose first state passes the regime gate",
                     "epochs": a.epochs, "batch": 4096, "lr": 1e-3, "pair_generation_s": gen_s, "training_s": train_s,
                     "samples_per_s": n_pairs * a.epochs / train_s,
                     "tflops": n_pairs * a.epochs * 3 * flop_fwd / train_s / 1e12,
                     "loss_first_last_epoch": [losses[0], losses[-1]]},
        "inference_roofline": {"bound": "mfma", "pipe": "BF16 MFMA (dense 2.5 PFLOP/s)", "achieved": achieved,
                               "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / BF16_PEAK_TFLOPS,
                               "avg_launch_ms": inf_s * 1e3, "flops_per_launch": n * flop_fwd,
                               "kernel": "gemm_nt_bf16 (features, 5 GEMMs, decode + element projection)"},
    }


def batch_kernel_result(a, kind, precision, n=None, steps=None, warmup=None, cpu=True):
    """Batched rates or Jacobians over gri30 states resident in HBM."""
    import cantera_amd as ca
    from cantera_amd._lib import check, lib
    from cantera_amd.workloads import config2_states

    L = lib()
    sol = ca.Solution("gri30")
    n = n or (a.n if a.workload in ("gri30-rates", "gri30-jacobian") and a.n else 1_048_576)
    steps = steps or (a.steps if a.workload in ("gri30-rates", "gri30-jacobian") and a.steps else 20)
    warmup = a.warmup if warmup is None else warmup
    K = sol.n_species
    N = K + 2
    T, P, Y = config2_states(sol, n=n)
    jac = kind == "jacobian"
    f32 = precision == "fp32"
    out_doubles = n * N * N if jac else n * K
    d_in = L.cg_dev_alloc((2 + K) * n * 8)
    d_out = L.cg_dev_alloc(out_doubles * (4 if (jac and f32) else 8))
    for off, arr in ((0, T), (n, P), (2 * n, Y)):
        arr = np.ascontiguousarray(arr)
        check(L.cg_memcpy_h2d(d_in + off * 8, arr.ctypes.data_as(C.c_void_p), arr.nbytes))
    dT, dP, dY = d_in, d_in + n * 8, d_in + 2 * n * 8
    call = L.cg_jac_batch_device if jac else L.cg_rates_batch_device
    last = L.cg_jac_last_kernel_ms if jac else L.cg_rates_last_kernel_ms
    prec = ca.CG_FP32 if f32 else ca.CG_FP64
    for _ in range(max(1, warmup)):
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
        # (untimed): per state, max_k |x32 - x64| / max_k |x64|
        if jac:
            x32 = np.empty(out_doubles, dtype=np.float32)
            check(L.cg_memcpy_d2h(x32.ctypes.data_as(C.c_void_p), d_out, x32.nbytes))
            x32 = x32.astype(np.float64).reshape(n, N * N).T
            L.cg_dev_free(d_out)
            d_out = L.cg_dev_alloc(out_doubles * 8)
        else:
            x32 = np.empty((K, n))
            check(L.cg_memcpy_d2h(x32.ctypes.data_as(C.c_void_p), d_out, x32.nbytes))
        check(call(sol.handle, n, dT, dP, dY, d_out, ca.CG_FP64, None))
        check(L.cg_device_synchronize())
        x64 = np.empty(out_doubles) if jac else np.empty((K, n))
        check(L.cg_memcpy_d2h(x64.ctypes.data_as(C.c_void_p), d_out, x64.nbytes))
        if jac:
            x64 = x64.reshape(n, N * N).T
        e = np.abs(x32 - x64).max(axis=0) / (np.abs(x64).max(axis=0) + 1e-300)
        fp32_error = {"per_state_max_rel_to_largest_entry": {"max": float(e.max()),
                                                             "p99": float(np.percentile(e, 99)),
                                                             "median": float(np.median(e))}}
        del x32, x64
    fl, tr, _ = rhs_cost(sol.path)
    F_rates = fl + TRANS_FLOPS * tr
    avg_s = float(np.mean(kernel_ms)) / 1e3
    name = ("jac_batch_gri30" if jac else "rates_gri30") + ("_fp32" if f32 else "")
    if jac:
        # HBM-bound: T, P, Y in; the N x N Jacobian out
        es = 4 if f32 else 8
        alg_bytes = n * (8 * (2 + K) + es * N * N)
        roof = {"bound": "hbm", "achieved": alg_bytes / avg_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": alg_bytes / avg_s / 1e9 / HBM_PEAK_GBS, "traffic": measured_traffic(name, n),
                "kernel": "jac_batch_kernel", "avg_launch_ms": avg_s * 1e3, "bytes_per_launch": alg_bytes}
        metric = f"gri30 reactor Jacobians/s ({'FP32' if f32 else 'FP64'} output, 55x55 dense per state)"
    else:
        flops = n * F_rates
        roof = {"bound": "valu", "pipe": VALU_PIPE, "achieved": flops / avg_s / 1e12, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": flops / avg_s / 1e12 / FP64_PEAK_TFLOPS,
                "traffic": measured_traffic(name, n), "kernel": "rates_kernel", "avg_launch_ms": avg_s * 1e3,
                "flops_per_launch": flops, "hbm_bytes_algorithmic": n * 8 * (2 + 2 * K),
                "hbm_frac": n * 8 * (2 + 2 * K) / avg_s / 1e9 / HBM_PEAK_GBS}
        metric = "gri30 net production rates, states/s (" + ("FP32 rate constants" if f32 else "FP64") + ")"
    out = {
        "metric": metric, "value": n * steps / elapsed, "unit": "states/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": elapsed / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"gri30 states, {n} per launch: T U[300,2500] K, P 0.1-50 atm, CH4/O2/N2 majors at "
                               f"phi U[0.3,3], minors X 1e-12..1e-2; species-major SoA in HBM"},
        "roofline": roof,
    }
    if f32:
        out["dtype"] = ("f32 Jacobian (f64 state and rates inside)" if jac
                        else "f32 rate-constant transcendentals, f64 otherwise")
        out["fp32_error_vs_fp64"] = fp32_error
    if cpu and not a.no_cpu:
        from oracle.oracle import Mech
        om = Mech(sol.path)

        def run(S, th):
            S = min(S, n)
            Ys = np.ascontiguousarray(Y[:, :S])
            if jac:
                om.jacobian_batch(T[:S], P[:S], Ys, nthreads=th)
            else:
                om.rates_batch(T[:S], P[:S], Ys, nthreads=th)
            return S

        sa, so = (16384, 2048) if jac else (262144, 32768)
        out["cpu_baseline"] = cpu_baseline(
            run, "states/s", sa, so,
            lambda S: f"first {min(S, n)} of the {n} states, oracle "
                      f"{'analytic reactor Jacobian' if jac else 'net production rates'}, OpenMP schedule(static,1)")
    L.cg_dev_free(d_in)
    L.cg_dev_free(d_out)
    return out


def batch_kernels(a):
    kind = "jacobian" if a.workload == "gri30-jacobian" else "rates"
    out = batch_kernel_result(a, kind, a.precision if a.precision in ("fp64", "fp32") else "fp64")
    from cantera_amd._lib import check, lib
    copy_calib(a, lib(), check)
    print(json.dumps(out), flush=True)
    return 0


def mole_result(ca, L, check, sol, wl, d_tout, steps, n, el_dense, y64, cpu_args=None):
    """The same reactors as IdealGasConstPressureMoleReactors with the
    adaptive preconditioner and GMRES (the reference's path for large
    mechanisms: ReactorNet.preconditioner = AdaptivePreconditioner()), with the
    reference's default derivative settings and with third-body and falloff
    derivatives skipped; timed like the dense line, compared with its final
    states."""
    out = {}
    for skip in (False, True):
        mb = ca.MoleReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol, skip_third_bodies=skip,
                                 skip_falloff=skip)
        el, kms = _timed_advances(L, check, mb, d_tout, steps, n, log=f"mole+GMRES skip={skip}")
        st, ls = mb.stats, mb.linear_stats
        T, Y = mb.T, mb.Y
        out["skip_third_bodies_and_falloff" if skip else "default_derivatives"] = {
            "value": n * steps / el, "unit": "reactor-steps/s", "speedup_vs_dense_fp64": el_dense / el,
            "avg_launch_ms": float(np.mean(kms)), "all_reactors_integrated": bool((mb.status != 0).sum() == 0),
            "max_abs_dT_vs_dense_K": float(np.abs(T - y64[:, 1]).max()),
            "max_abs_dY_vs_dense": float(np.abs(Y - y64[:, 2:]).max()),
            "integrator": {k: float(np.mean(v)) for k, v in st.items() if k != "order"},
            "linear_solver": {k: float(np.mean(v)) for k, v in ls.items()},
        }
        del mb
    out["note"] = ("state [T, n_k]; the species' atol is atol times the initial total moles "
                   "(MoleReactor::updateDefaultTolerances), so the mole path takes fewer internal steps")
    if not getattr(cpu_args, "no_cpu", True):
        from oracle.oracle import Mech
        om = Mech(sol.path)

        def run(S, th):
            S = min(S, n)
            om.run_mole_reactors(wl.T0[:S], wl.P[:S], wl.Y0[:S], wl.dt[:S], steps, wl.rtol, wl.atol, nthreads=th)
            return S * steps

        out["cpu_baseline"] = cpu_baseline(
            run, "reactor-steps/s", 512, 32,
            lambda S: f"{min(S, n)} of the {n} reactors x {steps} steps, mole reactor with the adaptive "
                      f"preconditioner and GMRES (oracle, default derivative settings), OpenMP schedule(static,1)")
    return out


def ndodecane_result(a, steps=5, mixed=False, n=16384):
    """n-dodecane, 262,144 reactors x `steps` advances of 1e-6 s from t = 0 (the
    start-up step, the costliest of the 100): FP64, and with mixed=True the
    mixed form on a fresh batch (error against FP64)."""
    import cantera_amd as ca
    from cantera_amd._lib import check, lib
    from cantera_amd.workloads import config5

    L = lib()
    sol = ca.Solution("nDodecane")
    wl = config5(sol, n=n, nsteps=steps)
    touts = np.ascontiguousarray(np.outer(np.arange(1, steps + 1, dtype=np.float64), wl.dt))
    d_tout = _dev_array(L, check, touts)
    rb = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
    el, kms = _timed_advances(L, check, rb, d_tout, steps, n, log="nDodecane fp64")
    ok = int((rb.status != 0).sum()) == 0
    stats = rb.stats
    y64 = rb.state
    del rb
    gc.collect()
    W, _ = flop_model(sol.path, stats)
    avg_s = float(np.mean(kms)) / 1e3
    achieved = W / steps / avg_s / 1e12
    out = {
        "metric": "nDodecane reactor-steps/s (FP64 BDF)", "value": n * steps / el, "unit": "reactor-steps/s",
        "steps": steps, "ms_per_step": el / steps * 1e3, "dtype": "f64", "all_reactors_integrated": bool(ok),
        "config": {"workload": f"nDodecane ({sol.n_species} species, {sol.n_reactions} reactions) "
                               f"IdealGasConstPressureReactor, {n} reactors x {steps} advance steps of 1e-6 s from "
                               f"t = 0, T0 U[700,2200] K, c12h26/air phi U[0.3,3], P 20-60 atm, rtol 1e-8 atol 1e-15"},
        "roofline": {"bound": "valu", "pipe": LDS_LU_PIPE, "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": measured_traffic("bdf_advance_nDodecane", n, steps), "kernel": "bdf_advance_l2",
                     "avg_launch_ms": avg_s * 1e3, "flops_per_launch": W / steps},
        "integrator": {k: float(np.mean(v)) for k, v in stats.items() if k != "order"},
    }
    if mixed:
        out["mixed_precision"] = mixed_result(ca, L, check, sol, wl, d_tout, steps, n, el, y64)
    out["mole_gmres"] = mole_result(ca, L, check, sol, wl, d_tout, steps, n, el, y64, cpu_args=a)
    L.cg_dev_free(d_tout)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_reactors(sol.path, wl, steps, 512, 32)
    return out


def h2o2_result(a, n=1024, steps=100):
    """h2o2, 1,024 reactors to 1 ms with outputs every 1e-5 s and the net
    production rates evaluated at every recorded state (the kinetics1 sample's
    loop, samples/cxx/kinetics1/kinetics1.cpp:42-60, batched): per step one
    advance and one rate evaluation at the new states, both queued on the
    device; states and rates of every output kept in HBM."""
    import cantera_amd as ca
    from cantera_amd._lib import check, lib
    from cantera_amd.workloads import config1

    L = lib()
    sol = ca.Solution("h2o2")
    K = sol.n_species
    wl = config1(sol, n=n)
    touts = np.ascontiguousarray(np.outer(np.arange(1, steps + 1, dtype=np.float64), wl.dt))
    d_tout = _dev_array(L, check, touts)
    d_y = L.cg_dev_alloc(steps * n * (K + 2) * 8)
    d_w = L.cg_dev_alloc(steps * n * K * 8)
    warm = ca.ReactorBatch(sol, wl.T0[:64], wl.P[:64], wl.Y0[:64], rtol=wl.rtol, atol=wl.atol)
    warm.advance(wl.dt[:64])
    del warm
    rb = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol)
    check(L.cg_device_synchronize())
    t0 = time.perf_counter()
    for k in range(steps):
        check(L.cg_reactors_advance_device(rb.handle, C.c_void_p(d_tout + k * n * 8)))
        rb.state_to_device(d_y, k)
        rb.rates_to_device(d_w, k)
    check(L.cg_device_synchronize())
    el = time.perf_counter() - t0
    ok = int((rb.status != 0).sum()) == 0
    w_last = np.empty((K, n))
    check(L.cg_memcpy_d2h(w_last.ctypes.data_as(C.c_void_p), d_w + (steps - 1) * n * K * 8, w_last.nbytes))
    wsum = float(np.abs(w_last).sum())
    for p in (d_tout, d_y, d_w):
        L.cg_dev_free(p)
    out = {
        "metric": "h2o2 reactor-steps/s (FP64 BDF + net production rates at every output)",
        "value": n * steps / el, "unit": "reactor-steps/s", "steps": steps, "ms_per_step": el / steps * 1e3,
        "dtype": "f64", "all_reactors_integrated": bool(ok and np.isfinite(wsum)),
        "config": {"workload": f"h2o2 ({sol.n_species} species, {sol.n_reactions} reactions) "
                               f"IdealGasConstPressureReactor, {n} reactors x {steps} outputs every 1e-5 s, "
                               f"T0 U[1000,1800] K, H2/air phi U[0.5,2], 1 atm, rtol 1e-8 atol 1e-15; each step = "
                               f"advance + state copy + net production rates at the new states, all on the device"},
        "roofline": {"bound": "latency", "note": "1,024 reactors give at most one wavefront per SIMD on 256 CUs; "
                                                 "the integrator's dependent chains set the time", "traffic": None},
    }
    del rb
    if not a.no_cpu:
        from oracle.oracle import Mech
        om = Mech(sol.path)

        def run(S, th):
            S = min(S, n)
            y, _ = om.run_reactors(wl.T0[:S], wl.P[:S], wl.Y0[:S], wl.dt[:S], steps, wl.rtol, wl.atol,
                                   analytic=False, nthreads=th, keep_states=True)
            for k in range(1, steps + 1):
                om.rates_batch(y[:, k, 1], wl.P[:S], np.ascontiguousarray(y[:, k, 2:].T), nthreads=th)
            return S * steps

        out["cpu_baseline"] = cpu_baseline(
            run, "reactor-steps/s", n, 128,
            lambda S: f"{min(S, n)} reactors x {steps} outputs, CVODES-method BDF (oracle) + rates at each output, "
                      f"OpenMP schedule(static,1)")
    return out


def surrogate(a):
    """The surrogate target case alone (one JSON line)."""
    import cantera_amd as ca
    from cantera_amd.workloads import config3
    sol = ca.Solution("gri30")
    steps = a.steps or 20
    wl = config3(sol, n=a.n or 65536, nsteps=steps)
    block = surrogate_block(a, sol, wl, steps)
    out = {"metric": "gri30 reactor-steps/s at dt = 1e-5 s: mixed + guarded surrogate vs plain FP64",
           "value": block["mixed_plus_surrogate_reactor_steps_per_s"], "unit": "reactor-steps/s", "n_gpus": 1,
           "steps": steps, "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16 network (fp32 accumulate), f64 guard, f32 iteration matrix / f64 residual fallback",
           "data": "synthetic", "config": {"workload": block["workload"]}, "surrogate": block,
           "roofline": block["inference_roofline"]}
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
                    choices=["gri30-reactors", "gri30-rates", "gri30-jacobian", "ndodecane-reactors", "ndodecane-mole",
                             "h2o2-reactors",
                             "gri30-surrogate"],
                    help="gri30-reactors is the headline line; the others measure one workload alone")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--precision", choices=["fp64", "mixed", "fp32"], default="fp64",
                    help="reactors: fp64 or mixed (FP32 LU); gri30-rates/-jacobian: fp64 or fp32")
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--schedule", choices=["longest-first", "index"], default="longest-first")
    ap.add_argument("--matrix-storage", choices=["default", "lds", "global"], default="default")
    ap.add_argument("--graph", action="store_true",
                    help="reactor workloads: run the K steps as one captured HIP graph")
    ap.add_argument("--autotune", action="store_true",
                    help="reactor workloads: pick the integrator launch by timing the candidates first")
    ap.add_argument("--train-reactors", type=int, default=16384, help="surrogate: training reactors (x 100 steps)")
    ap.add_argument("--epochs", type=int, default=20, help="surrogate: training epochs")
    ap.add_argument("--const-volume", action="store_true",
                    help="reactor workloads: closed rigid IdealGasReactor instead of IdealGasConstPressureReactor")
    ap.add_argument("--no-mixed", action="store_true",
                    help="gri30-/ndodecane-reactors: skip the mixed-precision run after the FP64 line")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline: skip the surrogate block and the other workloads (profiling runs)")
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
    if a.workload == "ndodecane-mole":
        # dense mass-fraction integrator and the mole reactor with GMRES on the
        # same n-dodecane reactors (the default line's n-dodecane sub-result)
        print(json.dumps(ndodecane_result(a, steps=a.steps or 5, n=a.n or 16384)), flush=True)
        return 0
    return reactors(a)


if __name__ == "__main__":
    sys.exit(main())
