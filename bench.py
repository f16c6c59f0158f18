"""Benchmark: GRI-3.0 constant-pressure reactor integration throughput.

Workload: 65,536 independent gri30 IdealGasConstPressureReactors, T0 ~ U[900,
2000] K, CH4/air at phi ~ U[0.5, 3.0], P ~ 1 atm x exp(U[ln 1, ln 40]), each
advanced 100 times by its own dt ~ exp(U[ln 1e-6, ln 1e-4]) s, rtol 1e-8,
atol 1e-15, FP64.  A "step" is one ReactorNet::advance of every reactor by its
dt (65,536 reactor-steps).

Warm-up steps run on a separate batch with the same inputs, so the timed
region covers steps 1..K of a fresh batch from t = 0, with all inputs already
resident in HBM.  One JSON line is printed.
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

# Algorithmic cost convention: every FP64 add/sub/mul/div counts 1, an FMA 2,
# and each transcendental (exp, log, log10, exp10, pow) counts TRANS_FLOPS.
TRANS_FLOPS = 20.0


def rhs_cost(mech_path):
    """(flops, transcendentals) of one reactor RHS evaluation, from the tables."""
    K = I = 0
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
                        tr += 5                   # 3 exp (Fcent) + 2 log10 + exp10 - counted 5 (T2 term optional)
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


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "bdf_advance_gri30_65536_traffic.md")


def measured_traffic(n):
    """HBM bytes per integrator launch from the committed rocprofv3 --pmc summary.

    The summary holds FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the
    integrator kernel and of a plain copy kernel of known size; each counter is
    scaled by (copy bytes / copy counter).  None when the summary is absent or
    was taken at another batch size.
    """
    try:
        kv = {}
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_") and ":" in line:
                    k, v = line.split(":", 1)
                    kv[k.strip()] = float(v.split()[0])
        if int(kv["traffic_n_reactors"]) != n:
            return None
        return kv["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(wl, steps, sample, threads):
    from oracle.oracle import Mech
    om = Mech(os.path.join(ROOT, "cantera_amd", "mechanisms", "gri30.mech"))
    S = min(sample, len(wl.T0))
    t0 = time.perf_counter()
    _, st = om.run_reactors(wl.T0[:S], wl.P[:S], wl.Y0[:S], wl.dt[:S], steps, wl.rtol, wl.atol,
                            analytic=False, nthreads=threads, keep_states=False)
    el = time.perf_counter() - t0
    return {"value": S * steps / el, "unit": "reactor-steps/s", "cores": threads, "kind": "port",
            "sample": f"{S} of the {len(wl.T0)} reactors x {steps} steps, CVODES-method BDF with "
                      f"difference-quotient Jacobian (oracle), OpenMP schedule(static,1), {el:.1f} s"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--schedule", choices=["longest-first", "index"], default="longest-first")
    ap.add_argument("--cpu-sample", type=int, default=16384)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="after the timed region, run one device-to-device copy (blit kernel) of MIB MiB "
                         "(known HBM bytes, to scale FETCH_SIZE/WRITE_SIZE counters)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the reactors are independent, "
              "so N GPUs would each run their own shard).", file=sys.stderr)
        return 2

    import cantera_amd as ca
    from cantera_amd._lib import check, lib
    from cantera_amd.workloads import config3

    L = lib()
    sol = ca.Solution("gri30")
    wl = config3(sol, n=a.n, nsteps=a.steps)
    prec = ca.CG_FP64 if a.precision == "fp64" else ca.CG_MIXED
    n = a.n

    # all output times resident in HBM before the timed region
    touts = np.ascontiguousarray(np.outer(np.arange(1, a.steps + 1, dtype=np.float64), wl.dt))
    d_tout = L.cg_dev_alloc(touts.nbytes)
    check(L.cg_memcpy_h2d(d_tout, touts.ctypes.data_as(C.c_void_p), touts.nbytes))

    if a.warmup > 0:
        wb = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol, precision=prec)
        wb.set_options(0, a.waves_per_cu)
        for k in range(a.warmup):
            check(L.cg_reactors_advance_device(wb.handle, C.c_void_p(d_tout + k * n * 8)))
        check(L.cg_device_synchronize())
        del wb

    rb = ca.ReactorBatch(sol, wl.T0, wl.P, wl.Y0, rtol=wl.rtol, atol=wl.atol, precision=prec)
    rb.set_options(0, a.waves_per_cu)
    rb.set_schedule(a.schedule == "longest-first")
    check(L.cg_device_synchronize())
    kernel_ms = []
    t0 = time.perf_counter()
    for k in range(a.steps):
        check(L.cg_reactors_advance_device(rb.handle, C.c_void_p(d_tout + k * n * 8)))
        kernel_ms.append(rb.last_kernel_ms())
        print(f"step {k + 1}/{a.steps}: kernel {kernel_ms[-1]:.2f} ms", file=sys.stderr, flush=True)
    check(L.cg_device_synchronize())
    elapsed = time.perf_counter() - t0
    bad = int((rb.status != 0).sum())
    if bad:
        print(f"{bad} reactors failed to integrate", file=sys.stderr)
        return 1
    stats = rb.stats
    W, F_rhs = flop_model(sol.path, stats)
    avg_launch_s = float(np.mean(kernel_ms)) / 1e3
    achieved = W / a.steps / avg_launch_s / 1e12
    # HBM bytes: state records, saved Jacobian and factors per call
    value = n * a.steps / elapsed
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "reactor-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f64 state/residual, f32 iteration matrix",
        "data": "synthetic",
        "config": {
            "workload": f"gri30 (53 species, 325 reactions) IdealGasConstPressureReactor, {n} reactors x "
                        f"{a.steps} advance steps, per-reactor dt log-uniform in [1e-6, 1e-4] s, "
                        f"T0 U[900,2000] K, CH4/air phi U[0.5,3], P 1-40 atm, rtol 1e-8 atol 1e-15",
            "precision": a.precision,
            "parallelism": "1 GPU, one wavefront per reactor, persistent work-queue kernel",
            "schedule": a.schedule,
        },
        "roofline": {
            "bound": "mfma",
            "pipe": "FP64 (MI355X vector FMA and MFMA share the 78.6 TFLOP/s FP64 peak)",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": measured_traffic(n),
            "kernel": "bdf_advance_kernel",
            "avg_launch_ms": avg_launch_s * 1e3,
            "flops_per_launch": W / a.steps,
        },
        "integrator": {k: float(np.mean(v)) for k, v in stats.items() if k != "order"},
    }
    if not a.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        threads = max(1, min(threads, 16))
        out["cpu_baseline"] = cpu_baseline(wl, a.steps, a.cpu_sample, threads)
    L.cg_dev_free(d_tout)
    if a.copy_calib > 0:
        nb = a.copy_calib << 20
        src, dst = L.cg_dev_alloc(nb), L.cg_dev_alloc(nb)
        check(L.cg_memcpy_d2d(dst, src, nb))
        check(L.cg_device_synchronize())
        L.cg_dev_free(src)
        L.cg_dev_free(dst)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
