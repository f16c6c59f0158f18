// Guarded surrogate steps for reactor batches, entirely on the device.
//
// One call advances every reactor whose output time is exactly one surrogate
// step dt ahead.  Per call:
//   1. gather_kernel: each reactor's output state (reference order [m, T, Y])
//      and whether it is eligible (status >= 0, t_out - t = dt); reactors
//      with any other positive span go straight to the integrator's list;
//   2. the network predicts y1 for every row (GEMMs on the matrix cores);
//   3. guard_kernel checks each eligible prediction in FP64 and either keeps
//      it (the reactor restarts from y1 at t + dt, as CVodeReInit would after
//      a jump in state) or marks the reactor for the integrator;
//   4. the BDF kernel advances the marked reactors (BdfArgs::work_flag) in
//      the batch's longest-first order, skipping the others.
// No host synchronisation and no allocation happen per call.
//
// The guard (all FP64, the integrator's own right-hand side f):
//   * physical state: T in (200, 6000) K, Y_k >= -1e-9, |sum Y - 1| <= 1e-6,
//     elements conserved to 1e-6 relative;
//   * regime gate: T0 > gate_T and the radical pool (H, O, OH) is being
//     destroyed, sum_k f_k(y0) / W_k < 0.  Before this point (induction) a
//     perturbation of the radical pool shifts the ignition time, and the
//     shift is amplified by the ignition's steep rise; past it perturbations
//     relax;
//   * step error estimate from the trapezoidal defect
//     r = y1 - y0 - dt/2 (f(y0) + f(y1)): |r_T| <= tol_T for the temperature,
//     and for species |r_k| / (1 + dt/2 lambda_k) <= tol_Y, where lambda_k is
//     the species' loss frequency at y1 (the diagonal of I - dt/2 J for a
//     species consumed at first order), so that the defect of a fast species
//     is read as the error it implies rather than as that error amplified by
//     dt lambda_k.
#include <hip/hip_runtime.h>

#include <cstddef>

#include "bdf.hpp"
#include "wave_kinetics.hpp"

namespace cg {

namespace {

__device__ __forceinline__ int ref_index(int c, int K) { return c < K ? c + 2 : (c == K ? 1 : 0); }

// Per reactor (one 64-thread block): Y row (reference order) of its output
// state, P, and elig = 1 (one surrogate step ahead), 2 (integrate), 0 (nothing
// to do).
__global__ void gather_kernel(int n, const double* __restrict__ rec, size_t stride, int NP, int N, int K,
                              const double* __restrict__ tout, double dt, double* __restrict__ Y,
                              double* __restrict__ P, int* __restrict__ elig,
                              unsigned long long* __restrict__ counters) {
    const int r = blockIdx.x;
    if (r >= n) return;
    const double* rr = rec + static_cast<size_t>(r) * stride;
    for (int c = threadIdx.x; c < N; c += blockDim.x) Y[static_cast<size_t>(r) * N + ref_index(c, K)] = rr[6 * NP + c];
    if (threadIdx.x == 0) {
        const double* x = rr + 7 * NP;
        P[r] = x[C_P];
        const double span = tout[r] - x[C_TUSER];
        int e = 0;
        if (x[C_STATUS] >= 0.0 && span > 0.0) e = (fabs(span - dt) <= 1e-9 * dt) ? 1 : 2;
        elig[r] = e;
        if (e == 2) atomicAdd(&counters[2], 1ull);
        if (e == 1) atomicAdd(&counters[0], 1ull);
    }
}

// Record of reactor rr restarted at state y (internal order, lane-distributed)
// and time t: history reset to order 1, next step from the initial-step rule.
template <int NS>
__device__ __forceinline__ void restart_record(double* rr, int NP, int N, const double (&y)[NS], double t) {
    const int lane = lane_id();
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int c = lane + kWave * s;
        if (c < N) {
            rr[c] = y[s];                               // zn[0]
            for (int j = 1; j <= 5; j++) rr[j * NP + c] = 0.0;
            rr[6 * NP + c] = y[s];                      // output state
        }
    }
    if (lane == 0) {
        double* x = rr + 7 * NP;
        x[C_TN] = t;
        x[C_TUSER] = t;
        x[C_H] = 0.0; x[C_HPRIME] = 0.0; x[C_HSCALE] = 0.0; x[C_HU] = 0.0;
        x[C_ETA] = 1.0; x[C_ETAMAX] = 10000.0;
        x[C_Q] = 1; x[C_QPRIME] = 1; x[C_QWAIT] = 2; x[C_L] = 2;
        x[C_STARTED] = 0;
        x[C_NSTLJ] = -1.0e9;            // forces a fresh Jacobian at the first setup
        x[C_NSTLP] = -1.0e9;
        x[C_HAVE_LU] = 0.0;
        x[C_NSCON] = 0;
        for (int k = 0; k < 7; k++) x[C_TAU + k] = 0.0;
        for (int k = 0; k < 6; k++) x[C_TQ + k] = 0.0;
        x[C_GAMMA] = 0.0; x[C_GAMMAP] = 0.0; x[C_GAMRAT] = 1.0; x[C_CRATE] = 1.0; x[C_DELP] = 0.0;
        x[C_ACNRM] = 0.0; x[C_RL1] = 0.0; x[C_SAVED_TQ5] = 0.0;
    }
}

struct GuardArgs {
    DevMech m;                      // first: rhs_call reads it through kernarg_mech()
    int n;
    int* elig;
    const double* Ycur;             // n x N, reference order
    const double* Ynext;            // n x N, reference order (prediction)
    double* Fcur;                   // n x N, internal order: f(y0) when fvalid
    int* fvalid;
    const double* P;
    double* rec;
    size_t stride;
    int NP;
    const double* tout;
    double dt, tol_T, tol_Y, gate_T;
    int rad[3];                     // H, O, OH (-1: absent)
    unsigned long long* counters;   // [eligible, kept, integrated]
};

static_assert(offsetof(GuardArgs, m) == 0, "GuardArgs must begin with the mechanism");

template <int NS>
__global__ __launch_bounds__(64) void guard_kernel(GuardArgs a) {
    const DevMech& m = a.m;
    const int lane = lane_id();
    ScratchLayout L;
    int off = 0;
    L.yb = off;  off += kWave * NS;
    L.F = off;   off += kWave * NS;
    L.Y0 = off;  off += kWave * NS;
    L.F0 = off;  off += kWave * NS;
    L.C = off;   off += m.K + 1;
    L.MU = off;  off += m.K + 1;
    L.WD = off;  off += m.K;
    L.W2 = off;
    const WaveScratch S = scratch_view(L);
    const int N = m.N, K = m.K;
    for (int i = blockIdx.x; i < a.n; i += gridDim.x) {
        if (uni(a.elig[i]) != 1) continue;
        const double* y0r = a.Ycur + static_cast<size_t>(i) * N;
        const double* y1r = a.Ynext + static_cast<size_t>(i) * N;
        const double P = uni(a.P[i]);
        double y0[NS], y1[NS], f0[NS], f1[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int c = lane + kWave * s;
            y0[s] = (c < N) ? y0r[ref_index(c, K)] : 0.0;
            y1[s] = (c < N) ? y1r[ref_index(c, K)] : 0.0;
        }
        // physical range, sum of Y
        int bad = 0;
        double sum1 = 0.0;
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int c = lane + kWave * s;
            if (c < K) {
                if (!(y1[s] >= -1e-9)) bad = 1;
                sum1 += y1[s];
            } else if (c == K) {
                if (!(y1[s] > 200.0 && y1[s] < 6000.0)) bad = 1;
            }
        }
        sum1 = uni(wave_sum(sum1));
        if (!(fabs(sum1 - 1.0) <= 1e-6)) bad = 1;
        // element conservation
        for (int e = 0; e < m.E; e++) {
            double d = 0.0, t = 0.0;
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int c = lane + kWave * s;
                if (c < K) {
                    const double w = m.comp[c * m.E + e] * m.rmw[c];
                    d += w * (y1[s] - y0[s]);
                    t += w * y0[s];
                }
            }
            d = uni(wave_sum(d));
            t = uni(wave_sum(t));
            if (!(fabs(d) <= 1e-6 * t + 1e-300)) bad = 1;
        }
        // regime gate, first half: hot enough (no RHS needed to tell)
        double tl = 0.0;
#pragma unroll
        for (int s = 0; s < NS; s++)
            if (lane + kWave * s == K) tl = y0[s];
        const double T0 = uni(wave_sum(tl));
        if (!(T0 > a.gate_T) || a.rad[0] < 0) bad = 1;
        bad = __any(bad);
        double* fr = a.Fcur + static_cast<size_t>(i) * N;
        if (!bad) {
            // f(y0): kept from the previous kept step, else evaluated
            if (uni(a.fvalid[i])) {
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    const int c = lane + kWave * s;
                    f0[s] = (c < N) ? fr[c] : 0.0;
                }
            } else {
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    const int c = lane + kWave * s;
                    if (c < N) S.yb[c] = y0[s];
                }
                wave_sync();
                rhs_call<NS>(kernarg_mech(), L, P);
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    const int c = lane + kWave * s;
                    f0[s] = (c < N) ? S.F[c] : 0.0;
                }
                wave_sync();
            }
            // regime gate, second half: the radical pool is being destroyed
            double rsum = 0.0;
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int c = lane + kWave * s;
                if (c < K && (c == a.rad[0] || c == a.rad[1] || c == a.rad[2])) rsum += f0[s] * m.rmw[c];
            }
            rsum = uni(wave_sum(rsum));
            if (!(rsum < 0.0)) bad = 1;
        }
        bad = __any(bad);
        if (!bad) {
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int c = lane + kWave * s;
                if (c < N) S.yb[c] = y1[s];
            }
            wave_sync();
            rhs_call<NS>(kernarg_mech(), L, P);
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int c = lane + kWave * s;
                f1[s] = (c < N) ? S.F[c] : 0.0;
            }
            wave_sync();
            loss_frequency_call<NS>(kernarg_mech(), L, P);
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int c = lane + kWave * s;
                if (c <= K) {
                    const double r = fabs(y1[s] - y0[s] - 0.5 * a.dt * (f0[s] + f1[s]));
                    const double err = (c == K) ? r : r / (1.0 + 0.5 * a.dt * S.W2[c]);
                    if (!(err <= ((c == K) ? a.tol_T : a.tol_Y))) bad = 1;
                }
            }
            wave_sync();
            bad = __any(bad);
        }
        double* rr = a.rec + static_cast<size_t>(i) * a.stride;
        if (bad) {
            if (lane == 0) {
                a.fvalid[i] = 0;
                a.elig[i] = 2;
                atomicAdd(&a.counters[2], 1ull);
            }
        } else {
            restart_record<NS>(rr, a.NP, N, y1, uni(a.tout[i]));
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int c = lane + kWave * s;
                if (c < N) fr[c] = f1[s];
            }
            if (lane == 0) {
                a.fvalid[i] = 1;
                atomicAdd(&a.counters[1], 1ull);
            }
        }
    }
}

// Output states (reference order [m, T, Y]) of all reactors into Y (n x N).
__global__ void export_kernel(int n, const double* __restrict__ rec, size_t stride, int NP, int N, int K,
                              double* __restrict__ Y) {
    const int r = blockIdx.x;
    if (r >= n) return;
    const double* rr = rec + static_cast<size_t>(r) * stride;
    for (int c = threadIdx.x; c < N; c += blockDim.x) Y[static_cast<size_t>(r) * N + ref_index(c, K)] = rr[6 * NP + c];
}

__global__ void clear_ints_kernel(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace

size_t guard_lds_bytes(const DevMech& m, int ns) {
    const size_t d = static_cast<size_t>(kWave) * ns * 4 + 2 * (m.K + 1) + m.K + 2 * m.K;
    return d * sizeof(double);
}

const void* guard_kernel_ptr(int ns) {
    return ns == 1 ? reinterpret_cast<const void*>(&guard_kernel<1>) : reinterpret_cast<const void*>(&guard_kernel<2>);
}

hipError_t export_states(int n, const double* rec, size_t stride, int NP, int N, int K, double* Y, hipStream_t s) {
    hipLaunchKernelGGL(export_kernel, dim3(n), dim3(64), 0, s, n, rec, stride, NP, N, K, Y);
    return hipGetLastError();
}

hipError_t surrogate_clear(int* p, int n, hipStream_t s) {
    hipLaunchKernelGGL(clear_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t surrogate_gather(int n, const double* rec, size_t stride, int NP, int N, int K, const double* tout,
                            double dt, double* Y, double* P, int* elig, unsigned long long* counters, hipStream_t s) {
    hipLaunchKernelGGL(gather_kernel, dim3(n), dim3(64), 0, s, n, rec, stride, NP, N, K, tout, dt, Y, P, elig,
                       counters);
    return hipGetLastError();
}

hipError_t surrogate_guard(const DevMech& m, int ns, int n, int* elig, const double* Ycur, const double* Ynext,
                           double* Fcur, int* fvalid, const double* P, double* rec, size_t stride, int NP,
                           const double* tout, double dt, double tol_T, double tol_Y, double gate_T, const int* rad,
                           unsigned long long* counters, int grid, size_t lds, hipStream_t s) {
    GuardArgs a{};
    a.m = m;
    a.n = n;
    a.elig = elig;
    a.Ycur = Ycur;
    a.Ynext = Ynext;
    a.Fcur = Fcur;
    a.fvalid = fvalid;
    a.P = P;
    a.rec = rec;
    a.stride = stride;
    a.NP = NP;
    a.tout = tout;
    a.dt = dt;
    a.tol_T = tol_T;
    a.tol_Y = tol_Y;
    a.gate_T = gate_T;
    for (int j = 0; j < 3; j++) a.rad[j] = rad[j];
    a.counters = counters;
    if (ns == 1) hipLaunchKernelGGL((guard_kernel<1>), dim3(grid), dim3(64), lds, s, a);
    else hipLaunchKernelGGL((guard_kernel<2>), dim3(grid), dim3(64), lds, s, a);
    return hipGetLastError();
}

}  // namespace cg
