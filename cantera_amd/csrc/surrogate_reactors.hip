// Guarded surrogate steps for reactor batches.
//
// A reactor whose next output time is a whole number of surrogate steps dt
// away takes the steps from the MLP, each checked in FP64 before it is kept:
//   * the state stays physical: T in (200, 6000) K, Y_k >= -1e-9, |sum Y - 1|
//     <= 1e-6 after the step;
//   * elements are conserved: |sum_k a_ek (Y1_k - Y0_k) / W_k| <= 1e-6 of
//     sum_k a_ek Y0_k / W_k for every element e;
//   * the trapezoidal defect r = y1 - y0 - dt/2 (f(y0) + f(y1)) of the
//     reactor ODE (f from the same FP64 kinetics as the integrator) is small:
//     |r_T| <= dT_tol and |r_k| <= rtol_g * max(|Y0_k|, |Y1_k|) + atol_g.
// The first rejected step hands the reactor back to the BDF integrator,
// restarted from the last kept state; a reactor whose steps were all kept is
// restarted at its output time, as CVodeReInit would be after a jump in state.
#include <hip/hip_runtime.h>

#include "bdf.hpp"
#include "wave_kinetics.hpp"

namespace cg {

namespace {

__device__ __forceinline__ int ref_index(int c, int K) { return c < K ? c + 2 : (c == K ? 1 : 0); }

// Ycur[i] (reference order [m, T, Y]) = the output state of reactor list[i]; P too.
__global__ void gather_states_kernel(int n, const int* __restrict__ list, const double* __restrict__ rec,
                                     size_t stride, int NP, int N, int K, double* __restrict__ Y,
                                     double* __restrict__ P) {
    const int i = blockIdx.x;
    if (i >= n) return;
    const double* r = rec + static_cast<size_t>(list[i]) * stride;
    for (int c = threadIdx.x; c < N; c += blockDim.x) Y[static_cast<size_t>(i) * N + ref_index(c, K)] = r[6 * NP + c];
    if (threadIdx.x == 0) P[i] = r[7 * NP + C_P];
}

// One guard per row still in play (done[i] < nsub[i], not rejected).
// f(y0) of the first step is computed here; each kept step leaves f(y1) in
// Fcur (internal order) for the next, so a step costs one RHS evaluation.
// The mechanism is the first parameter: rhs_call reads it through
// kernarg_mech().
template <int NS>
__global__ __launch_bounds__(64) void guard_kernel(DevMech m, int n, const int* __restrict__ nsub,
                                                   int* __restrict__ done, int* __restrict__ rejected,
                                                   double* __restrict__ Ycur, const double* __restrict__ Ynext,
                                                   double* __restrict__ Fcur, const double* __restrict__ Pa,
                                                   double dt, double dT_tol, double rtol_g, double atol_g) {
    const int lane = lane_id();
    ScratchLayout L;
    int off = 0;
    L.yb = off;  off += kWave * NS;
    L.F = off;   off += kWave * NS;
    L.Y0 = off;  off += kWave * NS;
    L.F0 = off;  off += kWave * NS;
    L.C = off;   off += m.K;
    L.MU = off;  off += m.K;
    L.WD = off;  off += m.K;
    L.W2 = off;
    const WaveScratch S = scratch_view(L);
    const int N = m.N, K = m.K;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        if (uni(rejected[i]) || uni(done[i]) >= uni(nsub[i])) continue;
        const double* y0r = Ycur + static_cast<size_t>(i) * N;
        const double* y1r = Ynext + static_cast<size_t>(i) * N;
        const double P = uni(Pa[i]);
        double y0[NS], y1[NS], f0[NS], f1[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int c = lane + kWave * s;
            y0[s] = (c < N) ? y0r[ref_index(c, K)] : 0.0;
            y1[s] = (c < N) ? y1r[ref_index(c, K)] : 0.0;
        }
        // physical range
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
            double d = 0.0, a = 0.0;
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int c = lane + kWave * s;
                if (c < K) {
                    const double w = m.comp[c * m.E + e] * m.rmw[c];
                    d += w * (y1[s] - y0[s]);
                    a += w * y0[s];
                }
            }
            d = uni(wave_sum(d));
            a = uni(wave_sum(a));
            if (!(fabs(d) <= 1e-6 * a + 1e-300)) bad = 1;
        }
        bad = __any(bad);
        if (!bad) {
            // trapezoidal defect with the FP64 right-hand side
            double* fr = Fcur + static_cast<size_t>(i) * N;
            if (uni(done[i]) == 0) {
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
            } else {
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    const int c = lane + kWave * s;
                    f0[s] = (c < N) ? fr[c] : 0.0;
                }
            }
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
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int c = lane + kWave * s;
                if (c <= K) {
                    const double r = y1[s] - y0[s] - 0.5 * dt * (f0[s] + f1[s]);
                    const double tol = (c == K) ? dT_tol : rtol_g * fmax(fabs(y0[s]), fabs(y1[s])) + atol_g;
                    if (!(fabs(r) <= tol)) bad = 1;
                }
            }
            bad = __any(bad);
        }
        if (bad) {
            if (lane == 0) rejected[i] = 1;
        } else {
            double* yc = Ycur + static_cast<size_t>(i) * N;
            for (int c = lane; c < N; c += kWave) yc[c] = y1r[c];
            double* fr = Fcur + static_cast<size_t>(i) * N;
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int c = lane + kWave * s;
                if (c < N) fr[c] = f1[s];
            }
            if (lane == 0) done[i] += 1;
        }
    }
}

// Restart reactor list[i] at state Ycur[i] and time t0[i] + done[i] * dt
// (rows with done[i] == 0 are left untouched: their history is still valid).
__global__ void restart_kernel(int n, const int* __restrict__ list, const int* __restrict__ done,
                               const double* __restrict__ t0, double dt, const double* __restrict__ Y, double* rec,
                               size_t stride, int NP, int N, int K) {
    const int i = blockIdx.x;
    if (i >= n || done[i] == 0) return;
    double* r = rec + static_cast<size_t>(list[i]) * stride;
    const double t = t0[i] + done[i] * dt;
    for (int c = threadIdx.x; c < N; c += blockDim.x) {
        const double v = Y[static_cast<size_t>(i) * N + ref_index(c, K)];
        r[c] = v;                       // zn[0]
        for (int j = 1; j <= 5; j++) r[j * NP + c] = 0.0;
        r[6 * NP + c] = v;              // output state
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double* x = r + 7 * NP;
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

}  // namespace

size_t guard_lds_bytes(const DevMech& m, int ns) {
    const size_t d = static_cast<size_t>(kWave) * ns * 4 + 2 * m.K + m.K + 2 * m.K;
    return d * sizeof(double);
}

const void* guard_kernel_ptr(int ns) {
    return ns == 1 ? reinterpret_cast<const void*>(&guard_kernel<1>) : reinterpret_cast<const void*>(&guard_kernel<2>);
}

hipError_t surrogate_gather(int n, const int* list, const double* rec, size_t stride, int NP, int N, int K, double* Y,
                            double* P, hipStream_t s) {
    hipLaunchKernelGGL(gather_states_kernel, dim3(n), dim3(64), 0, s, n, list, rec, stride, NP, N, K, Y, P);
    return hipGetLastError();
}

hipError_t surrogate_guard(const DevMech& m, int ns, int n, const int* nsub, int* done, int* rejected, double* Ycur,
                           const double* Ynext, double* Fcur, const double* Pa, double dt, double dT_tol,
                           double rtol_g, double atol_g, int grid, size_t lds, hipStream_t s) {
    if (ns == 1)
        hipLaunchKernelGGL((guard_kernel<1>), dim3(grid), dim3(64), lds, s, m, n, nsub, done, rejected, Ycur, Ynext,
                           Fcur, Pa, dt, dT_tol, rtol_g, atol_g);
    else
        hipLaunchKernelGGL((guard_kernel<2>), dim3(grid), dim3(64), lds, s, m, n, nsub, done, rejected, Ycur, Ynext,
                           Fcur, Pa, dt, dT_tol, rtol_g, atol_g);
    return hipGetLastError();
}

hipError_t surrogate_restart(int n, const int* list, const int* done, const double* t0, double dt, const double* Y,
                             double* rec, size_t stride, int NP, int N, int K, hipStream_t s) {
    hipLaunchKernelGGL(restart_kernel, dim3(n), dim3(64), 0, s, n, list, done, t0, dt, Y, rec, stride, NP, N, K);
    return hipGetLastError();
}

}  // namespace cg
