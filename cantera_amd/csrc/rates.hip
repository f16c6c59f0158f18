// Batched net production rates (Kinetics::getNetProductionRates,
// src/kinetics/Kinetics.cpp:447-457) over many independent states.
//
// Layout: species-major SoA in HBM, Y[k*n + s] and wdot[k*n + s].  One
// wavefront evaluates one state at a time (grid-stride over states): lane k
// loads Y_k, the wavefront runs the species, reaction and production passes
// with its scratch in LDS, and lane k stores wdot_k.  Consecutive states go to
// neighbouring wavefronts, so the 64-byte lines of Y and wdot (8 states of one
// species) are shared through L2.  The LDS scratch is small (a few KB), so many
// wavefronts per CU hide the latency of the table loads and transcendentals.
// kF32: the rate constants' transcendentals in FP32 (see eval_reaction).
#include <hip/hip_runtime.h>

#include "bdf.hpp"
#include "wave_kinetics.hpp"

namespace cg {

namespace {

// At most 168 VGPRs, so that three wavefronts share a SIMD (latency-bound
// kernel; without the bound the allocator settles at 169 for NS = 1).
template <int NS, bool kF32>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void rates_kernel(
    DevMech m, long n, const double* __restrict__ T, const double* __restrict__ P, const double* __restrict__ Y,
    double* __restrict__ wdot) {
    extern __shared__ double lds_raw[];
    const int lane = lane_id();
    double* base = lds_raw;
    WaveScratch S{};
    S.yb = base;  base += kWave * NS;
    S.C = base;   base += m.K + 1;
    S.MU = base;  base += m.K + 1;
    S.WD = base;
    const XcdStates xs(n);
    for (long st = xs.lo; st < xs.hi; st += xs.step) {
        const double Ts = uni(T[st]);
        const double Ps = uni(P[st]);
        double yk[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int k = lane + kWave * s;
            yk[s] = (k < m.K) ? Y[static_cast<size_t>(k) * n + st] : 0.0;
        }
        GasState g;
        SpeciesRegs<NS> sr;
        species_pass<NS, true>(m, S, Ts, Ps, yk, g, sr);
        production_rates<NS, kF32>(m, S, g, sr);
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int k = lane + kWave * s;
            if (k < m.K) wdot[static_cast<size_t>(k) * n + st] = sr.wdot[s];
        }
        wave_sync();
    }
}

// Net production rates at the output states of a reactor batch (one record
// per reactor: the interpolated state [Y, T, m] after the 6 Nordsieck columns,
// then the control block with P, or V for a constant-volume reactor).
template <int NS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void reactor_rates_kernel(
    DevMech m, long n, const double* __restrict__ rec, size_t stride, int cv, double* __restrict__ wdot) {
    extern __shared__ double lds_raw[];
    const int lane = lane_id();
    double* base = lds_raw;
    WaveScratch S{};
    S.yb = base;  base += kWave * NS;
    S.C = base;   base += m.K + 1;
    S.MU = base;  base += m.K + 1;
    S.WD = base;
    constexpr int NP = kWave * NS;
    const XcdStates xs(n);
    for (long st = xs.lo; st < xs.hi; st += xs.step) {
        const double* rr = rec + static_cast<size_t>(st) * stride;
        const double Ts = uni(rr[6 * NP + m.K]);
        const double mass = uni(rr[6 * NP + m.K + 1]);
        const double pv = uni(rr[7 * NP + C_P]);
        double yk[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int k = lane + kWave * s;
            yk[s] = (k < m.K) ? rr[6 * NP + k] : 0.0;
        }
        GasState g;
        SpeciesRegs<NS> sr;
        species_pass<NS, true>(m, S, Ts, cv ? 0.0 : pv, yk, g, sr, cv ? mass / pv : 0.0);
        production_rates<NS, false>(m, S, g, sr);
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int k = lane + kWave * s;
            if (k < m.K) wdot[static_cast<size_t>(k) * n + st] = sr.wdot[s];
        }
        wave_sync();
    }
}

}  // namespace

hipError_t reactor_rates_launch(const DevMech& m, int ns, long n, const double* rec, size_t stride, int cv,
                                double* wdot, int grid, size_t lds, hipStream_t stream) {
    if (ns == 1)
        hipLaunchKernelGGL((reactor_rates_kernel<1>), dim3(grid), dim3(64), lds, stream, m, n, rec, stride, cv, wdot);
    else if (ns == 2)
        hipLaunchKernelGGL((reactor_rates_kernel<2>), dim3(grid), dim3(64), lds, stream, m, n, rec, stride, cv, wdot);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

// state slots per lane of the rate kernels: 1, 2, or 4 for 128 < K + 2 <= 256
static int rates_ns(int ns) { return ns <= 2 ? ns : 4; }

const void* reactor_rates_kernel_ptr(int ns) {
    return ns == 1 ? reinterpret_cast<const void*>(&reactor_rates_kernel<1>)
                   : reinterpret_cast<const void*>(&reactor_rates_kernel<2>);
}

size_t rates_lds_bytes(const DevMech& m, int ns) {
    const size_t d = static_cast<size_t>(kWave) * rates_ns(ns) + 2 * (m.K + 1) + m.K;
    return d * sizeof(double);
}

hipError_t rates_launch(const DevMech& m, int ns, bool f32, long n, const double* T, const double* P,
                        const double* Y, double* wdot, int grid, size_t lds, hipStream_t stream) {
    if (ns == 1 && !f32)
        hipLaunchKernelGGL((rates_kernel<1, false>), dim3(grid), dim3(64), lds, stream, m, n, T, P, Y, wdot);
    else if (ns == 1)
        hipLaunchKernelGGL((rates_kernel<1, true>), dim3(grid), dim3(64), lds, stream, m, n, T, P, Y, wdot);
    else if (ns == 2 && !f32)
        hipLaunchKernelGGL((rates_kernel<2, false>), dim3(grid), dim3(64), lds, stream, m, n, T, P, Y, wdot);
    else if (ns == 2)
        hipLaunchKernelGGL((rates_kernel<2, true>), dim3(grid), dim3(64), lds, stream, m, n, T, P, Y, wdot);
    else if (rates_ns(ns) == 4 && !f32)
        hipLaunchKernelGGL((rates_kernel<4, false>), dim3(grid), dim3(64), lds, stream, m, n, T, P, Y, wdot);
    else if (rates_ns(ns) == 4)
        hipLaunchKernelGGL((rates_kernel<4, true>), dim3(grid), dim3(64), lds, stream, m, n, T, P, Y, wdot);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

const void* rates_kernel_ptr(int ns, bool f32) {
    if (ns == 1) return f32 ? reinterpret_cast<const void*>(&rates_kernel<1, true>)
                            : reinterpret_cast<const void*>(&rates_kernel<1, false>);
    if (ns == 2) return f32 ? reinterpret_cast<const void*>(&rates_kernel<2, true>)
                            : reinterpret_cast<const void*>(&rates_kernel<2, false>);
    return f32 ? reinterpret_cast<const void*>(&rates_kernel<4, true>)
               : reinterpret_cast<const void*>(&rates_kernel<4, false>);
}

}  // namespace cg
