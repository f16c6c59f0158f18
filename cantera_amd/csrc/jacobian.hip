// Batched analytic reactor Jacobians d(ydot)/dy of IdealGasConstPressureReactor
// states, the quantity Reactor::finiteDifferenceJacobian approximates
// (src/zeroD/Reactor.cpp:233-276 builds it column by column from eval()).
//
// Input layout as for the batched rates: T[n], P[n], Y[k*n + s] (species-major).
// Output: J[s*N*N + r*N + c], row-major per state, in the reactor's state order
// [m, T, Y_0..Y_{K-1}].  ydot does not depend on the reactor mass (dY/dt and
// dT/dt scale with V/m = 1/rho), so the state is taken with m = 1 kg and the m
// row and column are zero.
//
// One wavefront per state: the Jacobian assembly (two RHS evaluations, at the
// state and with T + dT, then the per-reaction derivatives), in LDS when it
// fits, else in the state's output block.
#include <hip/hip_runtime.h>

#include "bdf.hpp"
#include "wave_jac.hpp"

namespace cg {

namespace {

// kLdsJ: assemble J in LDS (FP64 LDS atomics for the per-reaction terms), then
// write it out with coalesced stores; otherwise assemble it in place in HBM
// (L2 atomics) -- used when N x N doubles do not fit next to the scratch.
// The assembly is inlined here, where the scratch and the matrix are known to
// be LDS: as an out-of-line function taking them through pointer arguments it
// used flat loads, stores and atomics for every access, and its callee-saved
// registers went to scratch memory.
// OT: element type of the output (double, or float for CG_FP32: the FP64
// assembly rounded once at the store, which halves the HBM bytes written).
template <int NS, bool kLdsJ, typename OT>
__global__ __launch_bounds__(64) void jac_batch_kernel(DevMech m, long n, const double* __restrict__ T,
                                                       const double* __restrict__ P, const double* __restrict__ Y,
                                                       OT* __restrict__ J) {
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
    L.W2 = off;  off += 2 * m.K;
    const WaveScratch S = scratch_view(L);
    double* Jl = lds_base() + off;
    const int NN = m.N * m.N;
    const XcdStates xs(n);
    for (long st = xs.lo; st < xs.hi; st += xs.step) {
        const double Ts = T[st];
        const double Ps = uni(P[st]);
        double y[NS], f[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int i = lane + kWave * s;
            double v = 0.0;
            if (i < m.K) v = Y[static_cast<size_t>(i) * n + st];
            else if (i == m.K) v = Ts;
            else if (i == m.K + 1) v = 1.0;
            y[s] = v;
            if (i < m.N) S.yb[i] = v;
        }
        wave_sync();
#pragma unroll
        for (int s = 0; s < NS; s++) f[s] = 0.0;   // not read: the Jacobian takes its own base RHS
        OT* Jg = J + static_cast<size_t>(st) * NN;
        if constexpr (kLdsJ) {
            reactor_jacobian<NS, true, true>(m, S, Ps, y, f, Jl);
            for (int e = lane; e < NN; e += kWave) Jg[e] = static_cast<OT>(Jl[e]);
            wave_sync();
        } else {
            static_assert(sizeof(OT) == sizeof(double), "in-place assembly needs a double output");
            reactor_jacobian<NS, true, true>(m, S, Ps, y, f, Jg);
        }
    }
}

}  // namespace

// J goes to LDS for one state slot per lane (N <= 64); larger systems assemble in HBM
static bool jac_in_lds(int ns) { return ns == 1; }

size_t jac_lds_bytes(const DevMech& m, int ns) {
    size_t d = static_cast<size_t>(kWave) * ns * 4 + 2 * (m.K + 1) + m.K + 2 * m.K;
    if (jac_in_lds(ns)) d += static_cast<size_t>(m.N) * m.N;
    return d * sizeof(double);
}

hipError_t jac_launch(const DevMech& m, int ns, bool f32, long n, const double* T, const double* P, const double* Y,
                      void* J, int grid, size_t lds, hipStream_t stream) {
    if (ns == 1 && f32)
        hipLaunchKernelGGL((jac_batch_kernel<1, true, float>), dim3(grid), dim3(64), lds, stream, m, n, T, P, Y,
                           static_cast<float*>(J));
    else if (ns == 1)
        hipLaunchKernelGGL((jac_batch_kernel<1, true, double>), dim3(grid), dim3(64), lds, stream, m, n, T, P, Y,
                           static_cast<double*>(J));
    else if (ns == 2 && !f32)
        hipLaunchKernelGGL((jac_batch_kernel<2, false, double>), dim3(grid), dim3(64), lds, stream, m, n, T, P, Y,
                           static_cast<double*>(J));
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

const void* jac_kernel_ptr(int ns, bool f32) {
    if (ns == 1) return f32 ? reinterpret_cast<const void*>(&jac_batch_kernel<1, true, float>)
                            : reinterpret_cast<const void*>(&jac_batch_kernel<1, true, double>);
    return reinterpret_cast<const void*>(&jac_batch_kernel<2, false, double>);
}

}  // namespace cg
