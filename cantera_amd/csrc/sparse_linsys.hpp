// Linear-system policy of the mole-reactor integrator: the preconditioner
// M = I - gamma J on the fill pattern of sparse_lu.hpp, factored without
// pivoting in LDS (right-looking: per elimination step, the L column scaled by
// the pivot, then the step's independent updates a_ij -= l_ik u_kj spread
// over the lanes) and applied by the two triangular solves (column-oriented,
// one elimination step per wavefront barrier).  The factors persist in the
// reactor's global slot between calls, as the dense policies keep theirs.
#pragma once
#include "bdf.hpp"
#include "wave_kinetics.hpp"

namespace cg {

// the integrator kernel's one parameter in the kernel-argument segment
using KernargArgsSp = const __attribute__((address_space(4))) BdfArgs;
__device__ __forceinline__ KernargArgsSp* kernarg_bdf_args() {
    return (KernargArgsSp*)__builtin_amdgcn_kernarg_segment_ptr();
}

// GMRES scratch of the mole integrator kept after the factors in LDS
constexpr int kGmMaxl = 5;                      // SPGMR Krylov dimension (SUNDIALS default, no restarts)
// layout: Hessenberg (row i, column l at i * kGmMaxl + l), Givens rotations
// (c, s per column), then 8 counters (lin_iters, prec_solves, jt products,
// linear convergence failures, lin_rhs_evals)
constexpr int kGmH = 0;
constexpr int kGmRot = (kGmMaxl + 1) * kGmMaxl;
constexpr int kGmCnt = kGmRot + 2 * kGmMaxl;
constexpr int kGmDoubles = kGmCnt + 8;

template <int NS>
struct SparseLdsLinSys {
    double* lu;   // nnz values
    double* x;    // kWave * NS: solve workspace, elimination order
    double* gm;   // kGmDoubles: Hessenberg, rotations, least-squares rhs, counters

    __device__ __forceinline__ static int nnz() { return kernarg_bdf_args()->sp.nnz; }
    __device__ __forceinline__ static size_t lds_doubles(int nnz_) {
        return ((static_cast<size_t>(nnz_) + 1) & ~static_cast<size_t>(1)) + kWave * NS + kGmDoubles;
    }
    __device__ __forceinline__ void init(double* base, int) {
        lu = base;
        x = base + ((nnz() + 1) & ~1);
        gm = x + kWave * NS;
    }
    __device__ __forceinline__ void bind(double*, int*) {}
    __device__ __forceinline__ void restore(const double* save, const int*, int) {
        const int nz = nnz();
        for (int p = lane_id(); p < nz; p += kWave) lu[p] = save[p];
        wave_sync();
    }
    __device__ __forceinline__ void save(double* save, int*, int) {
        const int nz = nnz();
        for (int p = lane_id(); p < nz; p += kWave) save[p] = lu[p];
    }
    // M = I - gamma J (AdaptivePreconditioner: m_identity - m_gamma * J)
    __device__ __forceinline__ void form(const double* J, int N, double gamma) {
        KernargArgsSp* a = kernarg_bdf_args();
        const int nz = a->sp.nnz;
        const int lane = lane_id();
        wave_sync();
        for (int p = lane; p < nz; p += kWave) lu[p] = -gamma * J[p];
        wave_sync();
        for (int k = lane; k < N; k += kWave) lu[a->sp.diag[k]] += 1.0;
        wave_sync();
    }
    // false on a zero pivot
    __device__ __forceinline__ bool factor(int N) {
        KernargArgsSp* a = kernarg_bdf_args();
        const int lane = lane_id();
        const int* __restrict__ lptr = a->sp.lptr;
        const int* __restrict__ lslot = a->sp.lslot;
        const int* __restrict__ optr = a->sp.optr;
        const int* __restrict__ ops = a->sp.ops;
        const int* __restrict__ diag = a->sp.diag;
        for (int k = 0; k < N; k++) {
            const int l0 = lptr[k], l1 = lptr[k + 1];
            if (l0 == l1) continue;                // nothing below the pivot
            const double piv = lu[diag[k]];
            if (piv == 0.0) return false;
            for (int t = l0 + lane; t < l1; t += kWave) lu[lslot[t]] = lu[lslot[t]] / piv;
            wave_sync();
            const int o1 = optr[k + 1];
            for (int t = optr[k] + lane; t < o1; t += kWave) {
                const int ij = ops[3 * t], ik = ops[3 * t + 1], kj = ops[3 * t + 2];
                lu[ij] = lu[ij] - lu[ik] * lu[kj];
            }
            wave_sync();
        }
        return lu[diag[N - 1]] != 0.0;
    }
    // b <- M^{-1} b (lane-distributed, internal order)
    __device__ __forceinline__ void solve(double (&b)[NS], int N) {
        KernargArgsSp* a = kernarg_bdf_args();
        const int lane = lane_id();
        const int* __restrict__ iperm = a->sp.iperm;
        const int* __restrict__ lptr = a->sp.lptr;
        const int* __restrict__ lslot = a->sp.lslot;
        const int* __restrict__ lrow = a->sp.lrow;
        const int* __restrict__ uptr = a->sp.uptr;
        const int* __restrict__ uslot = a->sp.uslot;
        const int* __restrict__ urow = a->sp.urow;
        const int* __restrict__ diag = a->sp.diag;
        wave_sync();
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int i = lane + kWave * s;
            if (i < N) x[iperm[i]] = b[s];
        }
        wave_sync();
        for (int k = 0; k < N; k++) {
            const int l0 = lptr[k], l1 = lptr[k + 1];
            if (l0 == l1) continue;
            const double xk = x[k];
            for (int t = l0 + lane; t < l1; t += kWave) x[lrow[t]] -= lu[lslot[t]] * xk;
            wave_sync();
        }
        for (int k = N - 1; k >= 0; k--) {
            const double xk = x[k] / lu[diag[k]];
            wave_sync();
            if (lane == 0) x[k] = xk;
            const int u0 = uptr[k], u1 = uptr[k + 1];
            for (int t = u0 + lane; t < u1; t += kWave) x[urow[t]] -= lu[uslot[t]] * xk;
            wave_sync();
        }
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int i = lane + kWave * s;
            b[s] = (i < N) ? x[iperm[i]] : 0.0;
        }
        wave_sync();
    }
};

}  // namespace cg
