// IdealGasConstPressureMoleReactor on one wavefront: right-hand side and the
// preconditioner Jacobian of the mole-based state, internal order
// [n_0..n_{K-1}, T] (kmol, K) -- the reference's [T, n_1..n_K] with T moved
// last so that lane k owns species k as in the mass-fraction kernels.  With
// m.const_volume the constant-volume IdealGasMoleReactor, [n_0..n_{K-1}, T, V]
// (the reference's [T, V, n_1..n_K]).
#pragma once
#include "bdf.hpp"
#include "wave_jac.hpp"

namespace cg {

// IdealGasConstPressureMoleReactor::updateState / eval
// (src/zeroD/IdealGasConstPressureMoleReactor.cpp:38-104) over the left-hand
// side: mass sum_k n_k W_k, Y_k = n_k W_k / mass (Phase::setMolesNoTruncate,
// Phase.cpp:550-568), V = mass / rho(T, P); dn_k/dt = wdot_k V,
// dT/dt = -sum_k wdot_k h_k V / (mass c_p).  Constant volume
// (IdealGasMoleReactor::updateState / eval, IdealGasMoleReactor.cpp:97-166):
// rho = mass / V, u_k = h_k - R T and c_v in the energy equation, dV/dt = 0.
template <int NS>
__device__ __forceinline__ void mole_rhs(const DevMech& m, const WaveScratch& S, double P, const double (&y)[NS],
                                         double (&f)[NS], GasState& g, SpeciesRegs<NS>& sr, double& vol,
                                         double& totcp) {
    const int lane = lane_id();
    const int K = m.K;
    const bool cv = m.const_volume;
    const int N = K + 1 + (cv ? 1 : 0);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + kWave * s;
        if (i < N) S.yb[i] = y[s];
    }
    wave_sync();
    const double T = uni(S.yb[K]);
    const double V = cv ? uni(S.yb[K + 1]) : 0.0;
    double ms = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int k = lane + kWave * s;
        if (k < K) ms += y[s] * m.mw[k];
    }
    const double mass = uni(wave_sum(ms));
    double yk[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int k = lane + kWave * s;
        yk[s] = (k < K) ? y[s] * m.mw[k] / mass : 0.0;
    }
    species_pass<NS>(m, S, T, P, yk, g, sr, cv ? mass / V : 0.0);
    production_rates<NS>(m, S, g, sr);
    vol = cv ? V : mass / g.dens;
    const double hshift = cv ? 1.0 : 0.0;   // u_k / RT = h_k / RT - 1
    double e = 0.0, inner = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int k = lane + kWave * s;
        if (k < K) {
            e += sr.wdot[s] * ((sr.hrt[s] - hshift) * g.RT) * vol;
            inner += sr.ym[s] * sr.cpr[s];
        }
    }
    e = uni(wave_sum(e));
    inner = uni(wave_sum(inner));
    // m c_p, or m c_v
    totcp = mass * (kGasConstant * (g.mmw * inner - hshift) / g.mmw);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + kWave * s;
        f[s] = (i < K) ? sr.wdot[s] * vol : ((i == K) ? -e / totcp : 0.0);
    }
}

// Out-of-line mole RHS: state from S.yb, ydot to S.F (as rhs_call).
template <int NS>
__device__ __noinline__ void mole_rhs_call(KernargMech* mech, ScratchLayout L, double P) {
    const DevMech m = *(const DevMech*)mech;
    const WaveScratch S = scratch_view(L);
    const int lane = lane_id();
    const int N = m.K + 1 + (m.const_volume ? 1 : 0);
    double y[NS], f[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + kWave * s;
        y[s] = (i < N) ? S.yb[i] : 0.0;
    }
    GasState g;
    SpeciesRegs<NS> sr;
    double vol, totcp;
    mole_rhs<NS>(m, S, P, y, f, g, sr, vol, totcp);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + kWave * s;
        if (i < N) S.F[i] = f[s];
    }
    wave_sync();
}

// IdealGasConstPressureMoleReactor::getJacobianElements
// (IdealGasConstPressureMoleReactor.cpp:108-174) at the state in S.Y0 (the
// caller clips it at 0, AdaptivePreconditioner::stateAdjustment): species
// block d wdot/dC (netProductionRates_ddCi without the -n_i/(V N) dC_i/dn_j
// terms), T column by a one-sided difference with increment sqrt(eps) T, T row
// (c_p,j qdot - (m c_p) sum_k h_k J_kj) / (m c_p)^2 -- at constant volume
// with c_v,j, u_k and m c_v, and no V row or column
// (IdealGasMoleReactor::getJacobianElements, IdealGasMoleReactor.cpp:179-262).  The species block is
// accumulated densely in Jd (K x K, global) with atomics, the default-
// efficiency part of d[M]/dC per row in LDS (S.WD); the entries are then
// written into the value slots of the preconditioner's fill pattern (Js,
// zero at fill-only slots).
template <int NS>
__device__ __noinline__ void mole_jac_call(KernargMech* mech, ScratchLayout L, double P, double* __restrict__ Jd,
                                           double* __restrict__ Js, const SparseTables* sp) {
    const DevMech m = *(const DevMech*)mech;
    const WaveScratch S = scratch_view(L);
    const int lane = lane_id();
    const int K = m.K, N = K + 1 + (m.const_volume ? 1 : 0);
    const double hshift = m.const_volume ? 1.0 : 0.0;
    const int skip = sp->skip;
    double y[NS], y1[NS], f0[NS], f1[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + kWave * s;
        y[s] = (i < N) ? S.Y0[i] : 0.0;
        y1[s] = y[s];
    }
    GasState g;
    SpeciesRegs<NS> sr;
    double vol, totcp;
    const double Tval = bcast<NS>(y, K);
    const double dT = kSqrtEps * Tval;
#pragma unroll
    for (int s = 0; s < NS; s++)
        if (lane + kWave * s == K) y1[s] = Tval + dT;
    mole_rhs<NS>(m, S, P, y1, f1, g, sr, vol, totcp);
    mole_rhs<NS>(m, S, P, y, f0, g, sr, vol, totcp);   // base state left in LDS and g
    for (int i = lane; i < K * K; i += kWave) Jd[i] = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int k = lane + kWave * s;
        if (k < K) S.WD[k] = 0.0;
    }
    __threadfence();
    wave_sync();
    const double kpd = pdep_rate_lane(m, g);
    double* const acc = (skip & 1) ? nullptr : S.WD;
    for (int p = lane; p < m.I; p += kWave) scatter_reaction_derivs(m, p, g, S.C, S.MU, kpd, Jd, K, acc, skip);
    __threadfence();
    wave_sync();
    // h_k into MU (the scatter is done with the chemical potentials)
    double qd = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int k = lane + kWave * s;
        if (k < K) {
            const double hk = (sr.hrt[s] - hshift) * g.RT;
            S.MU[k] = hk;
            qd += hk * (sr.wdot[s] * vol);
        }
    }
    const double qdot = uni(wave_sum(qd));
    wave_sync();
    const double denom = 1.0 / (totcp * totcp);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int j = lane + kWave * s;
        if (j < K) {
            double hs = 0.0;
            for (int k = 0; k < K; k++) hs += S.MU[k] * (Jd[k * K + j] + (acc ? S.WD[k] : 0.0));
            S.W2[j] = (kGasConstant * (sr.cpr[s] - hshift) * qdot - totcp * hs) * denom;
        }
        if (j < N) S.F[j] = (f1[s] - f0[s]) / dT;
    }
    wave_sync();
    for (int p = lane; p < sp->nnz; p += kWave) {
        const int r = sp->srow[p], c = sp->scol[p];
        double v = 0.0;                       // V row and column, fill-only slots
        if (c == K) v = S.F[r];
        else if (c < K && r == K) v = S.W2[c];
        else if (c < K && r < K) v = Jd[r * K + c] + (acc ? S.WD[r] : 0.0);
        Js[p] = v;
    }
    __threadfence();
    wave_sync();
}

}  // namespace cg
