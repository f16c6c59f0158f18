// Analytic reactor Jacobian d(ydot)/dy for one reactor, assembled by one
// wavefront into an FP64 matrix in global memory (row-major, internal state
// order [Y_0..Y_{K-1}, T, m]).
//
// Y columns: exact chain rule through C_j = rho Y_j / W_j at fixed T -- at
// constant pressure the molar density P/RT is fixed; at constant volume rho is
// fixed and the total concentration, hence the default-efficiency part of [M],
// varies -- with dwdot/dC from mass action, third-body and
// falloff [M] dependence -- the structure of BulkKinetics::netProductionRates_ddCi
// (src/kinetics/BulkKinetics.cpp:761-811) and ThirdBodyCalc::derivatives
// (include/cantera/kinetics/ThirdBodyCalc.h:97-100), at fixed rate constants.
// T column: one-sided difference with increment sqrt(eps) T, as the reference's
// mole-reactor Jacobian does for temperature
// (src/zeroD/IdealGasConstPressureMoleReactor.cpp:126-152).
// Per-reaction contributions are scattered with FP64 global atomics (L2).
#pragma once
#include "wave_kinetics.hpp"

namespace cg {

constexpr double kSqrtEps = 1.4901161193847656e-08;  // sqrt(DBL_EPSILON)

__device__ __forceinline__ void mat_add(double* p, double v) { unsafeAtomicAdd(p, v); }

// d ropnet / d C_j contributions of reaction p, scattered into JC (K x K block
// of A).  All per-reaction lists use static indices (registers only).
// skip (derivative settings of BulkKinetics, BulkKinetics.cpp:266-281):
// bit 0 leaves out the third-body [M] terms, bit 1 the falloff d k/d[M]
// (the rate then counts as linear in [M], BulkKinetics.cpp:761-792).
__device__ __forceinline__ void scatter_reaction_derivs(const DevMech& m, int p, const GasState& g,
                                                        const double* __restrict__ C, const double* __restrict__ MU,
                                                        double kpd, double* A, int LD, double* acc, int skip = 0) {
    RxnRegs x;
    load_rxn(m, p, x);
    RxnVals v;
    eval_reaction(m, p, g, C, MU, x, kpd, v);
    const int type = (x.flags >> 1) & 3;
    const bool rev = x.flags & 1;
    const bool rgen = x.flags & 8, pgen = x.flags & 16;
    const double* gs = (x.gen >= 0) ? m.gend + 12 * x.gen : m.gend;
    const double kfM = (type == kThreeBody) ? v.kf * v.M : v.kf;
    constexpr int E = 2 * kMaxSide;
    // entries 0..3 reactants, 4..7 products
    int ek[E];
    double est[E];
    bool eval_[E];
#pragma unroll
    for (int a = 0; a < kMaxSide; a++) {
        eval_[a] = a < x.nr;
        ek[a] = x.rk[a];
        est[a] = rgen ? gs[4 + a] : 1.0;
        eval_[kMaxSide + a] = a < x.np;
        ek[kMaxSide + a] = x.pk[a];
        est[kMaxSide + a] = pgen ? gs[8 + a] : 1.0;
    }
    // net stoichiometric coefficient of each distinct species (first entry leads)
    double rnu[E];
    bool lead[E];
#pragma unroll
    for (int a = 0; a < E; a++) {
        double nu = 0.0;
        bool first = eval_[a];
#pragma unroll
        for (int b = 0; b < E; b++) {
            if (eval_[b] && ek[b] == ek[a]) {
                nu += (b < kMaxSide) ? -est[b] : est[b];
                if (b < a) first = false;
            }
        }
        rnu[a] = nu;
        lead[a] = first && nu != 0.0;
    }
    double ordr[kMaxSide], ordp[kMaxSide];
#pragma unroll
    for (int a = 0; a < kMaxSide; a++) {
        ordr[a] = rgen ? gs[a] : 1.0;
        ordp[a] = pgen ? gs[8 + a] : 1.0;
    }
    double pr_r = 1.0, pr_p = 1.0;
#pragma unroll
    for (int a = 0; a < kMaxSide; a++) {
        if (a < x.nr) pr_r *= (ordr[a] == 1.0) ? C[x.rk[a]] : pow(C[x.rk[a]], ordr[a]);
        if (rev && a < x.np) pr_p *= (ordp[a] == 1.0) ? C[x.pk[a]] : pow(C[x.pk[a]], ordp[a]);
    }
    auto scatter = [&](int j, double d) {
#pragma unroll
        for (int r = 0; r < E; r++)
            if (lead[r]) mat_add(&A[ek[r] * LD + j], rnu[r] * d);
    };
#pragma unroll
    for (int a = 0; a < kMaxSide; a++) {
        if (a < x.nr && ordr[a] != 0.0) {
            double d = kfM * ordr[a] * ((ordr[a] == 1.0) ? 1.0 : pow(C[x.rk[a]], ordr[a] - 1.0));
#pragma unroll
            for (int c = 0; c < kMaxSide; c++)
                if (c != a && c < x.nr) d *= (ordr[c] == 1.0) ? C[x.rk[c]] : pow(C[x.rk[c]], ordr[c]);
            scatter(x.rk[a], d);
        }
    }
    if (rev) {
#pragma unroll
        for (int a = 0; a < kMaxSide; a++) {
            if (a < x.np) {
                double d = -kfM * v.rkcn * ordp[a] * ((ordp[a] == 1.0) ? 1.0 : pow(C[x.pk[a]], ordp[a] - 1.0));
#pragma unroll
                for (int c = 0; c < kMaxSide; c++)
                    if (c != a && c < x.np) d *= (ordp[c] == 1.0) ? C[x.pk[c]] : pow(C[x.pk[c]], ordp[c]);
                scatter(x.pk[a], d);
            }
        }
    }
    // acc (constant volume): C_tot = sum_l C_l is not fixed, so d[M]/dC_l also
    // holds the default efficiency; that part, equal in every column, is
    // accumulated per species in acc[k] instead of being scattered K times
    const double tbd = acc ? m.tbdef[p] : 0.0;
    if (!(skip & 1) && type != kElementary && (x.tbn > 0 || tbd != 0.0)) {
        const double diff = pr_r - v.rkcn * pr_p;
        double q;
        if (type == kThreeBody || (skip & 2)) {
            q = v.kf * diff;
        } else {
            const double* fo = m.fod + 8 * x.fo;
            const double4 rd = reinterpret_cast<const double4*>(m.rd)[p];
            double rc_low = fo[0] * exp(fo[1] * g.logT - fo[2] * g.recipT);
            double rc_high = rd.x * exp(rd.y * g.logT - rd.z * g.recipT);
            double Pr = v.M * rc_low / (rc_high + kSmallNumber);
            double F = 1.0, dlogF = 0.0;
            if (type == kTroe && (x.flags & 256)) {
                // SRI: log10 F = log10(w) / (1 + lpr^2) + log10(d T^e)
                double w = fo[3] * exp(-fo[4] / g.T);
                if (fo[5] != 0.0) w += exp(-g.T / fo[5]);
                double lpr = log10(fmax(Pr, kSmallNumber));
                double u = 1.0 + lpr * lpr;
                F = pow(w, 1.0 / u) * (fo[6] * pow(g.T, fo[7]));
                dlogF = (Pr > kSmallNumber) ? -log10(w) * 2.0 * lpr / (u * u) : 0.0;
            } else if (type == kTroe) {
                double Fcent = (1.0 - fo[3]) * exp(-g.T * fo[4]) + fo[3] * exp(-g.T * fo[5]);
                if (fo[6] != 0.0) Fcent += exp(-fo[6] / g.T);
                Fcent += fo[7] * g.T;
                double lFc = log10(fmax(Fcent, kSmallNumber));
                double lpr = log10(fmax(Pr, kSmallNumber));
                double cc = -0.4 - 0.67 * lFc, nn = 0.75 - 1.27 * lFc;
                double xx = lpr + cc, den = nn - 0.14 * xx;
                double f1 = xx / den;
                double df1 = nn / (den * den);
                double w = 1.0 + f1 * f1;
                F = exp10(lFc / w);
                dlogF = (Pr > kSmallNumber) ? -lFc * 2.0 * f1 * df1 / (w * w) : 0.0;
            }
            if (x.flags & 512) {
                // k = k0 F/(1+Pr): dk/d[M] = k0 (k0/kinf) F/(1+Pr) (dlogF/Pr - 1/(1+Pr))
                const double r0 = rc_low / (rc_high + kSmallNumber);
                const double dl = (Pr > kSmallNumber) ? dlogF / Pr : 0.0;
                q = rc_low * r0 * F / (1.0 + Pr) * (dl - 1.0 / (1.0 + Pr)) * diff;
            } else {
                double k0p = rc_low * rc_high / (rc_high + kSmallNumber);
                q = k0p * F / (1.0 + Pr) * (1.0 / (1.0 + Pr) + dlogF) * diff;
            }
        }
        for (int t = 0; t < x.tbn; t++) scatter(m.tbk[x.tbo + t], q * m.tbe[x.tbo + t]);
        if (tbd != 0.0) {
#pragma unroll
            for (int r = 0; r < E; r++)
                if (lead[r]) atomicAdd(&acc[ek[r]], rnu[r] * q * tbd);
        }
    }
}

// Build J = d(ydot)/dy into the FP64 matrix J (global memory, N x N row-major).
// y/f0: state and RHS at the expansion point (lane-distributed, internal order).
// kRefOrder = false: J in the internal order [Y_0..Y_{K-1}, T, m];
// kRefOrder = true:  J in the reactor's own order [m, T, Y_0..Y_{K-1}]
// (IdealGasConstPressureReactor::getState, src/zeroD/IdealGasConstPressureReactor.cpp:16-26).
// kBaseF0 = true: f0 is not used; the RHS at y that this function evaluates
// anyway (to leave the species scratch at the expansion point) is the base of
// the temperature difference, which saves the caller an RHS evaluation.
// The per-reaction derivatives are accumulated with FP64 global atomics, so
// their summation order is not fixed.
template <int NS, bool kRefOrder = false, bool kBaseF0 = false>
__device__ __forceinline__ void reactor_jacobian(const DevMech& m, const WaveScratch& S, double P, const double (&y)[NS],
                                 const double (&f0)[NS], double* __restrict__ J) {
    const int lane = lane_id();
    const int K = m.K, N = m.N;
    // constant volume: u_k and c_v in the energy row, rho fixed in the species block
    const bool cv = m.const_volume;
    const double hshift = cv ? 1.0 : 0.0;
    // species block, temperature/mass rows and columns
    double* const JC = kRefOrder ? J + 2 * N + 2 : J;
    double* const rowT = kRefOrder ? J + N + 2 : J + K * N;
    double* const rowM = kRefOrder ? J + 2 : J + (K + 1) * N;
    double* const colT = kRefOrder ? J + 1 : J + K;
    double* const colM = kRefOrder ? J : J + K + 1;
    GasState g;
    SpeciesRegs<NS> sr;
    RhsScalars rs;
    // temperature column by a one-sided difference
    double y1[NS], f1[NS], dfdT[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) y1[s] = y[s];
    const double Tval = bcast<NS>(y, K);
    const double dT = kSqrtEps * Tval;
#pragma unroll
    for (int s = 0; s < NS; s++)
        if (lane + kWave * s == K) y1[s] = Tval + dT;
    reactor_rhs<NS>(m, S, P, y1, f1, g, sr, rs);
    // base state (species scratch, wdot)
    double fb[NS];
    reactor_rhs<NS>(m, S, P, y, fb, g, sr, rs);
#pragma unroll
    for (int s = 0; s < NS; s++) dfdT[s] = (f1[s] - (kBaseF0 ? fb[s] : f0[s])) / dT;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int j = lane + kWave * s;
        if (j < N)
            for (int i = 0; i < N; i++) J[i * N + j] = 0.0;
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int k = lane + kWave * s;
        if (k < K) S.WD[k] = 0.0;
    }
    __threadfence();
    wave_sync();
    const double kpd = pdep_rate_lane(m, g);
    double* const acc = cv ? S.WD : nullptr;
    for (int p = lane; p < m.I; p += kWave) scatter_reaction_derivs(m, p, g, S.C, S.MU, kpd, JC, N, acc);
    __threadfence();
    wave_sync();
    // D_k = sum_l JC[k][l] C_l ; keep wdot and D in W2, h_k in MU
    double hD = 0.0, H = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int k = lane + kWave * s;
        if (k < K) {
            double d = 0.0;
            for (int l = 0; l < K; l++) d += JC[k * N + l] * S.C[l];
            const double hk = (sr.hrt[s] - hshift) * g.RT;
            S.W2[k] = sr.wdot[s];
            S.W2[K + k] = d;
            hD += hk * d;
            H += hk * sr.wdot[s];
        }
    }
    hD = uni(wave_sum(hD));
    H = uni(wave_sum(H));
    wave_sync();
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int k = lane + kWave * s;
        if (k < K) S.MU[k] = (sr.hrt[s] - hshift) * g.RT;
    }
    wave_sync();
    const double rho = g.dens, Ct = g.ctot, Wbar = g.mmw, cp = rs.cp_mass;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int j = lane + kWave * s;
        if (j < K) {
            double hJ = 0.0;
            const double wj = m.mw[j];
            if (cv) {
                // rho fixed: dC_l/dY_j = rho / W_j delta_lj
                for (int k = 0; k < K; k++) {
                    const double jc = JC[k * N + j] + S.WD[k];
                    hJ += S.MU[k] * jc;
                    JC[k * N + j] = m.mw[k] / wj * jc;
                }
                const double cvj = kGasConstant * (sr.cpr[s] - 1.0) / wj;
                rowT[j] = -1.0 / (rho * cp) * (rho * hJ / wj - H * cvj / cp);
            } else {
                for (int k = 0; k < K; k++) hJ += S.MU[k] * JC[k * N + j];
                for (int k = 0; k < K; k++) {
                    const double jc = JC[k * N + j];
                    JC[k * N + j] = m.mw[k] / wj * (jc + (S.W2[k] - S.W2[K + k]) / Ct);
                }
                const double cpj = kGasConstant * sr.cpr[s] / wj;
                rowT[j] = -1.0 / (rho * cp) * ((rho * hJ - Wbar * hD + Wbar * H) / wj - H * cpj / cp);
            }
            rowM[j] = 0.0;
        }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int i = lane + kWave * s;
        if (i < N) {
            const int r = kRefOrder ? (i < K ? i + 2 : (i == K ? 1 : 0)) : i;
            colT[r * N] = (i == K + 1) ? 0.0 : dfdT[s];
            colM[r * N] = 0.0;
        }
    }
    __threadfence();
    wave_sync();
}

// Out-of-line Jacobian: expansion point in S.Y0, its RHS in S.F0.
template <int NS>
__device__ __noinline__ void jac_call(KernargMech* mech, ScratchLayout L, double P, double* J) {
    const DevMech m = *(const DevMech*)mech;
    const WaveScratch S = scratch_view(L);
    const int lane = lane_id();
    double y[NS], f0[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + kWave * s;
        y[s] = (i < m.N) ? S.Y0[i] : 0.0;
        f0[s] = (i < m.N) ? S.F0[i] : 0.0;
    }
    reactor_jacobian<NS>(m, S, P, y, f0, J);
}

// jac_call assembling into LDS at double offset offJ (N x N row-major), then
// copying the result to Jsave (global, the copy reused while the Jacobian is
// current).  The LDS pointer is formed here, so the assembly's accesses and
// atomics are LDS instructions.
template <int NS>
__device__ __noinline__ void jac_lds_call(KernargMech* mech, ScratchLayout L, double P, int offJ, double* Jsave) {
    const DevMech m = *(const DevMech*)mech;
    const WaveScratch S = scratch_view(L);
    const int lane = lane_id();
    double y[NS], f0[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + kWave * s;
        y[s] = (i < m.N) ? S.Y0[i] : 0.0;
        f0[s] = (i < m.N) ? S.F0[i] : 0.0;
    }
    double* J = lds_base() + offJ;
    reactor_jacobian<NS>(m, S, P, y, f0, J);
    const int NN = m.N * m.N;
    for (int e = lane; e < NN; e += kWave) Jsave[e] = J[e];
    wave_sync();
}

}  // namespace cg
