// Batched stiff integration of IdealGasConstPressureReactor, one wavefront per
// reactor, persistent kernel drawing reactors from a device work queue.
//
// Algorithm: variable-order (q <= 5), variable-step BDF in fixed-leading-
// coefficient Nordsieck form with a modified Newton corrector on a dense LU of
// M = I - gamma J -- the method CVODES runs for the reference
// (ReactorNet -> CVodesIntegrator, src/numerics/CVodesIntegrator.cpp:281-549;
// SUNDIALS 7.5 is not vendored, the method is restated from Hindmarsh et al.,
// ACM TOMS 31 (2005) and the CVODE user guide).  advance(tout) steps past tout
// and interpolates (CVodesIntegrator::integrate, :479-529); history, the saved
// Jacobian and the factored iteration matrix persist between calls.
// The step-control constants and the names of the step helpers (predict,
// rescale, set_eta, prepare_next_step, ...) follow SUNDIALS CVODE (cvode.c,
// Copyright (c) 2002-2025, Lawrence Livermore National Security and Southern
// Methodist University; BSD-3-Clause).
//
// Layout: the Nordsieck history and work vectors are lane-distributed
// registers (component i in lane i % 64, slot i / 64); the step-control block
// lives in the wavefront's LDS; the iteration matrix is held by the linear-
// system policy LS (registers or LDS).
//
// Precision modes: plain (FP64 everywhere) and mixed (iteration matrix, its
// LU and the triangular solves in FP32; Newton residual, state, error test and
// step control in FP64 -- an inexact Newton corrected by the FP64 residual).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "bdf.hpp"
#include "linsys.hpp"
#include "sparse_linsys.hpp"
#include "wave_jac.hpp"
#include "wave_mole.hpp"

namespace cg {

// rhs_call and jac_call read the mechanism through kernarg_mech(): the
// integrator kernels take one BdfArgs parameter, which must begin with it.
static_assert(offsetof(BdfArgs, mech) == 0, "BdfArgs must begin with the mechanism");

namespace bdfc {
constexpr int QMAX = 5;
constexpr int MAXCOR = 3;
constexpr int MXNEF = 7;
constexpr int MXNCF = 10;
constexpr int MXNEF1 = 3;
constexpr int SMALL_NEF = 2;
constexpr int LONG_WAIT = 10;
constexpr int SMALL_NST = 10;
constexpr double NLSCOEF = 0.1;
constexpr double ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2, ETAMIN = 0.1, ETACF = 0.25;
constexpr double ADDON = 1e-6, BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0, THRESH = 1.5;
constexpr double CRDOWN = 0.3, DGMAX_LSETUP = 0.3, RDIV = 2.0, DGMAX_JBAD = 0.2;
constexpr int MSBP = 20, MSBJ = 51;
constexpr double HLB_FACTOR = 100.0, HUB_FACTOR = 0.1, H_BIAS = 0.5;
constexpr int HIN_ITERS = 4;
constexpr double UROUND = 2.220446049250313e-16;
enum { FIRST_CALL = 0, PREV_CONV_FAIL = 1, PREV_ERR_FAIL = 2 };
enum { NO_FAILURES = 0, FAIL_BAD_J = 1, FAIL_OTHER = 2 };
}  // namespace bdfc

// step-control block (wave-uniform), kept in LDS
struct Ctl {
    double tn, h, hprime, hscale, eta, etamax, hu;
    double tau[bdfc::QMAX + 2], tq[6], l[bdfc::QMAX + 1];
    double rl1, gamma, gammap, gamrat, crate, delp, acnrm, saved_tq5, t_user, P;
    double nfe, nje, nsetups, nni, ncfn, netf, nfe_dq;
    int q, qprime, qwait, L, nst, nstlp, nstlj, nscon, started, convfail, jcur, lu_dirty;
};
constexpr size_t kCtlDoubles = (sizeof(Ctl) + 7) / 8;

template <int NS>
__device__ __forceinline__ double wrms(const double (&v)[NS], const double (&w)[NS], int N) {
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < NS; t++) {
        double x = v[t] * w[t];
        s += x * x;
    }
    return uni(sqrt(wave_sum(s) / N));
}

// the integrator kernel's one parameter, addressed in the kernel-argument
// segment (read with scalar loads wherever the index is run-time)
using KernargArgs = const __attribute__((address_space(4))) BdfArgs;
__device__ __forceinline__ KernargArgs* kernarg_args() {
    return (KernargArgs*)__builtin_amdgcn_kernarg_segment_ptr();
}

// kSens: also integrate the forward sensitivities dy/dp to the reaction-rate
// multipliers a.sens_rxn (CVODES staggered corrector, as the reference runs
// it: CV_STAGGERED, sensitivities out of the error test,
// src/numerics/CVodesIntegrator.cpp:263-279).  Each parameter's Nordsieck
// history is kept in global memory and updated alongside zn.
// kMole: IdealGasConstPressureMoleReactor (state [n_0..n_{K-1}, T]) with the
// adaptive preconditioner (LS = SparseLdsLinSys) inside CVODES' SPGMR linear
// solver (CVodesIntegrator.cpp:424-434, with ReactorNet::setPreconditioner,
// ReactorNet.cpp:330-339,881-908): the Newton systems are solved by GMRES,
// J v by difference quotients of the RHS, M = I - gamma J only preconditions.
// linear-system policies that can take a fresh Jacobian in their own matrix
// area (jac_in_place, jac_offset, form_in_place)
template <class LS, class = void>
struct has_jac_in_place : std::false_type {};
template <class LS>
struct has_jac_in_place<LS, std::void_t<decltype(&LS::form_in_place)>> : std::true_type {};

template <int NS, class LS, typename MT, bool kSens = false, bool kMole = false>
struct Integrator {
    const DevMech& m;
    const BdfArgs& a;
    WaveScratch S;
    ScratchLayout L;
    Ctl& c;
    LS ls;
    int N, lane;
    double rtol, atol;
    double zn[bdfc::QMAX + 1][NS];
    double ewt[NS], acor[NS], y[NS], ftemp[NS], tempv[NS];
    double* rec;
    size_t rid;
    double* sb = nullptr;   // this reactor's sensitivity block (kSens)
    double atol_sp = 0.0;   // kMole: absolute tolerance of the species' moles

    __device__ __forceinline__ Integrator(const DevMech& m_, const BdfArgs& a_, Ctl& c_) : m(m_), a(a_), c(c_) {}

    __device__ __forceinline__ int comp(int s) const { return lane + kWave * s; }

    __device__ __forceinline__ void load(size_t r) {
        rid = r;
        rec = a.rec + r * a.rec_stride;
        if constexpr (kSens) sb = a.sens + r * a.sens_stride;
        const int NP = kWave * NS;
#pragma unroll
        for (int j = 0; j <= bdfc::QMAX; j++)
#pragma unroll
            for (int s = 0; s < NS; s++) zn[j][s] = rec[j * NP + comp(s)];
        const double* x = rec + 7 * NP;
        if (lane == 0) {
            c.tn = x[C_TN]; c.h = x[C_H]; c.hprime = x[C_HPRIME]; c.hscale = x[C_HSCALE]; c.eta = x[C_ETA];
            c.etamax = x[C_ETAMAX]; c.hu = x[C_HU];
            for (int i = 0; i < bdfc::QMAX + 2; i++) c.tau[i] = x[C_TAU + i];
            for (int i = 0; i < 6; i++) c.tq[i] = x[C_TQ + i];
            c.rl1 = x[C_RL1]; c.gamma = x[C_GAMMA]; c.gammap = x[C_GAMMAP]; c.gamrat = x[C_GAMRAT];
            c.crate = x[C_CRATE]; c.delp = x[C_DELP]; c.acnrm = x[C_ACNRM]; c.saved_tq5 = x[C_SAVED_TQ5];
            c.q = static_cast<int>(x[C_Q]); c.qprime = static_cast<int>(x[C_QPRIME]);
            c.qwait = static_cast<int>(x[C_QWAIT]); c.L = static_cast<int>(x[C_L]);
            c.nst = static_cast<int>(x[C_NST]); c.nstlp = static_cast<int>(x[C_NSTLP]);
            c.nstlj = static_cast<int>(x[C_NSTLJ]); c.nscon = static_cast<int>(x[C_NSCON]);
            c.started = static_cast<int>(x[C_STARTED]);
            c.t_user = x[C_TUSER];
            c.nfe = x[C_NFE]; c.nje = x[C_NJE]; c.nsetups = x[C_NSETUPS]; c.nni = x[C_NNI];
            c.ncfn = x[C_NCFN]; c.netf = x[C_NETF]; c.nfe_dq = x[C_NFE_DQ];
            c.P = x[C_P];
            c.jcur = 0;
            c.lu_dirty = 0;
            if constexpr (kMole) {
                double* cnt = ls.gm + kGmCnt;
                for (int i = 0; i < 5; i++) cnt[i] = x[C_NLI + i];
            }
        }
        if constexpr (kMole) atol_sp = uni(x[C_ATOL_SP]);
        const bool have_lu = uni(x[C_HAVE_LU]) != 0.0;
        ls.bind(a.template lu<MT>() + r * a.lu_stride, a.piv_save + r * a.piv_stride);
        wave_sync();
        if (have_lu)
            ls.restore(a.template lu<MT>() + r * a.lu_stride, a.piv_save + r * a.piv_stride, N);
    }

    __device__ __forceinline__ void store(int status) {
        const int NP = kWave * NS;
#pragma unroll
        for (int j = 0; j <= bdfc::QMAX; j++)
#pragma unroll
            for (int s = 0; s < NS; s++) rec[j * NP + comp(s)] = zn[j][s];
        const bool dirty = uni(c.lu_dirty) != 0;
        if (lane == 0) {
            double* x = rec + 7 * NP;
            x[C_TN] = uni(c.tn); x[C_H] = uni(c.h); x[C_HPRIME] = uni(c.hprime); x[C_HSCALE] = uni(c.hscale); x[C_ETA] = uni(c.eta);
            x[C_ETAMAX] = uni(c.etamax); x[C_HU] = uni(c.hu);
            for (int i = 0; i < bdfc::QMAX + 2; i++) x[C_TAU + i] = uni(c.tau[i]);
            for (int i = 0; i < 6; i++) x[C_TQ + i] = uni(c.tq[i]);
            x[C_RL1] = uni(c.rl1); x[C_GAMMA] = uni(c.gamma); x[C_GAMMAP] = uni(c.gammap); x[C_GAMRAT] = uni(c.gamrat);
            x[C_CRATE] = uni(c.crate); x[C_DELP] = uni(c.delp); x[C_ACNRM] = uni(c.acnrm); x[C_SAVED_TQ5] = uni(c.saved_tq5);
            x[C_Q] = uni(c.q); x[C_QPRIME] = uni(c.qprime); x[C_QWAIT] = uni(c.qwait); x[C_L] = uni(c.L); x[C_NST] = uni(c.nst);
            x[C_NSTLP] = uni(c.nstlp); x[C_NSTLJ] = uni(c.nstlj); x[C_NSCON] = uni(c.nscon); x[C_STARTED] = uni(c.started);
            x[C_TUSER] = uni(c.t_user);
            x[C_NFE] = uni(c.nfe); x[C_NJE] = uni(c.nje); x[C_NSETUPS] = uni(c.nsetups); x[C_NNI] = uni(c.nni); x[C_NCFN] = uni(c.ncfn);
            x[C_NETF] = uni(c.netf); x[C_NFE_DQ] = uni(c.nfe_dq);
            x[C_STATUS] = status;
            if (dirty) x[C_HAVE_LU] = 1.0;
            if constexpr (kMole) {
                const double* cnt = ls.gm + kGmCnt;
                for (int i = 0; i < 5; i++) x[C_NLI + i] = cnt[i];
            }
        }
        if (dirty) ls.save(a.template lu<MT>() + rid * a.lu_stride, a.piv_save + rid * a.piv_stride, N);
    }

    // ls_count: a J v product of the linear solver (lin_rhs_evals, not nfe)
    __device__ __forceinline__ int f_eval(const double (&yy)[NS], double (&ff)[NS], bool ls_count = false) {
        if (lane == 0) {
            if constexpr (kMole) {
                if (ls_count) ls.gm[kGmCnt + 4] += 1;
                else c.nfe += 1;
            } else {
                c.nfe += 1;
            }
        }
#pragma unroll
        for (int s = 0; s < NS; s++)
            if (comp(s) < N) S.yb[comp(s)] = yy[s];
        wave_sync();
        if constexpr (kMole) mole_rhs_call<NS>(kernarg_mech(), L, uni(c.P));
        else rhs_call<NS>(kernarg_mech(), L, uni(c.P));
        int bad = 0;
#pragma unroll
        for (int s = 0; s < NS; s++) {
            ff[s] = (comp(s) < N) ? S.F[comp(s)] : 0.0;
            if (comp(s) < N && !isfinite(ff[s])) bad = 1;
        }
        wave_sync();
        return __any(bad) ? 1 : 0;
    }

    // ---- sensitivity storage: column j of parameter p, this lane's components
    __device__ __forceinline__ double* scol(int p, int j) const {
        return sb + (static_cast<size_t>(p) * kSensCols + j) * (kWave * NS);
    }
    __device__ __forceinline__ void sload(int p, int j, double (&v)[NS]) const {
        const double* col = scol(p, j);
#pragma unroll
        for (int s = 0; s < NS; s++) v[s] = col[comp(s)];
    }
    __device__ __forceinline__ void sstore(int p, int j, const double (&v)[NS]) const {
        double* col = scol(p, j);
#pragma unroll
        for (int s = 0; s < NS; s++) col[comp(s)] = v[s];
    }
    __device__ __forceinline__ void sload_all(int p, double (&z)[bdfc::QMAX + 1][NS]) const {
#pragma unroll
        for (int j = 0; j <= bdfc::QMAX; j++) sload(p, j, z[j]);
    }
    __device__ __forceinline__ void sstore_all(int p, const double (&z)[bdfc::QMAX + 1][NS]) const {
#pragma unroll
        for (int j = 0; j <= bdfc::QMAX; j++) sstore(p, j, z[j]);
    }
    // z[j] for a run-time j, chosen on values (as pick)
    __device__ __forceinline__ static void pick_of(const double (&z)[bdfc::QMAX + 1][NS], int j, double (&out)[NS]) {
#pragma unroll
        for (int s = 0; s < NS; s++) out[s] = z[0][s];
#pragma unroll
        for (int t = 1; t <= bdfc::QMAX; t++)
#pragma unroll
            for (int s = 0; s < NS; s++) out[s] = (t == j) ? z[t][s] : out[s];
    }

    // right-hand side with reaction rx at rate multiplier v (not counted in nfe)
    __device__ __forceinline__ int f_eval_pert(const double (&yy)[NS], double (&ff)[NS], int rx, double v) {
#pragma unroll
        for (int s = 0; s < NS; s++)
            if (comp(s) < N) S.yb[comp(s)] = yy[s];
        wave_sync();
        rhs_pert_call<NS>(kernarg_mech(), L, uni(c.P), rx, v);
        int bad = 0;
#pragma unroll
        for (int s = 0; s < NS; s++) {
            ff[s] = (comp(s) < N) ? S.F[comp(s)] : 0.0;
            if (comp(s) < N && !isfinite(ff[s])) bad = 1;
        }
        wave_sync();
        return __any(bad) ? 1 : 0;
    }

    // Sensitivity right-hand side J s + df/dp for parameter p at (yy, sv): the
    // centred difference quotient CVODES forms without a user sensitivity
    // function (cvSensRhs1InternalDQ, CENTERED1; parameter value and scale 1
    // as Reactor::addSensitivityReaction registers them, src/zeroD/
    // Reactor.cpp:278-289), [f(y + D s; k_p (1 + D)) - f(y - D s; k_p (1 - D))]
    // / 2D with D = min(sqrt(max(rtol, uround)), 1 / |s|_wrms), state weights.
    __device__ __forceinline__ int sens_rhs(int p, const double (&yy)[NS], const double (&sv)[NS], double (&fS)[NS]) {
        const double del = sqrt(fmax(rtol, bdfc::UROUND));
        const double nrm = wrms<NS>(sv, ewt, a.neq_norm);
        const double D = fmin(del, 1.0 / fmax(nrm, 1.0 / del));
        const int rx = kernarg_args()->sens_rxn[p];
        double yp[NS], fp[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) yp[s] = yy[s] + D * sv[s];
        if (f_eval_pert(yp, fp, rx, 1.0 + D)) return 1;
#pragma unroll
        for (int s = 0; s < NS; s++) yp[s] = yy[s] - D * sv[s];
        if (f_eval_pert(yp, fS, rx, 1.0 - D)) return 1;
#pragma unroll
        for (int s = 0; s < NS; s++) fS[s] = (fp[s] - fS[s]) / (2.0 * D);
        return 0;
    }

    // Staggered corrector, after the state's step passed its error test:
    // Newton on the (linear) BDF equation of each s_n with the state's
    // iteration matrix, to the state's convergence test on the sensitivity
    // weights 1 / (rtol_sens |s| + atol_sens); at most 10 iterations.  The
    // correction is left in column 6.
    __device__ __forceinline__ int sens_correct() {
        using namespace bdfc;
        const double gamma = uni(c.gamma), rl1 = uni(c.rl1), gamrat = uni(c.gamrat), tq4 = uni(c.tq[4]);
        for (int p = 0; p < a.nsens; p++) {
            double z0[NS], z1[NS], acs[NS], ews[NS], d[NS], f[NS];
            sload(p, 0, z0);
            sload(p, 1, z1);
#pragma unroll
            for (int s = 0; s < NS; s++) {
                ews[s] = (comp(s) < N) ? 1.0 / (a.rtol_sens * fabs(z0[s]) + a.atol_sens) : 0.0;
                acs[s] = 0.0;
            }
            double crateS = 1.0, delpS = 0.0;
            for (int m_it = 0; m_it < 10; m_it++) {
#pragma unroll
                for (int s = 0; s < NS; s++) d[s] = z0[s] + acs[s];
                if (sens_rhs(p, y, d, f)) return 1;
#pragma unroll
                for (int s = 0; s < NS; s++) d[s] = -(rl1 * z1[s] + acs[s] - gamma * f[s]);
                ls.solve(d, N);
                if (gamrat != 1.0) {
                    const double sc = 2.0 / (1.0 + gamrat);
#pragma unroll
                    for (int s = 0; s < NS; s++) d[s] *= sc;
                }
#pragma unroll
                for (int s = 0; s < NS; s++) acs[s] += d[s];
                const double del = wrms<NS>(d, ews, a.neq_norm);
                if (m_it > 0) crateS = fmax(CRDOWN * crateS, del / delpS);
                if (del * fmin(1.0, crateS) / tq4 <= 1.0) break;
                delpS = del;
            }
            sstore(p, 6, acs);
        }
        return 0;
    }

    __device__ __forceinline__ static int sp_nnz() { return kernarg_args()->sp.nnz; }
    __device__ __forceinline__ static double wave_dot(const double (&u)[NS], const double (&v)[NS]) {
        double t = 0.0;
#pragma unroll
        for (int s = 0; s < NS; s++) t += u[s] * v[s];
        return uni(wave_sum(t));
    }

    // z = (I - gamma J) v, J v by cvLsDQJtimes' difference quotient at the
    // Newton iterate (y, ftemp): sig = 1 / |v|_wrms, up to 3 tries with sig
    // quartered after a recoverable RHS failure.  0 ok, 1 recoverable.
    __device__ __forceinline__ int atimes(const double (&v)[NS], double (&z)[NS]) {
        double sig = 1.0 / wrms<NS>(v, ewt, a.neq_norm);
        int ret = 1;
        double yy[NS];
        for (int it = 0; it < 3; it++) {
#pragma unroll
            for (int s = 0; s < NS; s++) yy[s] = sig * v[s] + y[s];
            ret = f_eval(yy, z, true);
            if (ret == 0) break;
            sig *= 0.25;
        }
        if (ret) return 1;
        const double siginv = 1.0 / sig, gamma = uni(c.gamma);
#pragma unroll
        for (int s = 0; s < NS; s++) z[s] = -gamma * (siginv * (z[s] - ftemp[s])) + v[s];
        if (lane == 0) ls.gm[kGmCnt + 2] += 1;
        return 0;
    }

    // cvLsSolve with SPGMR (SUNDIALS, restated from the published algorithm:
    // GMRES with modified Gram-Schmidt and Givens rotations, zero initial
    // guess, right preconditioner, both scalings the error weights, tolerance
    // eplifac tq[4] sqrt(N) in the scaled 2-norm).  Solves (I - gamma J) x = b
    // in place.  0 converged (or residual reduced on the first Newton
    // iteration), 1 recoverable failure.
    __device__ __forceinline__ int gmres(double (&b)[NS], int mnewt) {
        double* H = ls.gm + kGmH;
        double* rot = ls.gm + kGmRot;
        double* cnt = ls.gm + kGmCnt;
        const double deltar = 0.05 * uni(c.tq[4]);
        const double bnorm = wrms<NS>(b, ewt, a.neq_norm);
        if (bnorm <= deltar) {
            if (mnewt > 0)
#pragma unroll
                for (int s = 0; s < NS; s++) b[s] = 0.0;
            return 0;
        }
        const double delta = deltar * sqrt(static_cast<double>(N));
        double V[kGmMaxl + 1][NS];
#pragma unroll
        for (int s = 0; s < NS; s++) V[0][s] = ewt[s] * b[s];
        const double r_norm = sqrt(wave_dot(V[0], V[0]));
        if (r_norm <= delta) {
#pragma unroll
            for (int s = 0; s < NS; s++) b[s] = 0.0;
            return 0;
        }
        const double beta = r_norm;
        double rho = r_norm, rotp = 1.0;
        const double rinv = 1.0 / r_norm;
#pragma unroll
        for (int s = 0; s < NS; s++) V[0][s] *= rinv;
        int krydim = 0;
        bool conv = false;
#pragma unroll
        for (int l = 0; l < kGmMaxl; l++) {
            wave_sync();
            if (lane == 0) cnt[0] += 1;
            krydim = l + 1;
            double z[NS], w[NS];
#pragma unroll
            for (int s = 0; s < NS; s++) z[s] = (comp(s) < N) ? V[l][s] / ewt[s] : 0.0;
            ls.solve(z, N);
            if (lane == 0) cnt[1] += 1;
            if (atimes(z, w)) {
                if (lane == 0) cnt[3] += 1;
                return 1;
            }
#pragma unroll
            for (int s = 0; s < NS; s++) w[s] *= ewt[s];
            // SUNModifiedGS, with one reorthogonalisation pass when the new
            // norm falls below 1e-3 of the old
            double h[kGmMaxl + 1];
            const double vk = sqrt(wave_dot(w, w));
#pragma unroll
            for (int i = 0; i <= l; i++) {
                h[i] = wave_dot(V[i], w);
#pragma unroll
                for (int s = 0; s < NS; s++) w[s] = -h[i] * V[i][s] + w[s];
            }
            double nn = sqrt(wave_dot(w, w));
            if ((1000.0 * vk + nn) == 1000.0 * vk) {
                double n2 = 0.0;
#pragma unroll
                for (int i = 0; i <= l; i++) {
                    const double np = wave_dot(V[i], w);
                    if ((1000.0 * h[i] + np) == 1000.0 * h[i]) continue;
                    h[i] += np;
#pragma unroll
                    for (int s = 0; s < NS; s++) w[s] = -np * V[i][s] + w[s];
                    n2 += np * np;
                }
                if (n2 != 0.0) {
                    const double np = nn * nn - n2;
                    nn = np > 0.0 ? sqrt(np) : 0.0;
                }
            }
            // SUNQRfact update: previous rotations on the new column, then one more
#pragma unroll
            for (int k = 0; k < l; k++) {
                const double t1 = h[k], t2 = h[k + 1], cr = rot[2 * k], sr = rot[2 * k + 1];
                h[k] = cr * t1 - sr * t2;
                h[k + 1] = sr * t1 + cr * t2;
            }
            double cr, sr;
            {
                const double t1 = h[l], t2 = nn;
                if (t2 == 0.0) {
                    cr = 1.0;
                    sr = 0.0;
                } else if (fabs(t2) >= fabs(t1)) {
                    const double t3 = t1 / t2;
                    sr = -1.0 / sqrt(1.0 + t3 * t3);
                    cr = -sr * t3;
                } else {
                    const double t3 = t2 / t1;
                    cr = 1.0 / sqrt(1.0 + t3 * t3);
                    sr = -cr * t3;
                }
                h[l] = cr * t1 - sr * t2;
            }
            wave_sync();
            if (lane == 0) {
                rot[2 * l] = cr;
                rot[2 * l + 1] = sr;
#pragma unroll
                for (int i = 0; i <= l; i++) H[i * kGmMaxl + l] = h[i];
            }
            wave_sync();
            if (h[l] == 0.0) {
                if (lane == 0) cnt[3] += 1;
                return 1;
            }
            rotp *= sr;
            rho = fabs(rotp * r_norm);
            if (rho <= delta) {
                conv = true;
                break;
            }
            if (l + 1 < kGmMaxl + 1) {
                const double ninv = 1.0 / nn;
#pragma unroll
                for (int s = 0; s < NS; s++) V[l + 1][s] = w[s] * ninv;
            }
        }
        // SUNQRsol: g = R^{-1} Q (r_norm e_1); x = sum_k g_k V_k
        double g[kGmMaxl + 1];
        g[0] = r_norm;
#pragma unroll
        for (int i = 1; i <= kGmMaxl; i++) g[i] = 0.0;
#pragma unroll
        for (int k = 0; k < kGmMaxl; k++) {
            if (k < krydim) {
                const double cr = rot[2 * k], sr = rot[2 * k + 1], t1 = g[k], t2 = g[k + 1];
                g[k] = cr * t1 - sr * t2;
                g[k + 1] = sr * t1 + cr * t2;
            }
        }
#pragma unroll
        for (int k = kGmMaxl - 1; k >= 0; k--) {
            if (k < krydim) {
                g[k] /= H[k * kGmMaxl + k];
#pragma unroll
                for (int i = 0; i < k; i++) g[i] -= g[k] * H[i * kGmMaxl + k];
            }
        }
        if (!conv && !(rho < beta)) {
            if (lane == 0) cnt[3] += 1;
            return 1;
        }
        double xc[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < kGmMaxl; k++)
                if (k < krydim) t += g[k] * V[k][s];
            xc[s] = (comp(s) < N) ? t / ewt[s] : 0.0;
        }
        ls.solve(xc, N);
        wave_sync();
        if (lane == 0) cnt[1] += 1;
#pragma unroll
        for (int s = 0; s < NS; s++) b[s] = xc[s];
        if (conv) return 0;
        if (lane == 0) cnt[3] += 1;   // SUNLS_RES_REDUCED
        return mnewt == 0 ? 0 : 1;
    }

    // absolute tolerance of component comp(s): the mole reactor's species
    // (MoleReactor::updateDefaultTolerances, MoleReactor.cpp:64-76) use atol_sp
    __device__ __forceinline__ double atol_of(int s) const {
        if constexpr (kMole) return (comp(s) >= m.K) ? atol : atol_sp;
        else return atol;
    }
    __device__ __forceinline__ void ewt_set(const double (&yy)[NS]) {
#pragma unroll
        for (int s = 0; s < NS; s++) ewt[s] = (comp(s) < N) ? 1.0 / (rtol * fabs(yy[s]) + atol_of(s)) : 0.0;
    }

    // linear-system setup: Jacobian reuse decision, M = I - gamma J, LU
    __device__ __forceinline__ int lsetup() {
        using namespace bdfc;
        const double gamma = uni(c.gamma);
        const double dgamma = fabs(gamma / uni(c.gammap) - 1.0);
        const int nst = uni(c.nst), convfail = uni(c.convfail);
        const bool jbad = (nst == 0) || (nst >= uni(c.nstlj) + MSBJ) ||
                          ((convfail == FAIL_BAD_J) && (dgamma < DGMAX_JBAD)) || (convfail == FAIL_OTHER);
        double* Js = a.j_save + rid * a.j_stride;
        bool formed = false;   // M formed in place over a Jacobian assembled in the matrix area
        wave_sync();
        if (!jbad) {
            if (lane == 0) c.jcur = 0;
        } else {
            if (lane == 0) {
                c.nje += 1;
                c.nstlj = nst;
                c.jcur = 1;
                c.nfe_dq += 2;
            }
            if constexpr (kMole) {
                // ReactorNet::preconditionerSetup: the state clipped at the
                // preconditioner's atol (0, AdaptivePreconditioner::stateAdjustment)
#pragma unroll
                for (int s = 0; s < NS; s++)
                    if (comp(s) < N) S.Y0[comp(s)] = fmax(zn[0][s], 0.0);
                wave_sync();
                const SparseTables* sp = (const SparseTables*)&kernarg_args()->sp;
                mole_jac_call<NS>(kernarg_mech(), L, uni(c.P), Js + sp_nnz(), Js, sp);
            } else {
#pragma unroll
                for (int s = 0; s < NS; s++)
                    if (comp(s) < N) {
                        S.Y0[comp(s)] = zn[0][s];
                        S.F0[comp(s)] = ftemp[s];
                    }
                wave_sync();
                if constexpr (has_jac_in_place<LS>::value) {
                    if (ls.jac_in_place(N)) {
                        jac_lds_call<NS>(kernarg_mech(), L, uni(c.P), ls.jac_offset(), Js);
                        ls.form_in_place(N, gamma);
                        formed = true;
                    }
                }
                if (!formed) jac_call<NS>(kernarg_mech(), L, uni(c.P), Js);
            }
        }
        if (!formed) ls.form(Js, N, gamma);
        const bool ok = ls.factor(N);
        if (lane == 0) {
            c.nsetups += 1;
            c.lu_dirty = 1;
            if (ok) {
                c.gamrat = 1.0;
                c.gammap = gamma;
                c.crate = 1.0;
                c.nstlp = nst;
            }
        }
        wave_sync();
        return ok ? 0 : 1;
    }

    __device__ __forceinline__ void set_bdf() {
        using namespace bdfc;
        const int q = uni(c.q);
        const double h = uni(c.h);
        double l[QMAX + 1];
        double alpha0, alpha0_hat, xi_inv, xistar_inv, hsum;
        l[0] = l[1] = xi_inv = xistar_inv = 1.0;
#pragma unroll
        for (int i = 2; i <= QMAX; i++) l[i] = 0.0;
        alpha0 = alpha0_hat = -1.0;
        hsum = h;
        if (q > 1) {
#pragma unroll
            for (int j = 2; j < QMAX; j++) {
                if (j < q) {
                    hsum += uni(c.tau[j - 1]);
                    xi_inv = h / hsum;
                    alpha0 -= 1.0 / j;
#pragma unroll
                    for (int i = j; i >= 1; i--) l[i] += l[i - 1] * xi_inv;
                }
            }
            alpha0 -= 1.0 / q;
            xistar_inv = -l[1] - alpha0;
            hsum += uni(c.tau[q - 1]);
            xi_inv = h / hsum;
            alpha0_hat = -l[1] - xi_inv;
#pragma unroll
            for (int i = QMAX; i >= 1; i--)
                if (i <= q) l[i] += l[i - 1] * xistar_inv;
        }
        double lq = 0.0;
#pragma unroll
        for (int i = 1; i <= QMAX; i++)
            if (i == q) lq = l[i];
        const double A1 = 1.0 - alpha0_hat + alpha0;
        const double A2 = 1.0 + q * A1;
        double tq1 = uni(c.tq[1]), tq3 = uni(c.tq[3]);
        const double tq2 = fabs(A1 / (alpha0 * A2));
        const double tq5 = fabs(A2 * xistar_inv / (lq * xi_inv));
        if (uni(c.qwait) == 1) {
            if (q > 1) {
                const double C = xistar_inv / lq;
                const double A3 = alpha0 + 1.0 / q;
                const double A4 = alpha0_hat + xi_inv;
                const double Cpinv = (1.0 - A4 + A3) / A3;
                tq1 = fabs(C * Cpinv);
            } else {
                tq1 = 1.0;
            }
            hsum += uni(c.tau[q]);
            xi_inv = h / hsum;
            const double A5 = alpha0 - 1.0 / (q + 1);
            const double A6 = alpha0_hat - xi_inv;
            const double Cppinv = (1.0 - A6 + A5) / A2;
            tq3 = fabs(Cppinv / (xi_inv * (q + 2) * A5));
        }
        wave_sync();
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i <= QMAX; i++) c.l[i] = l[i];
            c.tq[1] = tq1;
            c.tq[2] = tq2;
            c.tq[3] = tq3;
            c.tq[4] = NLSCOEF / tq2;
            c.tq[5] = tq5;
        }
        wave_sync();
    }

    __device__ __forceinline__ void predict() {
        const int q = uni(c.q);
        if (lane == 0) c.tn += uni(c.h);
#pragma unroll
        for (int k = 1; k <= bdfc::QMAX; k++)
#pragma unroll
            for (int j = bdfc::QMAX; j >= 1; j--)
                if (k <= q && j >= k && j <= q)
#pragma unroll
                    for (int s = 0; s < NS; s++) zn[j - 1][s] += zn[j][s];
        if constexpr (kSens) {
            for (int p = 0; p < a.nsens; p++) {
                double z[bdfc::QMAX + 1][NS];
                sload_all(p, z);
#pragma unroll
                for (int k = 1; k <= bdfc::QMAX; k++)
#pragma unroll
                    for (int j = bdfc::QMAX; j >= 1; j--)
                        if (k <= q && j >= k && j <= q)
#pragma unroll
                            for (int s = 0; s < NS; s++) z[j - 1][s] += z[j][s];
                sstore_all(p, z);
            }
        }
        wave_sync();
    }
    __device__ __forceinline__ void restore(double saved_t) {
        const int q = uni(c.q);
        if (lane == 0) c.tn = saved_t;
#pragma unroll
        for (int k = 1; k <= bdfc::QMAX; k++)
#pragma unroll
            for (int j = bdfc::QMAX; j >= 1; j--)
                if (k <= q && j >= k && j <= q)
#pragma unroll
                    for (int s = 0; s < NS; s++) zn[j - 1][s] -= zn[j][s];
        if constexpr (kSens) {
            for (int p = 0; p < a.nsens; p++) {
                double z[bdfc::QMAX + 1][NS];
                sload_all(p, z);
#pragma unroll
                for (int k = 1; k <= bdfc::QMAX; k++)
#pragma unroll
                    for (int j = bdfc::QMAX; j >= 1; j--)
                        if (k <= q && j >= k && j <= q)
#pragma unroll
                            for (int s = 0; s < NS; s++) z[j - 1][s] -= z[j][s];
                sstore_all(p, z);
            }
        }
        wave_sync();
    }
    __device__ __forceinline__ void rescale() {
        const int q = uni(c.q);
        const double eta = uni(c.eta);
        double factor = eta;
#pragma unroll
        for (int j = 1; j <= bdfc::QMAX; j++) {
            if (j <= q) {
#pragma unroll
                for (int s = 0; s < NS; s++) zn[j][s] *= factor;
                factor *= eta;
            }
        }
        if constexpr (kSens) {
            for (int p = 0; p < a.nsens; p++) {
                double z[bdfc::QMAX + 1][NS];
                sload_all(p, z);
                double fac = eta;
#pragma unroll
                for (int j = 1; j <= bdfc::QMAX; j++) {
                    if (j <= q) {
#pragma unroll
                        for (int s = 0; s < NS; s++) z[j][s] *= fac;
                        fac *= eta;
                    }
                }
                sstore_all(p, z);
            }
        }
        const double h = uni(c.hscale) * eta;
        wave_sync();
        if (lane == 0) {
            c.h = h;
            c.hscale = h;
            c.nscon = 0;
        }
        wave_sync();
    }

    // zn[j] for a run-time j.  Every candidate is read and the choice is made
    // on values: a branch per candidate lets the optimiser merge the reads into
    // one load through a pointer into zn, which keeps the whole history array
    // in scratch memory instead of registers.
    __device__ __forceinline__ void pick(int j, double (&out)[NS]) {
#pragma unroll
        for (int s = 0; s < NS; s++) out[s] = opaque(zn[0][s]);
#pragma unroll
        for (int t = 1; t <= bdfc::QMAX; t++)
#pragma unroll
            for (int s = 0; s < NS; s++) out[s] = (t == j) ? opaque(zn[t][s]) : out[s];
    }

    __device__ __forceinline__ void increase_bdf() {
        using namespace bdfc;
        const int q = uni(c.q);
        double ll[QMAX + 1];
#pragma unroll
        for (int i = 0; i <= QMAX; i++) ll[i] = 0.0;
        double alpha1, prod, xiold, alpha0, hsum, xi;
        ll[2] = alpha1 = prod = xiold = 1.0;
        alpha0 = -1.0;
        const double hscale = uni(c.hscale);
        hsum = hscale;
        if (q > 1) {
#pragma unroll
            for (int j = 1; j < QMAX; j++) {
                if (j < q) {
                    hsum += uni(c.tau[j + 1]);
                    xi = hsum / hscale;
                    prod *= xi;
                    alpha0 -= 1.0 / (j + 1);
                    alpha1 += 1.0 / xi;
#pragma unroll
                    for (int i = j + 2; i >= 2; i--) ll[i] = ll[i] * xiold + ll[i - 1];
                    xiold = xi;
                }
            }
        }
        const double A1 = (-alpha0 - alpha1) / prod;
        double zL[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) zL[s] = A1 * zn[QMAX][s];
#pragma unroll
        for (int j = 2; j <= QMAX; j++)
#pragma unroll
            for (int s = 0; s < NS; s++) zn[j][s] = (j == q + 1) ? zL[s] : opaque(zn[j][s]);
#pragma unroll
        for (int j = 2; j <= QMAX; j++)
            if (j <= q)
#pragma unroll
                for (int s = 0; s < NS; s++) zn[j][s] += ll[j] * zL[s];
        if constexpr (kSens) {
            for (int p = 0; p < a.nsens; p++) {
                double z[QMAX + 1][NS];
                sload_all(p, z);
#pragma unroll
                for (int s = 0; s < NS; s++) zL[s] = A1 * z[QMAX][s];
#pragma unroll
                for (int j = 2; j <= QMAX; j++)
#pragma unroll
                    for (int s = 0; s < NS; s++) z[j][s] = (j == q + 1) ? zL[s] : z[j][s];
#pragma unroll
                for (int j = 2; j <= QMAX; j++)
                    if (j <= q)
#pragma unroll
                        for (int s = 0; s < NS; s++) z[j][s] += ll[j] * zL[s];
                sstore_all(p, z);
            }
        }
    }
    __device__ __forceinline__ void decrease_bdf() {
        using namespace bdfc;
        const int q = uni(c.q);
        double ll[QMAX + 1];
#pragma unroll
        for (int i = 0; i <= QMAX; i++) ll[i] = 0.0;
        ll[2] = 1.0;
        double hsum = 0.0, xi;
        const double hscale = uni(c.hscale);
#pragma unroll
        for (int j = 1; j <= QMAX - 2; j++) {
            if (j <= q - 2) {
                hsum += uni(c.tau[j]);
                xi = hsum / hscale;
#pragma unroll
                for (int i = j + 2; i >= 2; i--) ll[i] = ll[i] * xi + ll[i - 1];
            }
        }
        double zq[NS];
        pick(q, zq);
#pragma unroll
        for (int j = 2; j < QMAX; j++)
            if (j < q)
#pragma unroll
                for (int s = 0; s < NS; s++) zn[j][s] = -ll[j] * zq[s] + zn[j][s];
        if constexpr (kSens) {
            for (int p = 0; p < a.nsens; p++) {
                double z[QMAX + 1][NS];
                sload_all(p, z);
                pick_of(z, q, zq);
#pragma unroll
                for (int j = 2; j < QMAX; j++)
                    if (j < q)
#pragma unroll
                        for (int s = 0; s < NS; s++) z[j][s] = -ll[j] * zq[s] + z[j][s];
                sstore_all(p, z);
            }
        }
    }

    // modified Newton iteration; 0 ok, 1 recoverable failure
    __device__ __forceinline__ int nls(int nflag) {
        using namespace bdfc;
        wave_sync();
        if (lane == 0) c.convfail = (nflag == FIRST_CALL || nflag == PREV_ERR_FAIL) ? NO_FAILURES : FAIL_OTHER;
        bool callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (uni(c.nst) == 0) ||
                         (uni(c.nst) >= uni(c.nstlp) + MSBP) || (fabs(uni(c.gamrat) - 1.0) > DGMAX_LSETUP);
        wave_sync();
        bool jbad = false;
        double delta[NS];
        for (;;) {
#pragma unroll
            for (int s = 0; s < NS; s++) {
                acor[s] = 0.0;
                y[s] = zn[0][s];
            }
            if (f_eval(y, ftemp)) return 1;
            double gamma = uni(c.gamma), rl1 = uni(c.rl1);
#pragma unroll
            for (int s = 0; s < NS; s++) delta[s] = -gamma * ftemp[s] + (rl1 * zn[1][s] + acor[s]);
            if (callSetup) {
                if (jbad && lane == 0) c.convfail = FAIL_BAD_J;
                wave_sync();
                if (lsetup()) return 1;
                callSetup = false;
            }
            const double tq4 = uni(c.tq[4]);
            int m_it = 0;
            int status;
            double crate = uni(c.crate), delp = uni(c.delp), gamrat = uni(c.gamrat);
            for (;;) {
                if constexpr (kMole) {
                    // iterative solve: no gamma-ratio scaling (CVLS scalesol is off)
#pragma unroll
                    for (int s = 0; s < NS; s++) delta[s] = -delta[s];
                    if (gmres(delta, m_it)) {
                        m_it++;
                        status = 1;
                        break;
                    }
                } else if constexpr (sizeof(MT) == 8) {
#pragma unroll
                    for (int s = 0; s < NS; s++) delta[s] = -delta[s];
                    ls.solve(delta, N);
                } else {
                    float bf[NS];
#pragma unroll
                    for (int s = 0; s < NS; s++) bf[s] = static_cast<float>(-delta[s]);
                    ls.solve(bf, N);
#pragma unroll
                    for (int s = 0; s < NS; s++) delta[s] = static_cast<double>(bf[s]);
                }
                if (!kMole && gamrat != 1.0) {
                    const double sc = 2.0 / (1.0 + gamrat);
#pragma unroll
                    for (int s = 0; s < NS; s++) delta[s] *= sc;
                }
#pragma unroll
                for (int s = 0; s < NS; s++) acor[s] += delta[s];
                const double del = wrms<NS>(delta, ewt, a.neq_norm);
                if (m_it > 0) crate = fmax(CRDOWN * crate, del / delp);
                const double dcon = del * fmin(1.0, crate) / tq4;
                m_it++;
                if (dcon <= 1.0) {
                    const double acnrm = (m_it == 1) ? del : wrms<NS>(acor, ewt, a.neq_norm);
                    if (lane == 0) c.acnrm = acnrm;
                    status = 0;
                    break;
                }
                if (m_it >= 2 && del > RDIV * delp) { status = 1; break; }
                delp = del;
                if (m_it >= MAXCOR) { status = 1; break; }
#pragma unroll
                for (int s = 0; s < NS; s++) y[s] = zn[0][s] + acor[s];
                if (f_eval(y, ftemp)) { status = 1; break; }
#pragma unroll
                for (int s = 0; s < NS; s++) delta[s] = -gamma * ftemp[s] + (rl1 * zn[1][s] + acor[s]);
            }
            wave_sync();
            if (lane == 0) {
                c.nni += m_it;
                c.crate = crate;
                c.delp = delp;
            }
            wave_sync();
            if (status == 0) {
                if (lane == 0) c.jcur = 0;
#pragma unroll
                for (int s = 0; s < NS; s++) y[s] = zn[0][s] + acor[s];
                wave_sync();
                return 0;
            }
            if (!uni(c.jcur)) {
                callSetup = true;
                jbad = true;
                continue;
            }
            return 1;
        }
    }

    // 0 pass, 2 retry, -1 fail
    __device__ __forceinline__ int do_error_test(int& nflag, double saved_t, int& nef, double& dsm) {
        using namespace bdfc;
        dsm = uni(c.acnrm) * uni(c.tq[2]);
        if (dsm <= 1.0) return 0;
        nef++;
        nflag = PREV_ERR_FAIL;
        wave_sync();
        if (lane == 0) c.netf += 1;
        restore(saved_t);
        if (nef == MXNEF) return -1;
        if (nef <= MXNEF1) {
            double eta = 1.0 / (pow(BIAS2 * dsm, 1.0 / uni(c.L)) + ADDON);
            eta = fmax(ETAMIN, eta);
            if (nef >= SMALL_NEF) eta = fmin(eta, ETAMXF);
            wave_sync();
            if (lane == 0) {
                c.etamax = 1.0;
                c.eta = eta;
            }
            wave_sync();
            rescale();
            return 2;
        }
        if (uni(c.q) > 1) {
            decrease_bdf();
            const int q = uni(c.q);
            wave_sync();
            if (lane == 0) {
                c.etamax = 1.0;
                c.eta = ETAMIN;
                c.L = q;
                c.q = q - 1;
                c.qwait = q;
            }
            wave_sync();
            rescale();
            return 2;
        }
        const double h = uni(c.h) * ETAMIN;
        wave_sync();
        if (lane == 0) {
            c.etamax = 1.0;
            c.eta = ETAMIN;
            c.h = h;
            c.hscale = h;
            c.qwait = LONG_WAIT;
            c.nscon = 0;
        }
        wave_sync();
        if (f_eval(zn[0], tempv)) return -1;
#pragma unroll
        for (int s = 0; s < NS; s++) zn[1][s] = h * tempv[s];
        if constexpr (kSens) {
            for (int p = 0; p < a.nsens; p++) {
                double z0[NS], f[NS];
                sload(p, 0, z0);
                if (sens_rhs(p, zn[0], z0, f)) return -1;
#pragma unroll
                for (int s = 0; s < NS; s++) f[s] = h * f[s];
                sstore(p, 1, f);
            }
        }
        return 2;
    }

    __device__ __forceinline__ void complete_step() {
        using namespace bdfc;
        const int q = uni(c.q);
        const double h = uni(c.h);
        double l[QMAX + 1];
#pragma unroll
        for (int j = 0; j <= QMAX; j++) l[j] = uni(c.l[j]);
#pragma unroll
        for (int j = 0; j <= QMAX; j++)
            if (j <= q)
#pragma unroll
                for (int s = 0; s < NS; s++) zn[j][s] += l[j] * acor[s];
        const int nst = uni(c.nst) + 1;
        const int qwait = uni(c.qwait) - 1;
        double tau[QMAX + 2];
#pragma unroll
        for (int i = 0; i < QMAX + 2; i++) tau[i] = uni(c.tau[i]);
#pragma unroll
        for (int i = QMAX; i >= 2; i--)
            if (i <= q) tau[i] = tau[i - 1];
        if (q == 1 && nst > 1) tau[2] = tau[1];
        tau[1] = h;
        if (qwait == 1 && q != QMAX) {
#pragma unroll
            for (int s = 0; s < NS; s++) zn[QMAX][s] = acor[s];
        }
        if constexpr (kSens) {
            for (int p = 0; p < a.nsens; p++) {
                double z[QMAX + 1][NS], acs[NS];
                sload_all(p, z);
                sload(p, 6, acs);
#pragma unroll
                for (int j = 0; j <= QMAX; j++)
                    if (j <= q)
#pragma unroll
                        for (int s = 0; s < NS; s++) z[j][s] += l[j] * acs[s];
                if (qwait == 1 && q != QMAX)
#pragma unroll
                    for (int s = 0; s < NS; s++) z[QMAX][s] = acs[s];
                sstore_all(p, z);
            }
        }
        const double tq5 = uni(c.tq[5]);
        wave_sync();
        if (lane == 0) {
            c.nst = nst;
            c.nscon += 1;
            c.hu = h;
#pragma unroll
            for (int i = 0; i < QMAX + 2; i++) c.tau[i] = tau[i];
            c.qwait = qwait;
            if (qwait == 1 && q != QMAX) c.saved_tq5 = tq5;
        }
        wave_sync();
    }

    __device__ __forceinline__ void set_eta(double eta, int qprime) {
        using namespace bdfc;
        double hprime;
        const double h = uni(c.h);
        if (eta < THRESH) {
            eta = 1.0;
            hprime = h;
        } else {
            eta = fmin(eta, uni(c.etamax));
            hprime = h * eta;
        }
        wave_sync();
        if (lane == 0) {
            c.eta = eta;
            c.hprime = hprime;
            c.qprime = qprime;
        }
        wave_sync();
    }

    __device__ __forceinline__ void prepare_next_step(double dsm) {
        using namespace bdfc;
        const int q = uni(c.q), L = uni(c.L);
        if (uni(c.etamax) == 1.0) {
            const int qwait = uni(c.qwait);
            wave_sync();
            if (lane == 0) {
                c.qwait = qwait > 2 ? qwait : 2;
                c.qprime = q;
                c.hprime = uni(c.h);
                c.eta = 1.0;
            }
            wave_sync();
            return;
        }
        const double etaq = 1.0 / (pow(BIAS2 * dsm, 1.0 / L) + ADDON);
        if (uni(c.qwait) != 0) {
            set_eta(etaq, q);
            return;
        }
        double etaqm1 = 0.0, etaqp1 = 0.0;
        if (q > 1) {
            double zq[NS];
            pick(q, zq);
            const double ddn = wrms<NS>(zq, ewt, a.neq_norm) * uni(c.tq[1]);
            etaqm1 = 1.0 / (pow(BIAS1 * ddn, 1.0 / q) + ADDON);
        }
        const double saved_tq5 = uni(c.saved_tq5);
        if (q != QMAX && saved_tq5 != 0.0) {
            const double cquot = (uni(c.tq[5]) / saved_tq5) * pow(uni(c.h) / uni(c.tau[2]), static_cast<double>(L));
#pragma unroll
            for (int s = 0; s < NS; s++) tempv[s] = -cquot * zn[QMAX][s] + acor[s];
            const double dup = wrms<NS>(tempv, ewt, a.neq_norm) * uni(c.tq[3]);
            etaqp1 = 1.0 / (pow(BIAS3 * dup, 1.0 / (L + 1)) + ADDON);
        }
        wave_sync();
        if (lane == 0) c.qwait = 2;
        wave_sync();
        const double etam = fmax(etaqm1, fmax(etaq, etaqp1));
        double eta;
        int qprime;
        if (etam < THRESH) {
            eta = 1.0;
            qprime = q;
        } else if (etam == etaq) {
            eta = etaq;
            qprime = q;
        } else if (etam == etaqm1) {
            eta = etaqm1;
            qprime = q - 1;
        } else {
            eta = etaqp1;
            qprime = q + 1;
#pragma unroll
            for (int s = 0; s < NS; s++) zn[QMAX][s] = acor[s];
            if constexpr (kSens) {
                for (int p = 0; p < a.nsens; p++) {
                    double acs[NS];
                    sload(p, 6, acs);
                    sstore(p, QMAX, acs);
                }
            }
        }
        set_eta(eta, qprime);
    }

    __device__ __forceinline__ int bdf_step() {
        using namespace bdfc;
        const double saved_t = uni(c.tn);
        int ncf = 0, nef = 0, nflag = FIRST_CALL;
        double dsm = 0.0;
        if (uni(c.nst) > 0 && uni(c.hprime) != uni(c.h)) {
            const int q = uni(c.q), qprime = uni(c.qprime);
            if (qprime != q) {
                if (qprime == q + 1) increase_bdf();
                else if (qprime == q - 1) decrease_bdf();
                wave_sync();
                if (lane == 0) {
                    c.q = qprime;
                    c.L = qprime + 1;
                    c.qwait = qprime + 1;
                }
                wave_sync();
            }
            rescale();
        }
        for (;;) {
            predict();
            set_bdf();
            {
                const double rl1 = 1.0 / uni(c.l[1]);
                const double gamma = uni(c.h) * rl1;
                const int nst = uni(c.nst);
                const double gammap = (nst == 0) ? gamma : uni(c.gammap);
                const double gamrat = (nst > 0) ? gamma / gammap : 1.0;
                wave_sync();
                if (lane == 0) {
                    c.rl1 = rl1;
                    c.gamma = gamma;
                    c.gammap = gammap;
                    c.gamrat = gamrat;
                }
                wave_sync();
            }
            const int ret = nls(nflag);
            if (ret != 0) {
                wave_sync();
                if (lane == 0) c.ncfn += 1;
                restore(saved_t);
                ncf++;
                if (ncf == MXNCF) return -3;
                wave_sync();
                if (lane == 0) {
                    c.etamax = 1.0;
                    c.eta = ETACF;
                }
                wave_sync();
                nflag = PREV_CONV_FAIL;
                rescale();
                continue;
            }
            const int e = do_error_test(nflag, saved_t, nef, dsm);
            if (e == 2) continue;
            if (e < 0) return -4;
            break;
        }
        if constexpr (kSens)
            if (sens_correct()) return -5;
        complete_step();
        prepare_next_step(dsm);
        const double etamax = (uni(c.nst) <= SMALL_NST) ? ETAMX2 : ETAMX3;
        wave_sync();
        if (lane == 0) c.etamax = etamax;
        wave_sync();
        return 0;
    }

    __device__ __forceinline__ int initial_h(double tout) {
        using namespace bdfc;
        const double tn = uni(c.tn);
        const double tdist = fabs(tout - tn);
        const double tround = UROUND * fmax(fabs(tn), fabs(tout));
        if (tdist < 2.0 * tround) return -1;
        const double hlb = HLB_FACTOR * tround;
        double hub_inv = 0.0;
#pragma unroll
        for (int s = 0; s < NS; s++)
            if (comp(s) < N) {
                const double t1 = HUB_FACTOR * fabs(zn[0][s]) + (rtol * fabs(zn[0][s]) + atol_of(s));
                hub_inv = fmax(hub_inv, fabs(zn[1][s]) / t1);
            }
        hub_inv = uni(wave_max(hub_inv));
        double hub = HUB_FACTOR * tdist;
        if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
        double hg = sqrt(hlb * hub);
        double h0;
        if (hub < hlb) {
            h0 = hg;
        } else {
            double hnew = hg;
            double yy[NS], ft[NS];
            for (int c1 = 1; c1 <= HIN_ITERS; c1++) {
                double yddnrm = 0.0;
                bool ok = false;
                for (int c2 = 1; c2 <= HIN_ITERS; c2++) {
#pragma unroll
                    for (int s = 0; s < NS; s++) yy[s] = hg * zn[1][s] + zn[0][s];
                    if (f_eval(yy, ft) == 0) {
#pragma unroll
                        for (int s = 0; s < NS; s++) ft[s] = (ft[s] - zn[1][s]) * (1.0 / hg);
                        yddnrm = wrms<NS>(ft, ewt, a.neq_norm);
                        ok = true;
                        break;
                    }
                    hg *= 0.2;
                }
                if (!ok) return -2;
                hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
                if (c1 == HIN_ITERS) break;
                const double hrat = hnew / hg;
                if (hrat > 0.5 && hrat < 2.0) break;
                if (c1 > 1 && hrat > 2.0) {
                    hnew = hg;
                    break;
                }
                hg = hnew;
            }
            h0 = H_BIAS * hnew;
            if (h0 < hlb) h0 = hlb;
            if (h0 > hub) h0 = hub;
        }
        wave_sync();
        if (lane == 0) c.h = h0;
        wave_sync();
        return 0;
    }

    __device__ __forceinline__ int start(double tout) {
        ewt_set(zn[0]);
        if (f_eval(zn[0], zn[1])) return -5;
        if (initial_h(tout)) return -6;
        const double h = uni(c.h);
#pragma unroll
        for (int s = 0; s < NS; s++) zn[1][s] *= h;
        if constexpr (kSens) {
            // s(t0) = 0 (CVodesIntegrator::sensInit); zS[1] = h (J 0 + df/dp)
            for (int p = 0; p < a.nsens; p++) {
                double zero[NS], f[NS];
#pragma unroll
                for (int s = 0; s < NS; s++) zero[s] = 0.0;
                for (int j = 0; j < kSensCols; j++) sstore(p, j, zero);
                if (sens_rhs(p, zn[0], zero, f)) return -5;
#pragma unroll
                for (int s = 0; s < NS; s++) f[s] = f[s] * h;
                sstore(p, 1, f);
            }
        }
        wave_sync();
        if (lane == 0) {
            c.hscale = h;
            c.hprime = h;
            c.started = 1;
        }
        wave_sync();
        return 0;
    }

    __device__ __forceinline__ int one_step() {
        if (uni(c.nst) > 0) ewt_set(zn[0]);
        const double nrm = wrms<NS>(zn[0], ewt, a.neq_norm);
        if (bdfc::UROUND * nrm > 1.0) return -10;
        return bdf_step();
    }

    // advance to tout; writes the interpolated state to out
    __device__ __forceinline__ int advance(double tout, double* out, int max_steps) {
        const double t_user = uni(c.t_user);
        if (tout == t_user) return 0;
        if (tout < t_user) return -20;
        if (!uni(c.started)) {
            const int r = start(tout);
            if (r) return r;
        }
        int nsteps = 0;
        while (uni(c.tn) < tout) {
            if (nsteps >= max_steps) return -22;
            const int ret = one_step();
            if (ret) return ret;
            nsteps++;
        }
        const double sarg = (tout - uni(c.tn)) / uni(c.h);
        const int q = uni(c.q);
        double v[NS];
        pick(q, v);
#pragma unroll
        for (int j = bdfc::QMAX - 1; j >= 0; j--)
            if (j < q)
#pragma unroll
                for (int s = 0; s < NS; s++) v[s] = zn[j][s] + sarg * v[s];
#pragma unroll
        for (int s = 0; s < NS; s++)
            if (comp(s) < N) out[comp(s)] = v[s];
        if constexpr (kSens) {
            // CVodeGetSensDky(tout, 0) into column 7
            for (int p = 0; p < a.nsens; p++) {
                double z[bdfc::QMAX + 1][NS];
                sload_all(p, z);
                pick_of(z, q, v);
#pragma unroll
                for (int j = bdfc::QMAX - 1; j >= 0; j--)
                    if (j < q)
#pragma unroll
                        for (int s = 0; s < NS; s++) v[s] = z[j][s] + sarg * v[s];
                sstore(p, 7, v);
            }
        }
        wave_sync();
        if (lane == 0) c.t_user = tout;
        wave_sync();
        return 0;
    }
};

// LDS per wavefront: state broadcast, species/reaction scratch, control block,
// then the linear-system area (LDS policy only).
__host__ __device__ inline size_t bdf_scratch_doubles(int K, int ns) {
    return static_cast<size_t>(kWave) * ns * 4 + (K + 1) + (K + 1) + K + 2 * K + kCtlDoubles;
}

template <int NS, class LS, typename MT, bool kSens = false, bool kMole = false>
__device__ __forceinline__ void bdf_advance_body(const BdfArgs& a) {
    const DevMech& m = a.mech;
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
    double* base = lds_base() + off;
    Ctl& c = *reinterpret_cast<Ctl*>(base);
    base += kCtlDoubles;
    Integrator<NS, LS, MT, kSens, kMole> it(m, a, c);
    it.S = scratch_view(L);
    it.L = L;
    it.lane = lane_id();
    it.N = kMole ? m.K + 1 + (m.const_volume ? 1 : 0) : m.N;
    it.rtol = a.rtol;
    it.atol = a.atol;
    it.ls.init(base, a.ld);
    const int nwork = a.n;
    for (;;) {
        int r = 0;
        if (it.lane == 0) r = atomicAdd(a.queue, 1);
        r = uni(__shfl(r, 0, kWave));
        if (r >= nwork) break;
        const int rid = a.order ? uni(a.order[r]) : r;
        if (a.work_flag && uni(a.work_flag[rid]) != a.work_value) continue;
        double* rec = a.rec + static_cast<size_t>(rid) * a.rec_stride;
        const double status = uni(rec[7 * kWave * NS + C_STATUS]);
        if (status < 0.0) continue;   // failed earlier: leave untouched
        it.load(static_cast<size_t>(rid));
        const double tout = uni(a.tout[rid]);
        const int ret = it.advance(tout, rec + 6 * kWave * NS, a.max_steps);
        it.store(ret < 0 ? ret : 0);
        wave_sync();
    }
}

}  // namespace cg
