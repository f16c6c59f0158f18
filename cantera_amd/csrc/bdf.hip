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
//
// Precision modes: plain (FP64 everywhere) and mixed (iteration matrix, its
// LU and the triangular solves in FP32; Newton residual, state, error test and
// step control in FP64 -- an inexact Newton corrected by the FP64 residual).
#include <hip/hip_runtime.h>

#include "bdf.hpp"
#include "wave_jac.hpp"
#include "wave_linalg.hpp"

namespace cg {

namespace {

constexpr int QMAX = 5;
constexpr int MXSTEP = 20000;
constexpr int MAXCOR = 3;
constexpr int MXNEF = 7;
constexpr int MXNCF = 10;
constexpr int MXNEF1 = 3;
constexpr int SMALL_NEF = 2;
constexpr int LONG_WAIT = 10;
constexpr int SMALL_NST = 10;
constexpr double NLSCOEF = 0.1;
constexpr double ETAMX1 = 10000.0, ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2, ETAMIN = 0.1, ETACF = 0.25;
constexpr double ADDON = 1e-6, BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0, THRESH = 1.5;
constexpr double CRDOWN = 0.3, DGMAX_LSETUP = 0.3, RDIV = 2.0, DGMAX_JBAD = 0.2;
constexpr int MSBP = 20, MSBJ = 51;
constexpr double HLB_FACTOR = 100.0, HUB_FACTOR = 0.1, H_BIAS = 0.5;
constexpr int HIN_ITERS = 4;
constexpr double UROUND = 2.220446049250313e-16;

enum { FIRST_CALL = 0, PREV_CONV_FAIL = 1, PREV_ERR_FAIL = 2 };
enum { NO_FAILURES = 0, FAIL_BAD_J = 1, FAIL_OTHER = 2 };

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

template <int NS, typename MT>
struct Integrator {
    const DevMech& m;
    const BdfArgs& a;
    WaveScratch S;
    MT* A;            // LDS iteration matrix
    int LD;
    int* pmap;        // LDS
    MT* bv;           // LDS solve scratch (matrix precision)
    int N, lane;
    double P, rtol, atol;
    // Nordsieck history and work vectors (lane-distributed)
    double zn[QMAX + 1][NS];
    double ewt[NS], acor[NS], y[NS], ftemp[NS], tempv[NS];
    // control (wave-uniform)
    double tn, h, hprime, hscale, eta, etamax, hu;
    double tau[QMAX + 2], tq[6], l[QMAX + 1];
    double rl1, gamma, gammap, gamrat, crate, delp, acnrm, saved_tq5;
    int q, qprime, qwait, L, nst, nstlp, nstlj, nscon, started;
    double t_user;
    double nfe, nje, nsetups, nni, ncfn, netf, nfe_dq;
    int convfail, jcur, lu_dirty;
    double* rec;
    size_t rid;

    __device__ __forceinline__ Integrator(const DevMech& m_, const BdfArgs& a_) : m(m_), a(a_) {}

    __device__ __forceinline__ int comp(int s) const { return lane + kWave * s; }

    __device__ __forceinline__ void load(size_t r) {
        rid = r;
        rec = a.rec + r * a.rec_stride;
        const int NP = kWave * NS;
#pragma unroll
        for (int j = 0; j <= QMAX; j++)
#pragma unroll
            for (int s = 0; s < NS; s++) zn[j][s] = rec[j * NP + comp(s)];
        const double* c = rec + 7 * NP;
        tn = c[C_TN]; h = c[C_H]; hprime = c[C_HPRIME]; hscale = c[C_HSCALE]; eta = c[C_ETA];
        etamax = c[C_ETAMAX]; hu = c[C_HU];
#pragma unroll
        for (int i = 0; i < QMAX + 2; i++) tau[i] = c[C_TAU + i];
#pragma unroll
        for (int i = 0; i < 6; i++) tq[i] = c[C_TQ + i];
        rl1 = c[C_RL1]; gamma = c[C_GAMMA]; gammap = c[C_GAMMAP]; gamrat = c[C_GAMRAT]; crate = c[C_CRATE];
        delp = c[C_DELP]; acnrm = c[C_ACNRM]; saved_tq5 = c[C_SAVED_TQ5];
        q = static_cast<int>(c[C_Q]); qprime = static_cast<int>(c[C_QPRIME]); qwait = static_cast<int>(c[C_QWAIT]);
        L = static_cast<int>(c[C_L]); nst = static_cast<int>(c[C_NST]); nstlp = static_cast<int>(c[C_NSTLP]);
        nstlj = static_cast<int>(c[C_NSTLJ]); nscon = static_cast<int>(c[C_NSCON]);
        started = static_cast<int>(c[C_STARTED]);
        t_user = c[C_TUSER];
        nfe = c[C_NFE]; nje = c[C_NJE]; nsetups = c[C_NSETUPS]; nni = c[C_NNI]; ncfn = c[C_NCFN]; netf = c[C_NETF];
        nfe_dq = c[C_NFE_DQ];
        P = c[C_P];
        jcur = 0;
        lu_dirty = 0;
        if (c[C_HAVE_LU] != 0.0) {
            // restore the factored iteration matrix
            const MT* src = a.template lu<MT>() + r * a.lu_stride;
            for (int i = 0; i < N; i++)
#pragma unroll
                for (int s = 0; s < NS; s++)
                    if (comp(s) < N) A[i * LD + comp(s)] = src[i * LD + comp(s)];
            const int* ps = a.piv_save + r * a.piv_stride;
#pragma unroll
            for (int s = 0; s < NS; s++)
                if (comp(s) < N) pmap[comp(s)] = ps[comp(s)];
            __syncthreads();
        }
    }

    __device__ __forceinline__ void store(int status) {
        const int NP = kWave * NS;
#pragma unroll
        for (int j = 0; j <= QMAX; j++)
#pragma unroll
            for (int s = 0; s < NS; s++) rec[j * NP + comp(s)] = zn[j][s];
        if (lane == 0) {
            double* c = rec + 7 * NP;
            c[C_TN] = tn; c[C_H] = h; c[C_HPRIME] = hprime; c[C_HSCALE] = hscale; c[C_ETA] = eta;
            c[C_ETAMAX] = etamax; c[C_HU] = hu;
            for (int i = 0; i < QMAX + 2; i++) c[C_TAU + i] = tau[i];
            for (int i = 0; i < 6; i++) c[C_TQ + i] = tq[i];
            c[C_RL1] = rl1; c[C_GAMMA] = gamma; c[C_GAMMAP] = gammap; c[C_GAMRAT] = gamrat; c[C_CRATE] = crate;
            c[C_DELP] = delp; c[C_ACNRM] = acnrm; c[C_SAVED_TQ5] = saved_tq5;
            c[C_Q] = q; c[C_QPRIME] = qprime; c[C_QWAIT] = qwait; c[C_L] = L; c[C_NST] = nst;
            c[C_NSTLP] = nstlp; c[C_NSTLJ] = nstlj; c[C_NSCON] = nscon; c[C_STARTED] = started;
            c[C_TUSER] = t_user;
            c[C_NFE] = nfe; c[C_NJE] = nje; c[C_NSETUPS] = nsetups; c[C_NNI] = nni; c[C_NCFN] = ncfn;
            c[C_NETF] = netf; c[C_NFE_DQ] = nfe_dq;
            c[C_STATUS] = status;
            if (lu_dirty) c[C_HAVE_LU] = 1.0;
        }
        if (lu_dirty) {
            MT* dst = a.template lu<MT>() + rid * a.lu_stride;
            for (int i = 0; i < N; i++)
#pragma unroll
                for (int s = 0; s < NS; s++)
                    if (comp(s) < N) dst[i * LD + comp(s)] = A[i * LD + comp(s)];
            int* ps = a.piv_save + rid * a.piv_stride;
#pragma unroll
            for (int s = 0; s < NS; s++)
                if (comp(s) < N) ps[comp(s)] = pmap[comp(s)];
        }
    }

    __device__ __forceinline__ int f_eval(const double (&yy)[NS], double (&ff)[NS]) {
        nfe += 1;
        GasState g;
        SpeciesRegs<NS> sr;
        RhsScalars rs;
        reactor_rhs<NS>(m, S, P, yy, ff, g, sr, rs);
        int bad = 0;
#pragma unroll
        for (int s = 0; s < NS; s++)
            if (comp(s) < N && !isfinite(ff[s])) bad = 1;
        return __any(bad) ? 1 : 0;
    }

    __device__ __forceinline__ void ewt_set(const double (&yy)[NS]) {
#pragma unroll
        for (int s = 0; s < NS; s++) ewt[s] = (comp(s) < N) ? 1.0 / (rtol * fabs(yy[s]) + atol) : 0.0;
    }

    // ---- linear system setup: Jacobian reuse decision, M = I - gamma J, LU ----
    __device__ __forceinline__ int lsetup() {
        const double dgamma = fabs(gamma / gammap - 1.0);
        const bool jbad = (nst == 0) || (nst >= nstlj + MSBJ) ||
                          ((convfail == FAIL_BAD_J) && (dgamma < DGMAX_JBAD)) || (convfail == FAIL_OTHER);
        double* Js = a.j_save + rid * a.j_stride;
        if (!jbad) {
            jcur = 0;
            for (int i = 0; i < N; i++)
#pragma unroll
                for (int s = 0; s < NS; s++)
                    if (comp(s) < N) A[i * LD + comp(s)] = static_cast<MT>(Js[i * N + comp(s)]);
        } else {
            nje += 1;
            nstlj = nst;
            jcur = 1;
            nfe_dq += 2;
            if constexpr (sizeof(MT) == 8) {
                reactor_jacobian<NS, MT>(m, S, P, zn[0], ftemp, A, LD);
                for (int i = 0; i < N; i++)
#pragma unroll
                    for (int s = 0; s < NS; s++)
                        if (comp(s) < N) Js[i * N + comp(s)] = A[i * LD + comp(s)];
            } else {
                // build in FP64 in the save buffer via the FP32 matrix?  The
                // Jacobian scatter accumulates in the matrix precision; the
                // saved copy keeps that precision widened to FP64.
                reactor_jacobian<NS, MT>(m, S, P, zn[0], ftemp, A, LD);
                for (int i = 0; i < N; i++)
#pragma unroll
                    for (int s = 0; s < NS; s++)
                        if (comp(s) < N) Js[i * N + comp(s)] = static_cast<double>(A[i * LD + comp(s)]);
            }
        }
        __syncthreads();
        const MT mg = static_cast<MT>(-gamma);
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int j = comp(s);
            if (j < N)
                for (int i = 0; i < N; i++) {
                    MT v = mg * A[i * LD + j];
                    if (i == j) v += MT(1);
                    A[i * LD + j] = v;
                }
        }
        __syncthreads();
        nsetups += 1;
        lu_dirty = 1;
        if (!lu_factor<NS, MT>(A, LD, N, pmap)) return 1;
        gamrat = 1.0;
        gammap = gamma;
        crate = 1.0;
        nstlp = nst;
        return 0;
    }

    __device__ __forceinline__ void set_bdf() {
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
                    hsum += tau[j - 1];
                    xi_inv = h / hsum;
                    alpha0 -= 1.0 / j;
#pragma unroll
                    for (int i = j; i >= 1; i--) l[i] += l[i - 1] * xi_inv;
                }
            }
            alpha0 -= 1.0 / q;
            xistar_inv = -l[1] - alpha0;
            double tqm1 = 0.0;
#pragma unroll
            for (int j = 1; j <= QMAX; j++)
                if (j == q - 1 + 0 && j >= 1) tqm1 = tau[j];
            hsum += tqm1;
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
        tq[2] = fabs(A1 / (alpha0 * A2));
        tq[5] = fabs(A2 * xistar_inv / (lq * xi_inv));
        if (qwait == 1) {
            if (q > 1) {
                const double C = xistar_inv / lq;
                const double A3 = alpha0 + 1.0 / q;
                const double A4 = alpha0_hat + xi_inv;
                const double Cpinv = (1.0 - A4 + A3) / A3;
                tq[1] = fabs(C * Cpinv);
            } else {
                tq[1] = 1.0;
            }
            double tq_ = 0.0;
#pragma unroll
            for (int j = 1; j <= QMAX + 1; j++)
                if (j == q) tq_ = tau[j];
            hsum += tq_;
            xi_inv = h / hsum;
            const double A5 = alpha0 - 1.0 / (q + 1);
            const double A6 = alpha0_hat - xi_inv;
            const double Cppinv = (1.0 - A6 + A5) / A2;
            tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
        }
        tq[4] = NLSCOEF / tq[2];
    }

    __device__ __forceinline__ void predict() {
        tn += h;
#pragma unroll
        for (int k = 1; k <= QMAX; k++)
#pragma unroll
            for (int j = QMAX; j >= 1; j--)
                if (k <= q && j >= k && j <= q)
#pragma unroll
                    for (int s = 0; s < NS; s++) zn[j - 1][s] += zn[j][s];
    }
    __device__ __forceinline__ void restore(double saved_t) {
        tn = saved_t;
#pragma unroll
        for (int k = 1; k <= QMAX; k++)
#pragma unroll
            for (int j = QMAX; j >= 1; j--)
                if (k <= q && j >= k && j <= q)
#pragma unroll
                    for (int s = 0; s < NS; s++) zn[j - 1][s] -= zn[j][s];
    }
    __device__ __forceinline__ void rescale() {
        double factor = eta;
#pragma unroll
        for (int j = 1; j <= QMAX; j++) {
            if (j <= q) {
#pragma unroll
                for (int s = 0; s < NS; s++) zn[j][s] *= factor;
                factor *= eta;
            }
        }
        h = hscale * eta;
        hscale = h;
        nscon = 0;
    }

    __device__ __forceinline__ void increase_bdf() {
        double ll[QMAX + 1];
#pragma unroll
        for (int i = 0; i <= QMAX; i++) ll[i] = 0.0;
        double alpha1, prod, xiold, alpha0, hsum, xi;
        ll[2] = alpha1 = prod = xiold = 1.0;
        alpha0 = -1.0;
        hsum = hscale;
        if (q > 1) {
#pragma unroll
            for (int j = 1; j < QMAX; j++) {
                if (j < q) {
                    hsum += tau[j + 1];
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
        // zn[L] = A1 * zn[QMAX] with L = q + 1
        double zL[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) zL[s] = A1 * zn[QMAX][s];
#pragma unroll
        for (int j = 2; j <= QMAX; j++)
            if (j == q + 1)
#pragma unroll
                for (int s = 0; s < NS; s++) zn[j][s] = zL[s];
#pragma unroll
        for (int j = 2; j <= QMAX; j++)
            if (j <= q)
#pragma unroll
                for (int s = 0; s < NS; s++) zn[j][s] += ll[j] * zL[s];
    }
    __device__ __forceinline__ void decrease_bdf() {
        double ll[QMAX + 1];
#pragma unroll
        for (int i = 0; i <= QMAX; i++) ll[i] = 0.0;
        ll[2] = 1.0;
        double hsum = 0.0, xi;
#pragma unroll
        for (int j = 1; j <= QMAX - 2; j++) {
            if (j <= q - 2) {
                hsum += tau[j];
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
    }
    __device__ __forceinline__ void pick(int j, double (&out)[NS]) {
#pragma unroll
        for (int s = 0; s < NS; s++) out[s] = zn[0][s];
#pragma unroll
        for (int t = 1; t <= QMAX; t++)
            if (t == j)
#pragma unroll
                for (int s = 0; s < NS; s++) out[s] = zn[t][s];
    }

    // modified Newton iteration; 0 ok, 1 recoverable failure, -1 fatal
    __device__ __forceinline__ int nls(int nflag) {
        convfail = (nflag == FIRST_CALL || nflag == PREV_ERR_FAIL) ? NO_FAILURES : FAIL_OTHER;
        bool callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (nst == 0) ||
                         (nst >= nstlp + MSBP) || (fabs(gamrat - 1.0) > DGMAX_LSETUP);
        bool jbad = false;
        double delta[NS];
        for (;;) {
#pragma unroll
            for (int s = 0; s < NS; s++) {
                acor[s] = 0.0;
                y[s] = zn[0][s];
            }
            if (f_eval(y, ftemp)) return 1;
#pragma unroll
            for (int s = 0; s < NS; s++) delta[s] = -gamma * ftemp[s] + (rl1 * zn[1][s] + acor[s]);
            if (callSetup) {
                if (jbad) convfail = FAIL_BAD_J;
                if (lsetup()) return 1;
                callSetup = false;
            }
            int m_it = 0;
            int status;
            for (;;) {
                nni += 1;
                if constexpr (sizeof(MT) == 8) {
#pragma unroll
                    for (int s = 0; s < NS; s++) delta[s] = -delta[s];
                    lu_solve<NS, MT, double>(A, LD, N, pmap, reinterpret_cast<double*>(bv), delta);
                } else {
                    float bf[NS];
#pragma unroll
                    for (int s = 0; s < NS; s++) bf[s] = static_cast<float>(-delta[s]);
                    lu_solve<NS, MT, float>(A, LD, N, pmap, reinterpret_cast<float*>(bv), bf);
#pragma unroll
                    for (int s = 0; s < NS; s++) delta[s] = static_cast<double>(bf[s]);
                }
                if (gamrat != 1.0) {
                    const double sc = 2.0 / (1.0 + gamrat);
#pragma unroll
                    for (int s = 0; s < NS; s++) delta[s] *= sc;
                }
#pragma unroll
                for (int s = 0; s < NS; s++) acor[s] += delta[s];
                const double del = wrms<NS>(delta, ewt, a.neq_norm);
                if (m_it > 0) crate = fmax(CRDOWN * crate, del / delp);
                const double dcon = del * fmin(1.0, crate) / tq[4];
                if (dcon <= 1.0) {
                    acnrm = (m_it == 0) ? del : wrms<NS>(acor, ewt, a.neq_norm);
                    status = 0;
                    break;
                }
                if (m_it >= 1 && del > RDIV * delp) { status = 1; break; }
                delp = del;
                m_it++;
                if (m_it >= MAXCOR) { status = 1; break; }
#pragma unroll
                for (int s = 0; s < NS; s++) y[s] = zn[0][s] + acor[s];
                if (f_eval(y, ftemp)) { status = 1; break; }
#pragma unroll
                for (int s = 0; s < NS; s++) delta[s] = -gamma * ftemp[s] + (rl1 * zn[1][s] + acor[s]);
            }
            if (status == 0) {
                jcur = 0;
#pragma unroll
                for (int s = 0; s < NS; s++) y[s] = zn[0][s] + acor[s];
                return 0;
            }
            if (!jcur) {
                callSetup = true;
                jbad = true;
                continue;
            }
            return 1;
        }
    }

    // 0 pass, 2 retry, -1 fail
    __device__ __forceinline__ int do_error_test(int& nflag, double saved_t, int& nef, double& dsm) {
        dsm = acnrm * tq[2];
        if (dsm <= 1.0) return 0;
        nef++;
        netf += 1;
        nflag = PREV_ERR_FAIL;
        restore(saved_t);
        if (nef == MXNEF) return -1;
        etamax = 1.0;
        if (nef <= MXNEF1) {
            eta = 1.0 / (pow(BIAS2 * dsm, 1.0 / L) + ADDON);
            eta = fmax(ETAMIN, eta);
            if (nef >= SMALL_NEF) eta = fmin(eta, ETAMXF);
            rescale();
            return 2;
        }
        if (q > 1) {
            eta = ETAMIN;
            decrease_bdf();
            L = q;
            q--;
            qwait = L;
            rescale();
            return 2;
        }
        eta = ETAMIN;
        h *= eta;
        hscale = h;
        qwait = LONG_WAIT;
        nscon = 0;
        if (f_eval(zn[0], tempv)) return -1;
#pragma unroll
        for (int s = 0; s < NS; s++) zn[1][s] = h * tempv[s];
        return 2;
    }

    __device__ __forceinline__ void complete_step() {
        nst++;
        nscon++;
        hu = h;
#pragma unroll
        for (int i = QMAX; i >= 2; i--)
            if (i <= q) tau[i] = tau[i - 1];
        if (q == 1 && nst > 1) tau[2] = tau[1];
        tau[1] = h;
#pragma unroll
        for (int j = 0; j <= QMAX; j++)
            if (j <= q)
#pragma unroll
                for (int s = 0; s < NS; s++) zn[j][s] += l[j] * acor[s];
        qwait--;
        if (qwait == 1 && q != QMAX) {
#pragma unroll
            for (int s = 0; s < NS; s++) zn[QMAX][s] = acor[s];
            saved_tq5 = tq[5];
        }
    }

    __device__ __forceinline__ void set_eta() {
        if (eta < THRESH) {
            eta = 1.0;
            hprime = h;
        } else {
            eta = fmin(eta, etamax);
            hprime = h * eta;
        }
    }

    __device__ __forceinline__ void prepare_next_step(double dsm) {
        if (etamax == 1.0) {
            qwait = qwait > 2 ? qwait : 2;
            qprime = q;
            hprime = h;
            eta = 1.0;
            return;
        }
        const double etaq = 1.0 / (pow(BIAS2 * dsm, 1.0 / L) + ADDON);
        if (qwait != 0) {
            eta = etaq;
            qprime = q;
            set_eta();
            return;
        }
        qwait = 2;
        double etaqm1 = 0.0, etaqp1 = 0.0;
        if (q > 1) {
            double zq[NS];
            pick(q, zq);
            const double ddn = wrms<NS>(zq, ewt, a.neq_norm) * tq[1];
            etaqm1 = 1.0 / (pow(BIAS1 * ddn, 1.0 / q) + ADDON);
        }
        if (q != QMAX && saved_tq5 != 0.0) {
            const double cquot = (tq[5] / saved_tq5) * pow(h / tau[2], static_cast<double>(L));
#pragma unroll
            for (int s = 0; s < NS; s++) tempv[s] = -cquot * zn[QMAX][s] + acor[s];
            const double dup = wrms<NS>(tempv, ewt, a.neq_norm) * tq[3];
            etaqp1 = 1.0 / (pow(BIAS3 * dup, 1.0 / (L + 1)) + ADDON);
        }
        const double etam = fmax(etaqm1, fmax(etaq, etaqp1));
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
        }
        set_eta();
    }

    __device__ __forceinline__ int bdf_step() {
        const double saved_t = tn;
        int ncf = 0, nef = 0, nflag = FIRST_CALL;
        double dsm = 0.0;
        if (nst > 0 && hprime != h) {
            if (qprime != q) {
                if (qprime == q + 1) increase_bdf();
                else if (qprime == q - 1) decrease_bdf();
                q = qprime;
                L = q + 1;
                qwait = L;
            }
            rescale();
        }
        for (;;) {
            predict();
            set_bdf();
            rl1 = 1.0 / l[1];
            gamma = h * rl1;
            if (nst == 0) gammap = gamma;
            gamrat = (nst > 0) ? gamma / gammap : 1.0;
            const int ret = nls(nflag);
            if (ret != 0) {
                ncfn += 1;
                restore(saved_t);
                if (ret < 0) return -2;
                ncf++;
                etamax = 1.0;
                if (ncf == MXNCF) return -3;
                eta = ETACF;
                nflag = PREV_CONV_FAIL;
                rescale();
                continue;
            }
            const int e = do_error_test(nflag, saved_t, nef, dsm);
            if (e == 2) continue;
            if (e < 0) return -4;
            break;
        }
        complete_step();
        prepare_next_step(dsm);
        etamax = (nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
        return 0;
    }

    __device__ __forceinline__ int initial_h(double tout) {
        const double tdist = fabs(tout - tn);
        const double tround = UROUND * fmax(fabs(tn), fabs(tout));
        if (tdist < 2.0 * tround) return -1;
        const double hlb = HLB_FACTOR * tround;
        double hub_inv = 0.0;
#pragma unroll
        for (int s = 0; s < NS; s++)
            if (comp(s) < N) {
                const double t1 = HUB_FACTOR * fabs(zn[0][s]) + (rtol * fabs(zn[0][s]) + atol);
                hub_inv = fmax(hub_inv, fabs(zn[1][s]) / t1);
            }
        hub_inv = uni(wave_max(hub_inv));
        double hub = HUB_FACTOR * tdist;
        if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
        double hg = sqrt(hlb * hub);
        if (hub < hlb) {
            h = hg;
            return 0;
        }
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
        double h0 = H_BIAS * hnew;
        if (h0 < hlb) h0 = hlb;
        if (h0 > hub) h0 = hub;
        h = h0;
        return 0;
    }

    __device__ __forceinline__ int start(double tout) {
        ewt_set(zn[0]);
        if (f_eval(zn[0], zn[1])) return -5;
        if (initial_h(tout)) return -6;
#pragma unroll
        for (int s = 0; s < NS; s++) zn[1][s] *= h;
        hscale = h;
        hprime = h;
        started = 1;
        return 0;
    }

    __device__ __forceinline__ int one_step() {
        if (nst > 0) ewt_set(zn[0]);
        const double nrm = wrms<NS>(zn[0], ewt, a.neq_norm);
        if (UROUND * nrm > 1.0) return -10;
        return bdf_step();
    }

    // advance to tout; writes the interpolated state to out
    __device__ __forceinline__ int advance(double tout, double* out, int max_steps) {
        if (tout == t_user) return 0;
        if (tout < t_user) return -20;
        if (!started) {
            const int r = start(tout);
            if (r) return r;
        }
        int nsteps = 0;
        while (tn < tout) {
            if (nsteps >= max_steps) return -22;
            const int ret = one_step();
            if (ret) return ret;
            nsteps++;
        }
        const double sarg = (tout - tn) / h;
        double v[NS];
        pick(q, v);
#pragma unroll
        for (int j = QMAX - 1; j >= 0; j--)
            if (j < q)
#pragma unroll
                for (int s = 0; s < NS; s++) v[s] = zn[j][s] + sarg * v[s];
#pragma unroll
        for (int s = 0; s < NS; s++)
            if (comp(s) < N) out[comp(s)] = v[s];
        t_user = tout;
        return 0;
    }
};

template <int NS, typename MT>
__global__ __launch_bounds__(64) void bdf_advance_kernel(BdfArgs a) {
    extern __shared__ double lds_raw[];
    const DevMech& m = a.mech;
    Integrator<NS, MT> it(m, a);
    it.lane = lane_id();
    it.N = m.N;
    it.rtol = a.rtol;
    it.atol = a.atol;
    it.LD = a.ld;
    double* base = lds_raw;
    it.S.yb = base;                 base += kWave * NS;
    it.S.C = base;                  base += m.K;
    it.S.MU = base;                 base += m.K;
    it.S.ROP = base;                base += m.I;
    it.S.SUB = base;                base += m.nsub;
    it.S.W2 = base;                 base += 2 * m.K;
    it.bv = reinterpret_cast<MT*>(base); base += kWave * NS;
    it.pmap = reinterpret_cast<int*>(base); base += (kWave * NS) / 2;
    it.A = reinterpret_cast<MT*>(base);
    for (;;) {
        int r = 0;
        if (it.lane == 0) r = atomicAdd(a.queue, 1);
        r = __shfl(r, 0, kWave);
        r = uni(r);
        if (r >= a.n) break;
        const int rid = a.order ? a.order[r] : r;
        double* rec = a.rec + static_cast<size_t>(rid) * a.rec_stride;
        const double status = rec[7 * kWave * NS + C_STATUS];
        if (status < 0.0) continue;   // failed earlier: leave untouched
        it.load(static_cast<size_t>(rid));
        const double tout = a.tout[rid];
        const int ret = it.advance(tout, rec + 6 * kWave * NS, a.max_steps);
        it.store(ret < 0 ? ret : 0);
        __syncthreads();
    }
}

__global__ void reset_queue_kernel(int* q) { *q = 0; }

}  // namespace

size_t bdf_lds_bytes(const DevMech& m, int ns, int ld, bool fp32) {
    const size_t es = fp32 ? 4 : 8;
    size_t d = static_cast<size_t>(kWave) * ns + m.K + m.K + m.I + m.nsub + 2 * m.K;
    size_t bytes = d * 8;
    bytes += kWave * ns * 8;                    // bv (sized for doubles)
    bytes += (kWave * ns / 2) * 8;              // pmap
    bytes += static_cast<size_t>(m.N) * ld * es;  // matrix
    return bytes;
}

hipError_t bdf_launch(const BdfArgs& a, int ns, bool fp32, int grid, size_t lds, hipStream_t stream) {
    hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, stream, a.queue);
    if (ns == 1) {
        if (fp32) hipLaunchKernelGGL((bdf_advance_kernel<1, float>), dim3(grid), dim3(64), lds, stream, a);
        else hipLaunchKernelGGL((bdf_advance_kernel<1, double>), dim3(grid), dim3(64), lds, stream, a);
    } else if (ns == 2) {
        if (fp32) hipLaunchKernelGGL((bdf_advance_kernel<2, float>), dim3(grid), dim3(64), lds, stream, a);
        else hipLaunchKernelGGL((bdf_advance_kernel<2, double>), dim3(grid), dim3(64), lds, stream, a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

const void* bdf_kernel_ptr(int ns, bool fp32) {
    if (ns == 1) return fp32 ? reinterpret_cast<const void*>(&bdf_advance_kernel<1, float>)
                             : reinterpret_cast<const void*>(&bdf_advance_kernel<1, double>);
    return fp32 ? reinterpret_cast<const void*>(&bdf_advance_kernel<2, float>)
                : reinterpret_cast<const void*>(&bdf_advance_kernel<2, double>);
}

}  // namespace cg
