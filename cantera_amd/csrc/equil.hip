// Batched chemical equilibrium of ideal-gas states at fixed (h, P) or (T, P)
// (ThermoPhase::equilibrate("HP" / "TP"), src/thermo/ThermoPhase.cpp, whose
// default solver for a single ideal-gas phase is the element-potential method
// of ChemEquil, src/equil/ChemEquil.cpp:329-337).
//
// Method: Gibbs-energy minimisation with element potentials, the Newton
// iteration of Gordon & McBride (NASA RP-1311, 1994, eqs. 2.24-2.27 for a gas
// phase) with its step control (eqs. 3.1-3.3).  Unknowns per iteration are the
// element potentials pi_e, dln n (n = total kmol per kg) and, for HP, dln T;
// species follow as dln n_j = -mu_j/RT + sum_e a_ej pi_e + dln n + h_j dln T.
// Elements absent from a state, and the species containing them, are left
// out.  Convergence: every correction below 1e-12 (species: relative to the
// total).
//
// Layout: inputs and outputs species-major as the batched rates (T[n], P[n],
// Y[k*n + s]).  One wavefront per state: lane k owns species k (slot s for
// k = lane + 64 s); the sums of the Newton matrix are wavefront reductions,
// the (E+2) x (E+2) system is solved by one lane in LDS.  Work per iteration
// is O(K E^2), HBM traffic per state the state in and out (compute-bound,
// FP64 VALU).
#include <hip/hip_runtime.h>

#include "bdf.hpp"
#include "wave_kinetics.hpp"

namespace cg {

namespace {

constexpr int kEqMaxE = 8;
constexpr int kEqMaxNu = kEqMaxE + 2;
constexpr double kSize = 18.420681;       // ln(1e8): trace-species threshold of eq. 3.2
constexpr double kTraceLn = 9.2103404;    // ln(1e4)

template <int NS>
__global__ __launch_bounds__(64) void equil_kernel(DevMech m, long n, int hp, const double* __restrict__ Tin,
                                                   const double* __restrict__ Pin, const double* __restrict__ Yin,
                                                   double* __restrict__ Tout, double* __restrict__ Yout,
                                                   int* __restrict__ info, int max_iter) {
    __shared__ double As[kEqMaxNu * kEqMaxNu];
    __shared__ double xs[kEqMaxNu];
    const int lane = lane_id();
    const int K = m.K, E = m.E;
    const XcdStates xst(n);
    for (long st = xst.lo; st < xst.hi; st += xst.step) {
        const double T0 = uni(Tin[st]);
        const double P = uni(Pin[st]);
        double y[NS], tt[6];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int k = lane + kWave * s;
            y[s] = (k < K) ? Yin[static_cast<size_t>(k) * n + st] : 0.0;
        }
        // element amounts [kmol/kg] and specific enthalpy of the input state
        tt[0] = T0; tt[1] = T0 * T0; tt[2] = tt[1] * T0; tt[3] = tt[2] * T0; tt[4] = 1.0 / T0; tt[5] = log(T0);
        double hl = 0.0;
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int k = lane + kWave * s;
            if (k < K) {
                double cp, h, sR, g;
                nasa7(m.nasa + 16 * k, T0, tt, cp, h, sR, g);
                hl += y[s] * m.rmw[k] * h;
            }
        }
        const double h0 = uni(wave_sum(hl)) * kGasConstant * T0;
        double b0[kEqMaxE];
        double bmax = 0.0;
#pragma unroll
        for (int e = 0; e < kEqMaxE; e++) {
            double v = 0.0;
            if (e < E) {
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    const int k = lane + kWave * s;
                    if (k < K) v += m.comp[k * E + e] * y[s] * m.rmw[k];
                }
                v = uni(wave_sum(v));
            }
            b0[e] = v;
            bmax = fmax(bmax, v);
        }
        unsigned emask = 0;
#pragma unroll
        for (int e = 0; e < kEqMaxE; e++)
            if (e < E && b0[e] > 1e-14 * bmax) emask |= 1u << e;
        const int ne = __builtin_popcount(emask);
        const int nu = ne + 1 + (hp ? 1 : 0);
        bool act[NS];
        double lnn[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int k = lane + kWave * s;
            bool a = k < K;
            if (a)
                for (int e = 0; e < E; e++)
                    if (!((emask >> e) & 1u) && m.comp[k * E + e] != 0.0) a = false;
            act[s] = a;
            lnn[s] = a ? log(0.1 / K) : 0.0;
        }
        double lnntot = log(0.1);
        double T = hp ? 3800.0 : T0;
        const double lnP = log(P / m.pref);
        int status = -1, it = 0;
        for (; it < max_iter; it++) {
            tt[0] = T; tt[1] = T * T; tt[2] = tt[1] * T; tt[3] = tt[2] * T; tt[4] = 1.0 / T; tt[5] = log(T);
            double nj[NS], mu[NS], hj[NS], cpj[NS];
            double l_sumn = 0.0, l_sh = 0.0, l_scp = 0.0, l_shh = 0.0, l_snmu = 0.0, l_snhmu = 0.0;
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int k = lane + kWave * s;
                nj[s] = mu[s] = hj[s] = cpj[s] = 0.0;
                if (act[s]) {
                    double cp, h, sR, g;
                    nasa7(m.nasa + 16 * k, T, tt, cp, h, sR, g);
                    hj[s] = h;
                    cpj[s] = cp;
                    mu[s] = g + (lnn[s] - lnntot) + lnP;
                    nj[s] = exp(lnn[s]);
                    l_sumn += nj[s];
                    l_sh += nj[s] * h;
                    l_scp += nj[s] * cp;
                    l_shh += nj[s] * h * h;
                    l_snmu += nj[s] * mu[s];
                    l_snhmu += nj[s] * h * mu[s];
                }
            }
            const double sumn = uni(wave_sum(l_sumn)), sh = uni(wave_sum(l_sh)), scp = uni(wave_sum(l_scp));
            const double shh = uni(wave_sum(l_shh)), snmu = uni(wave_sum(l_snmu)), snhmu = uni(wave_sum(l_snhmu));
            const double ntot = exp(lnntot);
            // element rows: sum_j a_qj a_rj n_j, b_q, sum_j a_qj n_j h_j, sum_j a_qj n_j mu_j
            int pq = 0;
            for (int q = 0; q < E; q++) {
                if (!((emask >> q) & 1u)) continue;
                double lb = 0.0, lh = 0.0, lm = 0.0;
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    const int k = lane + kWave * s;
                    if (act[s]) {
                        const double an = m.comp[k * E + q] * nj[s];
                        lb += an;
                        lh += an * hj[s];
                        lm += an * mu[s];
                    }
                }
                const double bq = uni(wave_sum(lb)), hq = uni(wave_sum(lh)), mq = uni(wave_sum(lm));
                int pr = pq;
                for (int r = q; r < E; r++) {
                    if (!((emask >> r) & 1u)) continue;
                    double lr = 0.0;
#pragma unroll
                    for (int s = 0; s < NS; s++) {
                        const int k = lane + kWave * s;
                        if (act[s]) lr += m.comp[k * E + q] * m.comp[k * E + r] * nj[s];
                    }
                    const double rqr = uni(wave_sum(lr));
                    if (lane == 0) {
                        As[pq * nu + pr] = rqr;
                        As[pr * nu + pq] = rqr;
                    }
                    pr++;
                }
                if (lane == 0) {
                    As[pq * nu + ne] = bq;
                    As[ne * nu + pq] = bq;
                    xs[pq] = b0[q] - bq + mq;
                    if (hp) {
                        As[pq * nu + ne + 1] = hq;
                        As[(ne + 1) * nu + pq] = hq;
                    }
                }
                pq++;
            }
            if (lane == 0) {
                As[ne * nu + ne] = sumn - ntot;
                xs[ne] = ntot - sumn + snmu;
                if (hp) {
                    As[ne * nu + ne + 1] = sh;
                    As[(ne + 1) * nu + ne] = sh;
                    As[(ne + 1) * nu + ne + 1] = scp + shh;
                    xs[ne + 1] = h0 / (kGasConstant * T) - sh + snhmu;
                }
                // Gaussian elimination with partial pivoting (nu <= 10)
                int sing = 0;
                for (int c = 0; c < nu; c++) {
                    int p = c;
                    for (int r = c + 1; r < nu; r++)
                        if (fabs(As[r * nu + c]) > fabs(As[p * nu + c])) p = r;
                    if (As[p * nu + c] == 0.0) { sing = 1; break; }
                    if (p != c) {
                        for (int k = 0; k < nu; k++) {
                            const double t = As[c * nu + k]; As[c * nu + k] = As[p * nu + k]; As[p * nu + k] = t;
                        }
                        const double t = xs[c]; xs[c] = xs[p]; xs[p] = t;
                    }
                    for (int r = c + 1; r < nu; r++) {
                        const double f = As[r * nu + c] / As[c * nu + c];
                        for (int k = c; k < nu; k++) As[r * nu + k] -= f * As[c * nu + k];
                        xs[r] -= f * xs[c];
                    }
                }
                if (!sing)
                    for (int c = nu - 1; c >= 0; c--) {
                        double v = xs[c];
                        for (int k = c + 1; k < nu; k++) v -= As[c * nu + k] * xs[k];
                        xs[c] = v / As[c * nu + c];
                    }
                else
                    xs[0] = __builtin_nan("");
            }
            wave_sync();
            if (!(xs[0] == xs[0])) { status = -3; break; }
            const double dlnN = uni(xs[ne]), dlnT = hp ? uni(xs[ne + 1]) : 0.0;
            double d[NS];
            double l_big = 0.0, l_conv = 0.0, l_lam2 = 1.0;
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int k = lane + kWave * s;
                d[s] = 0.0;
                if (act[s]) {
                    double v = -mu[s] + dlnN + hj[s] * dlnT;
                    int p = 0;
                    for (int e = 0; e < E; e++)
                        if ((emask >> e) & 1u) v += m.comp[k * E + e] * xs[p++];
                    d[s] = v;
                    const double lnx = lnn[s] - lnntot;
                    if (lnx > -kSize) {
                        l_big = fmax(l_big, v);
                        l_conv = fmax(l_conv, nj[s] / sumn * fabs(v));
                    } else if (v >= 0.0 && v - dlnN != 0.0) {
                        l_lam2 = fmin(l_lam2, fabs((-lnx - kTraceLn) / (v - dlnN)));
                    }
                }
            }
            wave_sync();
            const double big = fmax(fmax(5.0 * fabs(dlnT), 5.0 * fabs(dlnN)), uni(wave_max(l_big)));
            const double conv = fmax(fmax(fabs(dlnT), fabs(dlnN)), uni(wave_max(l_conv)));
            const double lam2 = -uni(wave_max(-l_lam2));
            double lam = fmin(1.0, lam2);
            if (big > 2.0) lam = fmin(lam, 2.0 / big);
#pragma unroll
            for (int s = 0; s < NS; s++)
                if (act[s]) lnn[s] += lam * d[s];
            lnntot += lam * dlnN;
            if (hp) T = fmin(fmax(T * exp(lam * dlnT), 200.0), 6000.0);
            if (lam == 1.0 && conv < 1e-12) { status = 0; it++; break; }
        }
        double yo[NS], ysum = 0.0;
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int k = lane + kWave * s;
            yo[s] = act[s] ? exp(lnn[s]) / m.rmw[k] : 0.0;
            ysum += yo[s];
        }
        ysum = uni(wave_sum(ysum));
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int k = lane + kWave * s;
            if (k < K) Yout[static_cast<size_t>(k) * n + st] = yo[s] / ysum;
        }
        if (lane == 0) {
            Tout[st] = T;
            if (info) info[st] = status == 0 ? it : status;
        }
        wave_sync();
    }
}

}  // namespace

hipError_t equil_launch(const DevMech& m, int ns, int hp, long n, const double* T, const double* P, const double* Y,
                        double* Tout, double* Yout, int* info, int max_iter, int grid, hipStream_t stream) {
    if (m.E > kEqMaxE) return hipErrorInvalidValue;
    if (ns == 1)
        hipLaunchKernelGGL((equil_kernel<1>), dim3(grid), dim3(64), 0, stream, m, n, hp, T, P, Y, Tout, Yout, info,
                           max_iter);
    else if (ns == 2)
        hipLaunchKernelGGL((equil_kernel<2>), dim3(grid), dim3(64), 0, stream, m, n, hp, T, P, Y, Tout, Yout, info,
                           max_iter);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

const void* equil_kernel_ptr(int ns) {
    return ns == 1 ? reinterpret_cast<const void*>(&equil_kernel<1>) : reinterpret_cast<const void*>(&equil_kernel<2>);
}

}  // namespace cg
