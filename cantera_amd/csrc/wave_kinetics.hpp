// Wavefront-cooperative gas kinetics for one state at a time: lanes over
// species for thermo and concentrations, lanes over reactions for rate
// constants, equilibrium constants and rates of progress, and a balanced
// sub-row gather for the net production rates.  Shared by the batched rate
// kernel and the per-reactor BDF integrator.
//
// Reference algorithm (file:line in the Cantera sources):
//   BulkKinetics::updateROP            src/kinetics/BulkKinetics.cpp:457-545
//   Kinetics::getNetProductionRates    src/kinetics/Kinetics.cpp:447-457
//   StoichManagerN multiply rules      include/cantera/kinetics/StoichManager.h:147-366
//   ThirdBodyCalc::update/multiply     include/cantera/kinetics/ThirdBodyCalc.h:75-91
//   ArrheniusRate::evalFromStruct      include/cantera/kinetics/Arrhenius.h:196-198
//   FalloffRate::evalFromStruct        include/cantera/kinetics/Falloff.h:154-176
//   TroeRate::updateTemp / F           src/kinetics/Falloff.cpp:283-300
//   NasaPoly1/NasaPoly2                include/cantera/thermo/NasaPoly1.h:83-122, NasaPoly2.h:105-111
//   IdealGasConstPressureReactor::eval src/zeroD/IdealGasConstPressureReactor.cpp:46-110
//
// Internal state order on the device is [Y_0 .. Y_{K-1}, T, m]: component i
// lives in lane i % 64, register slot i / 64, so species k is owned by the
// lane that also owns Y_k (no cross-lane moves between thermo and state).
#pragma once
#include <hip/hip_runtime.h>

#include "mech.hpp"

namespace cg {

constexpr double kGasConstant = 6.02214076e26 * 1.380649e-23;  // ct_defs.h:84-123
constexpr double kSmallNumber = 1e-300;
constexpr double kBigNumber = 1e300;
constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x) & (kWave - 1); }

// 64-bit DPP move (both halves) with the given control; all lanes valid for
// quad_perm and row_ror controls.
template <int kCtrl>
__device__ __forceinline__ double dpp_f64(double x) {
    const long long u = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(u & 0xffffffffll), kCtrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(u >> 32), kCtrl, 0xf, 0xf, false);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// The value unchanged, but opaque to the optimiser: it cannot trace it back to
// the load that produced it (used to stop selects between array elements from
// being turned into a load through a selected pointer).
__device__ __forceinline__ double opaque(double x) {
    asm("" : "+v"(x));
    return x;
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(x & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(x >> 32), l);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// 64-bit DPP move (both halves) with the given control.
template <int kCtrl>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long x) {
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(x & 0xffffffffull), kCtrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(x >> 32), kCtrl, 0xf, 0xf, false);
    return (static_cast<unsigned long long>(static_cast<unsigned int>(hi)) << 32) | static_cast<unsigned int>(lo);
}

__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long x, int l) {
    const unsigned int lo = __builtin_amdgcn_readlane(static_cast<int>(x & 0xffffffffull), l);
    const unsigned int hi = __builtin_amdgcn_readlane(static_cast<int>(x >> 32), l);
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// Wavefront max of an unsigned 64-bit key: within each row of 16 lanes by
// quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_ror:4, row_ror:8; then over
// the four rows through lanes 0, 16, 32, 48.  Returns the (uniform) maximum.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long x) {
    unsigned long long y;
    y = dpp_u64<0xB1>(x);  x = y > x ? y : x;
    y = dpp_u64<0x4E>(x);  x = y > x ? y : x;
    y = dpp_u64<0x124>(x); x = y > x ? y : x;
    y = dpp_u64<0x128>(x); x = y > x ? y : x;
    unsigned long long m = readlane_u64(x, 0);
    const unsigned long long m1 = readlane_u64(x, 16), m2 = readlane_u64(x, 32), m3 = readlane_u64(x, 48);
    m = m1 > m ? m1 : m;
    m = m2 > m ? m2 : m;
    m = m3 > m ? m3 : m;
    return m;
}

// Wavefront sum, the same value in every lane: within each row of 16 lanes by
// DPP (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_ror:4, row_ror:8), then the
// four row sums through lanes 0, 16, 32, 48 (fixed order, so every lane and
// every call with the same inputs gives the same result).
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x124>(v);
    v += dpp_f64<0x128>(v);
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}
__device__ __forceinline__ float wave_sumf(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
    v = fmax(v, dpp_f64<0xB1>(v));
    v = fmax(v, dpp_f64<0x4E>(v));
    v = fmax(v, dpp_f64<0x124>(v));
    v = fmax(v, dpp_f64<0x128>(v));
    return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)), fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
}

// Order LDS traffic between the lanes of one wavefront.  Every kernel here
// runs one wavefront per workgroup, so the workgroup barrier is a cheap,
// compiler-visible ordering point (a wavefront-scope fence alone does not stop
// plain LDS loads from moving above another lane's stores).
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// make a value that all lanes hold identically explicitly wave-uniform
__device__ __forceinline__ double uni(double x) {
    long long v = __double_as_longlong(x);
    int lo = __builtin_amdgcn_readfirstlane(static_cast<int>(v & 0xffffffffLL));
    int hi = __builtin_amdgcn_readfirstlane(static_cast<int>(v >> 32));
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uni(float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); }

// broadcast element k of a lane-distributed vector (k wave-uniform)
template <int NS, typename T>
__device__ __forceinline__ T bcast(const T (&v)[NS], int k) {
    const int s = k >> 6;
    T src = v[0];
#pragma unroll
    for (int t = 1; t < NS; t++)
        if (s == t) src = v[t];
    if constexpr (sizeof(T) == 8) {
        long long x = __double_as_longlong(src);
        int lo = __builtin_amdgcn_readlane(static_cast<int>(x & 0xffffffffLL), k & 63);
        int hi = __builtin_amdgcn_readlane(static_cast<int>(x >> 32), k & 63);
        return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
    } else {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(src), k & 63));
    }
}

// States handled by this workgroup when a batch of n states, stored
// species-major (the 8 states of a 64-byte line are neighbours), is spread over
// the grid.  Workgroups are dispatched to the 8 XCDs round-robin
// (blockIdx % 8), and each XCD has its own L2: XCD x takes the contiguous range
// [x n / 8, (x + 1) n / 8), its workgroups stride through it, so the states
// sharing a line meet in one L2 instead of eight.  Every state is visited
// exactly once for any grid of at least 8 workgroups (plain grid-stride order
// otherwise).
struct XcdStates {
    long lo, hi, step;
    __device__ __forceinline__ XcdStates(long n) {
        const long G = gridDim.x, b = blockIdx.x;
        if (G < 8) {
            lo = b; hi = n; step = G;
            return;
        }
        const long x = b % 8, j = b / 8;
        const long gx = (G - x + 7) / 8;     // workgroups on XCD x
        const long c0 = n * x / 8, c1 = n * (x + 1) / 8;
        lo = c0 + j;
        hi = c1;
        step = gx;
    }
};

// Per-wave LDS scratch for one state evaluation.
struct WaveScratch {
    double* yb;    // N (state broadcast)
    double* C;     // K + 1 concentrations (C[K] = 1: the unused-slot sentinel of RxnI)
    double* MU;    // K + 1 standard chemical potentials [J/kmol] (MU[K] = 0)
    double* WD;    // K production-rate accumulators
    double* W2;    // 2K (Jacobian: wdot, D)
    double* F;     // N (right-hand side out)
    double* Y0;    // N (Jacobian expansion point)
    double* F0;    // N (right-hand side at Y0)
};

// The same scratch as offsets (in doubles) from the start of dynamic LDS, for
// passing to out-of-line device functions: pointers rebuilt from the LDS base
// inside the callee keep the LDS address space (ds_* instructions).
struct ScratchLayout {
    int yb, C, MU, WD, W2, F, Y0, F0;
};

__device__ __forceinline__ double* lds_base() {
    extern __shared__ double lds_raw[];
    return lds_raw;
}

__device__ __forceinline__ WaveScratch scratch_view(const ScratchLayout& L) {
    double* b = lds_base();
    WaveScratch S;
    S.yb = b + L.yb; S.C = b + L.C; S.MU = b + L.MU; S.WD = b + L.WD;
    S.W2 = b + L.W2; S.F = b + L.F; S.Y0 = b + L.Y0; S.F0 = b + L.F0;
    return S;
}

// Scalar thermodynamic state derived from (T, P, Y): setMassFractions_NoNorm +
// setState_TP (Phase.cpp:362-372; ThermoPhase.cpp:187-198; IdealGasPhase.h:338-354).
// c + a*b with the product rounded before the sum (no fused multiply-add):
// [M] = d C_tot + sum_k (e_k - d) C_k is accumulated exactly as the reference
// does (ThirdBodyCalc::update, ThirdBodyCalc.h:75-83), so a nearly cancelling
// [M] (a bath of an efficiency-0 species) rounds the same way.
__device__ __forceinline__ double add_mul_rn(double c, double a, double b) {
#pragma clang fp contract(off)
    return c + a * b;
}

struct GasState {
    double T, P, mmw, RT, dens, ctot, pressure, logT, recipT, rrt, logC0, tmp;
};

// Values one lane keeps for the species it owns (slot s = species lane + 64 s).
template <int NS>
struct SpeciesRegs {
    double ym[NS], hrt[NS], cpr[NS], wdot[NS];
};

template <typename R>
__device__ __forceinline__ void nasa7(const double* __restrict__ c16, double T, const double* tt, R& cp, R& h, R& s, R& g) {
    const double* c = (T <= c16[0]) ? c16 + 1 : c16 + 8;
    double ct0 = c[0];
    double ct1 = c[1] * tt[0];
    double ct2 = c[2] * tt[1];
    double ct3 = c[3] * tt[2];
    double ct4 = c[4] * tt[3];
    double cpv = ct0 + ct1 + ct2 + ct3 + ct4;
    double hv = ct0 + 0.5 * ct1 + 1.0 / 3.0 * ct2 + 0.25 * ct3 + 0.2 * ct4 + c[5] * tt[4];
    double sv = ct0 * tt[5] + ct1 + 0.5 * ct2 + 1.0 / 3.0 * ct3 + 0.25 * ct4 + c[6];
    cp = cpv;
    h = hv;
    s = sv;
    g = hv - sv;
}

// Per-reaction results.
struct RxnVals {
    double kf;     // forward rate constant (with falloff)
    double M;      // third-body concentration (1 if none)
    double rkcn;   // 1/Kc (0 if irreversible)
    double ropf, ropr;
};

// Reaction record unpacked into registers (static indexing only, so nothing
// spills to scratch).
struct RxnRegs {
    int flags, nr, np, tbo, tbn, fo, gen;
    int rk[kMaxSide], pk[kMaxSide];
};

__device__ __forceinline__ void load_rxn(const DevMech& m, int p, RxnRegs& x) {
    const int4* q = reinterpret_cast<const int4*>(m.ri + p);
    const int4 a = q[0], b = q[1], c = q[2];
    x.flags = a.x; x.nr = a.y; x.np = a.z; x.tbo = a.w;
    x.tbn = b.x; x.fo = b.y; x.gen = b.z;
    x.rk[0] = static_cast<short>(c.x & 0xffff); x.rk[1] = static_cast<short>(c.x >> 16);
    x.rk[2] = static_cast<short>(c.y & 0xffff); x.rk[3] = static_cast<short>(c.y >> 16);
    x.pk[0] = static_cast<short>(c.z & 0xffff); x.pk[1] = static_cast<short>(c.z >> 16);
    x.pk[2] = static_cast<short>(c.w & 0xffff); x.pk[3] = static_cast<short>(c.w >> 16);
}

// Transcendentals of the rate constants.  kF32: evaluated in FP32 (hardware
// exp2/log2 based) on FP64 arguments rounded to FP32, widened back to FP64;
// arguments whose result would leave the FP32 range take the FP64 function.
// The equilibrium constants never use these (their exp stays FP64).
template <bool kF32>
__device__ __forceinline__ double rate_exp(double arg) {
    if constexpr (kF32) {
        if (fabs(arg) < 80.0) return static_cast<double>(expf(static_cast<float>(arg)));
    }
    return exp(arg);
}
template <bool kF32>
__device__ __forceinline__ double rate_log10(double x) {
    if constexpr (kF32) {
        if (x > 1e-35 && x < 1e35) return static_cast<double>(log10f(static_cast<float>(x)));
    }
    return log10(x);
}
template <bool kF32>
__device__ __forceinline__ double rate_exp10(double x) {
    if constexpr (kF32) {
        if (fabs(x) < 35.0) return static_cast<double>(exp10f(static_cast<float>(x)));
    }
    return exp10(x);
}

// PLOG rate constant (PlogRate::updateFromStruct / evalFromStruct,
// include/cantera/kinetics/PlogRate.h:110-164).  Record r: ng, ln P[ng],
// first[ng], end[ng], then (A, b, Ea/R) per entry; groups 0 and ng-1 are the
// copies at ln P = -1000 and +1000.  The interval is [ln P[i-1], ln P[i]) with
// i the first group above ln P; each end takes ln(1e-300 + sum of the group's
// Arrhenius terms), and ln k is interpolated linearly in ln P.
__device__ __forceinline__ double plog_rate(const double* __restrict__ r, double logP, double logT, double recipT) {
    const int ng = static_cast<int>(r[0]);
    const double* lp = r + 1;
    const double* e = r + 1 + 3 * ng;
    int i = 1;
    while (i < ng - 1 && !(lp[i] > logP)) i++;
    double lk[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int q = i - 1 + h;
        double k = 1e-300;
        for (int j = static_cast<int>(r[1 + ng + q]); j < static_cast<int>(r[1 + 2 * ng + q]); j++)
            k += e[3 * j] * exp(e[3 * j + 1] * logT - e[3 * j + 2] * recipT);
        lk[h] = log(k);
    }
    const double rdp = 1.0 / (lp[i] - lp[i - 1]);
    return exp(lk[0] + (lk[1] - lk[0]) * (logP - lp[i - 1]) * rdp);
}

// Chebyshev rate constant (ChebyshevRate::updateFromStruct / evalFromStruct,
// include/cantera/kinetics/ChebyshevRate.h:133-171).  Record r: TrNum, TrDen,
// PrNum, PrDen, nT, nP, coefficients (nT x nP, row-major).  Each row is first
// reduced over the pressure polynomials, then the temperature series is summed
// with the same three-term recurrence, row by row.
__device__ __forceinline__ double cheb_rate(const double* __restrict__ r, double log10P, double recipT) {
    const int nT = static_cast<int>(r[4]), nP = static_cast<int>(r[5]);
    const double* c = r + 6;
    const double Pr = (2 * log10P + r[2]) * r[3];
    const double Tr = (2 * recipT + r[0]) * r[1];
    double logk = 0.0, Tm1 = Tr, Tn = 1.0;
    for (int i = 0; i < nT; i++) {
        double dot = c[i * nP];
        double Pm1 = Pr, Pn = 1.0;
        for (int j = 1; j < nP; j++) {
            const double Pp1 = 2 * Pr * Pn - Pm1;
            dot += Pp1 * c[i * nP + j];
            Pm1 = Pn;
            Pn = Pp1;
        }
        if (i == 0) {
            logk = dot;
        } else {
            const double Tp1 = 2 * Tr * Tn - Tm1;
            logk += Tp1 * dot;
            Tm1 = Tn;
            Tn = Tp1;
        }
    }
    return pow(10.0, logk);
}

// Rate constants of the PLOG and Chebyshev reactions at the gas state.  They
// come first in permuted order and number at most 64, so lane j evaluates
// reaction j once per state, before the reaction loop; the loop's first pass
// (lane j, reaction j) takes the value from the register.  Other lanes, and
// mechanisms without such reactions, return 0.
__device__ __forceinline__ double pdep_rate_lane(const DevMech& m, const GasState& g) {
    double k = 0.0;
    if (m.n_pd > 0) {
        const int lane = lane_id();
        if (lane < m.n_pd) {
            const int flags = m.ri[lane].flags;
            const int fo = m.ri[lane].fo;
            k = (flags & 64) ? plog_rate(m.pdd + fo, log(g.pressure), g.logT, g.recipT)
                             : cheb_rate(m.pdd + fo, log10(g.pressure), g.recipT);
        }
    }
    return k;
}

// SRI falloff function at reduced pressure pr (SriRate::updateTemp and F,
// Falloff.cpp:389-403); fo[3..7] = a, b, c, d, e.  Out of line: SRI reactions
// are rare, and inlined this code enlarges every caller's register frame.
__device__ __noinline__ double sri_falloff(const double* __restrict__ fo, double T, double pr) {
    double w = fo[3] * exp(-fo[4] / T);
    if (fo[5] != 0.0) w += exp(-T / fo[5]);
    const double lpr = log10(fmax(pr, kSmallNumber));
    return pow(w, 1.0 / (1.0 + lpr * lpr)) * (fo[6] * pow(T, fo[7]));
}

// Evaluate reaction p (permuted index) at the gas state; C and MU live in LDS.
// kpd: this lane's pdep_rate_lane value (used for flags bit 6/7 only).
// kF32: rate constants (Arrhenius, falloff centre and blending) through the
// FP32 transcendentals above; everything else FP64.
// perturb: the reaction's rate multiplier (m_perturb, Kinetics::setMultiplier),
// applied to the forward rate of progress before the reverse one is formed.
template <bool kF32 = false>
__device__ __forceinline__ void eval_reaction(const DevMech& m, int p, const GasState& g,
                                              const double* __restrict__ C, const double* __restrict__ MU,
                                              const RxnRegs& x, double kpd, RxnVals& o, double perturb = 1.0) {
    const double4 rd = reinterpret_cast<const double4*>(m.rd)[p];
    const int type = (x.flags >> 1) & 3;
    const bool rev = x.flags & 1;
    double M = 1.0;
    if (x.flags & 32) {
        // The first kTbPad efficiencies are loaded with no test against the
        // list length (the table ends with kTbPad padding entries), so their
        // table loads, then their concentration loads, are all in flight
        // together; entries past the reaction's own list are not added.  The
        // rest follow in groups of 4.
        double sum = 0.0;
        {
            int kk[kTbPad];
            double ee[kTbPad], cc[kTbPad];
#pragma unroll
            for (int b = 0; b < kTbPad; b++) {
                kk[b] = m.tbk[x.tbo + b];
                ee[b] = m.tbe[x.tbo + b];
            }
#pragma unroll
            for (int b = 0; b < kTbPad; b++) cc[b] = opaque(C[kk[b]]);
#pragma unroll
            for (int b = 0; b < kTbPad; b++)
                if (b < x.tbn) sum = add_mul_rn(sum, ee[b], cc[b]);
        }
        for (int j0 = kTbPad; j0 < x.tbn; j0 += 4) {
            int kk[4];
            double ee[4];
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const bool ok = j0 + b < x.tbn;
                kk[b] = ok ? m.tbk[x.tbo + j0 + b] : 0;
                ee[b] = ok ? m.tbe[x.tbo + j0 + b] : 0.0;
            }
            double cc[4];
#pragma unroll
            for (int b = 0; b < 4; b++) cc[b] = C[kk[b]];
#pragma unroll
            for (int b = 0; b < 4; b++)
                if (j0 + b < x.tbn) sum = add_mul_rn(sum, ee[b], cc[b]);
        }
        M = add_mul_rn(sum, m.tbdef[p], g.ctot);
    }
    double kf;
    if (type <= kThreeBody) {
        // PLOG and Chebyshev constants (FP64 in both precisions) come from kpd
        kf = (x.flags & 192) ? kpd : rd.x * rate_exp<kF32>(rd.y * g.logT - rd.z * g.recipT);
    } else {
        const double* fo = m.fod + 8 * x.fo;
        const bool sri = x.flags & 256;
        double work = 0.0;
        if (type == kTroe && !sri) {
            // Troe (fo[7] = 0) and Tsang (Fcent = A + B T)
            double Fcent = (1.0 - fo[3]) * rate_exp<kF32>(-g.T * fo[4]) + fo[3] * rate_exp<kF32>(-g.T * fo[5]);
            if (fo[6] != 0.0) Fcent += rate_exp<kF32>(-fo[6] / g.T);
            Fcent += fo[7] * g.T;
            work = rate_log10<kF32>(fmax(Fcent, kSmallNumber));
        }
        double rc_low = fo[0] * rate_exp<kF32>(fo[1] * g.logT - fo[2] * g.recipT);
        double rc_high = rd.x * rate_exp<kF32>(rd.y * g.logT - rd.z * g.recipT);
        double pr = M * rc_low / (rc_high + kSmallNumber);
        double F = 1.0;
        if (type == kTroe && sri) {
            F = sri_falloff(fo, g.T, pr);
        } else if (type == kTroe) {
            double lpr = rate_log10<kF32>(fmax(pr, kSmallNumber));
            double cc = -0.4 - 0.67 * work;
            double nn = 0.75 - 1.27 * work;
            double f1 = (lpr + cc) / (nn - 0.14 * (lpr + cc));
            double lgf = work / (1.0 + f1 * f1);
            F = rate_exp10<kF32>(lgf);
        }
        // FalloffRate::evalFromStruct (Falloff.h:154-176): Pr/(1+Pr) F k_inf, or
        // F/(1+Pr) k_0 for a chemically activated reaction
        const double fac = F / (1.0 + pr);
        kf = (x.flags & 512) ? fac * rc_low : (pr * fac) * rc_high;
    }
    o.kf = kf;
    o.M = M;
    // inverse equilibrium constant (BulkKinetics.cpp:465-484)
    double rkcn = 0.0;
    if (rev) {
        double dp = 0.0;
        if (x.flags & 16) {
            const double* gs = m.gend + 12 * x.gen;
#pragma unroll
            for (int j = 0; j < kMaxSide; j++)
                if (j < x.np) dp += gs[8 + j] * MU[x.pk[j]];
        } else {
            dp = MU[x.pk[0]];
            if (x.np > 1) dp += MU[x.pk[1]];
            if (x.np > 2) dp += MU[x.pk[2]];
        }
        double delta = dp;
        if (x.flags & 8) {
            const double* gs = m.gend + 12 * x.gen;
#pragma unroll
            for (int j = 0; j < kMaxSide; j++)
                if (j < x.nr) delta -= gs[4 + j] * MU[x.rk[j]];
        } else {
            double dr = MU[x.rk[0]];
            if (x.nr > 1) dr += MU[x.rk[1]];
            if (x.nr > 2) dr += MU[x.rk[2]];
            delta -= dr;
        }
        rkcn = fmin(exp(delta * g.rrt - rd.w * g.logC0), kBigNumber);
    }
    o.rkcn = rkcn;
    double ropf = kf * perturb;
    if (type == kThreeBody) ropf *= M;
    double ropr = ropf * rkcn;
    // concentration products with the reference's negative-concentration rules
    if (x.flags & 8) {
        const double* gs = m.gend + 12 * x.gen;
#pragma unroll
        for (int j = 0; j < kMaxSide; j++) {
            if (j < x.nr) {
                double order = gs[j];
                if (order != 0.0) {
                    double c = C[x.rk[j]];
                    ropf = (c > 0.0) ? ropf * pow(c, order) : 0.0;
                }
            }
        }
    } else {
        double a = C[x.rk[0]];
        if (x.nr == 1) {
            ropf *= a;
        } else if (x.nr == 2) {
            double b = C[x.rk[1]];
            ropf = (a < 0 && b < 0) ? 0.0 : ropf * (a * b);
        } else {
            double b = C[x.rk[1]], d = C[x.rk[2]];
            ropf = ((a < 0 && (b < 0 || d < 0)) || (b < 0 && d < 0)) ? 0.0 : ropf * (a * b * d);
        }
    }
    if (rev) {
        if (x.flags & 16) {
            const double* gs = m.gend + 12 * x.gen;
#pragma unroll
            for (int j = 0; j < kMaxSide; j++) {
                if (j < x.np) {
                    double c = C[x.pk[j]];
                    ropr = (c > 0.0) ? ropr * pow(c, gs[8 + j]) : 0.0;
                }
            }
        } else {
            double a = C[x.pk[0]];
            if (x.np == 1) {
                ropr *= a;
            } else if (x.np == 2) {
                double b = C[x.pk[1]];
                ropr = (a < 0 && b < 0) ? 0.0 : ropr * (a * b);
            } else {
                double b = C[x.pk[1]], d = C[x.pk[2]];
                ropr = ((a < 0 && (b < 0 || d < 0)) || (b < 0 && d < 0)) ? 0.0 : ropr * (a * b * d);
            }
        }
    }
    o.ropf = ropf;
    o.ropr = ropr;
}

// Net rate of progress of a plain mass-action reaction (RxnI flag bit 10:
// elementary, kRep sides of at most three entries): the same quantities as
// eval_reaction without its tests -- unused side entries point at the
// sentinel species (C = 1, MU = 0), which leaves every sum, product and
// negative-concentration rule of the C1/C2/C3 cases unchanged.  pf: rate
// multiplier of the forward rate of progress.
template <bool kF32 = false>
__device__ __forceinline__ double plain_ropnet(const DevMech& m, int p, const GasState& g,
                                              const double* __restrict__ C, const double* __restrict__ MU,
                                              const RxnRegs& x, double pf = 1.0) {
    const double4 rd = reinterpret_cast<const double4*>(m.rd)[p];
    const double kf = rd.x * rate_exp<kF32>(rd.y * g.logT - rd.z * g.recipT);
    const double dr = (MU[x.rk[0]] + MU[x.rk[1]]) + MU[x.rk[2]];
    const double dp = (MU[x.pk[0]] + MU[x.pk[1]]) + MU[x.pk[2]];
    const double delta = dp - dr;
    const bool rev = x.flags & 1;
    const double rkcn = rev ? fmin(exp(delta * g.rrt - rd.w * g.logC0), kBigNumber) : 0.0;
    double ropf = kf * pf;
    double ropr = ropf * rkcn;
    {
        const double a = C[x.rk[0]], b = C[x.rk[1]], d = C[x.rk[2]];
        ropf = ((a < 0 && (b < 0 || d < 0)) || (b < 0 && d < 0)) ? 0.0 : ropf * (a * b * d);
    }
    {
        const double a = C[x.pk[0]], b = C[x.pk[1]], d = C[x.pk[2]];
        ropr = rev ? (((a < 0 && (b < 0 || d < 0)) || (b < 0 && d < 0)) ? 0.0 : ropr * (a * b * d)) : 0.0;
    }
    return ropf - ropr;
}

// Gas state + per-species thermo, concentrations and chemical potentials.
// Y values: lane owns species lane + 64 s in yk[s].  Leaves C and MU in LDS.
// rho > 0: the state is (T, rho) instead of (T, P) (ThermoPhase::setState_TD);
// the pressure is then R T rho / Wbar.
// kSeqSum: 1/Wbar = sum_k Y_k/W_k summed in species order, as
// Phase::setMassFractions_NoNorm does (Phase.cpp:362-372), through S.WD (the
// batched rates and Jacobians, whose results are compared bit for bit);
// otherwise a wavefront tree sum (the integrator's right-hand side).
template <int NS, bool kSeqSum = false>
__device__ __forceinline__ void species_pass(const DevMech& m, const WaveScratch& S, double T, double P,
                                             const double (&yk)[NS], GasState& g, SpeciesRegs<NS>& sr,
                                             double rho = 0.0) {
    const int lane = lane_id();
    double sum = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int k = lane + kWave * s;
        sr.ym[s] = (k < m.K) ? yk[s] * m.rmw[k] : 0.0;
        sum += sr.ym[s];
    }
    if constexpr (kSeqSum) {
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int k = lane + kWave * s;
            if (k < m.K) S.WD[k] = sr.ym[s];
        }
        wave_sync();
        sum = 0.0;
        for (int k = 0; k < m.K; k++) sum += S.WD[k];
        wave_sync();
    } else {
        sum = uni(wave_sum(sum));
    }
    g.T = T;
    g.mmw = 1.0 / sum;
    g.RT = T * kGasConstant;
    g.dens = (rho > 0.0) ? rho : P * g.mmw / g.RT;
    g.ctot = g.dens / g.mmw;
    g.pressure = kGasConstant * g.ctot * T;
    g.P = (rho > 0.0) ? g.pressure : P;
    g.logT = log(T);
    g.recipT = 1.0 / T;
    g.rrt = 1.0 / g.RT;
    g.logC0 = log(g.pressure / g.RT);
    g.tmp = log(g.pressure / m.pref) * g.RT;
    double tt[6];
    tt[0] = T;
    tt[1] = T * T;
    tt[2] = tt[1] * T;
    tt[3] = tt[2] * T;
    tt[4] = 1.0 / T;
    tt[5] = g.logT;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int k = lane + kWave * s;
        if (k < m.K) {
            double cp, h, sR, g0;
            nasa7(m.nasa + 16 * k, T, tt, cp, h, sR, g0);
            sr.cpr[s] = cp;
            sr.hrt[s] = h;
            double mu = g0 * g.RT;
            mu += g.tmp;
            S.MU[k] = mu;
            S.C[k] = sr.ym[s] * g.dens;
        } else {
            sr.cpr[s] = 0.0;
            sr.hrt[s] = 0.0;
            if (k == m.K) {
                S.MU[k] = 0.0;
                S.C[k] = 1.0;
            }
        }
    }
}

// Net production rates for the species each lane owns (after species_pass).
// Each reaction lane adds -nu' ropnet to its reactants and +nu'' ropnet to its
// products with LDS FP64 atomics (the species lists are already in registers),
// so no second pass over a gather table is needed; the summation order of a
// species' terms is the hardware's, not the reference's reaction order.
// kPert: reaction pert_rxn (permuted index) runs with rate multiplier pert_v.
template <int NS, bool kF32 = false, bool kPert = false>
__device__ __forceinline__ void production_rates(const DevMech& m, const WaveScratch& S, const GasState& g,
                                                 SpeciesRegs<NS>& sr, int pert_rxn = -1, double pert_v = 1.0) {
    const int lane = lane_id();
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int k = lane + kWave * s;
        if (k < m.K) S.WD[k] = 0.0;
    }
    wave_sync();
    const double kpd = pdep_rate_lane(m, g);
    // Lane j takes reactions j, j + 64, ...  When the reactions of this pass
    // and of the next one are all plain mass action (the elementary reactions
    // come last in the permuted order), the lane evaluates both in one
    // branch-free block -- two independent chains of loads and exponentials
    // for the scheduler to interleave -- and the loop advances two passes.
    for (int p = lane; p < m.I;) {
        RxnRegs x, x1;
        load_rxn(m, p, x);
        const bool two = p + kWave < m.I;
        const int p1 = two ? p + kWave : p;
        load_rxn(m, p1, x1);
        if (__all((x.flags & 1024) && (x1.flags & 1024))) {
            double pf0 = 1.0, pf1 = 1.0;
            if constexpr (kPert) {
                pf0 = (p == pert_rxn) ? pert_v : 1.0;
                pf1 = (p1 == pert_rxn) ? pert_v : 1.0;
            }
            const double r0 = plain_ropnet<kF32>(m, p, g, S.C, S.MU, x, pf0);
            const double r1 = plain_ropnet<kF32>(m, p1, g, S.C, S.MU, x1, pf1);
#pragma unroll
            for (int j = 0; j < 3; j++)
                if (j < x.nr) atomicAdd(&S.WD[x.rk[j]], -r0);
#pragma unroll
            for (int j = 0; j < 3; j++)
                if (j < x.np) atomicAdd(&S.WD[x.pk[j]], r0);
            if (two) {
#pragma unroll
                for (int j = 0; j < 3; j++)
                    if (j < x1.nr) atomicAdd(&S.WD[x1.rk[j]], -r1);
#pragma unroll
                for (int j = 0; j < 3; j++)
                    if (j < x1.np) atomicAdd(&S.WD[x1.pk[j]], r1);
            }
            p += 2 * kWave;
        } else {
            RxnVals v;
            if constexpr (kPert) eval_reaction<kF32>(m, p, g, S.C, S.MU, x, kpd, v, p == pert_rxn ? pert_v : 1.0);
            else eval_reaction<kF32>(m, p, g, S.C, S.MU, x, kpd, v);
            const double r = v.ropf - v.ropr;
            const double* gs = (x.gen >= 0) ? m.gend + 12 * x.gen : m.gend;
            const bool rgen = x.flags & 8, pgen = x.flags & 16;
#pragma unroll
            for (int j = 0; j < kMaxSide; j++)
                if (j < x.nr) atomicAdd(&S.WD[x.rk[j]], -(rgen ? gs[4 + j] : 1.0) * r);
#pragma unroll
            for (int j = 0; j < kMaxSide; j++)
                if (j < x.np) atomicAdd(&S.WD[x.pk[j]], (pgen ? gs[8 + j] : 1.0) * r);
            p += kWave;
        }
    }
    wave_sync();
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int k = lane + kWave * s;
        sr.wdot[s] = (k < m.K) ? S.WD[k] : 0.0;
    }
}

// Scalars of the reactor energy equation kept for the Jacobian.
struct RhsScalars {
    double mass, vol, cp_mass, H;   // H = sum_k h_k wdot_k [J/m^3/s] (constant volume: c_v and u_k)
};

// IdealGasConstPressureReactor::eval (src/zeroD/IdealGasConstPressureReactor.cpp)
// or, with m.const_volume, IdealGasReactor::eval (src/zeroD/IdealGasReactor.cpp:50-117,
// closed rigid adiabatic: m c_v dT/dt = -sum_k u_k wdot_k V, u_k = h_k - R T),
// divided by the left-hand sides as ReactorNet::eval does: ydot for the lane's
// components (internal order).  y holds [Y_0..Y_{K-1}, T, m] distributed; P is
// the pressure, or the volume of a constant-volume reactor.
template <int NS, bool kPert = false>
__device__ __forceinline__ void reactor_rhs(const DevMech& m, const WaveScratch& S, double P, const double (&y)[NS],
                                            double (&f)[NS], GasState& g, SpeciesRegs<NS>& sr, RhsScalars& rs,
                                            int pert_rxn = -1, double pert_v = 1.0) {
    const int lane = lane_id();
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int i = lane + kWave * s;
        if (i < m.N) S.yb[i] = y[s];
    }
    wave_sync();
    const double T = uni(S.yb[m.K]);
    const double mass = uni(S.yb[m.K + 1]);
    const bool cv = m.const_volume;
    species_pass<NS>(m, S, T, P, y, g, sr, cv ? mass / P : 0.0);
    production_rates<NS, false, kPert>(m, S, g, sr, pert_rxn, pert_v);
    const double vol = cv ? P : mass / g.dens;
    const double hshift = cv ? 1.0 : 0.0;   // u_k / RT = h_k / RT - 1
    double e = 0.0, inner = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int k = lane + kWave * s;
        if (k < m.K) {
            e += sr.wdot[s] * ((sr.hrt[s] - hshift) * g.RT) * vol;
            inner += sr.ym[s] * sr.cpr[s];
        }
    }
    e = uni(wave_sum(e));
    inner = uni(wave_sum(inner));
    // c_p (or c_v = (c_p,mole - R) / Wbar) per unit mass
    const double cp_mass = kGasConstant * (g.mmw * inner - hshift) / g.mmw;
    const double mcpdTdt = -e;
    rs.mass = mass;
    rs.vol = vol;
    rs.cp_mass = cp_mass;
    rs.H = e / vol;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int i = lane + kWave * s;
        double v = 0.0;
        if (i < m.K) v = (sr.wdot[s] * vol) * m.mw[i] / mass;
        else if (i == m.K) v = mcpdTdt / (mass * cp_mass);
        f[s] = v;
    }
}

// The mechanism of a kernel whose first parameter is a DevMech, or a struct
// that begins with one, addressed in the kernel-argument segment.  Out-of-line
// functions take this pointer rather than a reference to the kernel's copy of
// the argument: a reference would make the kernel keep that copy in scratch
// memory and read every field through it, while the kernel-argument segment is
// read with scalar loads.
using KernargMech = const __attribute__((address_space(4))) DevMech;
__device__ __forceinline__ KernargMech* kernarg_mech() {
    return (KernargMech*)__builtin_amdgcn_kernarg_segment_ptr();
}

// Out-of-line reactor RHS: state taken from S.yb (written by the caller),
// ydot written to S.F.  One copy of the kinetics code serves every call site of
// the integrator (smaller code, and the caller keeps little live across it).
template <int NS>
__device__ __noinline__ void rhs_call(KernargMech* mech, ScratchLayout L, double P) {
    const DevMech m = *(const DevMech*)mech;
    const WaveScratch S = scratch_view(L);
    const int lane = lane_id();
    double y[NS], f[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + kWave * s;
        y[s] = (i < m.N) ? S.yb[i] : 0.0;
    }
    GasState g;
    SpeciesRegs<NS> sr;
    RhsScalars rs;
    reactor_rhs<NS>(m, S, P, y, f, g, sr, rs);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + kWave * s;
        if (i < m.N) S.F[i] = f[s];
    }
    wave_sync();
}

// rhs_call with reaction pert_rxn (permuted index) at rate multiplier pert_v:
// the perturbed right-hand sides of the sensitivity difference quotients.
template <int NS>
__device__ __noinline__ void rhs_pert_call(KernargMech* mech, ScratchLayout L, double P, int pert_rxn, double pert_v) {
    const DevMech m = *(const DevMech*)mech;
    const WaveScratch S = scratch_view(L);
    const int lane = lane_id();
    double y[NS], f[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + kWave * s;
        y[s] = (i < m.N) ? S.yb[i] : 0.0;
    }
    GasState g;
    SpeciesRegs<NS> sr;
    RhsScalars rs;
    reactor_rhs<NS, true>(m, S, P, y, f, g, sr, rs, pert_rxn, pert_v);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + kWave * s;
        if (i < m.N) S.F[i] = f[s];
    }
    wave_sync();
}

// Loss frequencies lambda_k = d_k / C_k [1/s] at the state in S.yb (internal
// order, pressure P), left in S.W2[0..K): d_k = sum_i nu'_ki ropf_i +
// nu''_ki ropr_i is the destruction rate of species k
// (Kinetics::getDestructionRates, src/kinetics/Kinetics.cpp:435-445).  Used
// by the surrogate guard to turn a defect of species k into an error estimate.
template <int NS>
__device__ __noinline__ void loss_frequency_call(KernargMech* mech, ScratchLayout L, double P) {
    const DevMech m = *(const DevMech*)mech;
    const WaveScratch S = scratch_view(L);
    const int lane = lane_id();
    double y[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + kWave * s;
        y[s] = (i < m.K) ? S.yb[i] : 0.0;
    }
    const double T = S.yb[m.K];
    GasState g;
    SpeciesRegs<NS> sr;
    species_pass<NS>(m, S, T, P, y, g, sr);
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int k = lane + kWave * s;
        if (k < m.K) S.WD[k] = 0.0;
    }
    wave_sync();
    const double kpd = pdep_rate_lane(m, g);
    for (int p = lane; p < m.I; p += kWave) {
        RxnVals v;
        RxnRegs x;
        load_rxn(m, p, x);
        eval_reaction(m, p, g, S.C, S.MU, x, kpd, v);
        const double* gs = (x.gen >= 0) ? m.gend + 12 * x.gen : m.gend;
        const bool rgen = x.flags & 8, pgen = x.flags & 16;
#pragma unroll
        for (int j = 0; j < kMaxSide; j++)
            if (j < x.nr) atomicAdd(&S.WD[x.rk[j]], (rgen ? gs[4 + j] : 1.0) * v.ropf);
#pragma unroll
        for (int j = 0; j < kMaxSide; j++)
            if (j < x.np) atomicAdd(&S.WD[x.pk[j]], (pgen ? gs[8 + j] : 1.0) * v.ropr);
    }
    wave_sync();
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int k = lane + kWave * s;
        if (k < m.K) S.W2[k] = S.WD[k] / fmax(S.C[k], 1e-300);
    }
    wave_sync();
}

}  // namespace cg
