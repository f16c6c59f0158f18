// Iteration-matrix storage and dense LU for one reactor per wavefront.
//
// RegLinSys: the matrix lives in registers, row i in lane i (N <= NMAX <= 64).
//   Partial pivoting is implicit: no row moves; the lane chosen at step k is
//   marked as the k-th pivot row, and the triangular solves run in pivot order.
//   Every register index is static (all loops over columns are unrolled), so
//   nothing spills to scratch.  The arithmetic is the same as a row-swapping
//   LU (same multipliers, same updates), only the storage differs.
// LdsRegLinSys: the register solves of RegLinSys on factors kept in LDS
//   between calls (row i in lane i, column-major so a lane's row loads are
//   conflict-free); the factorisation runs on the matrix cores
//   (lr_factor_mfma: the MFMA-blocked LU of LdsLinSys on a row-major copy,
//   then the rows stored in the register layout), the solves are out-of-line
//   functions that load the factors into registers.  Used for 32 < N <= 64.
// LdsLinSys: the matrix lives in LDS (row-major, leading dimension LD), used
//   for N > 64 (two lanes' worth of rows per lane); factored by panels of 4
//   columns with the trailing rank-4 updates on the matrix cores
//   (lu_factor_mfma).
//
// Both read J from the FP64 saved-Jacobian buffer and form M = I - gamma J in
// the matrix precision MT (double: plain mode; float: mixed mode).
#pragma once
#include "wave_linalg.hpp"

namespace cg {

// Entry j of this lane's row of a row-major N x N matrix (lane = row), zero
// outside the leading N x N block.  The load itself is unconditional (indices
// clamped into the block) and its value passes through an empty asm: written
// as a select on the loaded value, the compiler moved each load under the
// condition with a wait right after it, which serialised the NMAX loads of a
// row (the iteration-matrix assembly cost ~39k cycles for N = 55 that way).
template <typename T>
__device__ __forceinline__ T row_entry(const T* __restrict__ M, int N, int lane, int j) {
    T x = M[min(lane, N - 1) * N + min(j, N - 1)];
    if constexpr (sizeof(T) == 8) x = opaque(x);
    else asm("" : "+v"(x));
    return (lane < N && j < N) ? x : T(0);
}

template <typename T>
__device__ __forceinline__ T shfl_t(T v, int src) {
    if constexpr (sizeof(T) == 8) {
        const long long x = __double_as_longlong(v);
        const int lo = __shfl(static_cast<int>(x & 0xffffffffLL), src, kWave);
        const int hi = __shfl(static_cast<int>(x >> 32), src, kWave);
        return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
    } else {
        return __shfl(v, src, kWave);
    }
}

template <int NMAX, typename MT>
struct RegLinSys {
    static constexpr int NS = 1;
    MT a[NMAX];
    int step;      // pivot step of this lane's row (NMAX if not a pivot row yet)
    int pivlane;   // lane k: row (lane) chosen as pivot at step k
    MT rdiag;      // pivot rows: 1 / U_kk of their step

    __device__ __forceinline__ void init(void*, int) {}
    __device__ __forceinline__ void bind(MT*, int*) {}

    __device__ __forceinline__ void form(const double* __restrict__ J, int N, double gamma) {
        const int lane = lane_id();
#pragma unroll
        for (int j = 0; j < NMAX; j++) {
            // 1 - gamma J_jj as one fused multiply-add, as the compiler formed it
            // when the product and the sum sat in one branch
            const double v =
                (lane < N && j < N) ? fma(-gamma, row_entry(J, N, lane, j), lane == j ? 1.0 : 0.0) : 0.0;
            a[j] = static_cast<MT>(v);
        }
    }

    // Partial pivoting without row moves: at step k the row (lane) with the
    // largest |entry| in column k among rows not yet used becomes the k-th
    // pivot row.  The search packs |entry| (upper 58 bits) and the row into one
    // 64-bit key and takes a wavefront max with DPP moves, so ties and
    // near-ties (equal leading 58 bits) go to the lowest row.  The pivot row is
    // broadcast with v_readlane in blocks of 8 columns; blocks entirely right of
    // column k are updated without selects, only the block holding column k
    // selects per column.  Each pivot row keeps 1/U_kk for the back substitution.
    __device__ __forceinline__ bool factor(int N) {
        // N arrives in a VGPR when this is called through an out-of-line
        // function: made uniform, every "k < N" below is a scalar branch
        N = uni(N);
        const int lane = lane_id();
        bool pivoted = lane >= N;
        step = NMAX;
        pivlane = lane;
        rdiag = MT(1);
        MT colk = a[0];
        for (int k = 0; k < N; k++) {
            const unsigned long long key =
                pivoted ? 0ull
                        : ((static_cast<unsigned long long>(__double_as_longlong(fabs(static_cast<double>(colk)))) &
                            ~63ull) |
                           static_cast<unsigned long long>(63 - lane));
            const int p = 63 - static_cast<int>(wave_max_u64(key) & 63ull);
            const MT piv = readlane_t(colk, p);
            if (piv == MT(0)) return false;
            const bool keep = pivoted || lane == p;   // rows that are (now) pivot rows
            const MT rp = MT(1) / piv;
            if (lane == p) { pivoted = true; step = k; rdiag = rp; }
            if (lane == k) pivlane = p;
            const MT l = colk * rp;
            const MT lm = keep ? MT(0) : l;
#pragma unroll
            for (int j0 = 0; j0 < NMAX; j0 += 8) {
                constexpr int B = 8;
                if (j0 > k) {
                    MT pr[B];
#pragma unroll
                    for (int b = 0; b < B; b++)
                        if (j0 + b < NMAX) pr[b] = readlane_t(a[j0 + b], p);
#pragma unroll
                    for (int b = 0; b < B; b++)
                        if (j0 + b < NMAX) a[j0 + b] = a[j0 + b] - lm * pr[b];
                } else if (j0 + B > k) {
                    MT pr[B];
#pragma unroll
                    for (int b = 0; b < B; b++)
                        if (j0 + b < NMAX) pr[b] = readlane_t(a[j0 + b], p);
#pragma unroll
                    for (int b = 0; b < B; b++) {
                        const int j = j0 + b;
                        if (j < NMAX) {
                            const MT upd = a[j] - lm * pr[b];
                            const MT atk = keep ? a[j] : l;
                            a[j] = (j > k) ? upd : ((j == k) ? atk : a[j]);
                        }
                    }
                }
            }
#pragma unroll
            for (int j0 = 0; j0 < NMAX; j0 += 8) {
                if (j0 <= k + 1 && k + 1 < j0 + 8) {
#pragma unroll
                    for (int b = 0; b < 8; b++)
                        if (j0 + b < NMAX && j0 + b == k + 1) colk = a[j0 + b];
                }
            }
        }
        // rows into pivot order (lane k holds the k-th pivot row), so the
        // triangular solves index lanes by step with no lookups
#pragma unroll
        for (int j = 0; j < NMAX; j++) a[j] = shfl_t(a[j], pivlane);
        rdiag = shfl_t(rdiag, pivlane);
        step = lane;
        return true;
    }

    // b: this lane's component of the right-hand side; returns the solution
    // component of the same lane.  With the rows in pivot order, P b is one
    // shuffle and step k of both substitutions reads lane k.
    template <typename V>
    __device__ __forceinline__ V solve(V b, int N) const {
        // All NMAX steps run, with no test against N: outside the leading
        // N x N block the factors are exact zeros and rdiag is 1 (form,
        // restore and factor keep them so), and b is zeroed on lanes >= N, so
        // steps k >= N leave b unchanged.  (A branch on k < N per step made
        // the compiler keep all NMAX conditions in VGPR lanes.)
        const int lane = lane_id();
        b = (lane < N) ? b : V(0);
        b = shfl_t(b, pivlane);
        // The multipliers are masked off the dependency chain (l_ik for rows
        // below k, u_ik for rows above), so a step is readlane -> FMA.  Lane k
        // holds the final numerator of x_k once step k is reached and is not
        // updated afterwards, so x = b / U_kk for all lanes at the end.
#pragma unroll
        for (int k = 0; k < NMAX; k++) {
            const V lk = (lane > k) ? static_cast<V>(a[k]) : V(0);
            const V yk = readlane_t(b, k);
            b = b - lk * yk;
        }
        const V rd = static_cast<V>(rdiag);
#pragma unroll
        for (int k = NMAX - 1; k >= 0; k--) {
            const V uk = (lane < k) ? static_cast<V>(a[k]) : V(0);
            const V xk = readlane_t(b * rd, k);
            b = b - uk * xk;
        }
        return b * rd;
    }

    template <typename V, int S>
    __device__ __forceinline__ void solve(V (&b)[S], int N) const {
        b[0] = solve(b[0], N);
    }

    __device__ __forceinline__ void save(MT* __restrict__ dst, int* __restrict__ pdst, int N) const {
        const int lane = lane_id();
        if (lane < N) {
#pragma unroll
            for (int j = 0; j < NMAX; j++)
                if (j < N) dst[lane * N + j] = a[j];
        }
        dst[static_cast<size_t>(N) * N + lane] = rdiag;
        pdst[lane] = step;
        pdst[kWave + lane] = pivlane;
    }

    __device__ __forceinline__ void restore(const MT* __restrict__ src, const int* __restrict__ psrc, int N) {
        const int lane = lane_id();
#pragma unroll
        for (int j = 0; j < NMAX; j++) a[j] = row_entry(src, N, lane, j);
        rdiag = src[static_cast<size_t>(N) * N + lane];
        step = psrc[lane];
        pivlane = psrc[kWave + lane];
    }
};

// Out-of-line LU of the iteration matrix that LdsRegLinSys::form leaves
// row-major in the same LDS area (leading dimension LD): factored by panels
// with the trailing updates on the matrix cores (lu_factor_mfma, same pivots
// and, entry by entry, the same fused multiply-adds as the register LU), then
// stored in the layout the register LU leaves for lr_solve -- A[j * 64 + i] =
// (PA = LU)(i, j) with row i the i-th pivot row, 1/U_ii after the NMAX
// columns, step[i] = i, pivlane[i] = the original row of pivot row i.
template <int NMAX, typename MT>
__device__ __noinline__ int lr_factor_mfma(int offA, int N, int LD) {
    N = uni(N);
    LD = uni(LD);
    MT* A = reinterpret_cast<MT*>(reinterpret_cast<char*>(lds_base()) + offA);
    int* stp = reinterpret_cast<int*>(A + (NMAX + 1) * kWave);
    int* pvl = stp + kWave;
    if (!lu_factor_mfma<1, MT>(A, LD, N, pvl)) return 0;
    const int lane = lane_id();
    const int row = min(lane, N - 1) * LD;
    MT a[NMAX];
#pragma unroll
    for (int j = 0; j < NMAX; j++) a[j] = load_or_zero(A, row + min(j, N - 1), lane < N && j < N);
    const MT dg = load_or_zero(A, row + min(lane, N - 1), lane < N);
    const int pv = pvl[lane];
    wave_sync();
#pragma unroll
    for (int j = 0; j < NMAX; j++) A[j * kWave + lane] = a[j];
    A[NMAX * kWave + lane] = (lane < N) ? MT(1) / dg : MT(1);
    stp[lane] = lane;
    pvl[lane] = (lane < N) ? pv : lane;
    wave_sync();
    return 1;
}

template <int NMAX, typename MT, typename V>
__device__ __noinline__ V lr_solve(int offA, V b, int N) {
    const MT* A = reinterpret_cast<const MT*>(reinterpret_cast<const char*>(lds_base()) + offA);
    const int* stp = reinterpret_cast<const int*>(A + (NMAX + 1) * kWave);
    const int* pvl = stp + kWave;
    RegLinSys<NMAX, MT> r;
    const int lane = lane_id();
#pragma unroll
    for (int j = 0; j < NMAX; j++) r.a[j] = A[j * kWave + lane];
    r.rdiag = A[NMAX * kWave + lane];
    r.step = stp[lane];
    r.pivlane = pvl[lane];
    return r.solve(b, N);
}

template <int NMAX, typename MT>
struct LdsRegLinSys {
    static constexpr int NS = 1;
    MT* A;        // LDS, (NMAX + 1) x 64: column j at A + 64 j, then 1/U_kk
    int* piv;     // LDS, step[64] then pivlane[64]
    int offA;     // byte offset of A from the start of dynamic LDS
    int LD;       // leading dimension of the row-major matrix that form writes

    __device__ __forceinline__ void init(void* lds, int ld) {
        LD = ld;
        A = static_cast<MT*>(lds);
        piv = reinterpret_cast<int*>(A + (NMAX + 1) * kWave);
        offA = static_cast<int>(static_cast<char*>(lds) - reinterpret_cast<char*>(lds_base()));
    }
    __device__ __forceinline__ void bind(MT*, int*) {}

    // M = I - gamma J written row-major (leading dimension LD) for the
    // factorisation: lane j takes column j, so each row of J is one coalesced
    // load; rows in blocks of 8 keep 8 loads in flight
    __device__ __forceinline__ void form(const double* __restrict__ J, int N, double gamma) {
        const int lane = lane_id();
        if (lane < N) {
            int i = 0;
            for (; i + 8 <= N; i += 8) {
                double x[8];
#pragma unroll
                for (int b = 0; b < 8; b++) x[b] = J[(i + b) * N + lane];
#pragma unroll
                for (int b = 0; b < 8; b++)
                    A[(i + b) * LD + lane] = static_cast<MT>(fma(-gamma, x[b], (i + b == lane) ? 1.0 : 0.0));
            }
            for (; i < N; i++)
                A[i * LD + lane] = static_cast<MT>(fma(-gamma, J[i * N + lane], (i == lane) ? 1.0 : 0.0));
        }
        wave_sync();
    }

    // A fresh Jacobian can be assembled in this matrix area itself (FP64, and
    // M's row stride equal to N, the Jacobian's): the assembly then uses LDS
    // atomics instead of global ones, and M = I - gamma J is formed in place.
    // The area holds (NMAX + 1) x 64 >= NMAX^2 doubles.
    __device__ __forceinline__ bool jac_in_place(int N) const { return sizeof(MT) == 8 && LD == N; }
    __device__ __forceinline__ int jac_offset() const { return offA / static_cast<int>(sizeof(double)); }
    __device__ __forceinline__ void form_in_place(int N, double gamma) {
        static_assert(NMAX * NMAX <= (NMAX + 1) * kWave, "the Jacobian must fit the matrix area");
        const int lane = lane_id();
        double* J = reinterpret_cast<double*>(A);
        if (lane < N) {
            int i = 0;
            for (; i + 8 <= N; i += 8) {
                double x[8];
#pragma unroll
                for (int b = 0; b < 8; b++) x[b] = J[(i + b) * N + lane];
#pragma unroll
                for (int b = 0; b < 8; b++) J[(i + b) * N + lane] = fma(-gamma, x[b], (i + b == lane) ? 1.0 : 0.0);
            }
            for (; i < N; i++) J[i * N + lane] = fma(-gamma, J[i * N + lane], (i == lane) ? 1.0 : 0.0);
        }
        wave_sync();
    }

    __device__ __forceinline__ bool factor(int N) { return lr_factor_mfma<NMAX, MT>(offA, N, LD) != 0; }

    template <typename V, int S>
    __device__ __forceinline__ void solve(V (&b)[S], int N) const {
        b[0] = lr_solve<NMAX, MT, V>(offA, b[0], N);
    }

    __device__ __forceinline__ void save(MT* __restrict__ dst, int* __restrict__ pdst, int) const {
        const int lane = lane_id();
#pragma unroll
        for (int j = 0; j <= NMAX; j++) dst[j * kWave + lane] = A[j * kWave + lane];
        pdst[lane] = piv[lane];
        pdst[kWave + lane] = piv[kWave + lane];
    }

    // The NMAX + 1 saved columns in blocks of 16: a block's loads are issued
    // together, and the pointers step by a block so that every load has an
    // immediate offset.  (Fully unrolled, the compiler kept one address
    // register per column for the whole kernel and waited for each pair of
    // loads in turn: 29 round trips to memory per reactor.)
    __device__ __forceinline__ void restore(const MT* __restrict__ src, const int* __restrict__ psrc, int) {
        constexpr int kCols = NMAX + 1, kB = 16;
        const int lane = lane_id();
        const MT* s = src + lane;
        MT* d = A + lane;
#pragma unroll 1
        for (int j0 = 0; j0 + kB <= kCols; j0 += kB) {
            MT x[kB];
#pragma unroll
            for (int b = 0; b < kB; b++) x[b] = s[b * kWave];
#pragma unroll
            for (int b = 0; b < kB; b++) d[b * kWave] = x[b];
            s += kB * kWave;
            d += kB * kWave;
        }
        {
            constexpr int kR = kCols % kB;
            MT x[kR > 0 ? kR : 1];
#pragma unroll
            for (int b = 0; b < kR; b++) x[b] = s[b * kWave];
            const int p0 = psrc[lane], p1 = psrc[kWave + lane];
#pragma unroll
            for (int b = 0; b < kR; b++) d[b * kWave] = x[b];
            piv[lane] = p0;
            piv[kWave + lane] = p1;
        }
        wave_sync();
    }
};

// Out-of-line register LU / solves of a matrix kept in global memory (the
// reactor's own saved-factor slot), layout A[j * 64 + i] = M(i, j), followed by
// step[64] and pivlane[64] in the pivot slot.  Each lane reads and writes only
// its own row, so no fence is needed between the factorisation and the solves.
template <int NMAX, typename MT>
__device__ __noinline__ int gr_factor(MT* __restrict__ A, int* __restrict__ piv, int N) {
    RegLinSys<NMAX, MT> r;
    const int lane = lane_id();
#pragma unroll
    for (int j = 0; j < NMAX; j++) r.a[j] = A[j * kWave + lane];
    const bool ok = r.factor(N);
#pragma unroll
    for (int j = 0; j < NMAX; j++) A[j * kWave + lane] = r.a[j];
    A[NMAX * kWave + lane] = r.rdiag;
    piv[lane] = r.step;
    piv[kWave + lane] = r.pivlane;
    return ok ? 1 : 0;
}

template <int NMAX, typename MT, typename V>
__device__ __noinline__ V gr_solve(const MT* __restrict__ A, const int* __restrict__ piv, V b, int N) {
    RegLinSys<NMAX, MT> r;
    const int lane = lane_id();
#pragma unroll
    for (int j = 0; j < NMAX; j++) r.a[j] = A[j * kWave + lane];
    r.rdiag = A[NMAX * kWave + lane];
    r.step = piv[lane];
    r.pivlane = piv[kWave + lane];
    return r.solve(b, N);
}

// M = I - gamma J into the global slot (column j of M at A + 64 j)
template <int NMAX, typename MT>
__device__ __noinline__ void gr_form(const double* __restrict__ J, MT* __restrict__ A, int N, double gamma) {
    N = uni(N);
    const int lane = lane_id();
#pragma unroll
    for (int j = 0; j < NMAX; j++) {
        const double v =
            (lane < N && j < N) ? fma(-gamma, row_entry(J, N, lane, j), lane == j ? 1.0 : 0.0) : 0.0;
        A[j * kWave + lane] = static_cast<MT>(v);
    }
}

template <int NMAX, typename MT>
struct GlobalRegLinSys {
    static constexpr int NS = 1;
    MT* A = nullptr;     // the reactor's factor slot ((NMAX + 1) x 64, last column 1/U_kk)
    int* piv = nullptr;  // the reactor's pivot slot (128 ints)

    __device__ __forceinline__ void init(void*, int) {}
    __device__ __forceinline__ void bind(MT* a, int* p) { A = a; piv = p; }

    __device__ __forceinline__ void form(const double* __restrict__ J, int N, double gamma) {
        gr_form<NMAX, MT>(J, A, N, gamma);
    }

    __device__ __forceinline__ bool factor(int N) { return gr_factor<NMAX, MT>(A, piv, N) != 0; }

    template <typename V, int S>
    __device__ __forceinline__ void solve(V (&b)[S], int N) const {
        b[0] = gr_solve<NMAX, MT, V>(A, piv, b[0], N);
    }

    // the factors already live in the reactor's slot
    __device__ __forceinline__ void save(MT*, int*, int) const {}
    __device__ __forceinline__ void restore(const MT*, const int*, int) {}
};

template <int NS_, typename MT>
struct LdsLinSys {
    static constexpr int NS = NS_;
    MT* A;        // LDS, N x LD
    int LD;
    int* pmap;    // LDS
    MT* bv;       // LDS

    __device__ __forceinline__ void bind(MT*, int*) {}
    __device__ __forceinline__ void init(void* lds, int ld) {
        LD = ld;
        char* p = static_cast<char*>(lds);
        bv = reinterpret_cast<MT*>(p);
        p += kWave * NS * sizeof(double);
        pmap = reinterpret_cast<int*>(p);
        p += kWave * NS * sizeof(int);
        A = reinterpret_cast<MT*>(p);
    }

    // rows in blocks of 8 so that the block's loads of J are in flight together
    __device__ __forceinline__ void form(const double* __restrict__ J, int N, double gamma) {
        const int lane = lane_id();
        const MT mg = static_cast<MT>(-gamma);
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int j = lane + kWave * s;
            if (j < N) {
                int i = 0;
                for (; i + 8 <= N; i += 8) {
                    double x[8];
#pragma unroll
                    for (int b = 0; b < 8; b++) x[b] = J[(i + b) * N + j];
#pragma unroll
                    for (int b = 0; b < 8; b++) {
                        MT v = mg * static_cast<MT>(x[b]);
                        if (i + b == j) v += MT(1);
                        A[(i + b) * LD + j] = v;
                    }
                }
                for (; i < N; i++) {
                    MT v = mg * static_cast<MT>(J[i * N + j]);
                    if (i == j) v += MT(1);
                    A[i * LD + j] = v;
                }
            }
        }
        wave_sync();
    }

    // A fresh FP64 Jacobian can be assembled in this matrix area (N x N
    // row-major, inside the N x LD area): LDS atomics instead of global ones,
    // then M = I - gamma J formed in place, each row moved to stride LD.
    __device__ __forceinline__ bool jac_in_place(int) const { return sizeof(MT) == 8; }
    __device__ __forceinline__ int jac_offset() const {
        return static_cast<int>((reinterpret_cast<char*>(A) - reinterpret_cast<char*>(lds_base())) / sizeof(double));
    }
    // Rows from the last to the first, 8 at a time, every load of a block
    // before any store: a row's new place (stride LD >= N) overlaps only the
    // old places of rows below it, which have been read already.
    __device__ __forceinline__ void form_in_place(int N, double gamma) {
        const int lane = lane_id();
        double* J = reinterpret_cast<double*>(A);
        int i1 = N;   // rows i1 .. N-1 are done
        while (i1 > 0) {
            const int i0 = max(i1 - 8, 0);
            double x[8][NS];
#pragma unroll
            for (int b = 0; b < 8; b++)
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    const int i = min(i0 + b, N - 1), j = min(lane + kWave * s, N - 1);
                    x[b][s] = J[i * N + j];
                }
            wave_sync();
#pragma unroll
            for (int b = 0; b < 8; b++)
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    const int i = i0 + b, j = lane + kWave * s;
                    if (i < i1 && j < N) {
                        // as form: (-gamma) J, plus 1 on the diagonal
                        double v = -gamma * x[b][s];
                        if (i == j) v += 1.0;
                        J[i * LD + j] = v;
                    }
                }
            wave_sync();
            i1 = i0;
        }
    }

    __device__ __forceinline__ bool factor(int N) { return lu_factor_mfma<NS, MT>(A, LD, N, pmap); }

    template <typename V>
    __device__ __forceinline__ void solve(V (&b)[NS], int N) const {
        lu_solve<NS, MT, V>(A, LD, N, pmap, reinterpret_cast<V*>(bv), b);
    }

    __device__ __forceinline__ void save(MT* __restrict__ dst, int* __restrict__ pdst, int N) const {
        const int lane = lane_id();
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int j = lane + kWave * s;
                if (j < N) dst[i * N + j] = A[i * LD + j];
            }
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int j = lane + kWave * s;
            if (j < N) pdst[j] = pmap[j];
        }
    }

    // Rows in blocks of 8 with the column index clamped instead of tested,
    // so that a block's 8 x NS loads are issued together (a load under the
    // j < N test was waited for before the next one was issued).
    __device__ __forceinline__ void restore(const MT* __restrict__ src, const int* __restrict__ psrc, int N) {
        const int lane = lane_id();
        int jc[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) jc[s] = min(lane + kWave * s, N - 1);
        int i = 0;
        for (; i + 8 <= N; i += 8) {
            MT x[8][NS];
#pragma unroll
            for (int b = 0; b < 8; b++)
#pragma unroll
                for (int s = 0; s < NS; s++) x[b][s] = src[(i + b) * N + jc[s]];
#pragma unroll
            for (int b = 0; b < 8; b++)
#pragma unroll
                for (int s = 0; s < NS; s++)
                    if (lane + kWave * s < N) A[(i + b) * LD + lane + kWave * s] = x[b][s];
        }
        for (; i < N; i++)
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const MT v = src[i * N + jc[s]];
                if (lane + kWave * s < N) A[i * LD + lane + kWave * s] = v;
            }
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int j = lane + kWave * s;
            const int v = psrc[jc[s]];
            if (j < N) pmap[j] = v;
        }
        wave_sync();
    }
};

}  // namespace cg
