// Iteration-matrix storage and dense LU for one reactor per wavefront.
//
// RegLinSys: the matrix lives in registers, row i in lane i (N <= NMAX <= 64).
//   Partial pivoting is implicit: no row moves; the lane chosen at step k is
//   marked as the k-th pivot row, and the triangular solves run in pivot order.
//   Every register index is static (all loops over columns are unrolled), so
//   nothing spills to scratch.  The arithmetic is the same as a row-swapping
//   LU (same multipliers, same updates), only the storage differs.
// LdsLinSys: the matrix lives in LDS (row-major, leading dimension LD), used
//   for N > 64 (two lanes' worth of rows per lane).
//
// Both read J from the FP64 saved-Jacobian buffer and form M = I - gamma J in
// the matrix precision MT (double: plain mode; float: mixed mode).
#pragma once
#include "wave_linalg.hpp"

namespace cg {

template <typename T>
__device__ __forceinline__ T readlane_t(T v, int l) {
    if constexpr (sizeof(T) == 8) {
        long long x = __double_as_longlong(v);
        int lo = __builtin_amdgcn_readlane(static_cast<int>(x & 0xffffffffLL), l);
        int hi = __builtin_amdgcn_readlane(static_cast<int>(x >> 32), l);
        return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
    } else {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
    }
}

template <int NMAX, typename MT>
struct RegLinSys {
    static constexpr int NS = 1;
    MT a[NMAX];
    int step;      // pivot step of this lane's row (NMAX if not a pivot row yet)
    int pivlane;   // lane k: row (lane) chosen as pivot at step k

    __device__ __forceinline__ void init(void*, int) {}

    __device__ __forceinline__ void form(const double* __restrict__ J, int N, double gamma) {
        const int lane = lane_id();
        const bool act = lane < N;
#pragma unroll
        for (int j = 0; j < NMAX; j++) {
            double v = 0.0;
            if (act && j < N) {
                v = -gamma * J[lane * N + j];
                if (lane == j) v += 1.0;
            }
            a[j] = static_cast<MT>(v);
        }
    }

    __device__ __forceinline__ bool factor(int N) {
        const int lane = lane_id();
        const bool act = lane < N;
        bool pivoted = !act;
        step = NMAX;
        pivlane = lane;
        MT colk = a[0];
        for (int k = 0; k < N; k++) {
            double best = pivoted ? -1.0 : fabs(static_cast<double>(colk));
            int bi = lane;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                double ob = __shfl_xor(best, off, kWave);
                int oi = __shfl_xor(bi, off, kWave);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            const int p = uni(bi);
            const MT piv = readlane_t(colk, p);
            if (piv == MT(0)) return false;
            const bool was = pivoted;
            if (lane == p) { pivoted = true; step = k; }
            if (lane == k) pivlane = p;
            const MT mult = MT(1) / piv;
            const MT l = colk * mult;
#pragma unroll
            for (int j = 0; j < NMAX; j++) {
                if (j < N) {
                    if (j == k) {
                        if (!was && lane != p) a[j] = l;
                    } else if (j > k) {
                        const MT pj = readlane_t(a[j], p);
                        if (!was && lane != p) a[j] -= l * pj;
                        if (j == k + 1) colk = a[j];
                    }
                }
            }
        }
        return true;
    }

    // b: this lane's component of the right-hand side; returns the solution
    template <typename V>
    __device__ __forceinline__ V solve(V b, int N) const {
        const int lane = lane_id();
#pragma unroll
        for (int k = 0; k < NMAX; k++) {
            if (k < N) {
                const int pk = __builtin_amdgcn_readlane(pivlane, k);
                const V yk = readlane_t(b, pk);
                if (step > k) b -= static_cast<V>(a[k]) * yk;
            }
        }
#pragma unroll
        for (int k = NMAX - 1; k >= 0; k--) {
            if (k < N) {
                const int pk = __builtin_amdgcn_readlane(pivlane, k);
                if (lane == pk) b = b / static_cast<V>(a[k]);
                const V xk = readlane_t(b, pk);
                if (step < k) b -= static_cast<V>(a[k]) * xk;
            }
        }
        return __shfl(b, pivlane, kWave);
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
        pdst[lane] = step;
        pdst[kWave + lane] = pivlane;
    }

    __device__ __forceinline__ void restore(const MT* __restrict__ src, const int* __restrict__ psrc, int N) {
        const int lane = lane_id();
#pragma unroll
        for (int j = 0; j < NMAX; j++) a[j] = (lane < N && j < N) ? src[lane * N + j] : MT(0);
        step = psrc[lane];
        pivlane = psrc[kWave + lane];
    }
};

template <int NS_, typename MT>
struct LdsLinSys {
    static constexpr int NS = NS_;
    MT* A;        // LDS, N x LD
    int LD;
    int* pmap;    // LDS
    MT* bv;       // LDS

    __device__ __forceinline__ void init(void* lds, int ld) {
        LD = ld;
        char* p = static_cast<char*>(lds);
        bv = reinterpret_cast<MT*>(p);
        p += kWave * NS * sizeof(double);
        pmap = reinterpret_cast<int*>(p);
        p += kWave * NS * sizeof(int);
        A = reinterpret_cast<MT*>(p);
    }

    __device__ __forceinline__ void form(const double* __restrict__ J, int N, double gamma) {
        const int lane = lane_id();
        const MT mg = static_cast<MT>(-gamma);
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int j = lane + kWave * s;
            if (j < N)
                for (int i = 0; i < N; i++) {
                    MT v = mg * static_cast<MT>(J[i * N + j]);
                    if (i == j) v += MT(1);
                    A[i * LD + j] = v;
                }
        }
        wave_sync();
    }

    __device__ __forceinline__ bool factor(int N) { return lu_factor<NS, MT>(A, LD, N, pmap); }

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

    __device__ __forceinline__ void restore(const MT* __restrict__ src, const int* __restrict__ psrc, int N) {
        const int lane = lane_id();
        for (int i = 0; i < N; i++)
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int j = lane + kWave * s;
                if (j < N) A[i * LD + j] = src[i * N + j];
            }
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int j = lane + kWave * s;
            if (j < N) pmap[j] = psrc[j];
        }
        wave_sync();
    }
};

}  // namespace cg
