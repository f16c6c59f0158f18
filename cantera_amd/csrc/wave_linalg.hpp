// Dense LU with partial pivoting and triangular solves for one matrix held in
// LDS, one wavefront per matrix.  Row i of the solution vector lives in lane
// i % 64, slot i / 64.  Pivot search is a wavefront arg-max with shuffles
// (first maximal entry wins, as in a column-oriented getrf).
#pragma once
#include "wave_kinetics.hpp"

namespace cg {

// v of lane l (l wave-uniform), for FP64 or FP32 values
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

// Factor A (N x N, leading dimension LD, row-major) in place: PA = LU with unit
// L below the diagonal.  pmap[i] = original row of permuted row i.  Returns
// false on an exactly zero pivot.
template <int NS, typename T>
__device__ bool lu_factor(T* A, int LD, int N, int* pmap) {
    N = uni(N);
    LD = uni(LD);
    const int lane = lane_id();
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int i = lane + kWave * s;
        if (i < N) pmap[i] = i;
    }
    wave_sync();
    for (int k = 0; k < N; k++) {
        // arg-max over rows i >= k as one 64-bit key: |A[i][k]| in the upper
        // 57 bits, 127 - i in the low 7 (ties and near-ties go to the lowest row)
        unsigned long long key = 0ull;
#pragma unroll
        for (int s = 0; s < NS; s++) {
            int i = lane + kWave * s;
            if (i >= k && i < N) {
                const unsigned long long kk =
                    (static_cast<unsigned long long>(__double_as_longlong(fabs(static_cast<double>(A[i * LD + k])))) &
                     ~127ull) |
                    static_cast<unsigned long long>(127 - i);
                key = kk > key ? kk : key;
            }
        }
        const int p = 127 - static_cast<int>(wave_max_u64(key) & 127ull);
        if (p < k || p >= N) return false;
        if (uni(A[p * LD + k]) == T(0)) return false;
        if (p != k) {
#pragma unroll
            for (int s = 0; s < NS; s++) {
                int j = lane + kWave * s;
                if (j < N) {
                    T a = A[k * LD + j];
                    A[k * LD + j] = A[p * LD + j];
                    A[p * LD + j] = a;
                }
            }
            if (lane == 0) {
                int t = pmap[k];
                pmap[k] = pmap[p];
                pmap[p] = t;
            }
        }
        wave_sync();
        const T mult = T(1) / A[k * LD + k];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            int i = lane + kWave * s;
            if (i > k && i < N) A[i * LD + k] *= mult;
        }
        wave_sync();
        // Trailing update, lane j owns column j: A[i][j] -= A[i][k] * A[k][j]
        // for i > k.  Rows are taken in blocks of kB so that the LDS loads of
        // a block (A[i][k] broadcast, A[i][j] per lane) are in flight together.
        constexpr int kB = 8;
#pragma unroll
        for (int s = 0; s < NS; s++) {
            int j = lane + kWave * s;
            if (j > k && j < N) {
                const T akj = A[k * LD + j];
                if (akj != T(0)) {
                    int i = k + 1;
                    for (; i + kB <= N; i += kB) {
                        T lik[kB], aij[kB];
#pragma unroll
                        for (int b = 0; b < kB; b++) {
                            lik[b] = A[(i + b) * LD + k];
                            aij[b] = A[(i + b) * LD + j];
                        }
#pragma unroll
                        for (int b = 0; b < kB; b++) A[(i + b) * LD + j] = aij[b] - lik[b] * akj;
                    }
                    for (; i < N; i++) A[i * LD + j] -= A[i * LD + k] * akj;
                }
            }
        }
        wave_sync();
    }
    return true;
}

// Solve (PA) x = P b in place: b holds the right-hand side distributed over
// lanes and returns x.  bv is an LDS scratch vector of length N.
template <int NS, typename T, typename V>
__device__ void lu_solve(const T* A, int LD, int N, const int* pmap, V* bv, V (&b)[NS]) {
    N = uni(N);
    LD = uni(LD);
    const int lane = lane_id();
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int i = lane + kWave * s;
        if (i < N) bv[i] = b[s];
    }
    wave_sync();
#pragma unroll
    for (int s = 0; s < NS; s++) {
        int i = lane + kWave * s;
        b[s] = (i < N) ? bv[pmap[i]] : V(0);
    }
    wave_sync();
    // Forward substitution with the unit lower factor, then back substitution.
    // Step k broadcasts b_k (lane k % 64, slot k / 64 -- the slot is a
    // compile-time index in each loop, so no select sits on the chain) and
    // updates the other rows with the column-k multipliers, loaded kB columns
    // ahead and masked to zero outside the triangle: a step is readlane -> FMA.
    constexpr int kB = 4;
#pragma unroll
    for (int t = 0; t < NS; t++) {
        const int kend = min(N - 1, kWave * (t + 1));
        for (int k0 = kWave * t; k0 < kend; k0 += kB) {
            T lv[kB][NS];
#pragma unroll
            for (int b2 = 0; b2 < kB; b2++)
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    const int i = lane + kWave * s;
                    const int k = k0 + b2;
                    lv[b2][s] = (i > k && i < N && k < kend) ? A[i * LD + k] : T(0);
                }
#pragma unroll
            for (int b2 = 0; b2 < kB; b2++) {
                const int k = k0 + b2;
                if (k < kend) {
                    const V bk = readlane_t(b[t], k - kWave * t);
#pragma unroll
                    for (int s = 0; s < NS; s++) b[s] -= static_cast<V>(lv[b2][s]) * bk;
                }
            }
        }
    }
    // Back substitution: 1/U_ii for every row first (in parallel); lane k holds
    // the final numerator of x_k when step k is reached and is not updated
    // afterwards, so x = b / U_ii for all rows at the end.
    V rd[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + kWave * s;
        rd[s] = (i < N) ? V(1) / static_cast<V>(A[i * LD + i]) : V(1);
    }
#pragma unroll
    for (int t = NS - 1; t >= 0; t--) {
        const int klo = kWave * t;
        for (int k0 = min(N, kWave * (t + 1)) - 1; k0 >= klo; k0 -= kB) {
            T uv[kB][NS];
#pragma unroll
            for (int b2 = 0; b2 < kB; b2++)
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    const int i = lane + kWave * s;
                    const int k = k0 - b2;
                    uv[b2][s] = (k >= klo && i < k) ? A[i * LD + k] : T(0);
                }
#pragma unroll
            for (int b2 = 0; b2 < kB; b2++) {
                const int k = k0 - b2;
                if (k >= klo) {
                    const V xk = readlane_t(b[t], k - klo) * readlane_t(rd[t], k - klo);
#pragma unroll
                    for (int s = 0; s < NS; s++) b[s] -= static_cast<V>(uv[b2][s]) * xk;
                }
            }
        }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) b[s] *= rd[s];
}

}  // namespace cg
