// Dense LU with partial pivoting and triangular solves for one matrix held in
// LDS, one wavefront per matrix.  Row i of the solution vector lives in lane
// i % 64, slot i / 64.  Pivot search is a wavefront arg-max with shuffles
// (first maximal entry wins, as in a column-oriented getrf).
#pragma once
#include <type_traits>

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

// One 16 x 16 trailing tile C -= L U (L: 16 x 4, U: 4 x 16) on the matrix
// cores: v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32.  Operand lanes (both
// checked on the hardware with an in-bounds probe): A(i, k) from lane i + 16 k,
// B(k, j) from lane j + 16 k; accumulator element (i, j) in lane j + 16 q,
// slot r, with i = q + 4 r (FP64) or i = 4 q + r (FP32), q = lane / 16.
typedef double mfma_d4 __attribute__((ext_vector_type(4)));
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ int mfma_acc_row(int lane, int r) {
    return (sizeof(T) == 8) ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
}

// A[idx] read unconditionally (idx must lie inside the matrix), zero where
// !keep.  The value passes through an empty asm statement: a plain select on a
// loaded value lets the compiler move the load under the condition, with a
// wait for it right there, which serialises every load of a tile.
template <typename T>
__device__ __forceinline__ T load_or_zero(const T* A, int idx, bool keep) {
    T x = A[idx];
    if constexpr (sizeof(T) == 8) x = opaque(x);
    else asm("" : "+v"(x));
    return keep ? x : T(0);
}

// Row r of a lane-distributed column (r wave-uniform).
template <int NS, typename T>
__device__ __forceinline__ T read_row(const T (&v)[NS], int r) {
    T x = T(0);
#pragma unroll
    for (int s = 0; s < NS; s++)
        if (r / kWave == s) x = readlane_t(v[s], r % kWave);
    return x;
}

// Operands of trailing tile (i0, j0) for the rank-4 update C -= L U,
// L = A[rows][k0 .. k0+3], U = A[k0 .. k0+3][cols]: addresses are clamped into
// the matrix and values outside it replaced by zero, so the loads need no
// branches.
template <typename T>
struct MfmaTile {
    using V4 = typename std::conditional<sizeof(T) == 8, mfma_d4, mfma_f4>::type;
    V4 c;
    T lneg, u;
    __device__ __forceinline__ void load(const T* A, int LD, int N, int k0, int i0, int j0) {
        const int lane = lane_id();
        const int kq = lane >> 4, lr = lane & 15;
        const int i = i0 + lr, j = j0 + lr;
        const int jc = min(j, N - 1);
        lneg = -load_or_zero(A, min(i, N - 1) * LD + k0 + kq, i < N);
        u = load_or_zero(A, (k0 + kq) * LD + jc, j < N);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int ic = i0 + mfma_acc_row<T>(lane, r);
            c[r] = load_or_zero(A, min(ic, N - 1) * LD + jc, ic < N && j < N);
        }
    }
    __device__ __forceinline__ void update() {
        if constexpr (sizeof(T) == 8) c = __builtin_amdgcn_mfma_f64_16x16x4f64(lneg, u, c, 0, 0, 0);
        else c = __builtin_amdgcn_mfma_f32_16x16x4f32(lneg, u, c, 0, 0, 0);
    }
    __device__ __forceinline__ void store(T* A, int LD, int N, int i0, int j0) const {
        const int lane = lane_id();
        const int j = j0 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int ic = i0 + mfma_acc_row<T>(lane, r);
            if (ic < N && j < N) A[ic * LD + j] = c[r];
        }
    }
};

// lu_factor blocked by panels of 4 columns: each panel column is factored with
// partial pivoting (full-row swaps; the pivot value and the column come from
// the registers of the arg-max search), the panel's 4 rows of U to the right
// are finished by the unit-lower 4 x 4 block, and the trailing matrix takes the
// rank-4 update A22 -= L21 U12 in 16 x 16 tiles on the matrix cores.  Same
// result layout as lu_factor (PA = LU in place, pmap); every entry receives
// the same fused multiply-adds in the same order as in lu_factor.
template <int NS, typename T>
__device__ bool lu_factor_mfma(T* A, int LD, int N, int* pmap) {
    N = uni(N);
    LD = uni(LD);
    const int lane = lane_id();
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + kWave * s;
        if (i < N) pmap[i] = i;
    }
    wave_sync();
    for (int k0 = 0; k0 < N; k0 += 4) {
        const int kend = min(k0 + 4, N);
        for (int k = k0; k < kend; k++) {
            T col[NS];
            unsigned long long key = 0ull;
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int i = lane + kWave * s;
                col[s] = load_or_zero(A, min(i, N - 1) * LD + k, i >= k && i < N);
                if (i >= k && i < N) {
                    const unsigned long long kk =
                        (static_cast<unsigned long long>(__double_as_longlong(fabs(static_cast<double>(col[s])))) &
                         ~127ull) |
                        static_cast<unsigned long long>(127 - i);
                    key = kk > key ? kk : key;
                }
            }
            const int p = 127 - static_cast<int>(wave_max_u64(key) & 127ull);
            if (p < k || p >= N) return false;
            const T piv = read_row<NS, T>(col, p);
            if (piv == T(0)) return false;
            const T akk = read_row<NS, T>(col, k);
            const T mult = T(1) / piv;
            if (p != k) {
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    const int j = lane + kWave * s;
                    if (j < N) {
                        const T a = A[k * LD + j];
                        A[k * LD + j] = A[p * LD + j];
                        A[p * LD + j] = a;
                    }
                }
                if (lane == 0) {
                    const int t = pmap[k];
                    pmap[k] = pmap[p];
                    pmap[p] = t;
                }
            }
            wave_sync();
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int i = lane + kWave * s;
                if (i > k && i < N) {
                    const T lik = ((i == p) ? akk : col[s]) * mult;
                    A[i * LD + k] = lik;
                    for (int c = k + 1; c < kend; c++) A[i * LD + c] -= lik * A[k * LD + c];
                }
            }
            wave_sync();
        }
        const int t0 = k0 + 4;
        if (t0 >= N) break;
        // U12: rows k0+1 .. k0+3 of the columns right of the panel
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const int j = lane + kWave * s;
            if (j >= t0 && j < N) {
                const T u0 = A[k0 * LD + j];
                const T u1 = A[(k0 + 1) * LD + j] - A[(k0 + 1) * LD + k0] * u0;
                const T u2 = A[(k0 + 2) * LD + j] - A[(k0 + 2) * LD + k0] * u0 - A[(k0 + 2) * LD + k0 + 1] * u1;
                const T u3 = A[(k0 + 3) * LD + j] - A[(k0 + 3) * LD + k0] * u0 - A[(k0 + 3) * LD + k0 + 1] * u1 -
                             A[(k0 + 3) * LD + k0 + 2] * u2;
                A[(k0 + 1) * LD + j] = u1;
                A[(k0 + 2) * LD + j] = u2;
                A[(k0 + 3) * LD + j] = u3;
            }
        }
        wave_sync();
        // A22 -= L21 U12 in 16 x 16 tiles: interior tiles (four at a time) with
        // per-lane offsets fixed for the panel and no masks, edge tiles masked
        const int nt = (N - t0 + 15) / 16, nfull = (N - t0) / 16;
        const int kq = lane >> 4, lr = lane & 15;
        int roff[4];
#pragma unroll
        for (int r = 0; r < 4; r++) roff[r] = mfma_acc_row<T>(lane, r) * LD + lr;
        const int uoff = (k0 + kq) * LD + lr;
        using V4 = typename MfmaTile<T>::V4;
        for (int ti = 0; ti < nt; ti++) {
            const int i0 = t0 + 16 * ti;
            if (ti < nfull) {
                const T lneg = -A[(i0 + lr) * LD + k0 + kq];
                const int rb = i0 * LD;
                int tj = 0;
                for (; tj + 3 < nfull; tj += 4) {
                    const int j0 = t0 + 16 * tj;
                    T u[4];
                    V4 c[4];
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        u[b] = A[uoff + j0 + 16 * b];
#pragma unroll
                        for (int r = 0; r < 4; r++) c[b][r] = A[rb + j0 + 16 * b + roff[r]];
                    }
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        if constexpr (sizeof(T) == 8) c[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(lneg, u[b], c[b], 0, 0, 0);
                        else c[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(lneg, u[b], c[b], 0, 0, 0);
                    }
#pragma unroll
                    for (int b = 0; b < 4; b++)
#pragma unroll
                        for (int r = 0; r < 4; r++) A[rb + j0 + 16 * b + roff[r]] = c[b][r];
                }
                for (; tj + 1 < nfull; tj += 2) {
                    const int j0 = t0 + 16 * tj;
                    const T u0 = A[uoff + j0], u1 = A[uoff + j0 + 16];
                    V4 c0, c1;
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        c0[r] = A[rb + j0 + roff[r]];
                        c1[r] = A[rb + j0 + 16 + roff[r]];
                    }
                    if constexpr (sizeof(T) == 8) {
                        c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(lneg, u0, c0, 0, 0, 0);
                        c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(lneg, u1, c1, 0, 0, 0);
                    } else {
                        c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(lneg, u0, c0, 0, 0, 0);
                        c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(lneg, u1, c1, 0, 0, 0);
                    }
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        A[rb + j0 + roff[r]] = c0[r];
                        A[rb + j0 + 16 + roff[r]] = c1[r];
                    }
                }
                for (; tj < nfull; tj++) {
                    const int j0 = t0 + 16 * tj;
                    const T u0 = A[uoff + j0];
                    V4 c0;
#pragma unroll
                    for (int r = 0; r < 4; r++) c0[r] = A[rb + j0 + roff[r]];
                    if constexpr (sizeof(T) == 8) c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(lneg, u0, c0, 0, 0, 0);
                    else c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(lneg, u0, c0, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; r++) A[rb + j0 + roff[r]] = c0[r];
                }
                if (nfull < nt) {
                    MfmaTile<T> e;
                    e.load(A, LD, N, k0, i0, t0 + 16 * nfull);
                    e.update();
                    e.store(A, LD, N, i0, t0 + 16 * nfull);
                }
            } else {
                for (int tj = 0; tj < nt; tj++) {
                    MfmaTile<T> e;
                    e.load(A, LD, N, k0, i0, t0 + 16 * tj);
                    e.update();
                    e.store(A, LD, N, i0, t0 + 16 * tj);
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
    // updates the other rows with the column-k multipliers, masked to zero
    // outside the triangle: a step is readlane -> FMA.  The multipliers come in
    // blocks of kB columns, each block's loads (in-bounds by clamping, masked
    // only where they are used) issued while the previous block's steps run,
    // so the LDS latency is off the chain.
    constexpr int kB = 8;
    auto column_block = [&](int k0, T (&v)[kB][NS]) {
#pragma unroll
        for (int b2 = 0; b2 < kB; b2++)
#pragma unroll
            for (int s = 0; s < NS; s++)
                v[b2][s] = A[min(lane + kWave * s, N - 1) * LD + min(max(k0 + b2, 0), N - 1)];
    };
#pragma unroll
    for (int t = 0; t < NS; t++) {
        const int kend = min(N - 1, kWave * (t + 1));
        T lv[kB][NS];
        column_block(kWave * t, lv);
        for (int k0 = kWave * t; k0 < kend; k0 += kB) {
            T nx[kB][NS];
            column_block(k0 + kB, nx);
#pragma unroll
            for (int b2 = 0; b2 < kB; b2++) {
                const int k = k0 + b2;
                if (k < kend) {
                    const V bk = readlane_t(b[t], k - kWave * t);
                    // slots s < t hold rows i < 64 t <= k: nothing to update
#pragma unroll
                    for (int s = t; s < NS; s++) {
                        const int i = lane + kWave * s;
                        const V l = (i > k && i < N) ? static_cast<V>(lv[b2][s]) : V(0);
                        b[s] -= l * bk;
                    }
                }
            }
#pragma unroll
            for (int b2 = 0; b2 < kB; b2++)
#pragma unroll
                for (int s = 0; s < NS; s++) lv[b2][s] = nx[b2][s];
        }
    }
    // Back substitution: 1/U_ii for every row first (in parallel); lane k holds
    // the final numerator of x_k when step k is reached and is not updated
    // afterwards, so x = b / U_ii for all rows at the end.
    V rd[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int i = lane + kWave * s;
        const int ic = min(i, N - 1);
        const T d = load_or_zero(A, ic * LD + ic, true);
        rd[s] = (i < N) ? V(1) / static_cast<V>(d) : V(1);
    }
    // columns k0, k0 - 1, ..., k0 - kB + 1 (clamped at 0)
    auto column_block_down = [&](int k0, T (&v)[kB][NS]) {
#pragma unroll
        for (int b2 = 0; b2 < kB; b2++)
#pragma unroll
            for (int s = 0; s < NS; s++)
                v[b2][s] = A[min(lane + kWave * s, N - 1) * LD + min(max(k0 - b2, 0), N - 1)];
    };
#pragma unroll
    for (int t = NS - 1; t >= 0; t--) {
        const int klo = kWave * t;
        T uv[kB][NS];
        column_block_down(min(N, kWave * (t + 1)) - 1, uv);
        for (int k0 = min(N, kWave * (t + 1)) - 1; k0 >= klo; k0 -= kB) {
            T nx[kB][NS];
            column_block_down(k0 - kB, nx);
#pragma unroll
            for (int b2 = 0; b2 < kB; b2++) {
                const int k = k0 - b2;
                if (k >= klo) {
                    const V xk = readlane_t(b[t], k - klo) * readlane_t(rd[t], k - klo);
                    // slots s > t hold rows i >= 64 (t + 1) > k: nothing to update
#pragma unroll
                    for (int s = 0; s <= t; s++) {
                        const int i = lane + kWave * s;
                        const V u = (i < k) ? static_cast<V>(uv[b2][s]) : V(0);
                        b[s] -= u * xk;
                    }
                }
            }
#pragma unroll
            for (int b2 = 0; b2 < kB; b2++)
#pragma unroll
                for (int s = 0; s < NS; s++) uv[b2][s] = nx[b2][s];
        }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) b[s] *= rd[s];
}

}  // namespace cg
