// MLP surrogate kernels: BF16 MFMA GEMM with fused epilogues, features and
// targets from FP64 reactor states, loss, Adam.  See surrogate.hpp.
#include "surrogate.hpp"

#include "wave_kinetics.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>

namespace cg {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;   // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                        // round to nearest even
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// GELU, tanh form (Hendrycks & Gimpel 2016), and its derivative.
constexpr float kGeluC = 0.7978845608028654f;   // sqrt(2/pi)
constexpr float kGeluA = 0.044715f;
__device__ __forceinline__ float gelu(float z) {
    const float t = tanhf(kGeluC * (z + kGeluA * z * z * z));
    return 0.5f * z * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad(float z) {
    const float t = tanhf(kGeluC * (z + kGeluA * z * z * z));
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * kGeluC * (1.0f + 3.0f * kGeluA * z * z);
}

enum Epi { kFwdGelu = 0, kFwdLinear = 1, kStoreF32 = 2, kBwdGelu = 3 };

struct EpiArgs {
    const float* bias;   // kFwdGelu / kFwdLinear: [N]
    float* zf;           // kFwdGelu: pre-activation out; kFwdLinear / kStoreF32: fp32 out
    uint16_t* hb;        // kFwdGelu: gelu(z) as bf16; kBwdGelu: acc * gelu'(zin) as bf16
    const float* zin;    // kBwdGelu: pre-activation of the layer below
    int ldc;
};

constexpr int kTM = 64, kTN = 64, kTK = 32, kLdsK = kTK + 8;

// C(m, n) = sum_k A(m, k) B(n, k) over BF16 operands, FP32 accumulation.
// A(m, k) = TA ? A[k * lda + m] : A[m * lda + k];  B(n, k) = TB ? B[k * ldb + n] : B[n * ldb + k].
// 256 threads: a 64 x 64 tile of C per workgroup, 32 x 32 per wavefront as
// 2 x 2 MFMA 16x16x16 tiles; K staged through LDS in steps of 32.  Out-of-range
// rows / columns / k read as zero and are not stored.
// kStoreF32 may split K over gridDim.z (chunks of kchunk, a multiple of kTK):
// each split adds its partial sums to the cleared output with FP32 atomics.
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                                 const uint16_t* __restrict__ B, int ldb, EpiArgs e, int kchunk) {
    __shared__ uint16_t As[kTM][kLdsK];
    __shared__ uint16_t Bs[kTN][kLdsK];
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.x * kTM, n0 = blockIdx.y * kTN;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int kbeg = blockIdx.z * kchunk, kend = min(K, kbeg + kchunk);
    for (int k0 = kbeg; k0 < kend; k0 += kTK) {
        // stage A (64 x 32) and B (64 x 32), 8 elements per thread each
        if (!TA) {
            const int r = t >> 2, kk = (t & 3) * 8;
            const int m = m0 + r;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int k = k0 + kk + q;
                As[r][kk + q] = (m < M && k < K) ? A[static_cast<size_t>(m) * lda + k] : 0;
            }
        } else {
            const int kk = t >> 3, rr = (t & 7) * 8;
            const int k = k0 + kk;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int m = m0 + rr + q;
                As[rr + q][kk] = (m < M && k < K) ? A[static_cast<size_t>(k) * lda + m] : 0;
            }
        }
        if (!TB) {
            const int r = t >> 2, kk = (t & 3) * 8;
            const int n = n0 + r;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int k = k0 + kk + q;
                Bs[r][kk + q] = (n < N && k < K) ? B[static_cast<size_t>(n) * ldb + k] : 0;
            }
        } else {
            const int kk = t >> 3, rr = (t & 7) * 8;
            const int k = k0 + kk;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int n = n0 + rr + q;
                Bs[rr + q][kk] = (n < N && k < K) ? B[static_cast<size_t>(k) * ldb + n] : 0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kTK; ks += 16) {
            bf16x4 a[2], b[2];
            const int kq = ks + 4 * (lane >> 4);
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const uint16_t* pa = &As[wm + 16 * i + (lane & 15)][kq];
                const uint16_t* pb = &Bs[wn + 16 * i + (lane & 15)][kq];
                a[i] = bf16x4{static_cast<short>(pa[0]), static_cast<short>(pa[1]), static_cast<short>(pa[2]),
                              static_cast<short>(pa[3])};
                b[i] = bf16x4{static_cast<short>(pb[0]), static_cast<short>(pb[1]), static_cast<short>(pb[2]),
                              static_cast<short>(pb[3])};
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // D(16x16) lane layout: row 4 * (lane / 16) + r, column lane % 16
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                const int n = n0 + wn + 16 * j + (lane & 15);
                if (m >= M || n >= N) continue;
                const float v = acc[i][j][r];
                const size_t o = static_cast<size_t>(m) * e.ldc + n;
                if (EPI == kFwdGelu) {
                    const float z = v + e.bias[n];
                    if (e.zf) e.zf[o] = z;      // pre-activation kept for training only
                    e.hb[o] = f2bf(gelu(z));
                } else if (EPI == kFwdLinear) {
                    e.zf[o] = v + e.bias[n];
                } else if (EPI == kStoreF32) {
                    if (gridDim.z > 1) atomicAdd(&e.zf[o], v);
                    else e.zf[o] = v;
                } else {
                    e.hb[o] = f2bf(v * gelu_grad(e.zin[o]));
                }
            }
}

// Features of row idx[r] (or r): T, ln P, ln max(Y_k, 1e-30), normalised and
// clipped to [-8, 8]; padding columns zero.
__global__ void features_kernel(int rows, const int* __restrict__ idx, int N, int K, const double* __restrict__ y,
                                const double* __restrict__ P, const double* __restrict__ norm, int in_p,
                                uint16_t* __restrict__ X) {
    const int r = blockIdx.x;
    if (r >= rows) return;
    const long s = idx ? idx[r] : r;
    const double* yr = y + s * N;
    for (int j = threadIdx.x; j < in_p; j += blockDim.x) {
        double f = 0.0;
        if (j == 0) f = yr[1];
        else if (j == 1) f = log(P[s]);
        else if (j < 2 + K) f = log(fmax(yr[j], 1e-30));
        float v = 0.f;
        if (j < 2 + K) {
            double z = (f - norm[j]) / norm[in_p + j];
            v = static_cast<float>(fmin(fmax(z, -8.0), 8.0));
        }
        X[static_cast<size_t>(r) * in_p + j] = f2bf(v);
    }
}

// Targets: the increments d = (T1 - T0, Y1_k - Y0_k), compressed as asinh(d / s)
// (linear below the scale s, logarithmic above, so small and large increments
// get comparable relative accuracy), then standardised.
__global__ void targets_kernel(int rows, const int* __restrict__ idx, int N, int K, const double* __restrict__ y0,
                               const double* __restrict__ y1, const double* __restrict__ norm, int in_p, int out_p,
                               float* __restrict__ T) {
    const int r = blockIdx.x;
    if (r >= rows) return;
    const long s = idx ? idx[r] : r;
    const double* mu = norm + 2 * in_p;
    const double* sd = mu + out_p;
    const double* sc = sd + out_p;
    for (int j = threadIdx.x; j < out_p; j += blockDim.x) {
        float v = 0.f;
        if (j <= K) {
            const double d = y1[s * N + 1 + j] - y0[s * N + 1 + j];
            v = static_cast<float>((asinh(d / sc[j]) - mu[j]) / sd[j]);
        }
        T[static_cast<size_t>(r) * out_p + j] = v;
    }
}

// Mean-squared error over the n_out valid outputs; gradient as BF16 operand.
__global__ void loss_kernel(int rows, int n_out, int out_p, const float* __restrict__ out,
                            const float* __restrict__ tgt, uint16_t* __restrict__ g, double* __restrict__ loss) {
    __shared__ double part[256];
    double acc = 0.0;
    const double scale = 1.0 / (static_cast<double>(rows) * n_out);
    const size_t total = static_cast<size_t>(rows) * out_p;
    for (size_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int j = static_cast<int>(i % out_p);
        float gv = 0.f;
        if (j < n_out) {
            const float d = out[i] - tgt[i];
            acc += static_cast<double>(d) * d;
            gv = static_cast<float>(2.0 * scale) * d;
        }
        g[i] = f2bf(gv);
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, part[0] * scale);
}

// Column sums of a BF16 gradient operand: bias gradient.
// db[c] += sum over a block of 64 rows of g[r][c] (db cleared by the caller):
// 256 threads as 4 row groups x 64 columns, coalesced 128-byte row segments,
// partial sums combined in LDS and added with one atomic per column and block.
constexpr int kColsumRows = 64;
__global__ __launch_bounds__(256) void colsum_kernel(int rows, int cols, const uint16_t* __restrict__ g,
                                                     float* __restrict__ db) {
    __shared__ float part[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int rg = threadIdx.x >> 6;
    const int r0 = blockIdx.y * kColsumRows;
    float s = 0.f;
    if (c < cols)
        for (int r = r0 + rg; r < min(rows, r0 + kColsumRows); r += 4) s += bf2f(g[static_cast<size_t>(r) * cols + c]);
    part[rg][threadIdx.x & 63] = s;
    __syncthreads();
    if (rg == 0 && c < cols) atomicAdd(&db[c], (part[0][threadIdx.x] + part[1][threadIdx.x]) + (part[2][threadIdx.x] + part[3][threadIdx.x]));
}

// Adam (Kingma & Ba 2015) with bias correction, then the BF16 weight copy.
__global__ void adam_kernel(size_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                            float* __restrict__ m2, uint16_t* __restrict__ wb, float lr, float c1, float c2) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (size_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.f - b1) * gi;
        const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        const float wi = w[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
        w[i] = wi;
        wb[i] = f2bf(wi);
    }
}

__global__ void to_bf16_kernel(size_t n, const float* __restrict__ w, uint16_t* __restrict__ wb) {
    for (size_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * blockDim.x)
        wb[i] = f2bf(w[i]);
}

// y1 = y0 + denormalised output; m carried over.  With the element table
// (comp != nullptr) the mass-fraction increments are projected onto the
// element-conserving subspace: with B_ek = a_ek / W_k and weights
// D_k = max(|Y0_k|, |Y0_k + dY_k|) + 1e-12, dY <- dY - D B^T (B D B^T)^-1 B dY,
// the correction of least D^-1-weighted norm, so species near zero hardly move
// and sum_k dY_k = 0 follows from element conservation.
// One wavefront per row, lanes over species: the E x E Gram matrix B D B^T and
// B dY are wavefront sums; every lane then solves the (symmetric positive
// definite) system by elimination; an element absent from the mechanism's
// species has a zero row and gets lambda = 0.
constexpr int kMaxK = 128, kMaxE = 8, kSlots = kMaxK / 64;
__global__ __launch_bounds__(64) void apply_kernel(int rows, int N, int K, const double* __restrict__ y0,
                                                   const float* __restrict__ out, const double* __restrict__ norm,
                                                   int in_p, int out_p, const double* __restrict__ comp,
                                                   const double* __restrict__ rmw, int E, double* __restrict__ y1) {
    const int r = blockIdx.x;
    if (r >= rows) return;
    const int lane = threadIdx.x;
    const double* mu = norm + 2 * in_p;
    const double* sd = mu + out_p;
    const double* sc = sd + out_p;
    const double* y0r = y0 + static_cast<size_t>(r) * N;
    double* y1r = y1 + static_cast<size_t>(r) * N;
    const float* o = out + static_cast<size_t>(r) * out_p;
    // j = 0 (mass) and j = 1 (T)
    if (lane == 0) y1r[0] = y0r[0];
    // the network output is FP32, so the decode is too (sinhf); the increments
    // and the projection below are FP64
    if (lane == 1) y1r[1] = y0r[1] + sc[0] * static_cast<double>(sinhf(static_cast<float>(o[0] * sd[0] + mu[0])));
    double dY[kSlots], D[kSlots];
#pragma unroll
    for (int t = 0; t < kSlots; t++) {
        const int k = lane + 64 * t;
        dY[t] = 0.0;
        D[t] = 0.0;
        if (k < K) {
            dY[t] = sc[1 + k] * static_cast<double>(sinhf(static_cast<float>(o[1 + k] * sd[1 + k] + mu[1 + k])));
            D[t] = fmax(fabs(y0r[2 + k]), fabs(y0r[2 + k] + dY[t])) + 1e-12;
        }
    }
    const bool project = comp && E >= 1 && E <= kMaxE;
    double lam[kMaxE];
#pragma unroll
    for (int a = 0; a < kMaxE; a++) lam[a] = 0.0;
    // B_ek of the lane's species
    double Bk[kSlots][kMaxE];
#pragma unroll
    for (int t = 0; t < kSlots; t++) {
        const int k = lane + 64 * t;
#pragma unroll
        for (int a = 0; a < kMaxE; a++) Bk[t][a] = (project && k < K && a < E) ? comp[k * E + a] * rmw[k] : 0.0;
    }
    if (project) {
        double G[kMaxE][kMaxE + 1];
#pragma unroll
        for (int a = 0; a < kMaxE; a++) {
#pragma unroll
            for (int b = 0; b <= kMaxE; b++) G[a][b] = 0.0;
        }
#pragma unroll
        for (int a = 0; a < kMaxE; a++) {
            if (a >= E) continue;
#pragma unroll
            for (int b = a; b < kMaxE; b++) {
                if (b >= E) continue;
                double part = 0.0;
#pragma unroll
                for (int t = 0; t < kSlots; t++) part += Bk[t][a] * D[t] * Bk[t][b];
                G[a][b] = G[b][a] = wave_sum(part);
            }
            double v = 0.0;
#pragma unroll
            for (int t = 0; t < kSlots; t++) v += Bk[t][a] * dY[t];
            G[a][kMaxE] = wave_sum(v);
        }
        // elimination without pivoting (SPD); zero rows give lambda = 0
        bool live[kMaxE];
        double rp[kMaxE];
#pragma unroll
        for (int c = 0; c < kMaxE; c++) {
            live[c] = c < E && G[c][c] > 1e-300;
            rp[c] = live[c] ? 1.0 / G[c][c] : 0.0;
            if (!live[c]) continue;
#pragma unroll
            for (int a = c + 1; a < kMaxE; a++) {
                if (a >= E) continue;
                const double f = G[a][c] * rp[c];
#pragma unroll
                for (int b = c; b <= kMaxE; b++)
                    if (b < E || b == kMaxE) G[a][b] -= f * G[c][b];
            }
        }
#pragma unroll
        for (int c = kMaxE - 1; c >= 0; c--) {
            if (!live[c]) continue;
            double v = G[c][kMaxE];
#pragma unroll
            for (int b = c + 1; b < kMaxE; b++)
                if (b < E) v -= G[c][b] * lam[b];
            lam[c] = v * rp[c];
        }
    }
#pragma unroll
    for (int t = 0; t < kSlots; t++) {
        const int k = lane + 64 * t;
        if (k >= K) continue;
        double c = 0.0;
#pragma unroll
        for (int a = 0; a < kMaxE; a++) c += Bk[t][a] * lam[a];
        y1r[2 + k] = y0r[2 + k] + (dY[t] - D[t] * c);
    }
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Forward-layer GEMM, C = A B^T with A (M x K) and B (N x K) both row-major,
// on the gfx950 v_mfma_f32_16x16x32_bf16 (A/B lane layout: row lane % 16,
// k = 8 (lane / 16) .. +7; D: row 4 (lane / 16) + r, column lane % 16).
// 256 threads, a 128 x BN tile per workgroup, (64 x BN/2) per wavefront; K in
// steps of 32 through LDS with 16-byte global loads and LDS accesses, the next
// step's tile loaded into registers while the current one is multiplied.
// Requires K % 32 == 0, N % BN == 0, lda, ldb % 8 == 0 and 16-byte aligned A,
// B (checked by the launcher); rows >= M read zero and are not stored.
template <int BN, int EPI>
__global__ __launch_bounds__(256) void gemm_nt_bf16(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                                    const uint16_t* __restrict__ B, int ldb, EpiArgs e) {
    constexpr int BM = 128, BK = 32, LK = BK + 8, WN = BN / 2, TJ = WN / 16, QB = BN / 64;
    __shared__ __attribute__((aligned(16))) uint16_t As[BM * LK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[BN * LK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int wm = (w >> 1) * 64, wn = (w & 1) * WN;
    f32x4 acc[4][TJ];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < TJ; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // chunk c = t + 256 q of a tile: row c / 4, k offset 8 (c % 4)
    const int kc = (t & 3) * 8, r0 = t >> 2, r1 = (t + 256) >> 2;
    const bool a0ok = m0 + r0 < M, a1ok = m0 + r1 < M;
    const uint16_t* pa0 = A + static_cast<size_t>(a0ok ? m0 + r0 : 0) * lda + kc;
    const uint16_t* pa1 = A + static_cast<size_t>(a1ok ? m0 + r1 : 0) * lda + kc;
    const uint16_t* pb0 = B + static_cast<size_t>(n0 + r0) * ldb + kc;
    const uint16_t* pb1 = B + static_cast<size_t>(n0 + (QB == 2 ? r1 : r0)) * ldb + kc;
    const uint4 zero{0u, 0u, 0u, 0u};
    uint4 ra0 = a0ok ? *reinterpret_cast<const uint4*>(pa0) : zero;
    uint4 ra1 = a1ok ? *reinterpret_cast<const uint4*>(pa1) : zero;
    uint4 rb0 = *reinterpret_cast<const uint4*>(pb0);
    uint4 rb1 = (QB == 2) ? *reinterpret_cast<const uint4*>(pb1) : zero;
    for (int k0 = 0; k0 < K; k0 += BK) {
        *reinterpret_cast<uint4*>(&As[r0 * LK + kc]) = ra0;
        *reinterpret_cast<uint4*>(&As[r1 * LK + kc]) = ra1;
        *reinterpret_cast<uint4*>(&Bs[r0 * LK + kc]) = rb0;
        if (QB == 2) *reinterpret_cast<uint4*>(&Bs[r1 * LK + kc]) = rb1;
        __syncthreads();
        if (k0 + BK < K) {
            const int k1 = k0 + BK;
            ra0 = a0ok ? *reinterpret_cast<const uint4*>(pa0 + k1) : zero;
            ra1 = a1ok ? *reinterpret_cast<const uint4*>(pa1 + k1) : zero;
            rb0 = *reinterpret_cast<const uint4*>(pb0 + k1);
            if (QB == 2) rb1 = *reinterpret_cast<const uint4*>(pb1 + k1);
        }
        bf16x8 a[4], b[TJ];
        const int ko = 8 * (lane >> 4);
#pragma unroll
        for (int i = 0; i < 4; i++) a[i] = *reinterpret_cast<const bf16x8*>(&As[(wm + 16 * i + (lane & 15)) * LK + ko]);
#pragma unroll
        for (int j = 0; j < TJ; j++) b[j] = *reinterpret_cast<const bf16x8*>(&Bs[(wn + 16 * j + (lane & 15)) * LK + ko]);
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < TJ; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < TJ; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
                const int n = n0 + wn + 16 * j + (lane & 15);
                if (m >= M) continue;
                const float v = acc[i][j][r];
                const size_t o = static_cast<size_t>(m) * e.ldc + n;
                if (EPI == kFwdGelu) {
                    const float z = v + e.bias[n];
                    if (e.zf) e.zf[o] = z;      // pre-activation kept for training only
                    e.hb[o] = f2bf(gelu(z));
                } else {
                    e.zf[o] = v + e.bias[n];
                }
            }
}

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <bool TA, bool TB, int EPI>
void gemm(int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb, const EpiArgs& e,
          hipStream_t s) {
    if constexpr (!TA && !TB && (EPI == kFwdGelu || EPI == kFwdLinear)) {
        const bool aligned = (reinterpret_cast<uintptr_t>(A) % 16 == 0) && (reinterpret_cast<uintptr_t>(B) % 16 == 0) &&
                             lda % 8 == 0 && ldb % 8 == 0 && K % 32 == 0;
        if (aligned && N % 128 == 0) {
            hipLaunchKernelGGL((gemm_nt_bf16<128, EPI>), dim3((M + 127) / 128, N / 128), dim3(256), 0, s, M, N, K, A,
                               lda, B, ldb, e);
            check(hipGetLastError(), "surrogate GEMM launch");
            return;
        }
        if (aligned && N % 64 == 0) {
            hipLaunchKernelGGL((gemm_nt_bf16<64, EPI>), dim3((M + 127) / 128, N / 64), dim3(256), 0, s, M, N, K, A,
                               lda, B, ldb, e);
            check(hipGetLastError(), "surrogate GEMM launch");
            return;
        }
    }
    dim3 grid((M + kTM - 1) / kTM, (N + kTN - 1) / kTN);
    int kchunk = K;
    if (EPI == kStoreF32) {
        // weight gradients: few output tiles and a long K (the batch rows);
        // split K so that the grid has about 256 workgroups
        const int tiles = static_cast<int>(grid.x * grid.y);
        const int splits = std::max(1, std::min(256 / std::max(tiles, 1), K / 256));
        if (splits > 1) {
            kchunk = ((K + splits - 1) / splits + kTK - 1) / kTK * kTK;
            grid.z = (K + kchunk - 1) / kchunk;
            check(hipMemsetAsync(e.zf, 0, static_cast<size_t>(M) * e.ldc * sizeof(float), s), "hipMemsetAsync");
        }
    }
    hipLaunchKernelGGL((gemm_bf16<TA, TB, EPI>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, e, kchunk);
    check(hipGetLastError(), "surrogate GEMM launch");
}

int pad64(int x) { return (x + 63) / 64 * 64; }

}  // namespace

SurrogateNet::SurrogateNet(int K, double dt, int hidden, int depth, uint64_t seed) {
    if (K < 1 || hidden < 16 || hidden % 16 || depth < 1 || depth > 16 || !(dt > 0))
        throw std::runtime_error("surrogate: bad shape (hidden must be a positive multiple of 16, 1 <= depth <= 16)");
    if (K > kMaxK) throw std::runtime_error("surrogate: at most 128 species");
    s_.K = K;
    s_.n_in = 2 + K;
    s_.n_out = 1 + K;
    s_.in_p = pad64(s_.n_in);
    s_.out_p = pad64(s_.n_out);
    s_.hidden = hidden;
    s_.depth = depth;
    dt_ = dt;
    rng_ = seed;
    size_t off = 0;
    for (int l = 0; l < s_.layers(); l++) {
        oW_.push_back(off);
        off += static_cast<size_t>(s_.fan_out(l)) * s_.fan_in(l);
        oB_.push_back(off);
        off += s_.fan_out(l);
    }
    n_par_ = off;
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases; padded
    // rows of the last layer zero (their gradients are zero, so they stay zero)
    std::vector<float> h(n_par_, 0.f);
    std::mt19937_64 g(seed);
    for (int l = 0; l < s_.layers(); l++) {
        const int fi = s_.fan_in(l), fo = s_.fan_out(l);
        const int real_in = (l == 0) ? s_.n_in : fi;
        const int real_out = (l == s_.depth) ? s_.n_out : fo;
        std::uniform_real_distribution<float> u(-1.f / std::sqrt(static_cast<float>(real_in)),
                                                1.f / std::sqrt(static_cast<float>(real_in)));
        for (int o = 0; o < real_out; o++)
            for (int i = 0; i < real_in; i++) h[oW_[l] + static_cast<size_t>(o) * fi + i] = u(g);
    }
    check(hipMalloc(&par_, n_par_ * 4), "hipMalloc(surrogate)");
    check(hipMalloc(&grad_, n_par_ * 4), "hipMalloc(surrogate)");
    check(hipMalloc(&m1_, n_par_ * 4), "hipMalloc(surrogate)");
    check(hipMalloc(&m2_, n_par_ * 4), "hipMalloc(surrogate)");
    check(hipMalloc(&wb_, n_par_ * 2), "hipMalloc(surrogate)");
    check(hipMalloc(&norm_, norm_len() * sizeof(double)), "hipMalloc(surrogate)");
    check(hipMalloc(&loss_, sizeof(double)), "hipMalloc(surrogate)");
    check(hipMemcpy(par_, h.data(), n_par_ * 4, hipMemcpyHostToDevice), "hipMemcpy");
    check(hipMemset(m1_, 0, n_par_ * 4), "hipMemset");
    check(hipMemset(m2_, 0, n_par_ * 4), "hipMemset");
    std::vector<double> nm(norm_len(), 0.0);
    for (int j = 0; j < s_.in_p; j++) nm[s_.in_p + j] = 1.0;
    for (int j = 0; j < s_.out_p; j++) nm[2 * s_.in_p + s_.out_p + j] = 1.0;
    for (int j = 0; j < s_.out_p; j++) nm[2 * s_.in_p + 2 * s_.out_p + j] = 1.0;
    set_normalization(nm.data());
    sync_bf16(nullptr);
    check(hipDeviceSynchronize(), "surrogate init");
}

SurrogateNet::~SurrogateNet() {
    for (void* p : {static_cast<void*>(par_), static_cast<void*>(grad_), static_cast<void*>(m1_),
                    static_cast<void*>(m2_), static_cast<void*>(wb_), static_cast<void*>(norm_),
                    static_cast<void*>(loss_), static_cast<void*>(act_), static_cast<void*>(pre_),
                    static_cast<void*>(out_), static_cast<void*>(tgt_), static_cast<void*>(g0_),
                    static_cast<void*>(g1_), static_cast<void*>(idx_)})
        if (p) (void)hipFree(p);
}

void SurrogateNet::sync_bf16(hipStream_t s) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3(256), dim3(256), 0, s, n_par_, par_, wb_);
    check(hipGetLastError(), "surrogate bf16 copy");
}

void SurrogateNet::ensure_batch(int rows) {
    if (rows <= cap_) return;
    for (void* p : {static_cast<void*>(act_), static_cast<void*>(pre_), static_cast<void*>(out_),
                    static_cast<void*>(tgt_), static_cast<void*>(g0_), static_cast<void*>(g1_),
                    static_cast<void*>(idx_)})
        if (p) (void)hipFree(p);
    cap_ = rows;
    const size_t r = static_cast<size_t>(rows);
    const size_t wmax = std::max({s_.in_p, s_.hidden, s_.out_p});
    check(hipMalloc(&act_, r * (s_.in_p + static_cast<size_t>(s_.depth) * s_.hidden) * 2), "hipMalloc(surrogate batch)");
    check(hipMalloc(&pre_, r * static_cast<size_t>(s_.depth) * s_.hidden * 4), "hipMalloc(surrogate batch)");
    check(hipMalloc(&out_, r * s_.out_p * 4), "hipMalloc(surrogate batch)");
    check(hipMalloc(&tgt_, r * s_.out_p * 4), "hipMalloc(surrogate batch)");
    check(hipMalloc(&g0_, r * wmax * 2), "hipMalloc(surrogate batch)");
    check(hipMalloc(&g1_, r * wmax * 2), "hipMalloc(surrogate batch)");
    check(hipMalloc(&idx_, r * 4), "hipMalloc(surrogate batch)");
}

// H_0 = features (already in act_), H_{l+1} = gelu(H_l W_l^T + b_l), out = H_depth W^T + b
void SurrogateNet::forward(int rows, hipStream_t st) {
    const size_t r = static_cast<size_t>(rows);
    for (int l = 0; l <= s_.depth; l++) {
        const int fi = s_.fan_in(l), fo = s_.fan_out(l);
        const uint16_t* H = (l == 0) ? act_ : act_ + r * (s_.in_p + static_cast<size_t>(l - 1) * s_.hidden);
        EpiArgs e{};
        e.bias = par_ + oB_[l];
        e.ldc = fo;
        if (l < s_.depth) {
            e.zf = pre_ + r * static_cast<size_t>(l) * s_.hidden;
            e.hb = act_ + r * (s_.in_p + static_cast<size_t>(l) * s_.hidden);
            gemm<false, false, kFwdGelu>(rows, fo, fi, H, fi, wb_ + oW_[l], fi, e, st);
        } else {
            e.zf = out_;
            gemm<false, false, kFwdLinear>(rows, fo, fi, H, fi, wb_ + oW_[l], fi, e, st);
        }
    }
}

// loss on out_/tgt_, gradient into grad_, optional Adam step
double SurrogateNet::backward_and_loss(int rows, bool update, double lr, hipStream_t st) {
    const size_t r = static_cast<size_t>(rows);
    check(hipMemsetAsync(loss_, 0, sizeof(double), st), "hipMemsetAsync");
    hipLaunchKernelGGL(loss_kernel, dim3(256), dim3(256), 0, st, rows, s_.n_out, s_.out_p, out_, tgt_, g0_, loss_);
    check(hipGetLastError(), "surrogate loss");
    uint16_t* G = g0_;
    uint16_t* Gn = g1_;
    for (int l = s_.depth; l >= 0; l--) {
        const int fi = s_.fan_in(l), fo = s_.fan_out(l);
        const uint16_t* H = (l == 0) ? act_ : act_ + r * (s_.in_p + static_cast<size_t>(l - 1) * s_.hidden);
        // dW_l = G^T H  (fo x fi)
        EpiArgs ew{};
        ew.zf = grad_ + oW_[l];
        ew.ldc = fi;
        gemm<true, true, kStoreF32>(fo, fi, rows, G, fo, H, fi, ew, st);
        check(hipMemsetAsync(grad_ + oB_[l], 0, static_cast<size_t>(fo) * sizeof(float), st), "hipMemsetAsync");
        hipLaunchKernelGGL(colsum_kernel, dim3((fo + 63) / 64, (rows + kColsumRows - 1) / kColsumRows), dim3(256), 0, st,
                           rows, fo, G, grad_ + oB_[l]);
        check(hipGetLastError(), "surrogate bias gradient");
        if (l > 0) {
            // G_{l-1} = (G W_l) * gelu'(Z_{l-1})  (rows x fi)
            EpiArgs eb{};
            eb.hb = Gn;
            eb.zin = pre_ + r * static_cast<size_t>(l - 1) * s_.hidden;
            eb.ldc = fi;
            gemm<false, true, kBwdGelu>(rows, fi, fo, G, fo, wb_ + oW_[l], fi, eb, st);
            std::swap(G, Gn);
        }
    }
    if (update) {
        adam_t_ += 1;
        const float c1 = 1.f - std::pow(0.9f, static_cast<float>(adam_t_));
        const float c2 = 1.f - std::pow(0.999f, static_cast<float>(adam_t_));
        hipLaunchKernelGGL(adam_kernel, dim3(512), dim3(256), 0, st, n_par_, par_, grad_, m1_, m2_, wb_,
                           static_cast<float>(lr), c1, c2);
        check(hipGetLastError(), "surrogate Adam");
    }
    double loss = 0.0;
    check(hipMemcpyAsync(&loss, loss_, sizeof(double), hipMemcpyDeviceToHost, st), "hipMemcpyAsync(loss)");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    return loss;
}

void SurrogateNet::fit_normalization(long n, const double* y0, const double* P, const double* y1, long sample) {
    const long S = std::min(n, sample);
    const int N = s_.K + 2;
    std::vector<double> a(static_cast<size_t>(S) * N), b(static_cast<size_t>(S) * N), p(S);
    check(hipMemcpy(a.data(), y0, a.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    check(hipMemcpy(b.data(), y1, b.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    check(hipMemcpy(p.data(), P, p.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    const int ip = s_.in_p, op = s_.out_p;
    std::vector<double> nm(norm_len(), 0.0), s1(ip, 0.0), s2(ip, 0.0), t1(op, 0.0), t2(op, 0.0), sq(op, 0.0);
    // scale of each increment: 1e-3 of its RMS over the sample
    for (long r = 0; r < S; r++)
        for (int j = 0; j < s_.n_out; j++) {
            const double d = b[static_cast<size_t>(r) * N + 1 + j] - a[static_cast<size_t>(r) * N + 1 + j];
            sq[j] += d * d;
        }
    std::vector<double> sc(op, 1.0);
    for (int j = 0; j < s_.n_out; j++) sc[j] = std::max(1e-3 * std::sqrt(sq[j] / S), 1e-300);
    for (long r = 0; r < S; r++) {
        const double* y = &a[static_cast<size_t>(r) * N];
        const double* z = &b[static_cast<size_t>(r) * N];
        for (int j = 0; j < s_.n_in; j++) {
            const double f = (j == 0) ? y[1] : (j == 1) ? std::log(p[r]) : std::log(std::max(y[j], 1e-30));
            s1[j] += f;
            s2[j] += f * f;
        }
        for (int j = 0; j < s_.n_out; j++) {
            const double d = std::asinh((z[1 + j] - y[1 + j]) / sc[j]);
            t1[j] += d;
            t2[j] += d * d;
        }
    }
    for (int j = 0; j < ip; j++) {
        const double mu = s1[j] / S, var = s2[j] / S - mu * mu;
        nm[j] = (j < s_.n_in) ? mu : 0.0;
        nm[ip + j] = (j < s_.n_in && var > 1e-24) ? std::sqrt(var) : 1.0;
    }
    for (int j = 0; j < op; j++) {
        const double mu = t1[j] / S, var = t2[j] / S - mu * mu;
        nm[2 * ip + j] = (j < s_.n_out) ? mu : 0.0;
        nm[2 * ip + op + j] = (j < s_.n_out && var > 1e-24) ? std::sqrt(var) : 1.0;
        nm[2 * ip + 2 * op + j] = sc[j];
    }
    set_normalization(nm.data());
}

double SurrogateNet::train_epoch(long n, const double* y0, const double* P, const double* y1, int batch, double lr,
                                 hipStream_t st) {
    if (batch < 1) throw std::runtime_error("surrogate: batch must be positive");
    ensure_batch(batch);
    std::vector<int> perm(n);
    for (long i = 0; i < n; i++) perm[i] = static_cast<int>(i);
    std::mt19937_64 g(rng_++);
    std::shuffle(perm.begin(), perm.end(), g);
    const int N = s_.K + 2;
    double total = 0.0;
    long nb = 0;
    for (long b0 = 0; b0 < n; b0 += batch) {
        const int rows = static_cast<int>(std::min<long>(batch, n - b0));
        check(hipMemcpyAsync(idx_, perm.data() + b0, rows * 4, hipMemcpyHostToDevice, st), "hipMemcpyAsync(idx)");
        hipLaunchKernelGGL(features_kernel, dim3(rows), dim3(64), 0, st, rows, idx_, N, s_.K, y0, P, norm_, s_.in_p,
                           act_);
        hipLaunchKernelGGL(targets_kernel, dim3(rows), dim3(64), 0, st, rows, idx_, N, s_.K, y0, y1, norm_, s_.in_p,
                           s_.out_p, tgt_);
        check(hipGetLastError(), "surrogate features");
        forward(rows, st);
        total += backward_and_loss(rows, true, lr, st);
        nb++;
    }
    return nb ? total / nb : 0.0;
}

double SurrogateNet::loss_grad(long n, const double* y0, const double* P, const double* y1, float* grad_host,
                               hipStream_t st) {
    const int rows = static_cast<int>(n);
    ensure_batch(rows);
    const int N = s_.K + 2;
    hipLaunchKernelGGL(features_kernel, dim3(rows), dim3(64), 0, st, rows, nullptr, N, s_.K, y0, P, norm_, s_.in_p,
                       act_);
    hipLaunchKernelGGL(targets_kernel, dim3(rows), dim3(64), 0, st, rows, nullptr, N, s_.K, y0, y1, norm_, s_.in_p,
                       s_.out_p, tgt_);
    check(hipGetLastError(), "surrogate features");
    forward(rows, st);
    const double loss = backward_and_loss(rows, false, 0.0, st);
    if (grad_host) check(hipMemcpy(grad_host, grad_, n_par_ * 4, hipMemcpyDeviceToHost), "hipMemcpy(grad)");
    return loss;
}

void SurrogateNet::predict(long n, const double* y0, const double* P, double* y1, hipStream_t st) {
    const int N = s_.K + 2;
    const long chunk = 65536;
    for (long c0 = 0; c0 < n; c0 += chunk) {
        const int rows = static_cast<int>(std::min(chunk, n - c0));
        ensure_batch(rows);
        hipLaunchKernelGGL(features_kernel, dim3(rows), dim3(64), 0, st, rows, nullptr, N, s_.K, y0 + c0 * N, P + c0,
                           norm_, s_.in_p, act_);
        check(hipGetLastError(), "surrogate features");
        forward(rows, st);
        hipLaunchKernelGGL(apply_kernel, dim3(rows), dim3(64), 0, st, rows, N, s_.K, y0 + c0 * N, out_, norm_,
                           s_.in_p, s_.out_p, comp_, rmw_, E_, y1 + c0 * N);
        check(hipGetLastError(), "surrogate apply");
    }
}

// ---- inference without allocation ------------------------------------------
//
// Workspace of predict_noalloc: one array per device, part of the code object,
// held until the process ends.  Its address on the current device comes from
// hipGetSymbolAddress (the array's name is only a host-side stand-in), once
// per device, and is kept per device.
constexpr size_t kNoAllocBytes = size_t{16} << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

void* scratch_on_current_device() {
    int dev = -1;
    check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("surrogate: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw std::runtime_error("surrogate: no address for the inference workspace");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

size_t SurrogateNet::noalloc_workspace_bytes() { return kNoAllocBytes; }

// Rows per pass through the workspace: features (in_p BF16), two hidden
// activations (BF16, ping-pong), the FP32 output (out_p); a multiple of 128
// (the GEMM row tile), so every buffer starts 256-byte aligned.
int SurrogateNet::noalloc_rows() const {
    const size_t per_row = static_cast<size_t>(s_.in_p) * 2 + 4 * static_cast<size_t>(s_.hidden) +
                           static_cast<size_t>(s_.out_p) * 4;
    const size_t rows = kNoAllocBytes / per_row / 128 * 128;
    return static_cast<int>(std::min<size_t>(rows, 1 << 20));
}

void SurrogateNet::predict_noalloc(long n, const double* y0, const double* P, double* y1, hipStream_t st) {
    const int R = noalloc_rows();
    if (R < 128)
        throw std::runtime_error("surrogate: network too wide for the 16 MiB inference workspace (fewer than 128 rows)");
    unsigned char* base = static_cast<unsigned char*>(scratch_on_current_device());
    const size_t r = static_cast<size_t>(R);
    uint16_t* feat = reinterpret_cast<uint16_t*>(base);
    uint16_t* ha = reinterpret_cast<uint16_t*>(base + r * s_.in_p * 2);
    uint16_t* hb = ha + r * s_.hidden;
    float* out = reinterpret_cast<float*>(base + r * s_.in_p * 2 + 4 * r * s_.hidden);
    const int N = s_.K + 2;
    for (long c0 = 0; c0 < n; c0 += R) {
        const int rows = static_cast<int>(std::min<long>(R, n - c0));
        hipLaunchKernelGGL(features_kernel, dim3(rows), dim3(64), 0, st, rows, nullptr, N, s_.K, y0 + c0 * N, P + c0,
                           norm_, s_.in_p, feat);
        check(hipGetLastError(), "surrogate features");
        const uint16_t* H = feat;
        uint16_t* next = ha;
        for (int l = 0; l <= s_.depth; l++) {
            const int fi = s_.fan_in(l), fo = s_.fan_out(l);
            EpiArgs e{};
            e.bias = par_ + oB_[l];
            e.ldc = fo;
            if (l < s_.depth) {
                e.zf = nullptr;
                e.hb = next;
                gemm<false, false, kFwdGelu>(rows, fo, fi, H, fi, wb_ + oW_[l], fi, e, st);
                H = next;
                next = (next == ha) ? hb : ha;
            } else {
                e.zf = out;
                gemm<false, false, kFwdLinear>(rows, fo, fi, H, fi, wb_ + oW_[l], fi, e, st);
            }
        }
        hipLaunchKernelGGL(apply_kernel, dim3(rows), dim3(64), 0, st, rows, N, s_.K, y0 + c0 * N, out, norm_,
                           s_.in_p, s_.out_p, comp_, rmw_, E_, y1 + c0 * N);
        check(hipGetLastError(), "surrogate apply");
    }
}

void SurrogateNet::set_elements(const double* comp, const double* rmw, int E) {
    if (E < 1 || E > kMaxE) throw std::runtime_error("surrogate: element projection supports 1 to 8 elements");
    comp_ = comp;
    rmw_ = rmw;
    E_ = E;
}

void SurrogateNet::get_params(float* host) const {
    check(hipMemcpy(host, par_, n_par_ * 4, hipMemcpyDeviceToHost), "hipMemcpy(params)");
}
void SurrogateNet::set_params(const float* host) {
    check(hipMemcpy(par_, host, n_par_ * 4, hipMemcpyHostToDevice), "hipMemcpy(params)");
    sync_bf16(nullptr);
    check(hipDeviceSynchronize(), "surrogate set_params");
}
void SurrogateNet::get_normalization(double* host) const {
    check(hipMemcpy(host, norm_, norm_len() * 8, hipMemcpyDeviceToHost), "hipMemcpy(norm)");
}
void SurrogateNet::set_normalization(const double* host) {
    check(hipMemcpy(norm_, host, norm_len() * 8, hipMemcpyHostToDevice), "hipMemcpy(norm)");
}

}  // namespace cg
