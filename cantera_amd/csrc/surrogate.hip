// MLP surrogate kernels: BF16 MFMA GEMM with fused epilogues, features and
// targets from FP64 reactor states, loss, Adam.  See surrogate.hpp.
#include "surrogate.hpp"

#include <algorithm>
#include <cmath>
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
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                                 const uint16_t* __restrict__ B, int ldb, EpiArgs e) {
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
    for (int k0 = 0; k0 < K; k0 += kTK) {
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
                    e.zf[o] = z;
                    e.hb[o] = f2bf(gelu(z));
                } else if (EPI == kFwdLinear) {
                    e.zf[o] = v + e.bias[n];
                } else if (EPI == kStoreF32) {
                    e.zf[o] = v;
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
__global__ void colsum_kernel(int rows, int cols, const uint16_t* __restrict__ g, float* __restrict__ db) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; r++) s += bf2f(g[static_cast<size_t>(r) * cols + c]);
    db[c] = s;
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
constexpr int kMaxK = 128, kMaxE = 8;
__global__ void apply_kernel(int rows, int N, int K, const double* __restrict__ y0, const float* __restrict__ out,
                             const double* __restrict__ norm, int in_p, int out_p, const double* __restrict__ comp,
                             const double* __restrict__ rmw, int E, double* __restrict__ y1) {
    __shared__ double dY[kMaxK], D[kMaxK];
    __shared__ double lam[kMaxE];
    const int r = blockIdx.x;
    if (r >= rows) return;
    const double* mu = norm + 2 * in_p;
    const double* sd = mu + out_p;
    const double* sc = sd + out_p;
    const double* y0r = y0 + static_cast<size_t>(r) * N;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        double d = 0.0;
        if (j >= 1)
            d = sc[j - 1] * sinh(static_cast<double>(out[static_cast<size_t>(r) * out_p + (j - 1)]) * sd[j - 1] + mu[j - 1]);
        if (j >= 2) {
            dY[j - 2] = d;
            D[j - 2] = fmax(fabs(y0r[j]), fabs(y0r[j] + d)) + 1e-12;
        }
        y1[static_cast<size_t>(r) * N + j] = y0r[j] + d;
    }
    __syncthreads();
    if (!comp || E < 1 || E > kMaxE) return;
    if (threadIdx.x == 0) {
        double G[kMaxE][kMaxE + 1];
        for (int a = 0; a < E; a++) {
            for (int b = 0; b < E; b++) {
                double g = 0.0;
                for (int k = 0; k < K; k++) g += comp[k * E + a] * rmw[k] * D[k] * comp[k * E + b] * rmw[k];
                G[a][b] = g;
            }
            double v = 0.0;
            for (int k = 0; k < K; k++) v += comp[k * E + a] * rmw[k] * dY[k];
            G[a][E] = v;
        }
        // Gaussian elimination with partial pivoting; elements absent from the
        // mechanism's species give zero rows and get lambda = 0
        int used[kMaxE];
        for (int a = 0; a < E; a++) used[a] = 0;
        for (int c = 0; c < E; c++) {
            int p = -1;
            double best = 0.0;
            for (int a = 0; a < E; a++)
                if (!used[a] && fabs(G[a][c]) > best) { best = fabs(G[a][c]); p = a; }
            lam[c] = 0.0;
            if (p < 0 || best < 1e-300) continue;
            used[p] = c + 1;
            for (int a = 0; a < E; a++) {
                if (a == p) continue;
                const double f = G[a][c] / G[p][c];
                for (int b = c; b <= E; b++) G[a][b] -= f * G[p][b];
            }
        }
        for (int a = 0; a < E; a++)
            if (used[a]) lam[used[a] - 1] = G[a][E] / G[a][used[a] - 1];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        double c = 0.0;
        for (int a = 0; a < E; a++) c += comp[k * E + a] * rmw[k] * lam[a];
        y1[static_cast<size_t>(r) * N + 2 + k] = y0r[2 + k] + (dY[k] - D[k] * c);
    }
}

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <bool TA, bool TB, int EPI>
void gemm(int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb, const EpiArgs& e,
          hipStream_t s) {
    dim3 grid((M + kTM - 1) / kTM, (N + kTN - 1) / kTN);
    hipLaunchKernelGGL((gemm_bf16<TA, TB, EPI>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, e);
    check(hipGetLastError(), "surrogate GEMM launch");
}

int pad64(int x) { return (x + 63) / 64 * 64; }

}  // namespace

SurrogateNet::SurrogateNet(int K, double dt, int hidden, int depth, uint64_t seed) {
    if (K < 1 || hidden < 16 || hidden % 16 || depth < 1 || depth > 16 || !(dt > 0))
        throw std::runtime_error("surrogate: bad shape (hidden must be a positive multiple of 16, 1 <= depth <= 16)");
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
        hipLaunchKernelGGL(colsum_kernel, dim3((fo + 255) / 256), dim3(256), 0, st, rows, fo, G, grad_ + oB_[l]);
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
