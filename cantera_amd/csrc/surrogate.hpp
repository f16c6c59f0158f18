// Learned one-step surrogate for constant-pressure reactors: an MLP that maps a
// reactor state (T, ln P, ln Y_k) to its increment (dT, dY_k) over a fixed step
// dt, trained on pairs produced by the FP64 integrator.  It has no counterpart
// in the reference; the reference's path is always the BDF integrator, which
// stays the fallback whenever the FP64 guard rejects a surrogate step.
//
// Arithmetic: BF16 weights and activations, FP32 accumulation on the matrix
// cores (v_mfma_f32_16x16x16_bf16), FP32 master weights and Adam moments.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "mech.hpp"

namespace cg {

struct SurrogateShape {
    int K;          // species
    int n_in;       // 2 + K features: T, ln P, ln max(Y_k, 1e-30)
    int n_out;      // 1 + K targets: dT, dY_k
    int in_p;       // n_in padded to a multiple of 64
    int out_p;      // n_out padded to a multiple of 64
    int hidden;     // width of the hidden layers
    int depth;      // number of hidden layers (GELU)
    int layers() const { return depth + 1; }
    int fan_in(int l) const { return l == 0 ? in_p : hidden; }
    int fan_out(int l) const { return l == depth ? out_p : hidden; }
};

class SurrogateNet {
public:
    SurrogateNet(int K, double dt, int hidden, int depth, uint64_t seed);
    ~SurrogateNet();
    SurrogateNet(const SurrogateNet&) = delete;
    SurrogateNet& operator=(const SurrogateNet&) = delete;

    const SurrogateShape& shape() const { return s_; }
    double dt() const { return dt_; }
    size_t n_params() const { return n_par_; }

    // Normalisation (feature and target means / standard deviations) from
    // pairs (y0 -> y1) on the device: y rows in the reactor's state order
    // [m, T, Y_0..Y_{K-1}], P per row.  Uses at most `sample` rows.
    void fit_normalization(long n, const double* y0, const double* P, const double* y1, long sample);
    // One pass over n pairs in minibatches of `batch` rows, in an order
    // drawn from the net's own generator; returns the mean loss.
    double train_epoch(long n, const double* y0, const double* P, const double* y1, int batch, double lr,
                       hipStream_t stream);
    // Loss and gradient of one batch (the rows given, no update): for checks.
    double loss_grad(long n, const double* y0, const double* P, const double* y1, float* grad_host,
                     hipStream_t stream);
    // y1 = y0 + net(y0) on device rows; m is carried over.
    void predict(long n, const double* y0, const double* P, double* y1, hipStream_t stream);
    // The same without any allocation: intermediate activations live in a
    // fixed 16 MiB device workspace (one per GPU, shared by every call on it,
    // so calls that overlap on one GPU race); rows go through it in passes of
    // noalloc_rows().  Throws if the network is too wide for 128 rows.
    void predict_noalloc(long n, const double* y0, const double* P, double* y1, hipStream_t stream);
    int noalloc_rows() const;
    static size_t noalloc_workspace_bytes();

    // element table (K x E) and 1/W_k of the mechanism (device pointers): with
    // them, predictions conserve elements exactly (projection)
    // (at most 8 elements)
    void set_elements(const double* comp, const double* rmw, int E);

    void get_params(float* host) const;
    void set_params(const float* host);
    // mu_in, sd_in (in_p each), mu_out, sd_out, scale_out (out_p each)
    size_t norm_len() const { return 2 * static_cast<size_t>(s_.in_p) + 3 * static_cast<size_t>(s_.out_p); }
    void get_normalization(double* host) const;
    void set_normalization(const double* host);

private:
    void ensure_batch(int rows);
    void forward(int rows, hipStream_t stream);
    double backward_and_loss(int rows, bool update, double lr, hipStream_t stream);
    void sync_bf16(hipStream_t stream);

    SurrogateShape s_{};
    double dt_ = 0.0;
    size_t n_par_ = 0;
    std::vector<size_t> oW_, oB_;   // parameter offsets per layer
    float* par_ = nullptr;          // FP32 master weights + biases
    float* grad_ = nullptr;
    float* m1_ = nullptr;
    float* m2_ = nullptr;
    uint16_t* wb_ = nullptr;        // BF16 weights (same offsets as par_)
    double* norm_ = nullptr;        // mu_in, sd_in, mu_out, sd_out, scale_out
    int adam_t_ = 0;
    const double* comp_ = nullptr;
    const double* rmw_ = nullptr;
    int E_ = 0;
    uint64_t rng_ = 0;
    // batch workspace
    int cap_ = 0;
    uint16_t* act_ = nullptr;       // H_0..H_depth (bf16), rows x width each
    float* pre_ = nullptr;          // Z_0..Z_{depth-1} (fp32)
    float* out_ = nullptr;          // rows x out_p
    float* tgt_ = nullptr;          // rows x out_p
    uint16_t* g0_ = nullptr;        // gradient operands (bf16), ping-pong
    uint16_t* g1_ = nullptr;
    int* idx_ = nullptr;
    double* loss_ = nullptr;
};

}  // namespace cg
