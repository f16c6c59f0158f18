// Host <-> device contract of the batched BDF integrator.
#pragma once
#include <hip/hip_runtime.h>

#include "mech.hpp"

namespace cg {

// Per-reactor control block layout (doubles), after 6 Nordsieck columns and the
// interpolated output vector, each 64*NS doubles.
enum CtlIndex {
    C_TN = 0, C_H, C_HPRIME, C_HSCALE, C_ETA, C_ETAMAX, C_HU,
    C_TAU = 7,            // 7 entries
    C_TQ = 14,            // 6 entries
    C_RL1 = 20, C_GAMMA, C_GAMMAP, C_GAMRAT, C_CRATE, C_DELP, C_ACNRM, C_SAVED_TQ5,
    C_Q = 28, C_QPRIME, C_QWAIT, C_L, C_NST, C_NSTLP, C_NSTLJ, C_NSCON, C_STARTED, C_TUSER,
    C_NFE = 38, C_NJE, C_NSETUPS, C_NNI, C_NCFN, C_NETF, C_STATUS, C_P, C_HAVE_LU, C_NFE_DQ,
    // mole-reactor kernels: linear-solver counters and the species' atol
    C_NLI = 48, C_NPS, C_NJT, C_NCFL, C_NFE_LS, C_ATOL_SP,
    C_COUNT = 64
};

// Tables of the mole-reactor preconditioner's sparse LU (device arrays, one
// set per batch; layout and meaning as SparseLUHost in sparse_lu.hpp).
struct SparseTables {
    int nnz;                                    // value slots of L + U
    int skip;                                   // bit 0: no third-body terms, bit 1: no falloff d k/d[M]
    const int *diag, *lptr, *lslot, *lrow, *optr, *ops, *uptr, *uslot, *urow, *iperm, *srow, *scol;
};

// most reaction-multiplier sensitivity parameters per reactor batch
constexpr int kMaxSens = 32;
// columns of a parameter's block: Nordsieck history 0..5, corrector 6, output 7
constexpr int kSensCols = 8;

struct BdfArgs {
    DevMech mech;
    int n;                // reactors
    int neq_norm;         // N used in the WRMS norm (K+2, includes the mass)
    int ld;               // LDS matrix leading dimension
    int max_steps;        // per advance call
    double rtol, atol;
    double* rec;          // n x rec_stride doubles
    size_t rec_stride;
    double* j_save;       // n x j_stride doubles (N x N row-major)
    size_t j_stride;
    void* lu_save_raw;    // n x lu_stride elements of the matrix type
    size_t lu_stride;
    int* piv_save;        // n x piv_stride
    size_t piv_stride;
    const double* tout;   // n, device
    const int* order;     // optional processing order (n), device
    const int* work_flag; // optional: only reactors r with work_flag[r] == work_value are advanced
    int work_value;
    int* queue;           // device work-queue head
    // forward sensitivities to reaction rate multipliers (sensitivity kernels only)
    int nsens;                      // parameters (<= kMaxSens)
    int sens_rxn[kMaxSens];         // permuted reaction index of each parameter
    double rtol_sens, atol_sens;
    double* sens;                   // n x sens_stride doubles: per parameter 8 columns of 64*NS
    size_t sens_stride;
    // mole reactor (IdealGasConstPressureMoleReactor, GMRES + sparse preconditioner)
    SparseTables sp;
    // typed view used by the kernel
    template <class T> __host__ __device__ T* lu() const { return static_cast<T*>(lu_save_raw); }
};

// kernel variant for a state size N: 16/32 = matrix in registers with that many
// columns, 56/64 = matrix in LDS factored/solved in registers, 2 = matrix in
// LDS with 2 state slots per lane (N <= 128), -1 = unsupported
int bdf_variant(int N);
// per-reactor storage (elements) of the saved factored matrix
size_t bdf_lu_elems(int N, int ld);
// global: for 32 < N <= 64, keep the iteration matrix in the reactor's
// global-memory factor slot instead of LDS
size_t bdf_lds_bytes(const DevMech& m, int ld, bool fp32, bool global);
// a.nsens > 0 selects the FP64 kernels that also integrate the sensitivities
hipError_t bdf_launch(const BdfArgs& a, bool fp32, bool global, int grid, size_t lds, hipStream_t stream);
const void* bdf_kernel_ptr(int N, bool fp32, bool global, bool sens = false);

// mole-reactor integrator: state [n_0..n_{K-1}, T], N = K + 1 <= 256
int bdf_mole_ns(int N);
size_t bdf_mole_lds_bytes(const DevMech& m, int nnz);
hipError_t bdf_mole_launch(const BdfArgs& a, int grid, size_t lds, hipStream_t stream);
const void* bdf_mole_kernel_ptr(int N);

// batched rate evaluation (species-major SoA in and out)
size_t rates_lds_bytes(const DevMech& m, int ns);
// f32: rate constants through FP32 transcendentals (CG_FP32)
hipError_t rates_launch(const DevMech& m, int ns, bool f32, long n, const double* T, const double* P,
                        const double* Y, double* wdot, int grid, size_t lds, hipStream_t stream);
const void* rates_kernel_ptr(int ns, bool f32);
// rates at the output states of a reactor batch (records as the integrator's)
hipError_t reactor_rates_launch(const DevMech& m, int ns, long n, const double* rec, size_t stride, int cv,
                                double* wdot, int grid, size_t lds, hipStream_t stream);
const void* reactor_rates_kernel_ptr(int ns);

// longest-first processing order for the integrator's work queue (kernels
// only, so it can be captured in a HIP graph); prev_nst, bucket, rank and
// order_out hold n ints, temp order_temp_bytes(n) bytes
size_t order_temp_bytes(int n);
hipError_t order_update(const double* rec, size_t rec_stride, int ctl_off, int n, int* prev_nst, int* bucket,
                        int* rank, int* order_out, void* temp, hipStream_t stream);

// guarded surrogate steps for reactor batches (all on the device)
size_t guard_lds_bytes(const DevMech& m, int ns);
const void* guard_kernel_ptr(int ns);
hipError_t surrogate_clear(int* p, int n, hipStream_t s);
// output states of all reactors, reference order [m, T, Y], n x N doubles
hipError_t export_states(int n, const double* rec, size_t stride, int NP, int N, int K, double* Y, hipStream_t s);
// elig[r]: 1 one surrogate step ahead, 2 integrate, 0 nothing to do; the
// guard turns the rejected 1s into 2s, and the integrator then advances the
// reactors with elig = 2 (BdfArgs::work_flag)
hipError_t surrogate_gather(int n, const double* rec, size_t stride, int NP, int N, int K, const double* tout,
                            double dt, double* Y, double* P, int* elig, unsigned long long* counters, hipStream_t s);
hipError_t surrogate_guard(const DevMech& m, int ns, int n, int* elig, const double* Ycur, const double* Ynext,
                           double* Fcur, int* fvalid, const double* P, double* rec, size_t stride, int NP,
                           const double* tout, double dt, double tol_T, double tol_Y, double gate_T, const int* rad,
                           unsigned long long* counters, int grid, size_t lds, hipStream_t s);

// batched reactor Jacobians (inputs as the rates, output N x N per state)
size_t jac_lds_bytes(const DevMech& m, int ns);
// f32: FP32 output (the FP64 assembly rounded at the store; needs ns == 1)
hipError_t jac_launch(const DevMech& m, int ns, bool f32, long n, const double* T, const double* P, const double* Y,
                      void* J, int grid, size_t lds, hipStream_t stream);
const void* jac_kernel_ptr(int ns, bool f32);

// batched chemical equilibrium (hp = 1: fixed h and P; 0: fixed T and P),
// inputs and outputs species-major; info[s] = iterations, or negative
hipError_t equil_launch(const DevMech& m, int ns, int hp, long n, const double* T, const double* P, const double* Y,
                        double* Tout, double* Yout, int* info, int max_iter, int grid, hipStream_t stream);
const void* equil_kernel_ptr(int ns);

}  // namespace cg
