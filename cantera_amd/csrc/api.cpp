// extern "C" boundary of libcantera_amd.so (declarations and reference
// citations in include/cantera_gpu.h).  Handles index owned objects; errors are
// reported as negative return codes plus a per-thread message, and no C++
// exception crosses the boundary.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/cantera_gpu.h"
#include "bdf.hpp"
#include "mech.hpp"
#include "sparse_lu.hpp"
#include "surrogate.hpp"

namespace {

using namespace cg;

thread_local std::string g_error;

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

int device_cu_count() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    }
    return cus;
}

struct MechObj {
    std::unique_ptr<DeviceMech> dm;
    hipEvent_t e0 = nullptr, e1 = nullptr;   // last rates kernel
    hipEvent_t j0 = nullptr, j1 = nullptr;   // last Jacobian kernel
    double last_ms = 0.0;
    int ns = 1;
    ~MechObj() {
        for (hipEvent_t e : {e0, e1, j0, j1})
            if (e) (void)hipEventDestroy(e);
    }
};

struct Batch {
    MechObj* mech = nullptr;
    int n = 0, N = 0, K = 0, ns = 1, ld = 0;
    bool fp32 = false;
    bool const_volume = false;   // IdealGasReactor (V = 1 m^3) instead of IdealGasConstPressureReactor
    double rtol = 1e-9, atol = 1e-15;
    double* rec = nullptr;
    double* jsave = nullptr;
    void* lusave = nullptr;
    int* piv = nullptr;
    double* tout = nullptr;
    int* queue = nullptr;
    int* lpt = nullptr;          // 5n ints: previous step counts, bucket, rank in bucket, (spare), order
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    bool lpt_on = true, have_order = false;
    bool global_matrix = false;  // 32 < N <= 64: iteration matrix in the global factor slot (else LDS)
    size_t rec_stride = 0, j_stride = 0, lu_stride = 0, piv_stride = 0;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0.0;
    int max_steps = 20000;
    int waves_per_cu = 0;
    size_t lds = 0;
    // multi-step advance replayed from one captured graph (cg_reactors_advance_graph)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_steps = 0;
    const double* graph_tout = nullptr;
    const int* graph_order = nullptr;
    // guarded surrogate steps: workspace allocated at the first such call
    double* sg_Y = nullptr;          // n x N current states (reference order)
    double* sg_Yn = nullptr;         // n x N predictions
    double* sg_F = nullptr;          // n x N f(y) of the last kept state (internal order)
    double* sg_P = nullptr;          // n
    int* sg_ints = nullptr;          // elig (n), fvalid (n)
    unsigned long long* sg_counters = nullptr;   // eligible, kept, integrated (accumulated)
    // forward sensitivities to reaction rate multipliers (ReactorNet defaults,
    // include/cantera/zeroD/ReactorNet.h: rtol 1e-4, atol 1e-6)
    int nsens = 0;
    int sens_rxn[kMaxSens] = {};     // permuted reaction index per parameter
    int sens_rxn_orig[kMaxSens] = {};
    double rtol_sens = 1e-4, atol_sens = 1e-6;
    double* sens = nullptr;          // n x sens_stride doubles
    size_t sens_stride = 0;
    bool launched = false;           // an integrator launch has been issued
    // mole-based reactor with the sparse preconditioner and GMRES
    bool mole = false;
    int* sp_tables = nullptr;        // device copy of the sparse-LU schedules
    SparseTables sp{};
    long sp_updates = 0;
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
        graph_steps = 0;
    }
    ~Batch() {
        drop_graph();
        for (void* p : {static_cast<void*>(rec), static_cast<void*>(jsave), lusave, static_cast<void*>(piv),
                        static_cast<void*>(tout), static_cast<void*>(queue), static_cast<void*>(lpt), sort_tmp,
                        static_cast<void*>(sg_Y), static_cast<void*>(sg_Yn), static_cast<void*>(sg_F),
                        static_cast<void*>(sg_P), static_cast<void*>(sg_ints), static_cast<void*>(sg_counters),
                        static_cast<void*>(sens), static_cast<void*>(sp_tables)})
            if (p) (void)hipFree(p);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

std::mutex g_mu;
std::vector<std::unique_ptr<MechObj>> g_mechs;
std::vector<std::unique_ptr<Batch>> g_batches;

struct SurrogateObj {
    MechObj* mech = nullptr;
    std::unique_ptr<SurrogateNet> net;
};
std::vector<std::unique_ptr<SurrogateObj>> g_surr;

SurrogateObj& get_surr(int32_t h) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (h < 0 || h >= static_cast<int32_t>(g_surr.size()) || !g_surr[h]) throw std::runtime_error("invalid surrogate handle");
    return *g_surr[h];
}

MechObj& get_mech(int32_t h) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (h < 0 || h >= static_cast<int32_t>(g_mechs.size()) || !g_mechs[h]) throw std::runtime_error("invalid mechanism handle");
    return *g_mechs[h];
}
Batch& get_batch(int32_t h) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (h < 0 || h >= static_cast<int32_t>(g_batches.size()) || !g_batches[h]) throw std::runtime_error("invalid reactor-batch handle");
    return *g_batches[h];
}

template <class F>
int32_t guarded(F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        g_error = e.what();
    } catch (...) {
        g_error = "unknown error";
    }
    return CG_ERR;
}

int grid_for(const void* kernel, size_t lds, int waves_per_cu, long work_items) {
    int per_cu = waves_per_cu;
    if (per_cu <= 0) {
        int nb = 0;
        check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 64, lds), "hipOccupancyMaxActiveBlocksPerMultiprocessor");
        per_cu = std::max(nb, 1);
    }
    long g = static_cast<long>(per_cu) * device_cu_count();
    return static_cast<int>(std::max<long>(1, std::min(g, work_items)));
}

constexpr size_t kLdsPerCu = 160 * 1024;

void set_lds_limit(const void* kernel, size_t lds) {
    if (lds > 65536)
        check(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)),
              "hipFuncSetAttribute(MaxDynamicSharedMemorySize)");
}

}  // namespace

extern "C" {

int32_t cg_version(void) { return 1; }

int32_t cg_last_error(char* buf, int32_t len) {
    if (buf && len > 0) {
        std::strncpy(buf, g_error.c_str(), static_cast<size_t>(len) - 1);
        buf[len - 1] = 0;
    }
    return static_cast<int32_t>(g_error.size());
}

int32_t cg_device_count(void) {
    return guarded([&] {
        int n = 0;
        check(hipGetDeviceCount(&n), "hipGetDeviceCount");
        return n;
    });
}

void* cg_dev_alloc(int64_t bytes) {
    void* p = nullptr;
    if (bytes <= 0) { g_error = "cg_dev_alloc: size must be positive"; return nullptr; }
    hipError_t e = hipMalloc(&p, static_cast<size_t>(bytes));
    if (e != hipSuccess) { g_error = std::string("hipMalloc: ") + hipGetErrorString(e); return nullptr; }
    return p;
}
int32_t cg_dev_free(void* p) {
    return guarded([&] { check(hipFree(p), "hipFree"); return 0; });
}
int32_t cg_memcpy_h2d(void* dst, const void* src, int64_t bytes) {
    return guarded([&] { check(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyHostToDevice), "hipMemcpy H2D"); return 0; });
}
int32_t cg_memcpy_d2h(void* dst, const void* src, int64_t bytes) {
    return guarded([&] { check(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToHost), "hipMemcpy D2H"); return 0; });
}
int32_t cg_memcpy_d2d(void* dst, const void* src, int64_t bytes) {
    return guarded([&] { check(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToDevice), "hipMemcpy D2D"); return 0; });
}
int32_t cg_device_synchronize(void) {
    return guarded([&] { check(hipDeviceSynchronize(), "hipDeviceSynchronize"); return 0; });
}

int32_t cg_mech_load(const char* path) {
    return guarded([&] {
        if (!path) throw std::runtime_error("null path");
        auto obj = std::make_unique<MechObj>();
        HostMech hm = load_mech(path);
        obj->dm = std::make_unique<DeviceMech>(hm);
        obj->ns = (obj->dm->view().N + 63) / 64;
        check(hipEventCreate(&obj->e0), "hipEventCreate");
        check(hipEventCreate(&obj->e1), "hipEventCreate");
        check(hipEventCreate(&obj->j0), "hipEventCreate");
        check(hipEventCreate(&obj->j1), "hipEventCreate");
        std::lock_guard<std::mutex> lk(g_mu);
        g_mechs.push_back(std::move(obj));
        return static_cast<int32_t>(g_mechs.size() - 1);
    });
}

int32_t cg_mech_free(int32_t h) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_mu);
        if (h < 0 || h >= static_cast<int32_t>(g_mechs.size()) || !g_mechs[h]) throw std::runtime_error("invalid mechanism handle");
        for (auto& b : g_batches)
            if (b && b->mech == g_mechs[h].get()) throw std::runtime_error("mechanism still used by a reactor batch");
        g_mechs[h].reset();
        return 0;
    });
}

int32_t cg_mech_nspecies(int32_t h) { return guarded([&] { return get_mech(h).dm->view().K; }); }
int32_t cg_mech_nreactions(int32_t h) { return guarded([&] { return get_mech(h).dm->view().I; }); }
int32_t cg_mech_nelements(int32_t h) { return guarded([&] { return get_mech(h).dm->view().E; }); }

int32_t cg_mech_species_index(int32_t h, const char* name) {
    return guarded([&] {
        const auto& names = get_mech(h).dm->host().names;
        for (size_t k = 0; k < names.size(); k++)
            if (names[k] == name) return static_cast<int32_t>(k);
        return -1;
    });
}

int32_t cg_mech_species_name(int32_t h, int32_t k, char* buf, int32_t len) {
    return guarded([&] {
        const auto& names = get_mech(h).dm->host().names;
        if (k < 0 || k >= static_cast<int32_t>(names.size())) throw std::runtime_error("species index out of range");
        if (buf && len > 0) {
            std::strncpy(buf, names[k].c_str(), static_cast<size_t>(len) - 1);
            buf[len - 1] = 0;
        }
        return static_cast<int32_t>(names[k].size());
    });
}

int32_t cg_mech_molecular_weights(int32_t h, int32_t len, double* mw) {
    return guarded([&] {
        const auto& w = get_mech(h).dm->host().mw;
        if (len < static_cast<int32_t>(w.size())) throw std::runtime_error("buffer too small");
        std::copy(w.begin(), w.end(), mw);
        return 0;
    });
}

int32_t cg_rates_batch_device(int32_t h, int64_t n, const double* T, const double* P, const double* Y,
                              double* wdot, int32_t precision, void* stream) {
    return guarded([&] {
        MechObj& mo = get_mech(h);
        if (precision != CG_FP64 && precision != CG_FP32)
            throw std::runtime_error("cg_rates_batch: precision must be CG_FP64 or CG_FP32");
        const bool f32 = precision == CG_FP32;
        if (n <= 0) return 0;
        if (!T || !P || !Y || !wdot) throw std::runtime_error("null array");
        const DevMech& m = mo.dm->view();
        const size_t lds = rates_lds_bytes(m, mo.ns);
        const void* kp = rates_kernel_ptr(mo.ns, f32);
        set_lds_limit(kp, lds);
        const int grid = grid_for(kp, lds, 0, n);
        hipStream_t s = static_cast<hipStream_t>(stream);
        check(hipEventRecord(mo.e0, s), "hipEventRecord");
        check(rates_launch(m, mo.ns, f32, n, T, P, Y, wdot, grid, lds, s), "rates kernel launch");
        check(hipEventRecord(mo.e1, s), "hipEventRecord");
        return 0;
    });
}

int32_t cg_rates_batch(int32_t h, int64_t n, const double* T, const double* P, const double* Y, double* wdot,
                       int32_t precision) {
    return guarded([&] {
        MechObj& mo = get_mech(h);
        if (n <= 0) return 0;
        if (!T || !P || !Y || !wdot) throw std::runtime_error("null array");
        const int K = mo.dm->view().K;
        const size_t bytes = static_cast<size_t>(n) * sizeof(double);
        double* d = nullptr;
        check(hipMalloc(&d, bytes * (2 + 2 * static_cast<size_t>(K))), "hipMalloc");
        std::unique_ptr<double, void (*)(double*)> guard(d, [](double* p) { (void)hipFree(p); });
        double *dT = d, *dP = d + n, *dY = d + 2 * n, *dW = d + (2 + static_cast<size_t>(K)) * n;
        check(hipMemcpy(dT, T, bytes, hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemcpy(dP, P, bytes, hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemcpy(dY, Y, bytes * K, hipMemcpyHostToDevice), "hipMemcpy");
        if (cg_rates_batch_device(h, n, dT, dP, dY, dW, precision, nullptr) < 0) throw std::runtime_error(g_error);
        check(hipMemcpy(wdot, dW, bytes * K, hipMemcpyDeviceToHost), "hipMemcpy");
        check(hipEventSynchronize(mo.e1), "hipEventSynchronize");
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, mo.e0, mo.e1), "hipEventElapsedTime");
        mo.last_ms = ms;
        return 0;
    });
}

double cg_rates_last_kernel_ms(int32_t h) {
    try {
        MechObj& mo = get_mech(h);
        check(hipEventSynchronize(mo.e1), "hipEventSynchronize");
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, mo.e0, mo.e1), "hipEventElapsedTime");
        mo.last_ms = ms;
        return ms;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1.0;
    }
}

int32_t cg_jac_batch_device(int32_t h, int64_t n, const double* T, const double* P, const double* Y, void* J,
                            int32_t precision, void* stream) {
    return guarded([&] {
        MechObj& mo = get_mech(h);
        if (precision != CG_FP64 && precision != CG_FP32)
            throw std::runtime_error("cg_jac_batch: precision must be CG_FP64 or CG_FP32");
        const bool f32 = precision == CG_FP32;
        if (f32 && mo.ns != 1) throw std::runtime_error("cg_jac_batch: CG_FP32 output needs K + 2 <= 64");
        if (mo.ns > 2) throw std::runtime_error("cg_jac_batch: mechanisms with more than 126 species are not supported");
        if (n <= 0) return 0;
        if (!T || !P || !Y || !J) throw std::runtime_error("null array");
        const DevMech& m = mo.dm->view();
        const size_t lds = jac_lds_bytes(m, mo.ns);
        const void* kp = jac_kernel_ptr(mo.ns, f32);
        set_lds_limit(kp, lds);
        const int grid = grid_for(kp, lds, 0, n);
        hipStream_t s = static_cast<hipStream_t>(stream);
        check(hipEventRecord(mo.j0, s), "hipEventRecord");
        check(jac_launch(m, mo.ns, f32, n, T, P, Y, J, grid, lds, s), "Jacobian kernel launch");
        check(hipEventRecord(mo.j1, s), "hipEventRecord");
        return 0;
    });
}

int32_t cg_jac_batch(int32_t h, int64_t n, const double* T, const double* P, const double* Y, void* J,
                     int32_t precision) {
    return guarded([&] {
        MechObj& mo = get_mech(h);
        if (n <= 0) return 0;
        if (!T || !P || !Y || !J) throw std::runtime_error("null array");
        const int K = mo.dm->view().K, N = mo.dm->view().N;
        const size_t bytes = static_cast<size_t>(n) * sizeof(double);
        const size_t jbytes = static_cast<size_t>(n) * N * N * (precision == CG_FP32 ? 4 : 8);
        double* d = nullptr;
        check(hipMalloc(&d, bytes * (2 + static_cast<size_t>(K)) + jbytes), "hipMalloc");
        std::unique_ptr<double, void (*)(double*)> guard(d, [](double* p) { (void)hipFree(p); });
        double *dT = d, *dP = d + n, *dY = d + 2 * n, *dJ = d + (2 + static_cast<size_t>(K)) * n;
        check(hipMemcpy(dT, T, bytes, hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemcpy(dP, P, bytes, hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemcpy(dY, Y, bytes * K, hipMemcpyHostToDevice), "hipMemcpy");
        if (cg_jac_batch_device(h, n, dT, dP, dY, dJ, precision, nullptr) < 0) throw std::runtime_error(g_error);
        check(hipMemcpy(J, dJ, jbytes, hipMemcpyDeviceToHost), "hipMemcpy");
        return 0;
    });
}

static int equil_mode(const char* XY) {
    if (!XY) throw std::runtime_error("null equilibrium mode");
    const std::string s(XY);
    if (s == "HP" || s == "PH") return 1;
    if (s == "TP" || s == "PT") return 0;
    throw std::runtime_error("equilibrium mode must be \"HP\" or \"TP\", got \"" + s + "\"");
}

int32_t cg_equilibrate_batch_device(int32_t h, int64_t n, const char* XY, const double* T, const double* P,
                                    const double* Y, double* T_out, double* Y_out, int32_t* info, int32_t max_iter,
                                    void* stream) {
    return guarded([&] {
        MechObj& mo = get_mech(h);
        const int hp = equil_mode(XY);
        if (n <= 0) return 0;
        if (!T || !P || !Y || !T_out || !Y_out) throw std::runtime_error("null array");
        if (max_iter < 1) throw std::runtime_error("max_iter must be positive");
        const DevMech& m = mo.dm->view();
        if (m.E > 8) throw std::runtime_error("equilibrium supports at most 8 elements");
        if (mo.ns > 2) throw std::runtime_error("equilibrium: mechanisms with more than 126 species are not supported");
        const void* kp = equil_kernel_ptr(mo.ns);
        const int grid = grid_for(kp, 0, 0, n);
        check(equil_launch(m, mo.ns, hp, n, T, P, Y, T_out, Y_out, info, max_iter, grid,
                           static_cast<hipStream_t>(stream)),
              "equilibrium kernel launch");
        return 0;
    });
}

int32_t cg_equilibrate_batch(int32_t h, int64_t n, const char* XY, const double* T, const double* P, const double* Y,
                             double* T_out, double* Y_out, int32_t* info) {
    return guarded([&] {
        MechObj& mo = get_mech(h);
        equil_mode(XY);
        if (n <= 0) return 0;
        if (!T || !P || !Y || !T_out || !Y_out) throw std::runtime_error("null array");
        const int K = mo.dm->view().K;
        for (int64_t s = 0; s < n; s++) {
            if (!(T[s] > 0.0) || !std::isfinite(T[s])) throw std::runtime_error("temperature must be positive and finite");
            if (!(P[s] > 0.0) || !std::isfinite(P[s])) throw std::runtime_error("pressure must be positive and finite");
        }
        const size_t b = static_cast<size_t>(n) * 8;
        double* d = nullptr;
        check(hipMalloc(&d, b * (3 + 2 * static_cast<size_t>(K)) + static_cast<size_t>(n) * 4), "hipMalloc");
        std::unique_ptr<double, void (*)(double*)> guard(d, [](double* p) { (void)hipFree(p); });
        double *dT = d, *dP = d + n, *dY = d + 2 * n, *dTo = dY + static_cast<size_t>(K) * n,
               *dYo = dTo + n;
        int* dI = reinterpret_cast<int*>(dYo + static_cast<size_t>(K) * n);
        check(hipMemcpy(dT, T, b, hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemcpy(dP, P, b, hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemcpy(dY, Y, b * K, hipMemcpyHostToDevice), "hipMemcpy");
        if (cg_equilibrate_batch_device(h, n, XY, dT, dP, dY, dTo, dYo, dI, 500, nullptr) < 0)
            throw std::runtime_error(g_error);
        check(hipDeviceSynchronize(), "equilibrium");
        check(hipMemcpy(T_out, dTo, b, hipMemcpyDeviceToHost), "hipMemcpy");
        check(hipMemcpy(Y_out, dYo, b * K, hipMemcpyDeviceToHost), "hipMemcpy");
        std::vector<int32_t> inf(n);
        check(hipMemcpy(inf.data(), dI, static_cast<size_t>(n) * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        int bad = 0;
        for (int64_t s = 0; s < n; s++) bad += inf[s] < 0;
        if (info) std::memcpy(info, inf.data(), static_cast<size_t>(n) * 4);
        if (bad) throw std::runtime_error(std::to_string(bad) + " state(s) did not reach equilibrium (see info)");
        return 0;
    });
}

double cg_jac_last_kernel_ms(int32_t h) {
    try {
        MechObj& mo = get_mech(h);
        check(hipEventSynchronize(mo.j1), "hipEventSynchronize");
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, mo.j0, mo.j1), "hipEventElapsedTime");
        return ms;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1.0;
    }
}

static int32_t create_batch(int32_t h, int64_t n, const double* T0, const double* P, const double* Y0, double rtol,
                            double atol, int32_t precision, bool const_volume) {
    return guarded([&] {
        MechObj& mo = get_mech(h);
        if (n <= 0 || n > (1LL << 31) - 1) throw std::runtime_error("reactor count out of range");
        if (!T0 || !P || !Y0) throw std::runtime_error("null array");
        if (!(rtol > 0) || !(atol > 0)) throw std::runtime_error("tolerances must be positive");
        if (precision != CG_FP64 && precision != CG_MIXED) throw std::runtime_error("unknown precision");
        const DevMech& m = mo.dm->view();
        const HostMech& hm = mo.dm->host();
        // Phase::setTemperature / Phase::setDensity reject non-positive values
        // (Phase.cpp); non-finite inputs are rejected here too
        for (int64_t s = 0; s < n; s++) {
            if (!(T0[s] > 0.0) || !std::isfinite(T0[s]))
                throw std::runtime_error("reactor " + std::to_string(s) + ": temperature must be positive and finite");
            if (!(P[s] > 0.0) || !std::isfinite(P[s]))
                throw std::runtime_error("reactor " + std::to_string(s) + ": pressure must be positive and finite");
            for (int k = 0; k < m.K; k++)
                if (!std::isfinite(Y0[s * m.K + k]))
                    throw std::runtime_error("reactor " + std::to_string(s) + ": non-finite mass fraction");
        }
        auto b = std::make_unique<Batch>();
        b->mech = &mo;
        b->n = static_cast<int>(n);
        b->K = m.K;
        b->N = m.N;
        b->ns = mo.ns;
        b->fp32 = precision == CG_MIXED;
        b->const_volume = const_volume;
        b->ld = (m.N % 2) ? m.N : m.N + 1;
        b->rtol = rtol;
        b->atol = atol;
        const int NP = 64 * b->ns;
        b->rec_stride = 7 * static_cast<size_t>(NP) + C_COUNT;
        b->j_stride = static_cast<size_t>(m.N) * m.N;
        b->lu_stride = bdf_lu_elems(m.N, b->ld);
        b->piv_stride = 2 * static_cast<size_t>(NP);
        if (bdf_variant(m.N) < 0) throw std::runtime_error("mechanisms with more than 126 species are not supported");
        const size_t es = b->fp32 ? 4 : 8;
        check(hipMalloc(&b->rec, b->rec_stride * n * 8), "hipMalloc(records)");
        check(hipMalloc(&b->jsave, b->j_stride * n * 8), "hipMalloc(jacobians)");
        check(hipMalloc(&b->lusave, b->lu_stride * n * es), "hipMalloc(factors)");
        check(hipMalloc(&b->piv, b->piv_stride * n * 4), "hipMalloc(pivots)");
        check(hipMalloc(&b->tout, n * 8), "hipMalloc(tout)");
        check(hipMalloc(&b->queue, 64), "hipMalloc(queue)");
        check(hipMalloc(&b->lpt, 5 * n * sizeof(int)), "hipMalloc(order)");
        check(hipMemset(b->lpt, 0, n * sizeof(int)), "hipMemset(order)");
        b->sort_tmp_bytes = order_temp_bytes(static_cast<int>(n));
        check(hipMalloc(&b->sort_tmp, b->sort_tmp_bytes > 0 ? b->sort_tmp_bytes : 16), "hipMalloc(sort)");
        check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreate(&b->e0), "hipEventCreate");
        check(hipEventCreate(&b->e1), "hipEventCreate");
        // initial records: zn[0] = y0 in device order [Y, T, m]; V = 1 m^3.  The
        // record's pressure slot holds the volume for a constant-volume reactor.
        std::vector<double> rec(b->rec_stride * n, 0.0);
        const double R = 6.02214076e26 * 1.380649e-23;
        for (int64_t s = 0; s < n; s++) {
            double* r = rec.data() + s * b->rec_stride;
            const double* y = Y0 + s * m.K;
            double sum = 0.0;
            for (int k = 0; k < m.K; k++) sum += y[k] * (1.0 / hm.mw[k]);
            const double mmw = 1.0 / sum;
            const double dens = P[s] * mmw / (T0[s] * R);
            for (int k = 0; k < m.K; k++) r[k] = y[k];
            r[m.K] = T0[s];
            r[m.K + 1] = dens * 1.0;
            for (int i = 0; i < m.N; i++) r[6 * NP + i] = r[i];   // yout = y0
            double* c = r + 7 * NP;
            c[C_Q] = 1; c[C_L] = 2; c[C_QWAIT] = 2; c[C_QPRIME] = 1; c[C_ETAMAX] = 10000.0;
            c[C_P] = const_volume ? 1.0 : P[s];
            c[C_ETA] = 1.0;
        }
        check(hipMemcpy(b->rec, rec.data(), rec.size() * 8, hipMemcpyHostToDevice), "hipMemcpy(records)");
        b->lds = bdf_lds_bytes(m, b->ld, b->fp32, b->global_matrix);
        set_lds_limit(bdf_kernel_ptr(m.N, b->fp32, b->global_matrix), b->lds);
        std::lock_guard<std::mutex> lk(g_mu);
        g_batches.push_back(std::move(b));
        return static_cast<int32_t>(g_batches.size() - 1);
    });
}

int32_t cg_reactors_create(int32_t h, int64_t n, const double* T0, const double* P, const double* Y0, double rtol,
                           double atol, int32_t precision) {
    return create_batch(h, n, T0, P, Y0, rtol, atol, precision, false);
}

int32_t cg_reactors_create_const_volume(int32_t h, int64_t n, const double* T0, const double* P, const double* Y0,
                                        double rtol, double atol, int32_t precision) {
    return create_batch(h, n, T0, P, Y0, rtol, atol, precision, true);
}

// Sparsity of the mole reactor's preconditioner Jacobian in the internal order
// [n_0..n_{K-1}, T] (constant volume: [n.., T, V], V coupled to nothing): T row
// and column, the diagonal, species of a reaction coupled to its reactants and
// products, and (unless skipped) to its explicit colliders and, for a default
// efficiency, to every species.
static std::vector<char> mole_pattern(const HostMech& hm, bool skip_tb, bool cv) {
    const int K = hm.K, N = K + 1 + (cv ? 1 : 0);
    std::vector<char> P(static_cast<size_t>(N) * N, 0);
    for (int i = 0; i < N; i++) {
        P[i * N + i] = 1;
        P[K * N + i] = 1;
        P[i * N + K] = 1;
    }
    for (const HostReaction& r : hm.rxn) {
        std::vector<int> sp(r.rk.begin(), r.rk.end());
        sp.insert(sp.end(), r.pk.begin(), r.pk.end());
        const bool tb = !skip_tb && (r.type == kThreeBody || r.type == kLindemann || r.type == kTroe ||
                                     r.type == kSri || r.type == kTsang);
        for (int k : sp) {
            for (int j : sp) P[k * N + j] = 1;
            if (tb) {
                for (int j : r.tbk) P[k * N + j] = 1;
                if (r.tbdef != 0.0)
                    for (int j = 0; j < K; j++) P[k * N + j] = 1;
            }
        }
    }
    return P;
}

int32_t cg_reactors_create_mole(int32_t h, int64_t n, const double* T0, const double* P, const double* Y0,
                                double rtol, double atol, int32_t flags) {
    return guarded([&] {
        MechObj& mo = get_mech(h);
        if (n <= 0 || n > (1LL << 31) - 1) throw std::runtime_error("reactor count out of range");
        if (!T0 || !P || !Y0) throw std::runtime_error("null array");
        if (!(rtol > 0) || !(atol > 0)) throw std::runtime_error("tolerances must be positive");
        if (flags < 0 || flags > 15)
            throw std::runtime_error("flags: 1 skip third bodies, 2 skip falloff, 4 user atol, 8 constant volume");
        const bool cv = (flags & 8) != 0;
        const DevMech& m = mo.dm->view();
        const HostMech& hm = mo.dm->host();
        const int K = m.K, N = K + 1 + (cv ? 1 : 0);
        if (bdf_mole_ns(N) < 0)
            throw std::runtime_error(std::string("mole-based reactors support at most ") + (cv ? "254" : "255") +
                                     " species");
        for (int64_t s = 0; s < n; s++) {
            if (!(T0[s] > 0.0) || !std::isfinite(T0[s]))
                throw std::runtime_error("reactor " + std::to_string(s) + ": temperature must be positive and finite");
            if (!(P[s] > 0.0) || !std::isfinite(P[s]))
                throw std::runtime_error("reactor " + std::to_string(s) + ": pressure must be positive and finite");
            for (int k = 0; k < K; k++)
                if (!std::isfinite(Y0[s * K + k]))
                    throw std::runtime_error("reactor " + std::to_string(s) + ": non-finite mass fraction");
        }
        // symbolic analysis of the preconditioner (host, once per batch)
        // and a device copy of its tables (the internal index layout is
        // this library's, so the permutation is applied to the reaction
        // species lists of the host mechanism: reactions in mechanism order)
        const SparseLUHost sl = sparse_lu_analyse(N, mole_pattern(hm, flags & 1, cv));
        auto b = std::make_unique<Batch>();
        b->mech = &mo;
        b->mole = true;
        b->const_volume = cv;
        b->n = static_cast<int>(n);
        b->K = K;
        b->N = N;
        b->ns = bdf_mole_ns(N);
        b->fp32 = false;
        b->rtol = rtol;
        b->atol = atol;
        b->ld = N;
        const int NP = 64 * b->ns;
        b->rec_stride = 7 * static_cast<size_t>(NP) + C_COUNT;
        b->j_stride = static_cast<size_t>(sl.nnz) + static_cast<size_t>(K) * K;
        b->lu_stride = static_cast<size_t>(sl.nnz);
        b->piv_stride = 1;
        b->sp_updates = sl.updates();
        // tables: diag, lptr, lslot, lrow, optr, ops, uptr, uslot, urow, iperm, srow, scol
        std::vector<const std::vector<int>*> parts = {&sl.diag, &sl.lptr, &sl.lslot, &sl.lrow, &sl.optr, &sl.ops,
                                                      &sl.uptr, &sl.uslot, &sl.urow, &sl.iperm, &sl.row, &sl.col};
        size_t total = 0;
        for (auto* v : parts) total += v->size() + 1;
        std::vector<int> flat;
        flat.reserve(total);
        std::vector<size_t> off;
        for (auto* v : parts) {
            off.push_back(flat.size());
            flat.insert(flat.end(), v->begin(), v->end());
            flat.push_back(0);
        }
        check(hipMalloc(&b->sp_tables, flat.size() * sizeof(int)), "hipMalloc(sparse tables)");
        check(hipMemcpy(b->sp_tables, flat.data(), flat.size() * sizeof(int), hipMemcpyHostToDevice),
              "hipMemcpy(sparse tables)");
        const int* d = b->sp_tables;
        b->sp.nnz = sl.nnz;
        b->sp.skip = flags & 3;
        b->sp.diag = d + off[0]; b->sp.lptr = d + off[1]; b->sp.lslot = d + off[2]; b->sp.lrow = d + off[3];
        b->sp.optr = d + off[4]; b->sp.ops = d + off[5]; b->sp.uptr = d + off[6]; b->sp.uslot = d + off[7];
        b->sp.urow = d + off[8]; b->sp.iperm = d + off[9]; b->sp.srow = d + off[10]; b->sp.scol = d + off[11];
        check(hipMalloc(&b->rec, b->rec_stride * n * 8), "hipMalloc(records)");
        check(hipMalloc(&b->jsave, b->j_stride * n * 8), "hipMalloc(jacobians)");
        check(hipMalloc(&b->lusave, b->lu_stride * n * 8), "hipMalloc(factors)");
        check(hipMalloc(&b->piv, b->piv_stride * n * 4), "hipMalloc(pivots)");
        check(hipMalloc(&b->tout, n * 8), "hipMalloc(tout)");
        check(hipMalloc(&b->queue, 64), "hipMalloc(queue)");
        check(hipMalloc(&b->lpt, 5 * n * sizeof(int)), "hipMalloc(order)");
        check(hipMemset(b->lpt, 0, n * sizeof(int)), "hipMemset(order)");
        b->sort_tmp_bytes = order_temp_bytes(static_cast<int>(n));
        check(hipMalloc(&b->sort_tmp, b->sort_tmp_bytes > 0 ? b->sort_tmp_bytes : 16), "hipMalloc(sort)");
        check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreate(&b->e0), "hipEventCreate");
        check(hipEventCreate(&b->e1), "hipEventCreate");
        // initial records: V = 1 m^3, n_k = rho V Y_k / W_k, internal [n.., T(, V)]
        std::vector<double> rec(b->rec_stride * n, 0.0);
        const double R = 6.02214076e26 * 1.380649e-23;
        for (int64_t s = 0; s < n; s++) {
            double* r = rec.data() + s * b->rec_stride;
            const double* y = Y0 + s * K;
            double sum = 0.0;
            for (int k = 0; k < K; k++) sum += y[k] * (1.0 / hm.mw[k]);
            const double mmw = 1.0 / sum;
            const double mass = P[s] * mmw / (T0[s] * R) * 1.0;
            double tot = 0.0;
            for (int k = 0; k < K; k++) {
                r[k] = mass * (1.0 / hm.mw[k]) * y[k];
                tot += r[k];
            }
            r[K] = T0[s];
            if (cv) r[K + 1] = 1.0;
            for (int i = 0; i < N; i++) r[6 * NP + i] = r[i];
            double* c = r + 7 * NP;
            c[C_Q] = 1; c[C_L] = 2; c[C_QWAIT] = 2; c[C_QPRIME] = 1; c[C_ETAMAX] = 10000.0;
            c[C_P] = P[s];
            c[C_ETA] = 1.0;
            c[C_ATOL_SP] = (flags & 4) ? atol : atol * tot;
        }
        check(hipMemcpy(b->rec, rec.data(), rec.size() * 8, hipMemcpyHostToDevice), "hipMemcpy(records)");
        b->lds = bdf_mole_lds_bytes(m, sl.nnz);
        if (b->lds > kLdsPerCu)
            throw std::runtime_error("the preconditioner's fill pattern (" + std::to_string(sl.nnz) +
                                     " entries) needs " + std::to_string(b->lds / 1024) +
                                     " KiB of LDS per reactor, more than the " + std::to_string(kLdsPerCu / 1024) +
                                     " KiB of a CU; skipping the third-body and falloff derivatives (flags 1 | 2) "
                                     "makes it sparser");
        set_lds_limit(bdf_mole_kernel_ptr(N), b->lds);
        std::lock_guard<std::mutex> lk(g_mu);
        g_batches.push_back(std::move(b));
        return static_cast<int32_t>(g_batches.size() - 1);
    });
}

int32_t cg_reactors_linear_stats(int32_t h, int64_t* stats) {
    return guarded([&] {
        Batch& b = get_batch(h);
        if (!stats) throw std::runtime_error("null output");
        check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
        const int NP = 64 * b.ns;
        std::vector<double> c(static_cast<size_t>(b.n) * C_COUNT);
        check(hipMemcpy2D(c.data(), C_COUNT * 8, b.rec + 7 * NP, b.rec_stride * 8, C_COUNT * 8, b.n, hipMemcpyDeviceToHost),
              "hipMemcpy2D(stats)");
        for (int s = 0; s < b.n; s++) {
            const double* x = c.data() + static_cast<size_t>(s) * C_COUNT;
            int64_t* o = stats + static_cast<size_t>(s) * 6;
            o[0] = static_cast<int64_t>(x[C_NLI]);
            o[1] = b.mole ? static_cast<int64_t>(x[C_NJE]) : 0;
            o[2] = static_cast<int64_t>(x[C_NPS]);
            o[3] = static_cast<int64_t>(x[C_NJT]);
            o[4] = static_cast<int64_t>(x[C_NFE_LS]);
            o[5] = static_cast<int64_t>(x[C_NCFL]);
        }
        return 0;
    });
}

int32_t cg_reactors_free(int32_t h) {
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_mu);
        if (h < 0 || h >= static_cast<int32_t>(g_batches.size()) || !g_batches[h]) throw std::runtime_error("invalid reactor-batch handle");
        (void)hipStreamSynchronize(g_batches[h]->stream);
        g_batches[h].reset();
        return 0;
    });
}

int32_t cg_reactors_set_options(int32_t h, int32_t max_steps, int32_t waves_per_cu) {
    return guarded([&] {
        Batch& b = get_batch(h);
        if (max_steps > 0) b.max_steps = max_steps;
        if (waves_per_cu >= 0) b.waves_per_cu = waves_per_cu;
        b.drop_graph();
        return 0;
    });
}

int32_t cg_reactors_set_matrix_storage(int32_t h, int32_t storage) {
    return guarded([&] {
        Batch& b = get_batch(h);
        if (storage < 0 || storage > 2) throw std::runtime_error("storage must be 0 (default), 1 (LDS) or 2 (global)");
        if (b.mole) throw std::runtime_error("matrix storage applies to mass-fraction reactor batches");
        b.global_matrix = storage == 2;
        b.drop_graph();
        const DevMech& m = b.mech->dm->view();
        b.lds = bdf_lds_bytes(m, b.ld, b.fp32, b.global_matrix);
        set_lds_limit(bdf_kernel_ptr(m.N, b.fp32, b.global_matrix, b.nsens > 0), b.lds);
        return 0;
    });
}

int32_t cg_reactors_set_schedule(int32_t h, int32_t longest_first) {
    return guarded([&] {
        Batch& b = get_batch(h);
        b.lpt_on = longest_first != 0;
        return 0;
    });
}

static BdfArgs advance_args(Batch& b) {
    BdfArgs a{};
    b.launched = true;
    a.mech = b.mech->dm->view();
    a.mech.const_volume = b.const_volume ? 1 : 0;
    a.n = b.n;
    // the reference's IdealGasReactor state also holds V (constant): it counts
    // in the WRMS norm's length and nowhere else
    a.neq_norm = b.N + (b.const_volume ? 1 : 0);
    a.ld = b.ld;
    a.max_steps = b.max_steps;
    a.rtol = b.rtol;
    a.atol = b.atol;
    a.rec = b.rec;
    a.rec_stride = b.rec_stride;
    a.j_save = b.jsave;
    a.j_stride = b.j_stride;
    a.lu_save_raw = b.lusave;
    a.lu_stride = b.lu_stride;
    a.piv_save = b.piv;
    a.piv_stride = b.piv_stride;
    a.tout = b.tout;
    int* order = b.lpt + 4 * static_cast<size_t>(b.n);
    a.order = (b.lpt_on && b.have_order) ? order : nullptr;
    a.queue = b.queue;
    a.nsens = b.nsens;
    for (int p = 0; p < b.nsens; p++) a.sens_rxn[p] = b.sens_rxn[p];
    a.rtol_sens = b.rtol_sens;
    a.atol_sens = b.atol_sens;
    a.sens = b.sens;
    a.sens_stride = b.sens_stride;
    a.sp = b.sp;
    if (b.mole) a.neq_norm = b.N;
    return a;
}

static const void* batch_kernel_ptr(const Batch& b) {
    return b.mole ? bdf_mole_kernel_ptr(b.N) : bdf_kernel_ptr(b.N, b.fp32, b.global_matrix, b.nsens > 0);
}

static hipError_t batch_launch(const Batch& b, const BdfArgs& a, int grid) {
    return b.mole ? bdf_mole_launch(a, grid, b.lds, b.stream) : bdf_launch(a, b.fp32, b.global_matrix, grid, b.lds, b.stream);
}

static int advance_grid(const Batch& b) {
    return grid_for(batch_kernel_ptr(b), b.lds, b.waves_per_cu, b.n);
}

// cost keys of the call just launched -> processing order of the next
static void update_order(Batch& b) {
    const size_t n = static_cast<size_t>(b.n);
    int* order = b.lpt + 4 * n;
    check(order_update(b.rec, b.rec_stride, 7 * 64 * b.ns, b.n, b.lpt, b.lpt + n, b.lpt + 2 * n, order, b.sort_tmp,
                       b.stream),
          "work-queue order update");
    b.have_order = true;
}

static void launch_advance(Batch& b) {
    const BdfArgs a = advance_args(b);
    const int grid = advance_grid(b);
    check(hipEventRecord(b.e0, b.stream), "hipEventRecord");
    check(batch_launch(b, a, grid), "integrator kernel launch");
    check(hipEventRecord(b.e1, b.stream), "hipEventRecord");
    update_order(b);
}

int32_t cg_reactors_advance_device(int32_t h, const double* t_out_dev) {
    return guarded([&] {
        Batch& b = get_batch(h);
        check(hipMemcpyAsync(b.tout, t_out_dev, static_cast<size_t>(b.n) * 8, hipMemcpyDeviceToDevice, b.stream),
              "hipMemcpyAsync(tout)");
        launch_advance(b);
        return 0;
    });
}

int32_t cg_reactors_advance_graph(int32_t h, int32_t n_steps, const double* t_out_dev) {
    return guarded([&] {
        Batch& b = get_batch(h);
        if (n_steps < 1) throw std::runtime_error("n_steps must be >= 1");
        if (!t_out_dev) throw std::runtime_error("null t_out");
        const int* order = (b.lpt_on && b.have_order) ? b.lpt + 4 * static_cast<size_t>(b.n) : nullptr;
        if (!b.graph_exec || b.graph_steps != n_steps || b.graph_tout != t_out_dev || b.graph_order != order) {
            b.drop_graph();
            // kernel arguments (including each step's t_out row) are fixed at
            // capture; the values behind the pointers are read at run time
            const int grid = advance_grid(b);
            const size_t nn = static_cast<size_t>(b.n);
            int* order_buf = b.lpt + 4 * nn;
            check(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            hipError_t err = hipSuccess;
            for (int k = 0; k < n_steps && err == hipSuccess; k++) {
                BdfArgs a = advance_args(b);
                // the first launch uses the order from before the call; each
                // later one the order its predecessor's step counts give
                a.order = (k == 0) ? order : (b.lpt_on ? order_buf : nullptr);
                a.tout = t_out_dev + static_cast<size_t>(k) * nn;
                err = batch_launch(b, a, grid);
                if (err == hipSuccess && b.lpt_on)
                    err = order_update(b.rec, b.rec_stride, 7 * 64 * b.ns, b.n, b.lpt, b.lpt + nn, b.lpt + 2 * nn,
                                       order_buf, b.sort_tmp, b.stream);
            }
            hipGraph_t g = nullptr;
            const hipError_t end = hipStreamEndCapture(b.stream, &g);
            check(err, "integrator kernel capture");
            check(end, "hipStreamEndCapture");
            b.graph = g;
            check(hipGraphInstantiate(&b.graph_exec, b.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
            b.graph_steps = n_steps;
            b.graph_tout = t_out_dev;
            b.graph_order = order;
        }
        check(hipEventRecord(b.e0, b.stream), "hipEventRecord");
        check(hipGraphLaunch(b.graph_exec, b.stream), "hipGraphLaunch");
        check(hipEventRecord(b.e1, b.stream), "hipEventRecord");
        if (b.lpt_on) b.have_order = true;
        return 0;
    });
}

int32_t cg_reactors_autotune(int32_t h, const double* t_probe_dev, int32_t* choice, double* times_ms,
                             int32_t max_candidates) {
    return guarded([&] {
        Batch& b = get_batch(h);
        if (!t_probe_dev) throw std::runtime_error("null t_probe");
        if (b.nsens > 0) throw std::runtime_error("autotune runs on batches without sensitivity parameters");
        if (b.mole) throw std::runtime_error("autotune runs on mass-fraction reactor batches only");
        const DevMech& m = b.mech->dm->view();
        const int v = bdf_variant(m.N);
        // (matrix storage, waves per CU) candidates; storage 1 = LDS, 2 = global
        std::vector<std::pair<int, int>> cand;
        if (v == 56 || v == 64) cand = {{1, 0}, {2, 0}, {2, 4}};
        else cand = {{0, 0}, {0, 4}, {0, 8}};
        // snapshot of everything an advance changes, restored before each trial
        const size_t n = static_cast<size_t>(b.n);
        const size_t es = b.fp32 ? 4 : 8;
        const size_t sizes[5] = {b.rec_stride * n * 8, b.j_stride * n * 8, b.lu_stride * n * es, b.piv_stride * n * 4,
                                 5 * n * sizeof(int)};
        void* live[5] = {b.rec, b.jsave, b.lusave, b.piv, b.lpt};
        void* snap[5] = {};
        struct Free {
            void** p;
            ~Free() { for (int i = 0; i < 5; i++) if (p[i]) (void)hipFree(p[i]); }
        } guard_snap{snap};
        check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
        for (int i = 0; i < 5; i++) {
            check(hipMalloc(&snap[i], sizes[i]), "hipMalloc(autotune snapshot)");
            check(hipMemcpyAsync(snap[i], live[i], sizes[i], hipMemcpyDeviceToDevice, b.stream), "hipMemcpyAsync");
        }
        const bool had_order = b.have_order;
        const bool launched0 = b.launched;   // trials are undone: not an initialisation
        auto restore = [&] {
            for (int i = 0; i < 5; i++)
                check(hipMemcpyAsync(live[i], snap[i], sizes[i], hipMemcpyDeviceToDevice, b.stream), "hipMemcpyAsync");
            b.have_order = had_order;
        };
        const bool global0 = b.global_matrix;
        const int wpc0 = b.waves_per_cu;
        int best = -1;
        double best_ms = 0.0;
        std::vector<std::pair<const void*, int>> seen;
        for (size_t c = 0; c < cand.size(); c++) {
            b.global_matrix = cand[c].first == 2;
            b.waves_per_cu = cand[c].second;
            b.lds = bdf_lds_bytes(m, b.ld, b.fp32, b.global_matrix);
            const void* kp = bdf_kernel_ptr(m.N, b.fp32, b.global_matrix);
            set_lds_limit(kp, b.lds);
            const int grid = advance_grid(b);
            double ms = -1.0;
            if (std::find(seen.begin(), seen.end(), std::make_pair(kp, grid)) == seen.end()) {
                seen.emplace_back(kp, grid);
                restore();
                check(hipMemcpyAsync(b.tout, t_probe_dev, n * 8, hipMemcpyDeviceToDevice, b.stream), "hipMemcpyAsync");
                launch_advance(b);
                check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
                float f = 0.f;
                check(hipEventElapsedTime(&f, b.e0, b.e1), "hipEventElapsedTime");
                ms = f;
                // the first candidate is the default launch; another replaces
                // it only when at least 3% faster (one trial's timing noise)
                if (best < 0 || ms < (best == 0 ? 0.97 : 1.0) * best_ms) { best = static_cast<int>(c); best_ms = ms; }
            }
            if (times_ms && static_cast<int>(c) < max_candidates) times_ms[c] = ms;
        }
        restore();
        check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
        if (best < 0) { b.global_matrix = global0; b.waves_per_cu = wpc0; }
        else { b.global_matrix = cand[best].first == 2; b.waves_per_cu = cand[best].second; }
        b.lds = bdf_lds_bytes(m, b.ld, b.fp32, b.global_matrix);
        set_lds_limit(bdf_kernel_ptr(m.N, b.fp32, b.global_matrix), b.lds);
        b.drop_graph();
        b.launched = launched0;
        if (choice) {
            choice[0] = b.global_matrix ? 2 : ((v == 56 || v == 64) ? 1 : 0);
            choice[1] = b.waves_per_cu;
        }
        return static_cast<int32_t>(cand.size());
    });
}

int32_t cg_reactors_add_sensitivity_reaction(int32_t h, int32_t rxn) {
    return guarded([&] {
        Batch& b = get_batch(h);
        const std::vector<int>& perm = b.mech->dm->perm();
        const int I = static_cast<int>(perm.size());
        if (rxn < 0 || rxn >= I)
            throw std::runtime_error("Reaction number out of range (" + std::to_string(rxn) + ")");
        if (b.launched)
            throw std::runtime_error("Sensitivity parameters cannot be added after the integrator has been initialized");
        if (b.nsens >= kMaxSens)
            throw std::runtime_error("at most " + std::to_string(kMaxSens) + " sensitivity parameters per batch");
        if (b.fp32) throw std::runtime_error("sensitivities are integrated in FP64 precision only");
        if (b.mole) throw std::runtime_error("sensitivities are not available for mole-based reactor batches");
        int pidx = -1;
        for (int i = 0; i < I; i++)
            if (perm[i] == rxn) pidx = i;
        if (pidx < 0) throw std::runtime_error("reaction not found in the device tables");
        const int NP = 64 * b.ns;
        const size_t stride = static_cast<size_t>(b.nsens + 1) * kSensCols * NP;
        double* buf = nullptr;
        check(hipMalloc(&buf, stride * b.n * 8), "hipMalloc(sensitivities)");
        check(hipMemset(buf, 0, stride * b.n * 8), "hipMemset(sensitivities)");
        if (b.sens) check(hipFree(b.sens), "hipFree(sensitivities)");
        b.sens = buf;
        b.sens_stride = stride;
        b.sens_rxn[b.nsens] = pidx;
        b.sens_rxn_orig[b.nsens] = rxn;
        b.nsens++;
        b.drop_graph();
        set_lds_limit(bdf_kernel_ptr(b.N, false, b.global_matrix, true), b.lds);
        return b.nsens - 1;
    });
}

int32_t cg_reactors_set_sensitivity_tolerances(int32_t h, double rtol, double atol) {
    return guarded([&] {
        Batch& b = get_batch(h);
        if (!(rtol > 0) || !(atol > 0)) throw std::runtime_error("tolerances must be positive");
        b.rtol_sens = rtol;
        b.atol_sens = atol;
        b.drop_graph();
        return 0;
    });
}

int32_t cg_reactors_get_sensitivities(int32_t h, int32_t normalized, double* S) {
    return guarded([&] {
        Batch& b = get_batch(h);
        if (!S) throw std::runtime_error("null output");
        check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
        const int NP = 64 * b.ns, N = b.N, K = b.K, np = b.nsens;
        if (np == 0) return 0;
        std::vector<double> sb(b.sens_stride * b.n), st(static_cast<size_t>(b.n) * N), ctl(static_cast<size_t>(b.n) * C_COUNT);
        check(hipMemcpy(sb.data(), b.sens, sb.size() * 8, hipMemcpyDeviceToHost), "hipMemcpy(sensitivities)");
        check(hipMemcpy2D(st.data(), N * 8, b.rec + 6 * NP, b.rec_stride * 8, N * 8, b.n, hipMemcpyDeviceToHost),
              "hipMemcpy2D(state)");
        check(hipMemcpy2D(ctl.data(), C_COUNT * 8, b.rec + 7 * NP, b.rec_stride * 8, C_COUNT * 8, b.n,
                          hipMemcpyDeviceToHost), "hipMemcpy2D(control)");
        // reference order [m, T, Y_0..] <- device order [Y_0.., T, m]
        auto dev = [&](int i) { return i == 0 ? K + 1 : (i == 1 ? K : i - 2); };
        for (int s = 0; s < b.n; s++) {
            const bool started = ctl[static_cast<size_t>(s) * C_COUNT + C_NST] > 0;
            for (int i = 0; i < N; i++) {
                const int d = dev(i);
                double den = st[static_cast<size_t>(s) * N + d];
                if (den == 0.0) den = 1e-300;   // SmallNumber
                for (int p = 0; p < np; p++) {
                    double v = started ? sb[s * b.sens_stride + (static_cast<size_t>(p) * kSensCols + 7) * NP + d] : 0.0;
                    if (normalized) v /= den;
                    S[(static_cast<size_t>(s) * N + i) * np + p] = v;
                }
            }
        }
        return 0;
    });
}

int32_t cg_reactors_status(int32_t h, int32_t* status) {
    return guarded([&] {
        Batch& b = get_batch(h);
        check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
        std::vector<double> st(b.n);
        const int NP = 64 * b.ns;
        check(hipMemcpy2D(st.data(), 8, b.rec + 7 * NP + C_STATUS, b.rec_stride * 8, 8, b.n, hipMemcpyDeviceToHost),
              "hipMemcpy2D(status)");
        int bad = 0;
        for (int s = 0; s < b.n; s++) {
            if (status) status[s] = static_cast<int32_t>(st[s]);
            if (st[s] < 0) bad++;
        }
        return bad;
    });
}

int32_t cg_reactors_advance(int32_t h, const double* t_out) {
    return guarded([&] {
        Batch& b = get_batch(h);
        if (!t_out) throw std::runtime_error("null t_out");
        check(hipMemcpyAsync(b.tout, t_out, static_cast<size_t>(b.n) * 8, hipMemcpyHostToDevice, b.stream),
              "hipMemcpyAsync(tout)");
        launch_advance(b);
        check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, b.e0, b.e1), "hipEventElapsedTime");
        b.last_ms = ms;
        int bad = cg_reactors_status(h, nullptr);
        if (bad < 0) throw std::runtime_error(g_error);
        if (bad > 0) throw std::runtime_error(std::to_string(bad) + " reactor(s) failed to integrate; see cg_reactors_status");
        return 0;
    });
}

int32_t cg_reactors_get_state(int32_t h, double* y) {
    return guarded([&] {
        Batch& b = get_batch(h);
        check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
        const int NP = 64 * b.ns;
        std::vector<double> buf(static_cast<size_t>(b.n) * b.N);
        check(hipMemcpy2D(buf.data(), b.N * 8, b.rec + 6 * NP, b.rec_stride * 8, b.N * 8, b.n, hipMemcpyDeviceToHost),
              "hipMemcpy2D(state)");
        if (b.mole) {
            // [T, n_1..n_K] (IdealGasConstPressureMoleReactor::getState) from
            // [n.., T]; constant volume [T, V, n_1..n_K] from [n.., T, V]
            const int s0 = b.const_volume ? 2 : 1;
            for (int s = 0; s < b.n; s++) {
                const double* src = buf.data() + static_cast<size_t>(s) * b.N;
                double* dst = y + static_cast<size_t>(s) * b.N;
                dst[0] = src[b.K];
                if (b.const_volume) dst[1] = src[b.K + 1];
                for (int k = 0; k < b.K; k++) dst[s0 + k] = src[k];
            }
            return 0;
        }
        for (int s = 0; s < b.n; s++) {
            const double* src = buf.data() + static_cast<size_t>(s) * b.N;
            double* dst = y + static_cast<size_t>(s) * b.N;
            dst[0] = src[b.K + 1];
            dst[1] = src[b.K];
            for (int k = 0; k < b.K; k++) dst[2 + k] = src[k];
        }
        return 0;
    });
}

int32_t cg_reactors_stats(int32_t h, int64_t* stats) {
    return guarded([&] {
        Batch& b = get_batch(h);
        check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
        const int NP = 64 * b.ns;
        std::vector<double> c(static_cast<size_t>(b.n) * C_COUNT);
        check(hipMemcpy2D(c.data(), C_COUNT * 8, b.rec + 7 * NP, b.rec_stride * 8, C_COUNT * 8, b.n, hipMemcpyDeviceToHost),
              "hipMemcpy2D(stats)");
        for (int s = 0; s < b.n; s++) {
            const double* x = c.data() + static_cast<size_t>(s) * C_COUNT;
            int64_t* o = stats + static_cast<size_t>(s) * 9;
            o[0] = static_cast<int64_t>(x[C_NST]);
            o[1] = static_cast<int64_t>(x[C_NFE]);
            o[2] = static_cast<int64_t>(x[C_NJE]);
            o[3] = static_cast<int64_t>(x[C_NSETUPS]);
            o[4] = static_cast<int64_t>(x[C_NNI]);
            o[5] = static_cast<int64_t>(x[C_NCFN]);
            o[6] = static_cast<int64_t>(x[C_NETF]);
            o[7] = static_cast<int64_t>(x[C_Q]);
            o[8] = static_cast<int64_t>(x[C_NFE_DQ]);
        }
        return 0;
    });
}

int32_t cg_reactors_get_control(int32_t h, double* ctl) {
    return guarded([&] {
        Batch& b = get_batch(h);
        if (!ctl) throw std::runtime_error("null array");
        check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
        const int NP = 64 * b.ns;
        check(hipMemcpy2D(ctl, C_COUNT * 8, b.rec + 7 * NP, b.rec_stride * 8, C_COUNT * 8, b.n, hipMemcpyDeviceToHost),
              "hipMemcpy2D(control)");
        return 0;
    });
}

double cg_reactors_last_kernel_ms(int32_t h) {
    try {
        Batch& b = get_batch(h);
        check(hipEventSynchronize(b.e1), "hipEventSynchronize");
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, b.e0, b.e1), "hipEventElapsedTime");
        b.last_ms = ms;
        return ms;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1.0;
    }
}

/* ---- learned surrogate ---------------------------------------------------- */

int32_t cg_surrogate_create(int32_t mech, double dt, int32_t hidden, int32_t depth, uint64_t seed) {
    return guarded([&] {
        MechObj& mo = get_mech(mech);
        if (mo.ns > 2) throw std::runtime_error("surrogate: mechanisms with more than 126 species are not supported");
        auto obj = std::make_unique<SurrogateObj>();
        obj->mech = &mo;
        obj->net = std::make_unique<SurrogateNet>(mo.dm->view().K, dt, hidden, depth, seed);
        obj->net->set_elements(mo.dm->view().comp, mo.dm->view().rmw, mo.dm->view().E);
        std::lock_guard<std::mutex> lk(g_mu);
        g_surr.push_back(std::move(obj));
        return static_cast<int32_t>(g_surr.size() - 1);
    });
}

int32_t cg_surrogate_free(int32_t s) {
    return guarded([&] {
        get_surr(s);
        std::lock_guard<std::mutex> lk(g_mu);
        g_surr[s].reset();
        return 0;
    });
}

int64_t cg_surrogate_nparams(int32_t s) {
    try {
        return static_cast<int64_t>(get_surr(s).net->n_params());
    } catch (const std::exception& e) {
        g_error = e.what();
        return CG_ERR;
    }
}

int32_t cg_surrogate_fit_normalization(int32_t s, int64_t n, const double* y0, const double* P, const double* y1) {
    return guarded([&] {
        if (n <= 0 || !y0 || !P || !y1) throw std::runtime_error("empty or null training pairs");
        get_surr(s).net->fit_normalization(n, y0, P, y1, 1 << 20);
        return 0;
    });
}

int32_t cg_surrogate_train_epoch(int32_t s, int64_t n, const double* y0, const double* P, const double* y1,
                                 int32_t batch, double lr, double* loss) {
    return guarded([&] {
        if (n <= 0 || !y0 || !P || !y1) throw std::runtime_error("empty or null training pairs");
        if (n > (1LL << 31) - 1) throw std::runtime_error("too many training pairs for one call");
        const double l = get_surr(s).net->train_epoch(n, y0, P, y1, batch, lr, nullptr);
        if (loss) *loss = l;
        return 0;
    });
}

int32_t cg_surrogate_loss_grad(int32_t s, int64_t n, const double* y0, const double* P, const double* y1,
                               double* loss, float* grad) {
    return guarded([&] {
        if (n <= 0 || n > (1 << 20) || !y0 || !P || !y1) throw std::runtime_error("1..2^20 pairs required");
        const double l = get_surr(s).net->loss_grad(n, y0, P, y1, grad, nullptr);
        if (loss) *loss = l;
        return 0;
    });
}

int32_t cg_surrogate_predict(int32_t s, int64_t n, const double* y0, const double* P, double* y1) {
    return guarded([&] {
        if (n <= 0) return 0;
        if (!y0 || !P || !y1) throw std::runtime_error("null array");
        get_surr(s).net->predict(n, y0, P, y1, nullptr);
        check(hipDeviceSynchronize(), "surrogate predict");
        return 0;
    });
}

int32_t cg_surrogate_predict_noalloc(int32_t s, int64_t n, const double* y0, const double* P, double* y1,
                                     void* stream) {
    return guarded([&] {
        if (n < 0) throw std::runtime_error("negative row count");
        if (n == 0) return 0;
        if (!y0 || !P || !y1) throw std::runtime_error("null array");
        get_surr(s).net->predict_noalloc(n, y0, P, y1, static_cast<hipStream_t>(stream));
        return 0;
    });
}

int32_t cg_surrogate_get_params(int32_t s, float* p) {
    return guarded([&] { get_surr(s).net->get_params(p); return 0; });
}
int32_t cg_surrogate_set_params(int32_t s, const float* p) {
    return guarded([&] { get_surr(s).net->set_params(p); return 0; });
}
int32_t cg_surrogate_get_normalization(int32_t s, double* v) {
    return guarded([&] { get_surr(s).net->get_normalization(v); return 0; });
}
int32_t cg_surrogate_set_normalization(int32_t s, const double* v) {
    return guarded([&] { get_surr(s).net->set_normalization(v); return 0; });
}

// guarded surrogate step for every reactor whose output time is one step dt
// ahead; the rest are integrated.  Stream-ordered on the batch stream.
static void advance_surrogate_device(Batch& b, SurrogateObj& so, const double* t_out_dev, double tol_T, double tol_Y,
                                     double gate_T) {
    if (so.mech != b.mech) throw std::runtime_error("surrogate and reactor batch use different mechanisms");
    if (b.const_volume) throw std::runtime_error("the surrogate path supports constant-pressure reactors only");
    if (b.nsens > 0) throw std::runtime_error("the surrogate path does not integrate sensitivities");
    if (b.mole) throw std::runtime_error("the surrogate path supports mass-fraction reactor batches only");
    if (!t_out_dev) throw std::runtime_error("null t_out");
    if (!(tol_T > 0) || !(tol_Y > 0)) throw std::runtime_error("guard tolerances must be positive");
    const int n = b.n, N = b.N, K = b.K, NP = 64 * b.ns;
    const size_t nn = static_cast<size_t>(n);
    if (!b.sg_Y) {
        check(hipMalloc(&b.sg_Y, nn * N * 8), "hipMalloc(surrogate states)");
        check(hipMalloc(&b.sg_Yn, nn * N * 8), "hipMalloc(surrogate predictions)");
        check(hipMalloc(&b.sg_F, nn * N * 8), "hipMalloc(surrogate rhs)");
        check(hipMalloc(&b.sg_P, nn * 8), "hipMalloc(surrogate pressures)");
        check(hipMalloc(&b.sg_ints, 2 * nn * 4), "hipMalloc(surrogate flags)");
        check(hipMalloc(&b.sg_counters, 3 * sizeof(unsigned long long)), "hipMalloc(surrogate counters)");
        check(surrogate_clear(b.sg_ints, 2 * n, b.stream), "surrogate clear");
        check(surrogate_clear(reinterpret_cast<int*>(b.sg_counters), 6, b.stream), "surrogate clear");
    }
    int* elig = b.sg_ints;
    int* fvalid = b.sg_ints + nn;
    const double dt = so.net->dt();
    const DevMech& m = b.mech->dm->view();
    const HostMech& hm = b.mech->dm->host();
    int rad[3];
    const char* names[3] = {"H", "O", "OH"};
    for (int j = 0; j < 3; j++) {
        rad[j] = -1;
        for (int k = 0; k < K; k++)
            if (hm.names[k] == names[j]) rad[j] = k;
    }
    if (rad[0] < 0 || rad[1] < 0 || rad[2] < 0) rad[0] = -1;   // no gate species: every step integrated
    if (t_out_dev != b.tout)
        check(hipMemcpyAsync(b.tout, t_out_dev, nn * 8, hipMemcpyDeviceToDevice, b.stream), "hipMemcpyAsync(tout)");
    check(surrogate_gather(n, b.rec, b.rec_stride, NP, N, K, b.tout, dt, b.sg_Y, b.sg_P, elig, b.sg_counters,
                           b.stream), "surrogate gather");
    so.net->predict(n, b.sg_Y, b.sg_P, b.sg_Yn, b.stream);
    const size_t lds = guard_lds_bytes(m, b.ns);
    const void* gk = guard_kernel_ptr(b.ns);
    set_lds_limit(gk, lds);
    const int grid = grid_for(gk, lds, 0, n);
    check(surrogate_guard(m, b.ns, n, elig, b.sg_Y, b.sg_Yn, b.sg_F, fvalid, b.sg_P, b.rec, b.rec_stride, NP, b.tout,
                          dt, tol_T, tol_Y, gate_T, rad, b.sg_counters, grid, lds, b.stream),
          "surrogate guard");
    // the rejected and the non-eligible reactors, in the batch's longest-first
    // order (a list built by atomics would hand them out in arrival order)
    BdfArgs a = advance_args(b);
    a.work_flag = elig;
    a.work_value = 2;
    check(hipEventRecord(b.e0, b.stream), "hipEventRecord");
    check(bdf_launch(a, b.fp32, b.global_matrix, advance_grid(b), b.lds, b.stream), "integrator kernel launch");
    check(hipEventRecord(b.e1, b.stream), "hipEventRecord");
    update_order(b);
}

int32_t cg_reactors_advance_surrogate_device(int32_t bh, int32_t sh, const double* t_out_dev, double tol_T,
                                             double tol_Y, double gate_T) {
    return guarded([&] {
        advance_surrogate_device(get_batch(bh), get_surr(sh), t_out_dev, tol_T, tol_Y, gate_T);
        return 0;
    });
}

int32_t cg_reactors_net_production_rates_device(int32_t bh, double* wdot_dev) {
    return guarded([&] {
        Batch& b = get_batch(bh);
        if (!wdot_dev) throw std::runtime_error("null rates array");
        if (b.mole) throw std::runtime_error("rates at the states of a mole-based batch are not available");
        const DevMech& m = b.mech->dm->view();
        const size_t lds = rates_lds_bytes(m, b.ns);
        const void* kp = reactor_rates_kernel_ptr(b.ns);
        set_lds_limit(kp, lds);
        const int grid = grid_for(kp, lds, 0, b.n);
        check(reactor_rates_launch(m, b.ns, b.n, b.rec, b.rec_stride, b.const_volume ? 1 : 0, wdot_dev, grid, lds,
                                   b.stream),
              "reactor rates kernel launch");
        return 0;
    });
}

int32_t cg_reactors_get_state_device(int32_t bh, double* y_dev) {
    return guarded([&] {
        Batch& b = get_batch(bh);
        if (!y_dev) throw std::runtime_error("null state array");
        if (b.mole) throw std::runtime_error("device state export is for mass-fraction reactor batches");
        check(export_states(b.n, b.rec, b.rec_stride, 64 * b.ns, b.N, b.K, y_dev, b.stream), "state export");
        return 0;
    });
}

int32_t cg_reactors_surrogate_counts(int32_t bh, int64_t* counts, int32_t reset) {
    return guarded([&] {
        Batch& b = get_batch(bh);
        unsigned long long c[3] = {0, 0, 0};
        if (b.sg_counters) {
            check(hipMemcpyAsync(c, b.sg_counters, sizeof(c), hipMemcpyDeviceToHost, b.stream), "hipMemcpyAsync");
            check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
            if (reset) check(surrogate_clear(reinterpret_cast<int*>(b.sg_counters), 6, b.stream), "surrogate clear");
        }
        if (counts)
            for (int j = 0; j < 3; j++) counts[j] = static_cast<int64_t>(c[j]);
        return 0;
    });
}

int32_t cg_reactors_advance_surrogate(int32_t bh, int32_t sh, const double* t_out, double tol_T, double tol_Y,
                                      double gate_T, int64_t* counts) {
    return guarded([&] {
        Batch& b = get_batch(bh);
        SurrogateObj& so = get_surr(sh);
        if (!t_out) throw std::runtime_error("null t_out");
        unsigned long long before[3] = {0, 0, 0};
        if (b.sg_counters) {
            check(hipMemcpyAsync(before, b.sg_counters, sizeof(before), hipMemcpyDeviceToHost, b.stream), "hipMemcpyAsync");
        }
        // t_out goes through the batch's own time buffer (copied in stream order)
        check(hipMemcpyAsync(b.tout, t_out, static_cast<size_t>(b.n) * 8, hipMemcpyHostToDevice, b.stream),
              "hipMemcpyAsync(t_out)");
        advance_surrogate_device(b, so, b.tout, tol_T, tol_Y, gate_T);
        unsigned long long after[3];
        check(hipMemcpyAsync(after, b.sg_counters, sizeof(after), hipMemcpyDeviceToHost, b.stream), "hipMemcpyAsync");
        check(hipStreamSynchronize(b.stream), "hipStreamSynchronize");
        if (counts)
            for (int j = 0; j < 3; j++) counts[j] = static_cast<int64_t>(after[j] - before[j]);
        int bad = cg_reactors_status(bh, nullptr);
        if (bad < 0) throw std::runtime_error(g_error);
        if (bad > 0) throw std::runtime_error(std::to_string(bad) + " reactor(s) failed to integrate; see cg_reactors_status");
        return 0;
    });
}

}  // extern "C"
