// Integrator dispatch: picks the kernel variant for the mechanism size.
#include "bdf_integrator.hpp"

namespace cg {

const void* bdf_kernel_r16(bool, bool);
const void* bdf_kernel_r32(bool, bool);
const void* bdf_kernel_lr56(bool, bool);
const void* bdf_kernel_lr64(bool, bool);
const void* bdf_kernel_l2(bool, bool);
const void* bdf_kernel_gr56(bool, bool);
const void* bdf_kernel_gr64(bool, bool);
hipError_t bdf_launch_r16(const BdfArgs&, bool, int, size_t, hipStream_t);
hipError_t bdf_launch_r32(const BdfArgs&, bool, int, size_t, hipStream_t);
hipError_t bdf_launch_lr56(const BdfArgs&, bool, int, size_t, hipStream_t);
hipError_t bdf_launch_lr64(const BdfArgs&, bool, int, size_t, hipStream_t);
hipError_t bdf_launch_l2(const BdfArgs&, bool, int, size_t, hipStream_t);
hipError_t bdf_launch_gr56(const BdfArgs&, bool, int, size_t, hipStream_t);
hipError_t bdf_launch_gr64(const BdfArgs&, bool, int, size_t, hipStream_t);

namespace {
__global__ void reset_queue_kernel(int* q) { *q = 0; }
}  // namespace

// 16, 32: iteration matrix in registers (one row per lane).
// 56, 64: iteration matrix kept in LDS, factored and solved in registers.
// 2: iteration matrix in LDS, state spread over 2 slots per lane.
int bdf_variant(int N) {
    if (N <= 16) return 16;
    if (N <= 32) return 32;
    if (N <= 56) return 56;
    if (N <= 64) return 64;
    if (N <= 128) return 2;
    return -1;
}

// register policies keep 64 reciprocal pivots after the factors
size_t bdf_lu_elems(int N, int ld) {
    const int v = bdf_variant(N);
    if (v == 56 || v == 64) return static_cast<size_t>(v + 1) * kWave;
    if (v == 16 || v == 32) return static_cast<size_t>(N) * ld + kWave;
    return static_cast<size_t>(N) * ld;
}

size_t bdf_lds_bytes(const DevMech& m, int ld, bool fp32, bool global) {
    const int v = bdf_variant(m.N);
    const int ns = (v == 2) ? 2 : 1;
    size_t bytes = bdf_scratch_doubles(m.K, ns) * sizeof(double);
    if (v == 2) {
        bytes += static_cast<size_t>(kWave) * ns * (sizeof(double) + sizeof(int));
        bytes += static_cast<size_t>(m.N) * ld * (fp32 ? 4 : 8);
    } else if ((v == 56 || v == 64) && !global) {
        bytes += static_cast<size_t>(v + 1) * kWave * (fp32 ? 4 : 8) + 2 * kWave * sizeof(int);
    }
    return bytes;
}

const void* bdf_kernel_ptr(int N, bool fp32, bool global, bool sens) {
    switch (bdf_variant(N)) {
        case 16: return bdf_kernel_r16(fp32, sens);
        case 32: return bdf_kernel_r32(fp32, sens);
        case 56: return global ? bdf_kernel_gr56(fp32, sens) : bdf_kernel_lr56(fp32, sens);
        case 64: return global ? bdf_kernel_gr64(fp32, sens) : bdf_kernel_lr64(fp32, sens);
        case 2: return bdf_kernel_l2(fp32, sens);
        default: return nullptr;
    }
}

hipError_t bdf_launch(const BdfArgs& a, bool fp32, bool global, int grid, size_t lds, hipStream_t stream) {
    hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, stream, a.queue);
    switch (bdf_variant(a.mech.N)) {
        case 16: return bdf_launch_r16(a, fp32, grid, lds, stream);
        case 32: return bdf_launch_r32(a, fp32, grid, lds, stream);
        case 56: return global ? bdf_launch_gr56(a, fp32, grid, lds, stream) : bdf_launch_lr56(a, fp32, grid, lds, stream);
        case 64: return global ? bdf_launch_gr64(a, fp32, grid, lds, stream) : bdf_launch_lr64(a, fp32, grid, lds, stream);
        case 2: return bdf_launch_l2(a, fp32, grid, lds, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cg
