// Integrator kernels of the mole-based reactor with the sparse preconditioner
// and GMRES (IdealGasConstPressureMoleReactor; N = K + 1 <= 128).
#include "bdf_integrator.hpp"

namespace cg {

namespace {
__global__ void reset_mole_queue(int* q) { *q = 0; }
}  // namespace

template <int NS>
__global__ __launch_bounds__(64) void bdf_advance_mole(BdfArgs a) {
    bdf_advance_body<NS, SparseLdsLinSys<NS>, double, false, true>(a);
}

int bdf_mole_ns(int N) { return N <= 64 ? 1 : (N <= 128 ? 2 : (N <= 256 ? 4 : -1)); }

size_t bdf_mole_lds_bytes(const DevMech& m, int nnz) {
    const int ns = bdf_mole_ns(m.K + 1);
    const size_t lu = (static_cast<size_t>(nnz) + 1) & ~static_cast<size_t>(1);
    return (bdf_scratch_doubles(m.K, ns) + lu + static_cast<size_t>(kWave) * ns + kGmDoubles) * sizeof(double);
}

const void* bdf_mole_kernel_ptr(int N) {
    switch (bdf_mole_ns(N)) {
        case 1: return reinterpret_cast<const void*>(&bdf_advance_mole<1>);
        case 2: return reinterpret_cast<const void*>(&bdf_advance_mole<2>);
        case 4: return reinterpret_cast<const void*>(&bdf_advance_mole<4>);
        default: return nullptr;
    }
}

hipError_t bdf_mole_launch(const BdfArgs& a, int grid, size_t lds, hipStream_t stream) {
    hipLaunchKernelGGL(reset_mole_queue, dim3(1), dim3(1), 0, stream, a.queue);
    switch (bdf_mole_ns(a.mech.K + 1)) {
        case 1: hipLaunchKernelGGL(bdf_advance_mole<1>, dim3(grid), dim3(64), lds, stream, a); break;
        case 2: hipLaunchKernelGGL(bdf_advance_mole<2>, dim3(grid), dim3(64), lds, stream, a); break;
        case 4: hipLaunchKernelGGL(bdf_advance_mole<4>, dim3(grid), dim3(64), lds, stream, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cg
