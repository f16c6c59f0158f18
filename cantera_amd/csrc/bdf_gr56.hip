// Integrator kernels with the iteration matrix kept in the reactor's global-memory slot, factored and solved in registers (32 < N <= 56).
#include "bdf_integrator.hpp"

namespace cg {

template <typename MT>
__global__ __launch_bounds__(64, 2) void bdf_advance_gr56(BdfArgs a) {
    bdf_advance_body<1, GlobalRegLinSys<56, MT>, MT>(a);
}

const void* bdf_kernel_gr56(bool fp32) {
    return fp32 ? reinterpret_cast<const void*>(&bdf_advance_gr56<float>)
                : reinterpret_cast<const void*>(&bdf_advance_gr56<double>);
}

hipError_t bdf_launch_gr56(const BdfArgs& a, bool fp32, int grid, size_t lds, hipStream_t stream) {
    if (fp32) hipLaunchKernelGGL(bdf_advance_gr56<float>, dim3(grid), dim3(64), lds, stream, a);
    else hipLaunchKernelGGL(bdf_advance_gr56<double>, dim3(grid), dim3(64), lds, stream, a);
    return hipGetLastError();
}

}  // namespace cg
