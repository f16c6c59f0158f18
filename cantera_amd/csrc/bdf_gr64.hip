// Integrator kernels with the iteration matrix kept in the reactor's global-memory slot, factored and solved in registers (56 < N <= 64).
#include "bdf_integrator.hpp"

namespace cg {

template <typename MT, bool kSens>
__global__ __launch_bounds__(64, 2) void bdf_advance_gr64(BdfArgs a) {
    bdf_advance_body<1, GlobalRegLinSys<64, MT>, MT, kSens>(a);
}

// sens: the FP64 kernel that also integrates sensitivities
const void* bdf_kernel_gr64(bool fp32, bool sens) {
    if (sens) return reinterpret_cast<const void*>(&bdf_advance_gr64<double, true>);
    return fp32 ? reinterpret_cast<const void*>(&bdf_advance_gr64<float, false>)
                : reinterpret_cast<const void*>(&bdf_advance_gr64<double, false>);
}

hipError_t bdf_launch_gr64(const BdfArgs& a, bool fp32, int grid, size_t lds, hipStream_t stream) {
    if (a.nsens > 0) hipLaunchKernelGGL((bdf_advance_gr64<double, true>), dim3(grid), dim3(64), lds, stream, a);
    else if (fp32) hipLaunchKernelGGL((bdf_advance_gr64<float, false>), dim3(grid), dim3(64), lds, stream, a);
    else hipLaunchKernelGGL((bdf_advance_gr64<double, false>), dim3(grid), dim3(64), lds, stream, a);
    return hipGetLastError();
}

}  // namespace cg
