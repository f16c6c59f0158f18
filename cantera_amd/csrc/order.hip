// Processing order of the persistent integrator's work queue: reactors sorted
// by the number of internal steps they took in the previous advance call, most
// expensive first (longest-processing-time-first list scheduling).  An advance
// call ends when its slowest wavefront finishes, so starting the stiff reactors
// first keeps the cheap ones for the end, where they fill the gaps.
//
// Runs on the batch stream after each integrator launch: one small kernel
// turns the step counters of the control blocks into keys, then a device radix
// sort (hipCUB) produces the order the next launch reads.  The order changes
// only which wavefront integrates which reactor, never a reactor's result.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>

#include "bdf.hpp"

namespace cg {

namespace {

constexpr int kKeyBits = 20;                       // step counts are capped below 2^20
constexpr int kKeyMax = (1 << kKeyBits) - 1;

__global__ void order_keys_kernel(const double* __restrict__ rec, size_t rec_stride, int ctl_off, int n,
                                  int* __restrict__ prev_nst, int* __restrict__ keys, int* __restrict__ idx) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int nst = static_cast<int>(rec[static_cast<size_t>(r) * rec_stride + ctl_off + C_NST]);
    const int d = nst - prev_nst[r];
    prev_nst[r] = nst;
    keys[r] = d < 0 ? 0 : (d > kKeyMax ? kKeyMax : d);
    idx[r] = r;
}

}  // namespace

size_t order_temp_bytes(int n) {
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, static_cast<const int*>(nullptr),
                                                       static_cast<int*>(nullptr), static_cast<const int*>(nullptr),
                                                       static_cast<int*>(nullptr), n, 0, kKeyBits);
    return bytes;
}

hipError_t order_update(const double* rec, size_t rec_stride, int ctl_off, int n, int* prev_nst, int* keys_in,
                        int* keys_out, int* idx_in, int* order_out, void* temp, size_t temp_bytes,
                        hipStream_t stream) {
    hipLaunchKernelGGL(order_keys_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, rec, rec_stride, ctl_off, n,
                       prev_nst, keys_in, idx_in);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipcub::DeviceRadixSort::SortPairsDescending(temp, temp_bytes, keys_in, keys_out, idx_in, order_out, n, 0,
                                                        kKeyBits, stream);
}

}  // namespace cg
