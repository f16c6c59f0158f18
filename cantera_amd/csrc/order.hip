// Processing order of the persistent integrator's work queue: reactors ordered
// by the number of internal steps they took in the previous advance call, most
// expensive first (longest-processing-time-first list scheduling).  An advance
// call ends when its slowest wavefront finishes, so starting the stiff reactors
// first keeps the cheap ones for the end, where they fill the gaps.
//
// Runs on the batch stream after each integrator launch as a counting sort on
// the step count (clamped to kBuckets - 1): clear the histogram, count and rank
// each reactor in its bucket (LDS then global atomics), scan, scatter.  Four
// kernels and no memset or library call, so the same sequence can be captured
// in a HIP graph.
// Reactors with equal counts land in atomic order; the order changes only
// which wavefront integrates which reactor, never a reactor's result.
#include <hip/hip_runtime.h>

#include "bdf.hpp"

namespace cg {

namespace {

constexpr int kBuckets = 1024;

__global__ void clear_hist_kernel(int* __restrict__ hist) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kBuckets) hist[i] = 0;
}

// bucket 0 holds the most expensive reactors.  Ranks within a bucket are
// taken in two levels: LDS atomics inside the workgroup, then one global
// atomic per non-empty bucket and workgroup reserves the workgroup's range
// (most reactors share a few step counts, so per-reactor global atomics
// would serialise on those buckets).
__global__ __launch_bounds__(256) void bucket_kernel(const double* __restrict__ rec, size_t rec_stride, int ctl_off,
                                                     int n, int* __restrict__ prev_nst, int* __restrict__ bucket,
                                                     int* __restrict__ rank, int* __restrict__ hist) {
    __shared__ int lh[kBuckets];
    const int t = threadIdx.x;
    for (int i = t; i < kBuckets; i += blockDim.x) lh[i] = 0;
    __syncthreads();
    const int r = blockIdx.x * blockDim.x + t;
    int b = 0, lr = 0;
    if (r < n) {
        const int nst = static_cast<int>(rec[static_cast<size_t>(r) * rec_stride + ctl_off + C_NST]);
        int d = nst - prev_nst[r];
        prev_nst[r] = nst;
        d = d < 0 ? 0 : (d > kBuckets - 1 ? kBuckets - 1 : d);
        b = kBuckets - 1 - d;
        lr = atomicAdd(&lh[b], 1);
    }
    __syncthreads();
    for (int i = t; i < kBuckets; i += blockDim.x) {
        const int c = lh[i];
        if (c > 0) lh[i] = atomicAdd(&hist[i], c);
    }
    __syncthreads();
    if (r < n) {
        bucket[r] = b;
        rank[r] = lh[b] + lr;
    }
}

// exclusive scan of the histogram: one workgroup of kBuckets threads
__global__ __launch_bounds__(kBuckets) void scan_hist_kernel(const int* __restrict__ hist, int* __restrict__ offs) {
    __shared__ int s[kBuckets];
    const int t = threadIdx.x;
    const int v = hist[t];
    s[t] = v;
    __syncthreads();
    for (int d = 1; d < kBuckets; d <<= 1) {
        const int add = t >= d ? s[t - d] : 0;
        __syncthreads();
        s[t] += add;
        __syncthreads();
    }
    offs[t] = s[t] - v;
}

__global__ void scatter_kernel(int n, const int* __restrict__ bucket, const int* __restrict__ rank,
                               const int* __restrict__ offs, int* __restrict__ order) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    order[offs[bucket[r]] + rank[r]] = r;
}

}  // namespace

size_t order_temp_bytes(int) { return 2 * kBuckets * sizeof(int); }

hipError_t order_update(const double* rec, size_t rec_stride, int ctl_off, int n, int* prev_nst, int* bucket,
                        int* rank, int* order_out, void* temp, hipStream_t stream) {
    int* hist = static_cast<int*>(temp);
    int* offs = hist + kBuckets;
    const dim3 g((n + 255) / 256), b(256);
    hipLaunchKernelGGL(clear_hist_kernel, dim3(kBuckets / 256), b, 0, stream, hist);
    hipLaunchKernelGGL(bucket_kernel, g, b, 0, stream, rec, rec_stride, ctl_off, n, prev_nst, bucket, rank, hist);
    hipLaunchKernelGGL(scan_hist_kernel, dim3(1), dim3(kBuckets), 0, stream, hist, offs);
    hipLaunchKernelGGL(scatter_kernel, g, b, 0, stream, n, bucket, rank, offs, order_out);
    return hipGetLastError();
}

}  // namespace cg
