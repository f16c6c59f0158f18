// Batched net production rates (Kinetics::getNetProductionRates,
// src/kinetics/Kinetics.cpp:447-457) over many independent states.
//
// Layout: species-major SoA in HBM, Y[k*n + s] and wdot[k*n + s].  A block is
// one wavefront that owns a tile of 64 states: the tile is read with coalesced
// loads into LDS, each state is then evaluated by the whole wavefront (lanes
// over species, then over reactions, then over gather sub-rows), its wdot is
// written over its Y column, and the tile is written back coalesced.
#include <hip/hip_runtime.h>

#include "bdf.hpp"
#include "wave_kinetics.hpp"

namespace cg {

namespace {

constexpr int kTile = 64;
constexpr int kTilePad = kTile + 1;

template <int NS>
__global__ __launch_bounds__(64) void rates_tile_kernel(DevMech m, long n, const double* __restrict__ T,
                                                        const double* __restrict__ P, const double* __restrict__ Y,
                                                        double* __restrict__ wdot) {
    extern __shared__ double lds_raw[];
    const int lane = lane_id();
    double* tile = lds_raw;                               // K x kTilePad
    double* base = tile + static_cast<size_t>(m.K) * kTilePad;
    WaveScratch S;
    S.yb = base;  base += kWave * NS;
    S.C = base;   base += m.K;
    S.MU = base;  base += m.K;
    S.ROP = base; base += m.I;
    S.SUB = base; base += m.nsub;
    S.W2 = nullptr;
    S.F = S.Y0 = S.F0 = nullptr;
    const long ntiles = (n + kTile - 1) / kTile;
    for (long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const long s0 = t * kTile;
        const int cnt = static_cast<int>(min(static_cast<long>(kTile), n - s0));
        const bool have = lane < cnt;
        const double Tl = have ? T[s0 + lane] : 300.0;
        const double Pl = have ? P[s0 + lane] : 101325.0;
        for (int k = 0; k < m.K; k++) tile[k * kTilePad + lane] = have ? Y[static_cast<size_t>(k) * n + s0 + lane] : 0.0;
        __syncthreads();
        for (int j = 0; j < cnt; j++) {
            const double Tj = uni(__shfl(Tl, j, kWave));
            const double Pj = uni(__shfl(Pl, j, kWave));
            double yk[NS];
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int k = lane + kWave * s;
                yk[s] = (k < m.K) ? tile[k * kTilePad + j] : 0.0;
            }
            GasState g;
            SpeciesRegs<NS> sr;
            species_pass<NS>(m, S, Tj, Pj, yk, g, sr);
            production_rates<NS>(m, S, g, sr);
#pragma unroll
            for (int s = 0; s < NS; s++) {
                const int k = lane + kWave * s;
                if (k < m.K) tile[k * kTilePad + j] = sr.wdot[s];
            }
        }
        __syncthreads();
        if (have)
            for (int k = 0; k < m.K; k++) wdot[static_cast<size_t>(k) * n + s0 + lane] = tile[k * kTilePad + lane];
        __syncthreads();
    }
}

}  // namespace

size_t rates_lds_bytes(const DevMech& m, int ns) {
    size_t d = static_cast<size_t>(m.K) * kTilePad + kWave * ns + 2 * m.K + m.I + m.nsub;
    return d * sizeof(double);
}

hipError_t rates_launch(const DevMech& m, int ns, long n, const double* T, const double* P, const double* Y,
                        double* wdot, int grid, size_t lds, hipStream_t stream) {
    if (ns == 1) hipLaunchKernelGGL((rates_tile_kernel<1>), dim3(grid), dim3(64), lds, stream, m, n, T, P, Y, wdot);
    else if (ns == 2) hipLaunchKernelGGL((rates_tile_kernel<2>), dim3(grid), dim3(64), lds, stream, m, n, T, P, Y, wdot);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

const void* rates_kernel_ptr(int ns) {
    return ns == 1 ? reinterpret_cast<const void*>(&rates_tile_kernel<1>)
                   : reinterpret_cast<const void*>(&rates_tile_kernel<2>);
}

}  // namespace cg
