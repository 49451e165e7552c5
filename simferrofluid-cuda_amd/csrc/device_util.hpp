// Device helpers shared by the gfx950 kernels: 64-lane wavefront reductions,
// a deterministic grid-wide reduction (last block finishes), and the 7-point
// operator applies in FP32 and FP64.
#pragma once
#include "sfp_internal.hpp"

namespace sfp {

constexpr int kBlock = 256;          // every kernel runs 256 threads = 4 wavefronts
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ int flat_tid() {
    return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}
__device__ __forceinline__ unsigned flat_bid() {
    return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}
__device__ __forceinline__ unsigned num_blocks() { return gridDim.x * gridDim.y * gridDim.z; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sums v[q] over the block; the totals are valid in thread 0.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N]) {
    __shared__ double sh[N][kWaves];
    const int tid = flat_tid();
#pragma unroll
    for (int q = 0; q < N; q++) v[q] = wave_sum(v[q]);
    if ((tid & 63) == 0) {
#pragma unroll
        for (int q = 0; q < N; q++) sh[q][tid >> 6] = v[q];
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < N; q++) {
            double s = sh[q][0];
#pragma unroll
            for (int w = 1; w < kWaves; w++) s += sh[q][w];
            v[q] = s;
        }
    }
    __syncthreads();
}

// Grid-wide sum of v[q]. Each block stores its partial sums; the block that
// arrives last adds all partials in block order (deterministic for a fixed
// launch shape) and returns true with the totals in thread 0's v[]. Every
// thread of every block must call it. `partials` holds N * num_blocks()
// doubles; `counter` must be 0 on entry and is 0 again on exit.
template <int N>
__device__ __forceinline__ bool grid_sum(double (&v)[N], double *partials, unsigned *counter) {
    __shared__ unsigned s_last;
    const unsigned nb = num_blocks(), bid = flat_bid();
    const int tid = flat_tid();
    block_sum<N>(v);
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < N; q++) partials[(size_t)q * nb + bid] = v[q];
        __threadfence();
        unsigned ticket = atomicAdd(counter, 1u);
        s_last = (ticket == nb - 1) ? 1u : 0u;
    }
    __syncthreads();
    if (s_last == 0) return false;
    __threadfence();
    double acc[N];
#pragma unroll
    for (int q = 0; q < N; q++) acc[q] = 0.0;
    for (unsigned b = tid; b < nb; b += kBlock) {
#pragma unroll
        for (int q = 0; q < N; q++) acc[q] += partials[(size_t)q * nb + b];
    }
    block_sum<N>(acc);
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < N; q++) v[q] = acc[q];
        atomicExch(counter, 0u);
    }
    return true;
}

// Strides of the z-fastest layout.
struct Strides {
    size_t sx, sy;
    __device__ __host__ explicit Strides(Dims d) : sx((size_t)d.ny * d.nz), sy((size_t)d.nz) {}
};

// -(off-diagonal part) of A x at cell c: sum of coupling * neighbour value.
template <class T, class Load>
__device__ __forceinline__ T offdiag_sum(const T *const cpl[3], Dims d, Strides s, int i, int j,
                                         int k, size_t c, Load xv) {
    T acc = 0;
    if (i > 0) acc += cpl[0][c] * (T)xv(c - s.sx);
    if (i + 1 < d.nx) acc += cpl[0][c + s.sx] * (T)xv(c + s.sx);
    if (j > 0) acc += cpl[1][c] * (T)xv(c - s.sy);
    if (j + 1 < d.ny) acc += cpl[1][c + s.sy] * (T)xv(c + s.sy);
    if (k > 0) acc += cpl[2][c] * (T)xv(c - 1);
    if (k + 1 < d.nz) acc += cpl[2][c + 1] * (T)xv(c + 1);
    return acc;
}

}  // namespace sfp
