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

// 1 / dg on active cells (dg > 0), 0 on inactive ones. Every kernel that
// scales by the inverse diagonal uses this one approximation.
__device__ __forceinline__ float rcp_or_zero(float dg) {
    return dg > 0.f ? __builtin_amdgcn_rcpf(dg) : 0.f;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sums v[q] over a block of T threads; the totals are valid in thread 0.
template <int N, int T = kBlock>
__device__ __forceinline__ void block_sum(double (&v)[N]) {
    constexpr int W = T / 64;
    __shared__ double sh[N][W];
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
            for (int w = 1; w < W; w++) s += sh[q][w];
            v[q] = s;
        }
    }
    __syncthreads();
}

// Block-level partial sums: block b stores its totals at partials[q*nb + b]
// (nb = gridDim.x). A one-block finalize kernel adds them in block order, so
// every reduction is deterministic for a given grid. (A last-block-done
// scheme needs a device-scope fence per block, which on a multi-XCD chip
// writes back L2 and serializes the blocks.)
template <int N>
__device__ __forceinline__ void block_partials(double (&v)[N], double *partials) {
    block_sum<N>(v);
    if (flat_tid() == 0) {
#pragma unroll
        for (int q = 0; q < N; q++) partials[(size_t)q * gridDim.x + blockIdx.x] = v[q];
    }
}

// Inside a one-block kernel of T threads: totals of nb block partials, valid
// in thread 0. Each thread keeps U running sums so U loads per sum are in
// flight at a time (tens of thousands of partials in a few microseconds);
// the order is fixed by nb and T.
// (stride: the distance between the N arrays, nb when 0)
template <int N, int T = kBlock, unsigned U = 8>
__device__ __forceinline__ void sum_partials(const double *partials, unsigned nb, double (&v)[N],
                                             size_t stride = 0) {
    if (stride == 0) stride = nb;
    static_assert(U == 8 || U == 16, "U is 8 or 16");
    double acc[N][U];
#pragma unroll
    for (int q = 0; q < N; q++)
#pragma unroll
        for (unsigned u = 0; u < U; u++) acc[q][u] = 0.0;
    unsigned b = flat_tid();
    for (; b + (U - 1) * T < nb; b += U * T) {
#pragma unroll
        for (unsigned u = 0; u < U; u++)
#pragma unroll
            for (int q = 0; q < N; q++) acc[q][u] += partials[(size_t)q * stride + b + u * T];
    }
    for (; b < nb; b += T)
#pragma unroll
        for (int q = 0; q < N; q++) acc[q][0] += partials[(size_t)q * stride + b];
#pragma unroll
    for (int q = 0; q < N; q++) {
        if constexpr (U == 16) {
#pragma unroll
            for (unsigned u = 0; u < 8; u++) acc[q][u] += acc[q][u + 8];
        }
        v[q] = ((acc[q][0] + acc[q][1]) + (acc[q][2] + acc[q][3])) +
               ((acc[q][4] + acc[q][5]) + (acc[q][6] + acc[q][7]));
    }
    block_sum<N, T>(v);
}

// Tile of a cell-parallel block: z index k (lanes), y index j (wavefronts)
// and the first of its cg.planes consecutive x planes. Blocks are taken in
// launch order (tiles of one plane, then the next chunk of planes); a
// renumbering that keeps each XCD on one contiguous slab measured slower for
// these stencils on MI355X, so none is applied.
__device__ __forceinline__ void cell_tile(CellGrid cg, int &k, int &j, int &i0) {
    const unsigned L = blockIdx.x;
    const unsigned per = cg.gx * cg.gy;
    const unsigned bz = L / per, rem = L - bz * per;
    const unsigned by = rem / cg.gx, bx = rem - by * cg.gx;
    k = (int)(bx * blockDim.x + threadIdx.x);
    j = (int)(by * blockDim.y + threadIdx.y);
    i0 = (int)bz * cg.planes;
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
