// Geometric-multigrid kernels for gfx950: FP32 Chebyshev smoothing with the
// coarse-grid correction fused into the first post-smoothing step,
// residual + restriction, FP64 Galerkin coarsening over 2x2x2 aggregates and
// the dense coarsest solve.
//
// Threads map to cells with z (the unit-stride axis of the reference layout)
// across the 64 lanes of a wavefront, y across the 4 wavefronts of a block,
// and x planes looped over with stride gridDim.z.
#include "device_util.hpp"
#include "kernels.hpp"

namespace sfp {

Launch cell_launch(Dims d, unsigned target_blocks) {
    Launch L;
    unsigned tz = 64;
    while (tz > 1 && (int)tz / 2 >= d.nz) tz /= 2;
    unsigned ty = kBlock / tz;
    unsigned gx = (d.nz + tz - 1) / tz;
    unsigned gy = (d.ny + ty - 1) / ty;
    unsigned per_plane = gx * gy;
    unsigned gz = target_blocks / per_plane;
    if (gz < 1) gz = 1;
    if (gz > (unsigned)d.nx) gz = d.nx;
    L.block = dim3(tz, ty, 1);
    L.grid = dim3(gx, gy, gz);
    return L;
}

namespace {

__device__ __forceinline__ float rcp_or_zero(float dg) {
    return dg > 0.f ? __builtin_amdgcn_rcpf(dg) : 0.f;
}

template <bool FROM64>
__global__ __launch_bounds__(kBlock) void k_smooth_first(Dims d, Op32 A, const double *r64,
                                                         float *b, float *x, float c0) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (k >= d.nz || j >= d.ny) return;
    for (int i = blockIdx.z; i < d.nx; i += gridDim.z) {
        const size_t c = d.idx(i, j, k);
        float bv;
        if constexpr (FROM64) {
            bv = (float)r64[c];
            b[c] = bv;
        } else {
            bv = b[c];
        }
        x[c] = c0 * rcp_or_zero(A.diag[c]) * bv;
    }
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_smooth(Dims d, Op32 A, const float *__restrict__ b,
                                                   const float *__restrict__ xin,
                                                   float *__restrict__ xout, float *dvec,
                                                   float cd, float cr,
                                                   const float *__restrict__ e, Dims dc) {
    constexpr bool kProlong = (MODE & SM_PROLONG) != 0;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (k >= d.nz || j >= d.ny) return;
    const Strides s(d);
    const Strides sc(dc);
    for (int i = blockIdx.z; i < d.nx; i += gridDim.z) {
        const size_t c = d.idx(i, j, k);
        const float dg = A.diag[c];
        float xc = xin[c];
        float xm[3], xp[3];
        xm[0] = i > 0 ? xin[c - s.sx] : 0.f;
        xp[0] = i + 1 < d.nx ? xin[c + s.sx] : 0.f;
        xm[1] = j > 0 ? xin[c - s.sy] : 0.f;
        xp[1] = j + 1 < d.ny ? xin[c + s.sy] : 0.f;
        xm[2] = k > 0 ? xin[c - 1] : 0.f;
        xp[2] = k + 1 < d.nz ? xin[c + 1] : 0.f;
        if constexpr (kProlong) {
            // piecewise-constant prolongation of the coarse correction e
            const size_t pc = dc.idx(i >> 1, j >> 1, k >> 1);
            const float ec = e[pc];
            if (dg > 0.f) xc += ec;
            if (i > 0) xm[0] += (i & 1) ? ec : e[pc - sc.sx];
            if (i + 1 < d.nx) xp[0] += (i & 1) ? e[pc + sc.sx] : ec;
            if (j > 0) xm[1] += (j & 1) ? ec : e[pc - sc.sy];
            if (j + 1 < d.ny) xp[1] += (j & 1) ? e[pc + sc.sy] : ec;
            if (k > 0) xm[2] += (k & 1) ? ec : e[pc - 1];
            if (k + 1 < d.nz) xp[2] += (k & 1) ? e[pc + 1] : ec;
        }
        float off = 0.f;
        off += A.cpl[0][c] * xm[0];
        if (i + 1 < d.nx) off += A.cpl[0][c + s.sx] * xp[0];
        off += A.cpl[1][c] * xm[1];
        if (j + 1 < d.ny) off += A.cpl[1][c + s.sy] * xp[1];
        off += A.cpl[2][c] * xm[2];
        if (k + 1 < d.nz) off += A.cpl[2][c + 1] * xp[2];
        const float r = b[c] - (dg * xc - off);
        float dprev = 0.f;
        if constexpr ((MODE & SM_D_FROM_X) != 0) dprev = xin[c];
        if constexpr ((MODE & SM_READ_D) != 0) dprev = dvec[c];
        const float dn = cd * dprev + cr * rcp_or_zero(dg) * r;
        xout[c] = xc + dn;
        if constexpr ((MODE & SM_WRITE_D) != 0) dvec[c] = dn;
    }
}

__global__ __launch_bounds__(kBlock) void k_restrict(Dims df, Op32 A, const float *__restrict__ b,
                                                     const float *__restrict__ x, Dims dc,
                                                     float *__restrict__ bc) {
    const int K = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    if (K >= dc.nz || J >= dc.ny) return;
    const Strides s(df);
    for (int I = blockIdx.z; I < dc.nx; I += gridDim.z) {
        float sum = 0.f;
        for (int di = 0; di < 2; di++) {
            const int i = 2 * I + di;
            if (i >= df.nx) break;
            for (int dj = 0; dj < 2; dj++) {
                const int j = 2 * J + dj;
                if (j >= df.ny) break;
                for (int dk = 0; dk < 2; dk++) {
                    const int k = 2 * K + dk;
                    if (k >= df.nz) break;
                    const size_t c = df.idx(i, j, k);
                    const float dg = A.diag[c];
                    if (!(dg > 0.f)) continue;
                    float off = 0.f;
                    if (i > 0) off += A.cpl[0][c] * x[c - s.sx];
                    if (i + 1 < df.nx) off += A.cpl[0][c + s.sx] * x[c + s.sx];
                    if (j > 0) off += A.cpl[1][c] * x[c - s.sy];
                    if (j + 1 < df.ny) off += A.cpl[1][c + s.sy] * x[c + s.sy];
                    if (k > 0) off += A.cpl[2][c] * x[c - 1];
                    if (k + 1 < df.nz) off += A.cpl[2][c + 1] * x[c + 1];
                    sum += b[c] - (dg * x[c] - off);
                }
            }
        }
        bc[dc.idx(I, J, K)] = sum;
    }
}

// Sum of the fine couplings along `axis` on the fine layer `layer` (the
// interface between two aggregates), over the 2x2 children in the other axes.
__device__ __forceinline__ double layer_coupling(Dims df, const double *cpl, int axis, int layer,
                                                 int I, int J, int K) {
    double sum = 0.0;
    int lo[3] = {2 * I, 2 * J, 2 * K};
    lo[axis] = layer;
    int hi[3] = {2 * I + 1, 2 * J + 1, 2 * K + 1};
    hi[axis] = layer;
    for (int i = lo[0]; i <= hi[0] && i < df.nx; i++)
        for (int j = lo[1]; j <= hi[1] && j < df.ny; j++)
            for (int k = lo[2]; k <= hi[2] && k < df.nz; k++) sum += cpl[df.idx(i, j, k)];
    return sum;
}

__global__ __launch_bounds__(kBlock) void k_coarsen(Dims df, Op64 Af, const float *diag32f,
                                                    Dims dc, double scale, double *diag64c,
                                                    MutFaces cpl64c, float *diag32c,
                                                    float *cpl32c0, float *cpl32c1,
                                                    float *cpl32c2) {
    const int K = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y * blockDim.y + threadIdx.y;
    if (K >= dc.nz || J >= dc.ny) return;
    const Strides s(df);
    float *cpl32c[3] = {cpl32c0, cpl32c1, cpl32c2};
    for (int I = blockIdx.z; I < dc.nx; I += gridDim.z) {
        bool active = false;
        double extra = 0.0;
        for (int i = 2 * I; i < 2 * I + 2 && i < df.nx; i++)
            for (int j = 2 * J; j < 2 * J + 2 && j < df.ny; j++)
                for (int k = 2 * K; k < 2 * K + 2 && k < df.nz; k++) {
                    const size_t c = df.idx(i, j, k);
                    if (!(diag32f[c] > 0.f)) continue;
                    active = true;
                    double off = Af.cpl[0][c] + Af.cpl[1][c] + Af.cpl[2][c];
                    if (i + 1 < df.nx) off += Af.cpl[0][c + s.sx];
                    if (j + 1 < df.ny) off += Af.cpl[1][c + s.sy];
                    if (k + 1 < df.nz) off += Af.cpl[2][c + 1];
                    extra += Af.diag[c] - off;
                }
        const int IC[3] = {I, J, K};
        double lower[3], upper[3];
        for (int a = 0; a < 3; a++) {
            lower[a] = IC[a] > 0 ? scale * layer_coupling(df, Af.cpl[a], a, 2 * IC[a], I, J, K) : 0.0;
            upper[a] = IC[a] + 1 < dc.n(a)
                           ? scale * layer_coupling(df, Af.cpl[a], a, 2 * IC[a] + 2, I, J, K)
                           : 0.0;
        }
        const size_t cc = dc.idx(I, J, K);
        if (active) {
            const double dg = scale * extra + lower[0] + upper[0] + lower[1] + upper[1] + lower[2] + upper[2];
            diag64c[cc] = dg;
            diag32c[cc] = dg > 0.0 ? (float)dg : 0.f;
        } else {
            diag64c[cc] = 1.0;
            diag32c[cc] = 0.f;
        }
        for (int a = 0; a < 3; a++) {
            const double v = active ? lower[a] : 0.0;
            cpl64c.a[a][cc] = v;
            cpl32c[a][cc] = (float)v;
        }
    }
}

// In-place Gauss-Jordan inversion of the dense coarsest matrix (SPD up to
// pure-Neumann blocks, whose zero pivots are dropped) in LDS, one block.
__global__ __launch_bounds__(kBlock) void k_coarse_invert(Dims d, Op64 A, float *ainv) {
    __shared__ double M[kMaxCoarse * kMaxCoarse];
    __shared__ double rowk[kMaxCoarse], colk[kMaxCoarse];
    __shared__ double scale_sh;
    const int n = (int)d.cells();
    const int tid = threadIdx.x;
    const Strides s(d);
    for (int e = tid; e < n * n; e += kBlock) M[e] = 0.0;
    __syncthreads();
    for (int c = tid; c < n; c += kBlock) {
        const int i = c / (d.ny * d.nz), j = (c / d.nz) % d.ny, k = c % d.nz;
        M[c * n + c] = A.diag[c];
        if (i > 0) M[c * n + (c - (int)s.sx)] = -A.cpl[0][c];
        if (i + 1 < d.nx) M[c * n + (c + (int)s.sx)] = -A.cpl[0][c + s.sx];
        if (j > 0) M[c * n + (c - (int)s.sy)] = -A.cpl[1][c];
        if (j + 1 < d.ny) M[c * n + (c + (int)s.sy)] = -A.cpl[1][c + s.sy];
        if (k > 0) M[c * n + (c - 1)] = -A.cpl[2][c];
        if (k + 1 < d.nz) M[c * n + (c + 1)] = -A.cpl[2][c + 1];
    }
    __syncthreads();
    if (tid == 0) {
        double m = 0.0;
        for (int c = 0; c < n; c++) m = fmax(m, fabs(M[c * n + c]));
        scale_sh = m;
    }
    __syncthreads();
    const double tiny = 1e-12 * scale_sh;
    for (int kk = 0; kk < n; kk++) {
        const double p = M[kk * n + kk];
        for (int t = tid; t < n; t += kBlock) {
            rowk[t] = M[kk * n + t];
            colk[t] = M[t * n + kk];
        }
        __syncthreads();
        if (fabs(p) <= tiny) {
            for (int t = tid; t < n; t += kBlock) {
                M[kk * n + t] = 0.0;
                M[t * n + kk] = 0.0;
            }
        } else {
            const double pinv = 1.0 / p;
            for (int e = tid; e < n * n; e += kBlock) {
                const int r = e / n, c = e % n;
                double v;
                if (r == kk && c == kk)
                    v = pinv;
                else if (r == kk)
                    v = rowk[c] * pinv;
                else if (c == kk)
                    v = -colk[r] * pinv;
                else
                    v = M[e] - colk[r] * rowk[c] * pinv;
                M[e] = v;
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < n * n; e += kBlock) ainv[e] = (float)M[e];
}

__global__ __launch_bounds__(kBlock) void k_coarse_solve(int n, const float *ainv, const float *b,
                                                         float *x) {
    __shared__ float bs[kMaxCoarse];
    const int tid = threadIdx.x;
    for (int t = tid; t < n; t += kBlock) bs[t] = b[t];
    __syncthreads();
    for (int r = tid; r < n; r += kBlock) {
        float acc = 0.f;
        for (int c = 0; c < n; c++) acc += ainv[r * n + c] * bs[c];
        x[r] = acc;
    }
}

}  // namespace

void launch_smooth_first(hipStream_t st, Dims d, Op32 A, const double *r64, float *b, float *x,
                         float c0) {
    Launch L = cell_launch(d);
    if (r64)
        hipLaunchKernelGGL(k_smooth_first<true>, L.grid, L.block, 0, st, d, A, r64, b, x, c0);
    else
        hipLaunchKernelGGL(k_smooth_first<false>, L.grid, L.block, 0, st, d, A, r64, b, x, c0);
    SFP_LAUNCH_CHECK();
}

void launch_smooth(hipStream_t st, int mode, Dims d, Op32 A, const float *b, const float *xin,
                   float *xout, float *dvec, float cd, float cr, const float *ecoarse) {
    Launch L = cell_launch(d);
    Dims dc = d.coarse();
#define SFP_SMOOTH_CASE(M)                                                                  \
    case M:                                                                                 \
        hipLaunchKernelGGL(k_smooth<M>, L.grid, L.block, 0, st, d, A, b, xin, xout, dvec, cd, \
                           cr, ecoarse, dc);                                                \
        break;
    switch (mode) {
        SFP_SMOOTH_CASE(SM_D_FROM_X)
        SFP_SMOOTH_CASE(SM_D_FROM_X | SM_WRITE_D)
        SFP_SMOOTH_CASE(SM_READ_D)
        SFP_SMOOTH_CASE(SM_READ_D | SM_WRITE_D)
        SFP_SMOOTH_CASE(SM_PROLONG)
        SFP_SMOOTH_CASE(SM_PROLONG | SM_WRITE_D)
        SFP_SMOOTH_CASE(0)
        SFP_SMOOTH_CASE(SM_WRITE_D)
        default:
            throw ArgError("unsupported smoother mode");
    }
#undef SFP_SMOOTH_CASE
    SFP_LAUNCH_CHECK();
}

void launch_restrict(hipStream_t st, Dims df, Op32 A, const float *b, const float *x, float *bc) {
    Dims dc = df.coarse();
    Launch L = cell_launch(dc);
    hipLaunchKernelGGL(k_restrict, L.grid, L.block, 0, st, df, A, b, x, dc, bc);
    SFP_LAUNCH_CHECK();
}

void launch_coarsen(hipStream_t st, Dims df, Op64 Af, const float *diag32f, Dims dc,
                    double scale, double *diag64c, double *const cpl64c[3], float *diag32c,
                    float *const cpl32c[3]) {
    Launch L = cell_launch(dc);
    MutFaces m{{cpl64c[0], cpl64c[1], cpl64c[2]}};
    hipLaunchKernelGGL(k_coarsen, L.grid, L.block, 0, st, df, Af, diag32f, dc, scale, diag64c, m,
                       diag32c, cpl32c[0], cpl32c[1], cpl32c[2]);
    SFP_LAUNCH_CHECK();
}

void launch_coarse_invert(hipStream_t st, Dims d, Op64 A, float *ainv) {
    if ((int)d.cells() > kMaxCoarse) throw ArgError("coarsest level too large for the dense solve");
    hipLaunchKernelGGL(k_coarse_invert, dim3(1), dim3(kBlock), 0, st, d, A, ainv);
    SFP_LAUNCH_CHECK();
}

void launch_coarse_solve(hipStream_t st, int n, const float *ainv, const float *b, float *x) {
    if (n > kMaxCoarse) throw ArgError("coarsest level too large for the dense solve");
    hipLaunchKernelGGL(k_coarse_solve, dim3(1), dim3(kBlock), 0, st, n, ainv, b, x);
    SFP_LAUNCH_CHECK();
}

}  // namespace sfp
