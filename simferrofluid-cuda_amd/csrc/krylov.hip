// FP64 kernels of the outer flexible CG (Notay's FCG with one stored
// direction): residual, matvec fused with its inner products, and the fused
// vector update. Inner products use the deterministic grid reduction of
// device_util.hpp; the scalar recurrences run in the last block so the host
// only reads ||r||^2.
#include "device_util.hpp"
#include "kernels.hpp"
#include "solver.hpp"

namespace sfp {
namespace {

__global__ __launch_bounds__(kBlock) void k_residual64(Dims d, Op64 A, const double *__restrict__ b,
                                                       const double *__restrict__ x,
                                                       double *__restrict__ r, double *partials,
                                                       unsigned *counter, KrylovScalars *sc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const bool in = k < d.nz && j < d.ny;
    const Strides s(d);
    double v[2] = {0.0, 0.0};
    if (in) {
        for (int i = blockIdx.z; i < d.nx; i += gridDim.z) {
            const size_t c = d.idx(i, j, k);
            const double bc = b[c];
            double rc = bc;
            if (x) {
                const double off =
                    offdiag_sum<double>(A.cpl, d, s, i, j, k, c, [&](size_t q) { return x[q]; });
                rc = bc - (A.diag[c] * x[c] - off);
            }
            r[c] = rc;
            v[0] += rc * rc;
            v[1] += bc * bc;
        }
    }
    if (grid_sum<2>(v, partials, counter) && flat_tid() == 0) {
        sc->rr = v[0];
        sc->bb = v[1];
    }
}

__global__ __launch_bounds__(kBlock) void k_fcg_matvec(Dims d, Op64 A, const float *__restrict__ z,
                                                       const double *__restrict__ r,
                                                       const double *__restrict__ qprev,
                                                       double *__restrict__ w, int first,
                                                       double *partials, unsigned *counter,
                                                       KrylovScalars *sc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const bool in = k < d.nz && j < d.ny;
    const Strides s(d);
    double v[3] = {0.0, 0.0, 0.0};  // z.w, z.q_prev, z.r
    if (in) {
        for (int i = blockIdx.z; i < d.nx; i += gridDim.z) {
            const size_t c = d.idx(i, j, k);
            const double zc = (double)z[c];
            const double off =
                offdiag_sum<double>(A.cpl, d, s, i, j, k, c, [&](size_t q) { return (double)z[q]; });
            const double wc = A.diag[c] * zc - off;
            w[c] = wc;
            v[0] += zc * wc;
            if (!first) v[1] += zc * qprev[c];
            v[2] += zc * r[c];
        }
    }
    if (grid_sum<3>(v, partials, counter) && flat_tid() == 0) {
        const double zw = v[0], zq = v[1], zr = v[2];
        double beta = 0.0, pq = zw, pr = zr;
        if (!first && sc->pq_prev != 0.0) {
            // p = z - beta p_prev with beta = z.q_prev / p_prev.q_prev makes p
            // A-orthogonal to p_prev; p.q and p.r follow from the dots above
            // (p_prev.r = 0 after the previous step, A symmetric).
            beta = zq / sc->pq_prev;
            pq = zw - beta * zq;
        }
        sc->zw = zw;
        sc->zq = zq;
        sc->zr = zr;
        sc->beta = beta;
        sc->pq = pq;
        sc->alpha = pq != 0.0 ? pr / pq : 0.0;
    }
}

__global__ __launch_bounds__(kBlock) void k_fcg_update(Dims d, const float *__restrict__ z,
                                                       const double *__restrict__ w,
                                                       double *__restrict__ p,
                                                       double *__restrict__ q,
                                                       double *__restrict__ x,
                                                       double *__restrict__ r, int first,
                                                       double *partials, unsigned *counter,
                                                       KrylovScalars *sc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const bool in = k < d.nz && j < d.ny;
    const double alpha = sc->alpha, beta = sc->beta;
    double v[1] = {0.0};
    if (in) {
        for (int i = blockIdx.z; i < d.nx; i += gridDim.z) {
            const size_t c = d.idx(i, j, k);
            const double zc = (double)z[c];
            double pc, qc;
            if (first) {
                pc = zc;
                qc = w[c];
            } else {
                pc = zc - beta * p[c];
                qc = w[c] - beta * q[c];
            }
            p[c] = pc;
            q[c] = qc;
            x[c] += alpha * pc;
            const double rc = r[c] - alpha * qc;
            r[c] = rc;
            v[0] += rc * rc;
        }
    }
    if (grid_sum<1>(v, partials, counter) && flat_tid() == 0) {
        sc->rr = v[0];
        sc->pq_prev = sc->pq;
    }
}

__global__ __launch_bounds__(kBlock) void k_jacobi(Dims d, Op32 A, const double *__restrict__ r,
                                                   float *__restrict__ z) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (k >= d.nz || j >= d.ny) return;
    for (int i = blockIdx.z; i < d.nx; i += gridDim.z) {
        const size_t c = d.idx(i, j, k);
        const float dg = A.diag[c];
        z[c] = dg > 0.f ? (float)(r[c] / (double)dg) : 0.f;
    }
}

__global__ __launch_bounds__(kBlock) void k_count_active(Dims d, const float *__restrict__ diag32,
                                                         double *partials, unsigned *counter,
                                                         KrylovScalars *sc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    double v[1] = {0.0};
    if (k < d.nz && j < d.ny) {
        for (int i = blockIdx.z; i < d.nx; i += gridDim.z)
            if (diag32[d.idx(i, j, k)] > 0.f) v[0] += 1.0;
    }
    if (grid_sum<1>(v, partials, counter) && flat_tid() == 0) sc->pad[0] = v[0];
}

}  // namespace

void launch_count_active(hipStream_t st, Dims d, const float *diag32, double *partials,
                         unsigned *counter, KrylovScalars *sc) {
    Launch L = cell_launch(d);
    hipLaunchKernelGGL(k_count_active, L.grid, L.block, 0, st, d, diag32, partials, counter, sc);
    SFP_LAUNCH_CHECK();
}

void launch_residual64(hipStream_t st, Dims d, Op64 A, const double *b, const double *x,
                       double *r, double *partials, unsigned *counter, KrylovScalars *sc) {
    Launch L = cell_launch(d);
    hipLaunchKernelGGL(k_residual64, L.grid, L.block, 0, st, d, A, b, x, r, partials, counter, sc);
    SFP_LAUNCH_CHECK();
}

void launch_fcg_matvec(hipStream_t st, Dims d, Op64 A, const float *z, const double *r,
                       const double *qprev, double *w, int first, double *partials,
                       unsigned *counter, KrylovScalars *sc) {
    Launch L = cell_launch(d);
    hipLaunchKernelGGL(k_fcg_matvec, L.grid, L.block, 0, st, d, A, z, r, qprev, w, first,
                       partials, counter, sc);
    SFP_LAUNCH_CHECK();
}

void launch_fcg_update(hipStream_t st, Dims d, const float *z, const double *w, double *p,
                       double *q, double *x, double *r, int first, double *partials,
                       unsigned *counter, KrylovScalars *sc) {
    Launch L = cell_launch(d);
    hipLaunchKernelGGL(k_fcg_update, L.grid, L.block, 0, st, d, z, w, p, q, x, r, first,
                       partials, counter, sc);
    SFP_LAUNCH_CHECK();
}

void launch_jacobi(hipStream_t st, Dims d, Op32 A, const double *r, float *z) {
    Launch L = cell_launch(d);
    hipLaunchKernelGGL(k_jacobi, L.grid, L.block, 0, st, d, A, r, z);
    SFP_LAUNCH_CHECK();
}

}  // namespace sfp
