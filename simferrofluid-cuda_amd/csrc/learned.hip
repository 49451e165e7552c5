// Glue between the FP64 Krylov solver and the U-Net: operator features,
// residual scaling and the masked network output.
#include "device_util.hpp"
#include "kernels.hpp"
#include "learned.hpp"

namespace sfp {
namespace {

// true inside a batch of FCG iterations that has stopped (KrylovScalars::stop)
__device__ __forceinline__ bool gated(const double *gate) { return gate && *gate != 0.0; }

__device__ __forceinline__ uint16_t to_bf16(float f) {
    const unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// Channels 1..3 of the network input (and zeros in 0, 4..7):
// log(diag/6), (diag - sum of the six couplings)/diag, clamp(phi/dx,-4,4)/4.
__global__ __launch_bounds__(kBlock) void k_features(CellGrid cg, Dims d, Op64 A, const float *diag32,
                                                     const double *phi, double inv_dx,
                                                     uint16_t *__restrict__ x8,
                                                     float *__restrict__ f3) {
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    if (k >= d.nz || j >= d.ny) return;
    const Strides s(d);
    for (int i = i0, ie = min(i0 + cg.planes, d.nx); i < ie; i++) {
        const size_t c = d.idx(i, j, k);
        float f[3] = {0.f, 0.f, 0.f};
        if (diag32[c] > 0.f) {
            const double dg = A.diag[c];
            double off = A.cpl[0][c] + A.cpl[1][c] + A.cpl[2][c];
            if (i + 1 < d.nx) off += A.cpl[0][c + s.sx];
            if (j + 1 < d.ny) off += A.cpl[1][c + s.sy];
            if (k + 1 < d.nz) off += A.cpl[2][c + 1];
            f[0] = (float)log(dg / 6.0);
            f[1] = (float)((dg - off) / dg);
            double p = phi ? phi[c] * inv_dx : 0.0;
            p = p < -4.0 ? -4.0 : (p > 4.0 ? 4.0 : p);
            f[2] = (float)(p * 0.25);
        }
        if (x8) {
            uint16_t *o = x8 + c * 8;
            o[0] = 0;
            o[1] = to_bf16(f[0]);
            o[2] = to_bf16(f[1]);
            o[3] = to_bf16(f[2]);
            o[4] = o[5] = o[6] = o[7] = 0;
        }
        if (f3) {
            f3[c * 3] = f[0];
            f3[c * 3 + 1] = f[1];
            f3[c * 3 + 2] = f[2];
        }
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_rms_partials(CellGrid cg, Dims d, const T *r,
                                                         const float *diag32, double *partials,
                                                         const double *gate) {
    if (gated(gate)) return;
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    double v[2] = {0.0, 0.0};
    if (k < d.nz && j < d.ny) {
        for (int i = i0, ie = min(i0 + cg.planes, d.nx); i < ie; i++) {
            const size_t c = d.idx(i, j, k);
            if (diag32[c] > 0.f) {
                const double rv = (double)r[c];
                v[0] += rv * rv;
                v[1] += 1.0;
            }
        }
    }
    block_partials<2>(v, partials);
}

__global__ __launch_bounds__(kBlock) void k_rms_finalize(unsigned nb, const double *partials,
                                                         KrylovScalars *sc, const double *gate) {
    if (gated(gate)) return;
    double v[2];
    sum_partials<2>(partials, nb, v);
    if (flat_tid() == 0) sc->pad[1] = (v[0] > 0.0 && v[1] > 0.0) ? 1.0 / sqrt(v[0] / v[1]) : 0.0;
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_pack_residual(CellGrid cg, Dims d, const T *r,
                                                          const float *diag32, const KrylovScalars *sc,
                                                          uint16_t *__restrict__ x8, const double *gate) {
    if (gated(gate)) return;
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    if (k >= d.nz || j >= d.ny) return;
    const double s = sc->pad[1];
    for (int i = i0, ie = min(i0 + cg.planes, d.nx); i < ie; i++) {
        const size_t c = d.idx(i, j, k);
        x8[c * 8] = diag32[c] > 0.f ? to_bf16((float)((double)r[c] * s)) : (uint16_t)0;
    }
}

// Mean of the fine level set over the (up to 8) children of each coarse cell.
__global__ __launch_bounds__(kBlock) void k_restrict_phi(CellGrid cg, Dims dc, Dims df, const double *phi,
                                                         double *__restrict__ out) {
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    if (k >= dc.nz || j >= dc.ny) return;
    for (int i = i0, ie = min(i0 + cg.planes, dc.nx); i < ie; i++) {
        double s = 0.0;
        int n = 0;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int e = 0; e < 2; e++) {
                    const int fi = 2 * i + a, fj = 2 * j + b, fk = 2 * k + e;
                    if (fi < df.nx && fj < df.ny && fk < df.nz) {
                        s += phi[df.idx(fi, fj, fk)];
                        n++;
                    }
                }
        out[dc.idx(i, j, k)] = s / n;
    }
}

// Partials of (z.b, z.Az) over the active cells, z taken as 0 on inactive ones.
__global__ __launch_bounds__(kBlock) void k_corr_dots(CellGrid cg, Dims d, Op32 A, const float *z,
                                                      const float *b, double *partials,
                                                      const double *gate) {
    if (gated(gate)) return;
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    double v[2] = {0.0, 0.0};
    if (k < d.nz && j < d.ny) {
        const Strides s(d);
        auto zv = [&](size_t c) { return A.diag[c] > 0.f ? z[c] : 0.f; };
        for (int i = i0, ie = min(i0 + cg.planes, d.nx); i < ie; i++) {
            const size_t c = d.idx(i, j, k);
            const float dg = A.diag[c];
            if (dg > 0.f) {
                const float zc = z[c];
                const float az = dg * zc - offdiag_sum<float>(A.cpl, d, s, i, j, k, c, zv);
                v[0] += (double)zc * b[c];
                v[1] += (double)zc * az;
            }
        }
    }
    block_partials<2>(v, partials);
}

__global__ __launch_bounds__(kBlock) void k_corr_finalize(unsigned nb, const double *partials,
                                                          KrylovScalars *sc, const double *gate) {
    if (gated(gate)) return;
    double v[2];
    sum_partials<2>(partials, nb, v);
    if (flat_tid() == 0) sc->pad[2] = v[1] > 0.0 ? v[0] / v[1] : 0.0;
}

__global__ __launch_bounds__(kBlock) void k_corr_scale(CellGrid cg, Dims d, const float *z,
                                                       const float *diag32, const KrylovScalars *sc,
                                                       float *__restrict__ e, const double *gate) {
    if (gated(gate)) return;
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    if (k >= d.nz || j >= d.ny) return;
    const float a = (float)sc->pad[2];
    for (int i = i0, ie = min(i0 + cg.planes, d.nx); i < ie; i++) {
        const size_t c = d.idx(i, j, k);
        e[c] = diag32[c] > 0.f ? a * z[c] : 0.f;
    }
}

__global__ __launch_bounds__(kBlock) void k_mask_output(CellGrid cg, Dims d, const float *zin,
                                                        const float *diag32, float *__restrict__ z) {
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    if (k >= d.nz || j >= d.ny) return;
    for (int i = i0, ie = min(i0 + cg.planes, d.nx); i < ie; i++) {
        const size_t c = d.idx(i, j, k);
        z[c] = diag32[c] > 0.f ? zin[c] : 0.f;
    }
}

}  // namespace

void launch_learned_features(hipStream_t st, Dims d, Op64 A, const float *diag32, const double *phi,
                             double spacing, uint16_t *x8, float *f3) {
    Launch L = cell_launch(d);
    hipLaunchKernelGGL(k_features, L.grid, L.block, 0, st, L.cg, d, A, diag32, phi, 1.0 / spacing, x8, f3);
    SFP_LAUNCH_CHECK();
}

template <class T>
void pack_residual(hipStream_t st, Dims d, const T *r, const float *diag32, double *partials,
                   KrylovScalars *sc, uint16_t *x8, const double *gate) {
    Launch L = cell_launch(d, kReducePlanes);
    hipLaunchKernelGGL(k_rms_partials<T>, L.grid, L.block, 0, st, L.cg, d, r, diag32, partials, gate);
    SFP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rms_finalize, dim3(1), dim3(kBlock), 0, st, L.blocks(), partials, sc, gate);
    SFP_LAUNCH_CHECK();
    Launch P = cell_launch(d);
    hipLaunchKernelGGL(k_pack_residual<T>, P.grid, P.block, 0, st, P.cg, d, r, diag32, sc, x8, gate);
    SFP_LAUNCH_CHECK();
}

void launch_pack_residual(hipStream_t st, Dims d, const double *r, const float *diag32,
                          double *partials, KrylovScalars *sc, uint16_t *x8) {
    pack_residual(st, d, r, diag32, partials, sc, x8, nullptr);
}

void launch_pack_residual32(hipStream_t st, Dims d, const float *r, const float *diag32,
                            double *partials, KrylovScalars *sc, uint16_t *x8, const double *gate) {
    pack_residual(st, d, r, diag32, partials, sc, x8, gate);
}

void launch_restrict_phi(hipStream_t st, Dims df, const double *phi, Dims dc, double *out) {
    Launch L = cell_launch(dc);
    hipLaunchKernelGGL(k_restrict_phi, L.grid, L.block, 0, st, L.cg, dc, df, phi, out);
    SFP_LAUNCH_CHECK();
}

void launch_scaled_correction(hipStream_t st, Dims d, Op32 A, const float *z, const float *b,
                              double *partials, KrylovScalars *sc, float *e, const double *gate) {
    Launch L = cell_launch(d, kReducePlanes);
    hipLaunchKernelGGL(k_corr_dots, L.grid, L.block, 0, st, L.cg, d, A, z, b, partials, gate);
    SFP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_corr_finalize, dim3(1), dim3(kBlock), 0, st, L.blocks(), partials, sc, gate);
    SFP_LAUNCH_CHECK();
    Launch P = cell_launch(d);
    hipLaunchKernelGGL(k_corr_scale, P.grid, P.block, 0, st, P.cg, d, z, A.diag, sc, e, gate);
    SFP_LAUNCH_CHECK();
}

void launch_mask_output(hipStream_t st, Dims d, const float *zin, const float *diag32, float *z) {
    Launch L = cell_launch(d);
    hipLaunchKernelGGL(k_mask_output, L.grid, L.block, 0, st, L.cg, d, zin, diag32, z);
    SFP_LAUNCH_CHECK();
}

}  // namespace sfp
