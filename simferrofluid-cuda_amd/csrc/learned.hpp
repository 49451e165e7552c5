// Launchers connecting the Krylov solver and the U-Net (learned.hip).
#pragma once
#include "sfp_internal.hpp"

#include <cstdint>

namespace sfp {

struct KrylovScalars;
// network input channels 1..3 into x8 ([N][8] BF16, others zeroed) and/or f3 ([N][3] FP32)
void launch_learned_features(hipStream_t st, Dims d, Op64 A, const float *diag32, const double *phi,
                             double spacing, uint16_t *x8, float *f3);
// x8[c][0] = r / RMS(r over active cells); the FP32 variant returns at once
// while *gate is nonzero (a stopped batch of FCG iterations)
void launch_pack_residual(hipStream_t st, Dims d, const double *r, const float *diag32,
                          double *partials, KrylovScalars *sc, uint16_t *x8);
void launch_pack_residual32(hipStream_t st, Dims d, const float *r, const float *diag32,
                            double *partials, KrylovScalars *sc, uint16_t *x8,
                            const double *gate = nullptr);
// out (coarse dc) = mean of phi over each cell's children on the fine grid df
void launch_restrict_phi(hipStream_t st, Dims df, const double *phi, Dims dc, double *out);
// e = c z on active cells with c = (z.b)/(z.Az), the A-norm optimal scale
// (sc->pad[2]); z is read as 0 on inactive cells; gated like the above
void launch_scaled_correction(hipStream_t st, Dims d, Op32 A, const float *z, const float *b,
                              double *partials, KrylovScalars *sc, float *e,
                              const double *gate = nullptr);
// z = zin on active cells, 0 elsewhere
void launch_mask_output(hipStream_t st, Dims d, const float *zin, const float *diag32, float *z);

}  // namespace sfp
