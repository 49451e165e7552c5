/*
 * sfp_test.h -- inspection hooks used by the parity tests. They expose the
 * assembled operators, the multigrid hierarchy and single kernels so each
 * can be checked against the oracle on its own. No counterpart in the
 * reference; not needed by a drop-in caller.
 */
#ifndef SFP_TEST_H
#define SFP_TEST_H
#include "sfp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* which: 0 = pressure system (last project), 1 = magnetic system (last solve) */
sfp_status sfp_test_num_levels(sfp_context ctx, int which, int *levels);
sfp_status sfp_test_level_dims(sfp_context ctx, int which, int level, int dims[3]);
/* FP64 operator of a level: diag and the 3 lower-face couplings (cells of that
 * level each); rhs (fine level only, may be NULL). Host pointers. */
sfp_status sfp_test_export_level(sfp_context ctx, int which, int level, double *diag,
                                 double *cpl_x, double *cpl_y, double *cpl_z, double *rhs);
/* FP32 operator of a level (diag is 0 on inactive cells). */
sfp_status sfp_test_export_level32(sfp_context ctx, int which, int level, float *diag,
                                   float *cpl_x, float *cpl_y, float *cpl_z);
/* w = A z with the fine FP64 operator (z given in FP32, as the solver feeds it). */
sfp_status sfp_test_apply64(sfp_context ctx, int which, const float *z, double *w);
/* z = B r, one application of the configured preconditioner. */
sfp_status sfp_test_precondition(sfp_context ctx, int which, const double *r, float *z);
/* r = rhs - A x with the fine FP64 operator, plus ||r||^2 and ||rhs||^2. */
sfp_status sfp_test_residual(sfp_context ctx, int which, const double *x, double *r,
                             double *rr, double *bb);
/* Solve A_l x = rhs with level `level`'s FP64 operator (0 = the assembled
 * system) by flexible CG with the V-cycle from that level down and the current
 * parameters, from x = 0 (host arrays of that level's cells). The system's
 * own right-hand side and solution (the warm start) are restored afterwards.
 * stats may be NULL. */
sfp_status sfp_test_solve(sfp_context ctx, int which, int level, const double *rhs, double *x,
                          sfp_stats *stats);
/* rc = the residual one V-cycle leaves for right-hand side r, restricted to
 * the first coarse level: the learned correction's input (host arrays: r
 * fine, rc coarse). */
sfp_status sfp_test_restrict_smoothed(sfp_context ctx, int which, const double *r, float *rc);
/* The last two post-smoothing steps on level `level` from iterate x1 and
 * direction d1 with right-hand side b (host FP32 arrays of that level's
 * cells; zero on its inactive cells), with Chebyshev coefficients (cdA, crA)
 * then (cdB, crB): x3_pair from the one-pass kernel, x3_steps from two single
 * steps. d1 NULL: the first two pre-smoothing passes from a stored first
 * iterate x1, which is also the previous direction. Either output may be
 * NULL. */
sfp_status sfp_test_smooth_pair(sfp_context ctx, int which, int level, const float *b,
                                const float *x1, const float *d1, float cdA, float crA,
                                float cdB, float crB, float *x3_pair, float *x3_steps);
/* The last pre-smoothing step and the restriction on level `level` (which
 * must have a coarser level): x2 = x1 + cd d1 + cr D^{-1} (b - A x1) and
 * bc = R (b - A x2), from the one-pass kernel (x2_fused, bc_fused) and from
 * the separate step and restriction (x2_steps, bc_steps). Host FP32 arrays:
 * b, x1, d1, x2_* of the level's cells, bc_* of the next level's. Any output
 * may be NULL. */
sfp_status sfp_test_smooth_restrict(sfp_context ctx, int which, int level, const float *b,
                                    const float *x1, const float *d1, float cd, float cr,
                                    float *x2_fused, float *bc_fused, float *x2_steps,
                                    float *bc_steps);
/* Run the V-cycle's small levels (at most 4096 cells) in one launch (enable
 * != 0, the default) or level by level. Results are the same bit for bit. */
sfp_status sfp_test_set_vcycle_tail(sfp_context ctx, int enable);
/* Run as one pass (mask bit 0) the last pre-smoothing step with the
 * restriction and (bit 1) the last two post-smoothing steps on large grids,
 * (bits 2 and 3) the same on small grids, (bit 4) the first two
 * pre-smoothing passes of a degree-3 cycle on large grids, or each as
 * separate launches. Results are the same bit for bit. */
sfp_status sfp_test_set_fused_pairs(sfp_context ctx, int mask); /* default 18 */
/* Let the stencil kernels of the magnetic system (FCG matvec, smoothing
 * steps, restriction) rebuild the FP64 and FP32 diagonals from the couplings
 * (each row's diagonal is their sum, bit for bit, away from the pinned cell)
 * instead of reading them, or read them. Results are the same bit for bit.
 * Takes effect at the next magnetic solve. */
sfp_status sfp_test_set_rowsum_diagonal(sfp_context ctx, int enable); /* default 1 */

#ifdef __cplusplus
}
#endif
#endif /* SFP_TEST_H */
