// Host-side launchers of the gfx950 kernels (defined in the .hip files).
#pragma once
#include "sfp_internal.hpp"

namespace sfp {

// Iteration kernels take an optional `gate` (&KrylovScalars::stop): when it
// points at a nonzero value the kernel returns at once (batched iterations
// after convergence). Null = always run.

// ---------------------------------------------------------------- multigrid (mg.hip)
// Chebyshev step modes.
enum SmoothMode : int {
    SM_FIRST = 1,      // input iterate is the zero-start first step c0 D^{-1} b, fused in
    SM_READ_D = 2,     // previous direction read from dvec
    SM_WRITE_D = 4,    // store the new direction
    SM_PROLONG = 8,    // input iterate is x + P e (coarse correction fused in)
    SM_XFIRST = 16     // input iterate is the zero-start first step c0 D^{-1} b, stored
                       // in xin (by the FCG kernels on the fine level, the restriction on
                       // the coarse ones); the previous direction equals it
};

// x = c0 * D^{-1} b from a zero start; at the fine level b is first rounded
// from the FP64 residual r64 (and stored to b).
void launch_smooth_first(hipStream_t s, Dims d, Op32 A, const double *r64, float *b, float *x,
                         float c0, const double *gate = nullptr);
// xout = x + d', d' = cd * d + cr * D^{-1} (b - A x) with x = xin, xin + P e,
// or the fused zero-start step c0 D^{-1} b (SM_FIRST; b from r64 when given).
void launch_smooth(hipStream_t s, int mode, Dims d, Op32 A, const double *r64, float *b,
                   const float *xin, float *xout, float *dvec, float c0, float cd, float cr,
                   const float *ecoarse, const double *gate = nullptr);
// Two steps in one pass: x2 = x1 + d2, d2 = cdA d1 + crA D^{-1} (b - A x1), then
// x3 = x2 + cdB d2 + crB D^{-1} (b - A x2); d2 is not stored. x3 must differ
// from x1. Same results as the two launch_smooth calls, bit for bit.
// With x1 == d1 (the stored zero-start iterate, which is also the previous
// direction) the pair is the first two steps of a cycle.
// pair_pays: grids on which it beats the two launches on MI355X (its 64 x 16
// tiles overlap by one cell per side; measured equal at 128^3, faster above).
inline bool pair_pays(Dims d) { return d.nz >= 64 && d.cells() > ((size_t)1 << 21); }
// Grids on which launch_smooth runs its steps (all but the fused first step)
// over the flat cell index: nz a multiple of 4, fewer than 2^32 cells.
bool flat_step(Dims d);
// Grids with z rows of 193..256 cells in groups of four run it over whole
// rows (k_smooth_pair_rows); rowact (null: all active) marks the z rows
// holding an active cell (launch_row_activity), whose loads the others skip.
void launch_smooth_pair(hipStream_t s, Dims d, Op32 A, const float *b, const float *x1, float *x3,
                        const float *d1, float cdA, float crA, float cdB, float crB,
                        const double *gate = nullptr, const unsigned char *rowact = nullptr);
bool row_pair(Dims d);
// rowact[i ny + j] = 1 when the z row (i, j) holds an active cell (diag32 > 0)
void launch_row_activity(hipStream_t s, Dims d, const float *diag32, unsigned char *rowact);
// The last pre-smoothing step and the restriction in one pass:
// x2 = x1 + cd d1 + cr D^{-1} (b - A x1), then bc = R (b - A x2) (the coarse
// right-hand side, as launch_restrict). Same results as launch_smooth
// (SM_READ_D) then launch_restrict, bit for bit; x2 must differ from x1.
void launch_smooth_restrict(hipStream_t s, Dims d, Op32 A, const float *b, const float *x1,
                            float *x2, const float *d1, float cd, float cr, float *bc,
                            const double *gate = nullptr);
// x += P e on active cells (diag32 > 0), in place.
void launch_prolong_add(hipStream_t s, Dims d, const float *diag32, const float *ecoarse, float *x,
                        const double *gate = nullptr);
// bc = R (b - A x): sum of the residuals of the (up to 8) children. With
// x0c, also x0c = c0 Dc^-1 bc (diag32c: the coarse FP32 diagonal), the first
// Chebyshev iterate of the coarse level's cycle, as the FCG update stores it
// on the fine level.
void launch_restrict(hipStream_t s, Dims df, Op32 A, const float *b, const float *x, float *bc,
                     const double *gate = nullptr, float *x0c = nullptr,
                     const float *diag32c = nullptr, float c0 = 0.f);
// Galerkin coarsening by 2x2x2 aggregation, scaled: writes the coarse FP64
// and FP32 operators, the couplings scaled by `scale` and each row's sum
// (diagonal minus couplings) by `rs_scale` (scale: scale * P^T A P).
void launch_coarsen(hipStream_t s, Dims df, Op64 Af, const float *diag32f, Dims dc,
                    double scale, double rs_scale, double *diag64c, double *const cpl64c[3],
                    float *diag32c,
                    float *const cpl32c[3]);
// Dense inverse of the coarsest operator (n = cells <= kMaxCoarse) into ainv (n*n floats).
constexpr int kMaxCoarse = 64;

// The coarse end of a V-cycle in one launch of one block: from level `first`
// (cells <= kTailMaxCells) down to the coarsest and back, with the per-cell
// arithmetic of the separate smoothing, restriction and dense-solve launches.
// The level's result is left in lv[0].x1 (tail_result); the other level
// vectors are not touched. The first level's b must already hold its
// right-hand side.
constexpr size_t kTailMaxCells = 4096;
constexpr int kTailMaxLevels = 8;
struct TailLevel {
    Dims d;
    Op32 A;
    float *b, *x0, *x1, *dv;
};
struct TailArgs {
    int nlev = 0;  // levels in the tail, the last the coarsest
    TailLevel lv[kTailMaxLevels];
    // Chebyshev degree and coefficients: [0] for the first level, [1] below it
    int deg[2] = {3, 3};
    float cd[2][8] = {}, cr[2][8] = {};
    const float *ainv = nullptr;
};
// The tail keeps 8 arrays of every level it covers in LDS (gfx950: up to
// 160 KiB per block).
constexpr size_t kTailLdsBytes = 149504;
inline size_t tail_lds_bytes(const TailArgs &t) {
    size_t b = 0;
    for (int l = 0; l < t.nlev; l++) b += 8 * sizeof(float) * t.lv[l].d.cells();
    return b;
}
void launch_vcycle_tail(hipStream_t s, const TailArgs &t, const double *gate = nullptr);
// The buffer that holds the tail's result for its first level.
inline float *tail_result(const TailArgs &t) { return t.lv[0].x1; }
void launch_coarse_invert(hipStream_t s, Dims d, Op64 A, float *ainv);
// On every level l < nlev: v[l][0..3][c] = 0 (and on level 0 w[c] = 0 when
// w is given) at the cells with diag32[l][c] == 0. start[l] is the first
// cell of level l in the launch (start[l + 1] = start[l] + n[l]).
constexpr int kClearMaxLevels = 16;
struct ClearArgs {
    int nlev = 0;
    unsigned start[kClearMaxLevels + 1] = {};
    unsigned n[kClearMaxLevels] = {};
    const float *diag32[kClearMaxLevels] = {};
    float *v[kClearMaxLevels][4] = {};
    double *w = nullptr;
};
void launch_clear_inactive(hipStream_t s, const ClearArgs &a);
void launch_coarse_solve(hipStream_t s, int n, const float *ainv, const float *b, float *x,
                         const double *gate = nullptr);

// ---------------------------------------------------------------- Krylov (krylov.hip)
// Reduction kernels cover this many x planes per block (fewer partials).
constexpr int kReducePlanes = 0;  // automatic
// Doubles of partial-sum scratch a reduction over grid d needs.
size_t reduction_partials(Dims d);
// Before a batch of gated iterations: scalars->stop = 0, ->iters = 0,
// ->stop_rr = stop_rr, ->stall_rr = stall_rr (0: no stall test),
// ->stalled = 0.
void launch_batch_begin(hipStream_t s, KrylovScalars *sc, double stop_rr, double stall_rr = 0.0);
// r = b - A x (x may be null = 0); scalars->rr = r.r, scalars->bb = b.b;
// r may be null (only the sums).
// With b32, also b32 = r rounded to FP32 and, with x0, x0 = c0 D32^-1 b32
// (diag32: the FP32 diagonal): the right-hand side and first Chebyshev
// iterate of a V-cycle on r, as the FCG update stores them.
void launch_residual64(hipStream_t s, Dims d, Op64 A, const double *b, const double *x,
                       double *r, double *partials, KrylovScalars *sc, float *b32 = nullptr,
                       float *x0 = nullptr, const float *diag32 = nullptr, float c0 = 0.f);
// w = A z; dots z.w, z.q_prev, z.r; the last block forms beta, p.q and alpha.
// active: the FP32 diagonal (0 = inactive cell) to skip all-inactive
// wavefronts, where z must be 0; null = every cell (any z).
void launch_fcg_matvec(hipStream_t s, Dims d, Op64 A, const float *active, const float *z,
                       const double *r, const double *qprev, double *w, int first,
                       double *partials, KrylovScalars *sc, const double *gate = nullptr,
                       hipEvent_t kernel_done = nullptr);
// The update of x in an FCG iteration: now (x += alpha p), later (left for
// the next iteration or launch_fcg_flush), or both the previous iteration's
// and this one's (x += alpha_prev p_prev + alpha p), so x is read and written
// every other iteration.
enum XUpdate : int { X_NOW = 0, X_LATER = 1, X_BOTH = 2 };
// p = z - beta p; q = w - beta q; x per xmode; r -= alpha q; rr = r.r
// (all-inactive wavefronts skipped when `active` is given); b32 (or null)
// receives r rounded to FP32 and x0 (or null) the first Chebyshev iterate
// c0 D^{-1} b32, with D the FP32 diagonal `active`.
void launch_fcg_update(hipStream_t s, Dims d, const float *active, const float *z,
                       const double *w, double *p, double *q, double *x, double *r, float *b32,
                       float *x0, float c0, int first, int xmode, double *partials,
                       KrylovScalars *sc, const double *gate = nullptr,
                       hipEvent_t kernel_done = nullptr);
// (kernel_done, if given, is recorded after the stencil kernel and before
// the one-block finalize that forms the scalars.)
// x += alpha p with the last iteration's alpha and direction (after an
// X_LATER iteration), over the update's groups of cells.
void launch_fcg_flush(hipStream_t s, Dims d, const float *active, const double *p, double *x,
                      const KrylovScalars *sc);
// z = D^{-1} r on active cells (Jacobi preconditioner).
void launch_jacobi(hipStream_t s, Dims d, Op32 A, const double *r, float *z,
                   const double *gate = nullptr);

// ---------------------------------------------------------------- assembly (assemble.hip)
struct Faces {
    const double *a[3];
};
struct MutFaces {
    double *a[3];
};
// Post-projection steps (post.hip).
// valid[f] (face grid axis) = fraction < 1 and a liquid cell on either side.
void launch_velocity_valid(hipStream_t s, Dims d, int axis, const double *phi, const double *frac,
                           unsigned char *valid);
// `steps` sweeps of Extrapolation::Solve on one grid: valid (the mask) and
// scratch (same size) are both overwritten.
void launch_extrapolate(hipStream_t s, Dims g, double *data, unsigned char *valid,
                        unsigned char *scratch, double clear, int steps);
// Collider::Enforce; old is the velocity before the call (not aliasing vel).
void launch_collider_enforce(hipStream_t s, Dims d, double spacing, const double origin[3], Faces old,
                             Faces frac, Faces normal, Faces colvel, MutFaces vel);
// Geometry of MagneticField's own StaggeredGrid (core/MagneticField.cpp:11-16).
struct MagGeom {
    double face_origin[3][3];
    double ls_origin[3];  // cell grid of the level set
    double spacing, inv;
};
MagGeom mag_geometry(Dims d, double spacing, const double origin[3]);

// Pressure::SetUnKnowns + BuildProjectionMatrix; liquid[c] = phi <= 0;
// *err set to 1 when a liquid cell has an open face on the grid boundary.
void launch_pressure_assemble(hipStream_t s, Dims d, const double *phi, Faces frac, Faces vel,
                              Faces us, const Faces *jump, double vol_error, double *diag64,
                              double *const cpl64[3], double *rhs, float *diag32,
                              float *const cpl32[3], unsigned char *liquid, int *err);
// Pressure::SetRightHandSide on the stored unknowns.
void launch_pressure_rhs(hipStream_t s, Dims d, const unsigned char *liquid, Faces frac,
                         Faces vel, Faces us, double vol_error, double *rhs);
// Pressure::ApplyProjection (p on the full grid, 0 at air cells).
void launch_pressure_apply(hipStream_t s, Dims d, const double *phi, Faces frac,
                           const Faces *jump, const unsigned char *liquid, const double *p,
                           MutFaces vel, int *err);
// MagneticField::BuildMatrix coefficients, matrix and AssembleRHS.
void launch_magnetic_mu(hipStream_t s, Dims d, const MagGeom &g, const double *phi, double chi,
                        double band, MutFaces mu);
// dirichlet: the opt-in box condition phi = 0 (no pinned cell) instead of
// the reference's Neumann box faces with cell 0 pinned.
void launch_magnetic_assemble(hipStream_t s, Dims d, Faces mu, Faces hext, double spacing,
                              double *diag64, double *const cpl64[3], double *rhs,
                              float *diag32, float *const cpl32[3], int dirichlet = 0);
// MagneticField::AssignTotalField.
void launch_magnetic_total_field(hipStream_t s, Dims d, double inv, const double *pot,
                                 Faces hext, MutFaces htot);
// MagneticField::CalculateVertexPressures.
void launch_vertex_pressures(hipStream_t s, Dims d, const MagGeom &g, Faces htot, double chi,
                             int nverts, const double *pos, const double *nrm, double *out);

}  // namespace sfp
