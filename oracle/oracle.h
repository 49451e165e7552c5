/*
 * FP64 CPU oracle for the two grid Poisson solves of SimFerrofluid.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the shipped library links, loads or
 * calls this code; only tests/, the smoke check and the CPU-baseline leg of
 * bench.py do, and only as the checker (or as the timed CPU baseline).
 *
 * Every function restates the reference algorithm in plain C++ with the
 * reference's evaluation order (no FMA contraction: built with
 * -ffp-contract=off), so assembled matrices and right-hand sides are
 * bit-identical to what core/Pressure.cpp and core/MagneticField.cpp produce.
 *
 * Layout: every array is z-fastest, index = z + nz*(y + ny*x)
 * (core/Grid.h:28). Face array a has extent (n + e_a) (core/StaggeredGrid.h:15-19).
 * Sparse matrices are CSR with int32 indices and column-sorted rows, which is
 * what Eigen::SparseMatrix<double,RowMajor>::setFromTriplets yields.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Planted mistakes for the oracle's closed-form tests: m selects one changed
 * expression (see the Mutation enum in assembly.cpp), 0 restores the
 * faithful restatement. Process-wide; the tests reset it after use. */
void orc_set_mutation(int m);

/* Pressure::SetUnKnowns (core/Pressure.cpp:104-123): liquid <=> phi <= 0,
 * rows numbered in linear-index order. grid2mat[cell] = row or -1.
 * Returns the number of unknowns. */
int orc_pressure_unknowns(int nx, int ny, int nz, const double *phi,
                          int *grid2mat, int *mat2grid);

/* Pressure::BuildProjectionMatrix (core/Pressure.cpp:38-78). jump may be NULL
 * (no pressure-jump callback); otherwise jump[a][face] is the value the
 * callback returns for that face. Returns nnz, or -1 when a liquid cell has an
 * open face (w > 0) on the grid boundary (the reference indexes outside the
 * grid there). */
int64_t orc_pressure_matrix(int nx, int ny, int nz, const double *phi,
                            const double *const frac[3], const double *const vel[3],
                            const double *const usolid[3], const double *const jump[3],
                            double vol_error, const int *grid2mat, const int *mat2grid, int n,
                            int *rowptr, int *col, double *val, double *rhs);

/* Pressure::SetRightHandSide (core/Pressure.cpp:80-102). */
void orc_pressure_rhs(int nx, int ny, int nz, const double *const frac[3],
                      const double *const vel[3], const double *const usolid[3],
                      double vol_error, const int *mat2grid, int n, double *rhs);

/* Pressure::ApplyProjection (core/Pressure.cpp:137-166). p is the compact
 * solution (one entry per unknown). Returns -1 if an open (w > 0) face lies on
 * the grid boundary, else 0. */
int orc_pressure_apply(int nx, int ny, int nz, const double *phi, const double *const frac[3],
                       const double *const jump[3], const int *grid2mat, const double *p,
                       double *const vel[3]);

/* MagneticField::BuildMatrix face coefficients (core/MagneticField.cpp:50-60,
 * core/MagneticField.h:44-54, core/TriLerp.h:22-61). origin/spacing are the
 * simulation StaggeredGrid's (the phi grid is its cell grid). */
void orc_magnetic_coeff(int nx, int ny, int nz, double spacing, const double origin[3],
                        const double *phi, double chi, double band, double *const mu[3]);

/* MagneticField::BuildMatrix matrix part (core/MagneticField.cpp:62-90).
 * bc = 0: the reference (box faces skipped, cell 0 pinned); bc = 1: the
 * library's opt-in Dirichlet box (phi = 0 on the box faces by a mirrored
 * ghost value: 2 mu_face on the diagonal; no pin), no reference counterpart. */
int64_t orc_magnetic_matrix(int nx, int ny, int nz, const double *const mu[3], int bc,
                            int *rowptr, int *col, double *val);

/* MagneticField::AssembleRHS (core/MagneticField.cpp:97-118); bc = 1 adds
 * the box faces' (mu - 1) H_ext fluxes and fills row 0. */
void orc_magnetic_rhs(int nx, int ny, int nz, double spacing, const double *const mu[3],
                      const double *const hext[3], int bc, double *rhs);

/* MagneticField::AssignTotalField (core/MagneticField.cpp:125-143). */
/* ||rhs - A x||_2^2 of the BuildMatrix system (core/MagneticField.cpp:62-90)
 * computed row by row without the CSR (grids whose matrix would not fit);
 * *bb receives ||rhs||_2^2. */
double orc_magnetic_residual(int nx, int ny, int nz, const double *const mu[3], int bc,
                             const double *rhs, const double *x, double *bb);

void orc_magnetic_total_field(int nx, int ny, int nz, double spacing, const double *potential,
                              const double *const hext[3], double *const htotal[3]);

/* MagneticField::CalculateVertexPressures (core/MagneticField.cpp:37-48). */
void orc_magnetic_vertex_pressures(int nx, int ny, int nz, double spacing, const double origin[3],
                                   const double *const htotal[3], double chi, int nverts,
                                   const double *positions, const double *normals,
                                   double *pressures);

/* ---- linear algebra ---- */
void orc_spmv(int n, const int *rowptr, const int *col, const double *val, const double *x,
              double *y);
/* ||b - A x||_2 / ||b||_2 (0 when b = 0 and x solves) */
double orc_rel_residual(int n, const int *rowptr, const int *col, const double *val,
                        const double *b, const double *x);

/* Krylov solvers. precond: 0 none, 1 Jacobi, 2 smoothed-aggregation AMG with
 * SPAI-0 relaxation (rebuilt on every call, as the reference's amgcl setup is).
 * x holds the initial guess on entry. Returns the iteration count; *rel gets
 * the final ||b-Ax||/||b||; *setup_s and *solve_s the wall times. */
int orc_cg(int n, const int *rowptr, const int *col, const double *val, const double *b,
           double *x, double tol, int maxit, int precond, double *rel, double *setup_s,
           double *solve_s);
int orc_bicgstab(int n, const int *rowptr, const int *col, const double *val, const double *b,
                 double *x, double tol, int maxit, int precond, double *rel, double *setup_s,
                 double *solve_s);

/* ---- U-Net of the learned preconditioner (unet.cpp) ----
 * Volumes are channels-last [batch][x][y][z][c]; input has 4 channels,
 * output 1; params as laid out in unet.cpp. Returns the masked MSE loss
 * against target (0 when target is NULL). */
long long orc_unet_param_count(void);
double orc_unet_forward(int batch, int nx, int ny, int nz, const double *params, const double *input,
                        const double *target, const double *mask, double *output);
/* Forward + backward of the masked MSE loss; grad has param_count entries. */
double orc_unet_gradients(int batch, int nx, int ny, int nz, const double *params, const double *input,
                          const double *target, const double *mask, double *grad);

/* ---- after the projection (core/Simulation.cpp:236-241) ----
 * Extrapolation mask of Simulation::ProjectVelocity: valid[a][face] = 1 when
 * fraction < 1 and a cell on either side has phi <= 0 (outside: air). */
void orc_velocity_valid(int nx, int ny, int nz, const double *phi, const double *const frac[3],
                        unsigned char *const valid[3]);
/* Extrapolation::Solve on the three face grids, in place on data. */
void orc_extrapolate_faces(int nx, int ny, int nz, double *const data[3],
                           const unsigned char *const valid[3], double clear, int steps);
/* Collider::Enforce, in place on vel; origin = StaggeredGrid::GetOrigin(). */
void orc_collider_enforce(int nx, int ny, int nz, double spacing, const double origin[3],
                          double *const vel[3], const double *const frac[3],
                          const double *const normal[3], const double *const colvel[3]);

/* Number of OpenMP threads the solvers use. */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
