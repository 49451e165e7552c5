/*
 * sfp.h -- C ABI of the MI355X (gfx950) grid Poisson solves that replace the
 * amgcl calls of SimFerrofluid's Pressure and MagneticField classes.
 *
 * Plain C types only: ints, doubles and pointers. Every 3-D array is in the
 * reference's z-fastest order, index = z + nz*(y + ny*x) (core/Grid.h:28);
 * face array a has extent (n + e_a) along its own axis
 * (core/StaggeredGrid.h:15-19). Host-pointer entry points copy to and from
 * the device; *_device entry points take device pointers and are ordered on
 * the context's HIP stream (sfp_stream).
 *
 * A context is bound to one grid (resolution, spacing, origin) and one GPU and
 * is not reentrant: calls on one context must not overlap, matching the
 * single-threaded use of Pressure/MagneticField in the reference
 * (demo/Driver.cpp:144).
 */
#ifndef SFP_H
#define SFP_H

#ifdef __cplusplus
extern "C" {
#endif

#define SFP_ABI_VERSION 4  /* 2: sfp_params gained the learned cycle's smoother (net_*);
                              3: magnetic_bc; 4: the pressure system's coarse smoother */

typedef enum sfp_status {
    SFP_OK = 0,
    SFP_ERR_ARGUMENT = 1,    /* null pointer, bad size or parameter */
    SFP_ERR_INPUT = 2,       /* an open (w > 0) face on the grid boundary next to liquid:
                                the reference indexes outside the grid there */
    SFP_ERR_HIP = 3,         /* HIP runtime error; see sfp_last_error */
    SFP_ERR_NO_DEVICE = 4,   /* no usable GPU */
    SFP_ERR_STATE = 5,       /* call out of order (e.g. reproject before project) */
    SFP_ERR_UNSUPPORTED = 6  /* e.g. learned preconditioner without a network */
} sfp_status;

typedef enum sfp_precond {
    SFP_PRECOND_MG = 0,      /* FP32 geometric-multigrid V-cycle inside FP64 flexible CG */
    SFP_PRECOND_LEARNED = 1, /* V-cycle plus a 3-D U-Net (BF16, MFMA) correction of the residual
                                it leaves, inside FP64 flexible CG, MG fallback */
    SFP_PRECOND_JACOBI = 2,  /* diagonal scaling (reference point for tests) */
    SFP_PRECOND_SMOOTH = 3   /* the learned cycle without its network (zero correction, so the
                                V-cycle plus extra post-smoothing): the ablation baseline */
} sfp_precond;

typedef struct sfp_params {
    double tol;          /* stop when ||b - A x||_2 / ||b||_2 <= tol (FP64) */
    int max_iters;       /* outer flexible-CG iteration cap */
    int precond;         /* sfp_precond */
    int mg_coarsest;     /* coarsen until every axis has <= this many cells (>= 2) */
    int cheby_degree;    /* Chebyshev smoother degree, pre and post (1..8) */
    double cheby_ratio;  /* smoothing interval [lambda_max/ratio, lambda_max] */
    double coarse_scale; /* coarse operator = scale * Galerkin(aggregation) */
    double stall_ratio;  /* learned preconditioner: fall back to MG when
                            ||r_k|| / ||r_{k-1}|| exceeds this */
    int verbose;         /* 1: print "{iters} iters" per solve like the reference */
    int coarse_cheby_degree;   /* Chebyshev degree on the levels below the finest (1..8) */
    double coarse_cheby_ratio; /* their smoothing interval ratio */
    /* The learned preconditioner's V-cycle smoother: the settings its network
       was trained under (recorded with the weights). 0 keeps the plain
       cycle's value above. After a fall-back the plain cycle runs with the
       settings above. */
    int net_cheby_degree;          /* 0 or 1..8 */
    double net_cheby_ratio;        /* 0 or > 1 */
    int net_coarse_cheby_degree;   /* 0 or 1..8 */
    double net_coarse_cheby_ratio; /* 0 or > 1 */
    int magnetic_bc;               /* sfp_magnetic_bc */
    /* The pressure system's Chebyshev smoother on its coarse levels (0: the
       coarse_cheby_* above, which the magnetic system uses). */
    int pressure_coarse_cheby_degree;   /* 0 or 1..8 (default 3) */
    double pressure_coarse_cheby_ratio; /* 0 or > 1 (default 4) */
} sfp_params;

/* Boundary condition of the magnetic potential on the box faces. */
typedef enum sfp_magnetic_bc {
    SFP_MAGNETIC_BC_REFERENCE = 0,    /* MagneticField::BuildMatrix's: box faces skipped (zero
                                         normal flux of grad phi), cell 0 pinned to 0 */
    SFP_MAGNETIC_BC_DIRICHLET_BOX = 1 /* phi = 0 on the box faces (mirrored ghost cells: a box
                                         face adds 2 mu to the diagonal and its (mu - 1) H_ext flux
                                         to the right-hand side); every cell an unknown. No
                                         counterpart in the reference: an opt-in variant */
} sfp_magnetic_bc;

typedef struct sfp_stats {
    int iterations;      /* outer iterations */
    int converged;       /* 1 when rel_residual <= tol */
    int fallback;        /* learned preconditioner: iterations done after falling back to MG */
    int precond_used;    /* sfp_precond the solve ended with */
    double rel_residual; /* final FP64 ||b - A x|| / ||b||, recomputed from x */
    double setup_ms;     /* assembly + hierarchy (GPU time) */
    double solve_ms;     /* Krylov iterations (GPU time) */
    double apply_ms;     /* velocity / field update (GPU time) */
    long long unknowns;  /* active rows of the system */
} sfp_stats;

typedef struct sfp_context_s *sfp_context;

/* ---- lifetime ---------------------------------------------------------- */

/* Create a context for an nx*ny*nz cell grid with the given spacing and
 * lower-corner origin (the simulation StaggeredGrid's GetOrigin()), on HIP
 * device `device`. Replaces the allocations of Pressure::Pressure
 * (core/Pressure.cpp:13-15) and MagneticField::MagneticField
 * (core/MagneticField.cpp:11-27). */
sfp_status sfp_create(sfp_context *out, int nx, int ny, int nz, double spacing,
                      const double origin[3], int device);
void sfp_destroy(sfp_context ctx);

void sfp_default_params(sfp_params *params);
sfp_status sfp_set_params(sfp_context ctx, const sfp_params *params);
sfp_status sfp_get_params(sfp_context ctx, sfp_params *params);

/* Text of the last error on this context (or of the last failed sfp_create
 * when ctx is NULL). */
const char *sfp_last_error(sfp_context ctx);
const char *sfp_status_string(sfp_status status);
/* The context's HIP stream, as a hipStream_t cast to void*. */
void *sfp_stream(sfp_context ctx);
/* Drop the magnetic solve's warm start (next solve starts from zero). */
sfp_status sfp_reset_warm_start(sfp_context ctx);

/* ---- pressure projection ---------------------------------------------- */

/* Pressure::Project (core/Pressure.cpp:17-26): unknowns = cells with
 * level_set <= 0 (SetUnKnowns :104-123), ghost-fluid matrix and divergence
 * right-hand side (BuildProjectionMatrix :38-78), solve from zero
 * (SolveLinearSystem :125-135), and the velocity update (ApplyProjection
 * :137-166), done in place on `velocity`.
 *   fraction[a], solid_velocity[a]: Collider::GetFraction() / Collider::Velocity
 *   jump[a]: per-face value of the pressure-jump callback (SetPressureJump,
 *            core/Pressure.h:20-22), or jump == NULL for no callback
 *   pressure: optional N^3 output, the solution on liquid cells and 0 elsewhere
 * Returns SFP_OK also when the solve did not reach tol (the reference accepts
 * non-convergence); see stats->converged. */
sfp_status sfp_pressure_project(sfp_context ctx, double *const velocity[3],
                                const double *level_set, const double *const fraction[3],
                                const double *const solid_velocity[3],
                                const double *const jump[3], double vol_error,
                                double *pressure, sfp_stats *stats);

/* Pressure::Reproject (core/Pressure.cpp:28-36): same unknowns and matrix as
 * the last sfp_pressure_project, new right-hand side without jump terms
 * (SetRightHandSide :80-102), warm start from the last pressure. */
sfp_status sfp_pressure_reproject(sfp_context ctx, double *const velocity[3],
                                  const double *level_set, const double *const fraction[3],
                                  const double *const solid_velocity[3],
                                  const double *const jump[3], double vol_error,
                                  double *pressure, sfp_stats *stats);

/* ---- FDM magnetostatic potential --------------------------------------- */

/* MagneticField::Solve (core/MagneticField.cpp:29-35): face permeability
 * mu = 1 + (1 - H_band(phi_face)) * chi (BuildMatrix :50-60), the pinned
 * 7-point mu-Laplacian (:62-90), right-hand side from (mu-1) H_ext
 * (AssembleRHS :97-118), a solve warm-started from the previous call
 * (SolveLinearSystem :120-123), and H = H_ext - grad(potential)
 * (AssignTotalField :125-143).
 *   potential: optional N^3 output; total_field: optional face outputs. */
sfp_status sfp_magnetic_solve(sfp_context ctx, const double *level_set,
                              const double *const external_field[3], double susceptibility,
                              double band_width, double *potential,
                              double *const total_field[3], sfp_stats *stats);

/* MagneticField::CalculateVertexPressures (core/MagneticField.cpp:37-48) on
 * the total field of the last sfp_magnetic_solve: positions and normals are
 * nverts x 3 doubles (x,y,z per vertex); pressures gets nverts doubles. */
sfp_status sfp_magnetic_vertex_pressures(sfp_context ctx, int nverts, const double *positions,
                                         const double *normals, double susceptibility,
                                         double *pressures);

/* ---- after the projection ---------------------------------------------- */

/* Extrapolation::Solve on the three face grids (core/Extrapolation.h:27-37,
 * core/Extrapolation.cpp:4-30): `steps` sweeps; in each, every invalid face
 * with at least one valid neighbour on its face grid (-x,+x,-y,+y,-z,+z)
 * takes their mean and becomes valid, every other invalid face is set to
 * clear_value. valid[a]: one byte per face of grid a, nonzero = valid (read
 * only). */
sfp_status sfp_extrapolate_faces(sfp_context ctx, double *const data[3],
                                 const unsigned char *const valid[3], double clear_value,
                                 int steps);

/* The velocity extrapolation of Simulation::ProjectVelocity
 * (core/Simulation.cpp:236-240): sfp_extrapolate_faces with clear value 0 and
 * valid = fraction < 1 and level_set <= 0 in a cell on either side of the
 * face (a side outside the grid counts as air). */
sfp_status sfp_extrapolate_velocity(sfp_context ctx, double *const velocity[3],
                                    const double *level_set, const double *const fraction[3],
                                    int steps);

/* Collider::Enforce (core/Collider.cpp:75-86): at every face whose solid
 * fraction is exactly 1, velocity becomes vel - ((vel - vel0) . n) n with vel,
 * vel0 = collider_velocity and n = normal (normalized) trilinearly
 * interpolated at the face (core/TriLerp.h), from the velocity before the
 * call. normal[a]: Collider::GetNormal(). */
sfp_status sfp_collider_enforce(sfp_context ctx, double *const velocity[3],
                                const double *const fraction[3], const double *const normal[3],
                                const double *const collider_velocity[3]);

/* Simulation::ProjectVelocity's three velocity passes (core/Simulation.cpp:
 * 236-241) with the velocity kept in HBM between them: sfp_pressure_project,
 * then sfp_extrapolate_velocity with `steps` sweeps, then
 * sfp_collider_enforce, each with its own semantics, on host arrays that are
 * uploaded once and downloaded once (velocity, and pressure when non-NULL).
 * Results are those of the three calls in a row, bit for bit. */
sfp_status sfp_projection_substep(sfp_context ctx, double *const velocity[3],
                                  const double *level_set, const double *const fraction[3],
                                  const double *const solid_velocity[3],
                                  const double *const jump[3], double vol_error, int steps,
                                  const double *const normal[3],
                                  const double *const collider_velocity[3], double *pressure,
                                  sfp_stats *stats);

/* ---- device-resident variants ------------------------------------------ */
/* Same semantics; every array pointer is a device pointer, work is enqueued
 * on sfp_stream(ctx) and the call returns after the solve finished. */
sfp_status sfp_pressure_project_device(sfp_context ctx, double *const velocity[3],
                                       const double *level_set, const double *const fraction[3],
                                       const double *const solid_velocity[3],
                                       const double *const jump[3], double vol_error,
                                       double *pressure, sfp_stats *stats);
sfp_status sfp_pressure_reproject_device(sfp_context ctx, double *const velocity[3],
                                         const double *level_set,
                                         const double *const fraction[3],
                                         const double *const solid_velocity[3],
                                         const double *const jump[3], double vol_error,
                                         double *pressure, sfp_stats *stats);
sfp_status sfp_magnetic_solve_device(sfp_context ctx, const double *level_set,
                                     const double *const external_field[3],
                                     double susceptibility, double band_width, double *potential,
                                     double *const total_field[3], sfp_stats *stats);
sfp_status sfp_extrapolate_velocity_device(sfp_context ctx, double *const velocity[3],
                                           const double *level_set,
                                           const double *const fraction[3], int steps);
sfp_status sfp_collider_enforce_device(sfp_context ctx, double *const velocity[3],
                                       const double *const fraction[3],
                                       const double *const normal[3],
                                       const double *const collider_velocity[3]);

/* ---- without allocation ------------------------------------------------ */
/* sfp_extrapolate_velocity_device without a context and without allocating
 * anything: no counterpart in the reference. For callers that cannot allocate
 * device memory (code under HIP graph capture, tight loops that forbid it).
 * velocity, level_set and fraction are device arrays of an nx*ny*nz grid;
 * the work is enqueued on `stream` (a hipStream_t; null = the default
 * stream) and the call returns without waiting. Its mask and sweep buffer
 * (two bytes per face of the largest face grid) live in one fixed device
 * array of 64 MiB that each GPU holds until the process ends: grids needing
 * more are rejected with SFP_ERR_ARGUMENT. Calls on one GPU share that array,
 * so calls that overlap in time on one GPU (other streams, other threads)
 * race; calls on one stream are ordered. Errors: sfp_last_error(NULL). */
sfp_status sfp_extrapolate_velocity_noalloc(int nx, int ny, int nz, double *const velocity[3],
                                            const double *level_set,
                                            const double *const fraction[3], int steps,
                                            void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SFP_H */
