// Drop-in Pressure (core/Pressure.cpp:1-167) over the gfx950 solver.
#include "Pressure.h"

#include "../../include/sfp.h"

#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>

namespace Pivot {

namespace {
[[noreturn]] void fail(sfp_context ctx, sfp_status st, char const *what) {
    std::fprintf(stderr, "%s failed: %s (%s)\n", what, sfp_status_string(st), sfp_last_error(ctx));
    std::exit(EXIT_FAILURE);
}
}  // namespace

Pressure::Pressure(StaggeredGrid const &sgrid) : m_SGrid{sgrid} {
    Vector3i const n = sgrid.GetResolution();
    Vector3d const o = sgrid.GetOrigin();
    double const origin[3] = {o[0], o[1], o[2]};
    sfp_status st = sfp_create(&m_Ctx, n[0], n[1], n[2], sgrid.GetSpacing(), origin, 0);
    if (st != SFP_OK) fail(nullptr, st, "sfp_create");
    sfp_params p;
    sfp_default_params(&p);
    p.tol = 1e-8;       // amgcl's default relative tolerance
    p.max_iters = 100;  // amgcl's default iteration cap
    st = sfp_set_params(m_Ctx, &p);
    if (st != SFP_OK) fail(m_Ctx, st, "sfp_set_params");
    m_Pressure.assign(sgrid.GetNumCells(), 0.0);
}

Pressure::~Pressure() { sfp_destroy(m_Ctx); }

void Pressure::Project(SGridData<double> &velocity, GridData<double> const &levelSet,
                       Collider const &collider, double volError) {
    Solve(velocity, levelSet, collider, volError, false);
}

void Pressure::Reproject(SGridData<double> &velocity, GridData<double> const &levelSet,
                         Collider const &collider, double volError) {
    Solve(velocity, levelSet, collider, volError, true);
}

bool Pressure::EvaluateJumps(SGridData<double> const &velocity, GridData<double> const &levelSet,
                             Collider const &collider) {
    // The callback is host code: evaluate it once per ghost-fluid face (open,
    // exactly one liquid neighbour), at theta = phi0 / (phi0 - phi1). The
    // reference calls it twice per such face: ApplyProjection passes this
    // theta (core/Pressure.cpp:154-159); BuildProjectionMatrix passes it for a
    // face on the cell's + side and 1 - phi1 / (phi1 - phi0) for one on its -
    // side (:61-64), equal in exact arithmetic but not always in the last bit.
    // A callback that ignores theta, like the simulation's own
    // (core/Simulation.cpp:216-235), gets identical values either way.
    bool const useJump = static_cast<bool>(m_PressureJump);
    if (useJump) {
        for (int axis = 0; axis < 3; axis++) {
            Grid const &fg = velocity[axis].GetGrid();
            Grid const &cg = levelSet.GetGrid();
            m_Jump[axis].assign(fg.GetNumVertices(), 0.0);
            ForEach(fg, [&](Vector3i const &face) {
                if (1 - collider.GetFraction()[axis][face] <= 0) return;
                Vector3i const c0 = StaggeredGrid::AdjCellOfFace(axis, face, 0);
                Vector3i const c1 = StaggeredGrid::AdjCellOfFace(axis, face, 1);
                if (!cg.IsValid(c0) || !cg.IsValid(c1)) return;
                double const phi0 = levelSet[c0], phi1 = levelSet[c1];
                if ((phi0 <= 0) == (phi1 <= 0)) return;
                double const theta = phi0 / (phi0 - phi1);
                m_Jump[axis][fg.IndexOf(face)] = m_PressureJump(axis, face, theta);
            });
        }
    }
    return useJump;
}

void Pressure::Solve(SGridData<double> &velocity, GridData<double> const &levelSet,
                     Collider const &collider, double volError, bool reproject) {
    bool const useJump = EvaluateJumps(velocity, levelSet, collider);
    double *vel[3] = {velocity[0].GetData().data(), velocity[1].GetData().data(),
                      velocity[2].GetData().data()};
    double const *frac[3] = {collider.GetFraction()[0].GetData().data(),
                             collider.GetFraction()[1].GetData().data(),
                             collider.GetFraction()[2].GetData().data()};
    double const *us[3] = {collider.Velocity[0].GetData().data(), collider.Velocity[1].GetData().data(),
                           collider.Velocity[2].GetData().data()};
    double const *jump[3] = {m_Jump[0].data(), m_Jump[1].data(), m_Jump[2].data()};
    sfp_stats stats;
    sfp_status const st =
        reproject ? sfp_pressure_reproject(m_Ctx, vel, levelSet.GetData().data(), frac, us,
                                           useJump ? jump : nullptr, volError, m_Pressure.data(), &stats)
                  : sfp_pressure_project(m_Ctx, vel, levelSet.GetData().data(), frac, us,
                                         useJump ? jump : nullptr, volError, m_Pressure.data(), &stats);
    if (st != SFP_OK) fail(m_Ctx, st, reproject ? "Pressure::Reproject" : "Pressure::Project");
    m_LastIters = stats.iterations;
    m_LastResidual = stats.rel_residual;
    if (stats.unknowns > 0) std::cout << std::setw(6) << stats.iterations << " iters";
}

void Pressure::ProjectExtrapolateEnforce(SGridData<double> &velocity, GridData<double> const &levelSet,
                                         Collider const &collider, double volError, int steps) {
    bool const useJump = EvaluateJumps(velocity, levelSet, collider);
    double *vel[3];
    double const *frac[3], *us[3], *jump[3], *nrm[3];
    for (int a = 0; a < 3; a++) {
        vel[a] = velocity[a].GetData().data();
        frac[a] = collider.GetFraction()[a].GetData().data();
        us[a] = collider.Velocity[a].GetData().data();
        jump[a] = m_Jump[a].data();
        nrm[a] = collider.GetNormal()[a].GetData().data();
    }
    sfp_stats stats;
    // Collider::Enforce's collider velocity is the same Collider::Velocity
    // the projection's solid faces move with (core/Collider.cpp:75-86)
    sfp_status const st = sfp_projection_substep(m_Ctx, vel, levelSet.GetData().data(), frac, us,
                                                 useJump ? jump : nullptr, volError, steps, nrm, us,
                                                 m_Pressure.data(), &stats);
    if (st != SFP_OK) fail(m_Ctx, st, "Pressure::Project + Extrapolation::Solve + Collider::Enforce");
    m_LastIters = stats.iterations;
    m_LastResidual = stats.rel_residual;
    if (stats.unknowns > 0) std::cout << std::setw(6) << stats.iterations << " iters";
}

}  // namespace Pivot
