// Drop-in MagneticField (core/MagneticField.cpp:1-144) over the gfx950 solver.
#include "MagneticField.h"

#include "../../include/sfp.h"

#include <cstdio>
#include <cstdlib>

namespace Pivot {

namespace {
[[noreturn]] void fail(sfp_context ctx, sfp_status st, char const *what) {
    std::fprintf(stderr, "%s failed: %s (%s)\n", what, sfp_status_string(st), sfp_last_error(ctx));
    std::exit(EXIT_FAILURE);
}
}  // namespace

MagneticField::MagneticField(StaggeredGrid const &sgrid, double susceptibility, double bandWidth)
    : m_Susceptibility(susceptibility), m_BandWidth(bandWidth) {
    Vector3i const n = sgrid.GetResolution();
    Vector3d const o = sgrid.GetOrigin();
    double const origin[3] = {o[0], o[1], o[2]};
    sfp_status st = sfp_create(&m_Ctx, n[0], n[1], n[2], sgrid.GetSpacing(), origin, 0);
    if (st != SFP_OK) fail(nullptr, st, "sfp_create");
    sfp_params p;
    sfp_default_params(&p);
    p.tol = 1e-6;       // the reference's prm.solver.tol (core/MagneticField.cpp:92-93)
    p.max_iters = 100;  // amgcl's default iteration cap
    st = sfp_set_params(m_Ctx, &p);
    if (st != SFP_OK) fail(m_Ctx, st, "sfp_set_params");
    m_Potential.assign(sgrid.GetNumCells(), 0.0);
    for (int a = 0; a < 3; a++) m_TotalField[a].assign(sgrid.GetFaceGrids()[a].GetNumVertices(), 0.0);
}

MagneticField::~MagneticField() { sfp_destroy(m_Ctx); }

void MagneticField::Solve(GridData<double> const &levelSet, SGridData<double> const &externalField) {
    double const *hext[3] = {externalField[0].GetData().data(), externalField[1].GetData().data(),
                             externalField[2].GetData().data()};
    double *htot[3] = {m_TotalField[0].data(), m_TotalField[1].data(), m_TotalField[2].data()};
    sfp_stats stats;
    sfp_status const st = sfp_magnetic_solve(m_Ctx, levelSet.GetData().data(), hext, m_Susceptibility,
                                             m_BandWidth, m_Potential.data(), htot, &stats);
    if (st != SFP_OK) fail(m_Ctx, st, "MagneticField::Solve");
    m_LastIters = stats.iterations;
    m_LastResidual = stats.rel_residual;
    std::printf("%6d iters (%.2e) ", stats.iterations, stats.rel_residual);
}

std::vector<double> MagneticField::CalculateVertexPressures(SurfaceMesh const &mesh) const {
    std::vector<double> pressures(mesh.size());
    if (mesh.size() == 0) return pressures;
    std::vector<double> pos(3 * (size_t)mesh.size()), nrm(3 * (size_t)mesh.size());
    for (int i = 0; i < mesh.size(); i++)
        for (int a = 0; a < 3; a++) {
            pos[3 * (size_t)i + a] = mesh.Positions[i][a];
            nrm[3 * (size_t)i + a] = mesh.Normals[i][a];
        }
    sfp_status const st = sfp_magnetic_vertex_pressures(m_Ctx, mesh.size(), pos.data(), nrm.data(),
                                                        m_Susceptibility, pressures.data());
    if (st != SFP_OK) fail(m_Ctx, st, "MagneticField::CalculateVertexPressures");
    return pressures;
}

}  // namespace Pivot
