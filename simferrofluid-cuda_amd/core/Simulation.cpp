// ProjectionStep: Simulation::UpdateMagneticPressure and
// Simulation::ProjectVelocity (core/Simulation.cpp:193-243).
#include "Simulation.h"

namespace Pivot {

ProjectionStep::ProjectionStep(StaggeredGrid const &sgrid, SGridData<double> &velocity,
                               GridData<double> const &levelSet, Collider const &collider,
                               Contour const &contour, Settings settings)
    : m_SGrid{sgrid}, m_Velocity{velocity}, m_LevelSet{levelSet}, m_Collider{collider},
      m_Contour{contour}, m_Settings{std::move(settings)}, m_Pressure(sgrid),
      m_MagneticField(sgrid, m_Settings.chi, sgrid.GetSpacing() * 2.0) {}

void ProjectionStep::UpdateMagneticPressure(double time) {
    m_MagneticField.SetSusceptibility(m_Settings.chi);
    m_MagneticField.SetBandWidth(m_LevelSet.GetGrid().GetSpacing() * 2.0);
    SGridData<double> externalField(m_SGrid.GetFaceGrids());
    for (int axis = 0; axis < 3; axis++) {
        Grid const &g = externalField[axis].GetGrid();
        ForEach(g, [&](Vector3i const &face) {
            Vector3d const pos = g.PositionOf(face);
            externalField[axis][face] = m_Settings.fieldApplied(pos, time)[axis];
        });
    }
    m_MagneticField.Solve(m_LevelSet, externalField);
    m_MagneticPressures = m_MagneticField.CalculateVertexPressures(m_Contour.GetMesh());
}

void ProjectionStep::ProjectVelocity(double dt, double time, double volError) {
    if (m_Settings.magneticEnabled) UpdateMagneticPressure(time);
    double const invSpacing = m_SGrid.GetInvSpacing();
    m_Pressure.SetPressureJump([&](int axis, Vector3i const &face, double) -> double {
        double pressure = 0;
        SurfaceMesh const &mesh = m_Contour.GetMesh();
        if (m_Settings.surfaceTensionEnabled && !m_Settings.semiImplicitSTEnabled) {
            int const index = m_Contour.VertexIndexOf(axis, face - Vector3i::Unit(axis));
            if (index >= 0)
                pressure += mesh.MeanCurvatures[index] * m_Settings.surfaceTensionCoeff /
                            m_Settings.liquidDensity * invSpacing * dt;
        }
        if (m_Settings.magneticEnabled) {
            int const index = m_Contour.VertexIndexOf(axis, face - Vector3i::Unit(axis));
            if (index >= 0)
                pressure -= m_MagneticPressures[index] / m_Settings.liquidDensity * invSpacing * dt;
        }
        return pressure;
    });
    // Pressure::Project, Extrapolation::Solve(velocity, 0, 6, fraction < 1 and
    // a liquid cell on either side) and Collider::Enforce, as the reference
    // calls them in a row (core/Simulation.cpp:236-241), with the velocity
    // kept on the GPU between the three (a side outside the cell grid counts
    // as air: the reference reads outside the grid there, which its
    // closed-box collider never allows)
    m_Pressure.ProjectExtrapolateEnforce(m_Velocity, m_LevelSet, m_Collider,
                                         m_Settings.semiImplicitSTEnabled ? 0 : volError, 6);
}

}  // namespace Pivot
