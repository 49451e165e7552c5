// The caller side of the two solves: the magnetic-pressure update and the
// velocity projection substep of the reference's Simulation
// (core/Simulation.cpp:193-243), over the drop-in Pressure, MagneticField,
// Extrapolation and Collider. The rest of the reference's Simulation
// (advection, reinitialization, meshing, surface tension, I/O) is not part of
// this library; ProjectionStep holds references to the state those produce,
// as Simulation holds the same objects as members.
#pragma once
#include "Collider.h"
#include "Contour.h"
#include "MagneticField.h"
#include "Pressure.h"

#include <functional>
#include <vector>

namespace Pivot {

class ProjectionStep {
  public:
    struct Settings {
        bool magneticEnabled = false;
        bool surfaceTensionEnabled = false;
        bool semiImplicitSTEnabled = false;
        double surfaceTensionCoeff = 0;
        double liquidDensity = 1;
        double chi = 0;  // Magnetic::GetChi()
        // applied field H(position, time) (Simulation::m_FieldApplied)
        std::function<Vector3d(Vector3d const &, double)> fieldApplied;
    };

    ProjectionStep(StaggeredGrid const &sgrid, SGridData<double> &velocity,
                   GridData<double> const &levelSet, Collider const &collider,
                   Contour const &contour, Settings settings);

    // Simulation::UpdateMagneticPressure, FDM branch (core/Simulation.cpp:193-212)
    void UpdateMagneticPressure(double time);

    // Simulation::ProjectVelocity (core/Simulation.cpp:214-242): magnetic
    // update, pressure jump from the mesh's curvature and magnetic pressure,
    // projection, velocity extrapolation and the wall condition.
    void ProjectVelocity(double dt, double time, double volError);

    std::vector<double> const &GetMagneticPressures() const { return m_MagneticPressures; }
    Pressure &GetPressure() { return m_Pressure; }
    MagneticField &GetMagneticField() { return m_MagneticField; }

  private:
    StaggeredGrid const &m_SGrid;
    SGridData<double> &m_Velocity;
    GridData<double> const &m_LevelSet;
    Collider const &m_Collider;
    Contour const &m_Contour;
    Settings m_Settings;
    Pressure m_Pressure;
    MagneticField m_MagneticField;
    std::vector<double> m_MagneticPressures;
};

}  // namespace Pivot
