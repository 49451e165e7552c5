// Drop-in replacement of the reference's MagneticField class
// (core/MagneticField.h:15-72) with the same public interface; the potential
// solve runs on the GPU through the C ABI in sfp.h instead of amgcl.
#pragma once
#include "GridData.h"
#include "SGridData.h"
#include "StaggeredGrid.h"
#include "SurfaceMesh.h"

#include <vector>

struct sfp_context_s;

namespace Pivot {

class MagneticField {
  public:
    MagneticField(StaggeredGrid const &sgrid, double susceptibility, double bandWidth);
    ~MagneticField();
    MagneticField(MagneticField const &) = delete;
    MagneticField &operator=(MagneticField const &) = delete;

    void SetSusceptibility(double susceptibility) { m_Susceptibility = susceptibility; }
    void SetBandWidth(double bandWidth) { m_BandWidth = bandWidth; }

    void Solve(GridData<double> const &levelSet, SGridData<double> const &externalField);

    std::vector<double> CalculateVertexPressures(SurfaceMesh const &mesh) const;

    // Potential and total field of the last solve (the reference keeps them
    // private: m_Potential, m_TotalField).
    std::vector<double> const &GetPotential() const { return m_Potential; }
    std::vector<double> const &GetTotalField(int axis) const { return m_TotalField[axis]; }
    int LastIterations() const { return m_LastIters; }
    double LastResidual() const { return m_LastResidual; }

  private:
    sfp_context_s *m_Ctx = nullptr;
    double m_Susceptibility;
    double m_BandWidth;
    std::vector<double> m_Potential;
    std::vector<double> m_TotalField[3];
    int m_LastIters = 0;
    double m_LastResidual = 0;
};

}  // namespace Pivot
