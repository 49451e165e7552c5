// The solid inputs the pressure projection reads from the reference's
// Collider (core/Collider.h:13,28): per-face solid fraction and solid
// velocity. Producing them (core/Collider.cpp:8-73) happens upstream of the
// solve; this holder carries them.
#pragma once
#include "SGridData.h"
#include "StaggeredGrid.h"

namespace Pivot {

class Collider {
  public:
    explicit Collider(StaggeredGrid const &sgrid)
        : Velocity(sgrid.GetFaceGrids()), m_Fraction(sgrid.GetFaceGrids()) {}

    SGridData<double> const &GetFraction() const { return m_Fraction; }
    SGridData<double> &Fraction() { return m_Fraction; }

    // Solid fraction 1 on the outermost face layer of every face grid: the
    // closed box the reference's domain collider produces.
    void SetBoxWalls() {
        for (int a = 0; a < 3; a++) {
            Grid const &g = m_Fraction[a].GetGrid();
            ForEach(g, [&](Vector3i const &f) {
                m_Fraction[a][f] = (f[a] == 0 || f[a] == g.GetSize()[a] - 1) ? 1.0 : 0.0;
            });
        }
    }

    SGridData<double> Velocity;

  private:
    SGridData<double> m_Fraction;
};

}  // namespace Pivot
