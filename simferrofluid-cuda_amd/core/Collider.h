// The parts of the reference's Collider (core/Collider.h:8-38) that the
// projection and its follow-up steps read: per-face solid fraction, solid
// velocity and wall normal, and Collider::Enforce (core/Collider.cpp:75-86),
// which runs on the GPU. Producing the fractions and normals
// (Collider::Finish, core/Collider.cpp:22-45: fast marching on the collider
// level set) happens upstream; this holder carries them.
#pragma once
#include "SGridData.h"
#include "StaggeredGrid.h"

namespace Pivot {

class Collider {
  public:
    explicit Collider(StaggeredGrid const &sgrid)
        : Velocity(sgrid.GetFaceGrids()), m_SGrid{sgrid}, m_Fraction(sgrid.GetFaceGrids()),
          m_Normal(sgrid.GetFaceGrids()) {}

    SGridData<double> const &GetFraction() const { return m_Fraction; }
    SGridData<double> const &GetNormal() const { return m_Normal; }
    SGridData<double> &Fraction() { return m_Fraction; }
    SGridData<double> &Normal() { return m_Normal; }

    // Solid fraction 1 on the outermost face layer of every face grid (the
    // closed box the reference's domain collider produces), and there the
    // wall's unit normal pointing into the domain.
    void SetBoxWalls() {
        for (int a = 0; a < 3; a++) {
            Grid const &g = m_Fraction[a].GetGrid();
            ForEach(g, [&](Vector3i const &f) {
                bool const lo = f[a] == 0, hi = f[a] == g.GetSize()[a] - 1;
                m_Fraction[a][f] = (lo || hi) ? 1.0 : 0.0;
                m_Normal[a][f] = lo ? 1.0 : (hi ? -1.0 : 0.0);
            });
        }
    }

    // Collider::Enforce (core/Collider.cpp:75-86): at faces with solid
    // fraction 1, remove the normal component of the fluid velocity relative
    // to the collider velocity (GPU, sfp_collider_enforce).
    void Enforce(SGridData<double> &fluidVelocity) const;

    SGridData<double> Velocity;

  private:
    StaggeredGrid const &m_SGrid;
    SGridData<double> m_Fraction;
    SGridData<double> m_Normal;
};

}  // namespace Pivot
