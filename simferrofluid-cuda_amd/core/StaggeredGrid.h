// StaggeredGrid (core/StaggeredGrid.h:8-58): cell grid at origin + dx/2, face
// grid a with extent N + e_a at origin + dx/2 (1 - e_a).
#pragma once
#include "Grid.h"

namespace Pivot {

class StaggeredGrid {
  public:
    StaggeredGrid(int boundaryWidth, double spacing, Vector3i resolution,
                  Vector3d center = Vector3d::Zero())
        : m_BoundaryWidth{boundaryWidth},
          m_Spacing{spacing},
          m_Resolution{resolution},
          m_Origin{center - resolution.cast<double>() * spacing / 2},
          m_CellGrid(spacing, resolution, m_Origin + Vector3d::Constant(spacing / 2)),
          m_FaceGrids{Grid(spacing, resolution + Vector3i(1, 0, 0), m_Origin + Vector3d(0, 1, 1) * spacing / 2),
                      Grid(spacing, resolution + Vector3i(0, 1, 0), m_Origin + Vector3d(1, 0, 1) * spacing / 2),
                      Grid(spacing, resolution + Vector3i(0, 0, 1), m_Origin + Vector3d(1, 1, 0) * spacing / 2)} {}

    int GetBoundaryWidth() const { return m_BoundaryWidth; }
    double GetSpacing() const { return m_Spacing; }
    double GetInvSpacing() const { return 1 / m_Spacing; }
    Vector3i GetResolution() const { return m_Resolution; }
    Vector3d GetOrigin() const { return m_Origin; }
    int GetNumCells() const { return m_CellGrid.GetNumVertices(); }
    Grid const &GetCellGrid() const { return m_CellGrid; }
    std::array<Grid, 3> const &GetFaceGrids() const { return m_FaceGrids; }

    static constexpr int FaceAxisOfCell(int ord) { return ord >> 1; }
    static constexpr int FaceSideOfCell(int ord) { return ord & 1 ? 1 : -1; }
    static std::pair<int, Vector3i> FaceOfCell(Vector3i const &cell, int ord) {
        return {FaceAxisOfCell(ord), cell + Vector3i::Unit(FaceAxisOfCell(ord)) * (ord & 1)};
    }
    static Vector3i AdjCellOfFace(int axis, Vector3i const &face, int ord) {
        return face - Vector3i::Unit(axis) * ((ord & 1) ^ 1);
    }

  private:
    int m_BoundaryWidth;
    double m_Spacing;
    Vector3i m_Resolution;
    Vector3d m_Origin;
    Grid m_CellGrid;
    std::array<Grid, 3> m_FaceGrids;
};

}  // namespace Pivot
