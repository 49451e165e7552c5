// The parts of the reference's Contour (core/Contour.h:8-35) that the
// pressure-jump callback reads: the surface mesh and, per grid edge, the
// index of the mesh vertex on it (-1 for none). Producing them (marching
// cubes, core/Contour.cpp) happens upstream; this holder carries them.
#pragma once
#include "GridData.h"
#include "SurfaceMesh.h"

#include <array>

namespace Pivot {

class Contour {
  public:
    // nodeGrid: the grid the contour is extracted from (the level set's cell
    // grid); edge grid a has one node fewer along a.
    explicit Contour(Grid const &nodeGrid)
        : m_NodeGrid{nodeGrid},
          m_EdgeGrids{edge_grid(nodeGrid, 0), edge_grid(nodeGrid, 1), edge_grid(nodeGrid, 2)},
          m_EdgeMark{GridData<int>(m_EdgeGrids[0], -1), GridData<int>(m_EdgeGrids[1], -1),
                     GridData<int>(m_EdgeGrids[2], -1)} {}

    SurfaceMesh const &GetMesh() const { return m_Mesh; }
    SurfaceMesh &GetMesh() { return m_Mesh; }

    // mesh vertex on the edge from node `edge` to node `edge + e_axis`
    int VertexIndexOf(int axis, Vector3i const &edge) const { return m_EdgeMark[axis][edge]; }
    GridData<int> &EdgeMark(int axis) { return m_EdgeMark[axis]; }
    Grid const &GetEdgeGrid(int axis) const { return m_EdgeGrids[axis]; }

  private:
    static Grid edge_grid(Grid const &g, int axis) {
        Vector3i n = g.GetSize();
        n[axis] -= 1;
        Vector3d o = g.GetOrigin();
        o[axis] += g.GetSpacing() / 2;
        return Grid(g.GetSpacing(), n, o);
    }

    Grid const &m_NodeGrid;
    std::array<Grid, 3> m_EdgeGrids;
    std::array<GridData<int>, 3> m_EdgeMark;
    SurfaceMesh m_Mesh;
};

}  // namespace Pivot
