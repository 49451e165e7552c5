// Grid index math with the reference's layout (core/Grid.h:6-47): z-fastest
// IndexOf = z + Nz*(y + Ny*x), PositionOf = origin + coord*spacing, neighbour
// ordinal ord -> axis ord>>1, side (ord&1 ? +1 : -1).
#pragma once
#include "Vector3.h"

namespace Pivot {

class Grid {
  public:
    Grid(double spacing, Vector3i const &size, Vector3d const &origin)
        : m_Spacing{spacing}, m_InvSpacing{1 / spacing}, m_Size{size}, m_Origin{origin} {}

    bool operator==(Grid const &rhs) const {
        return m_Spacing == rhs.m_Spacing && m_Size == rhs.m_Size && m_Origin == rhs.m_Origin;
    }

    double GetSpacing() const { return m_Spacing; }
    double GetInvSpacing() const { return m_InvSpacing; }
    Vector3i GetSize() const { return m_Size; }
    Vector3d GetOrigin() const { return m_Origin; }
    int GetNumVertices() const { return m_Size.prod(); }

    bool IsInside(Vector3i const &c, int offset = 0) const {
        for (int a = 0; a < 3; a++)
            if (c[a] < offset || c[a] >= m_Size[a] - offset) return false;
        return true;
    }
    bool IsValid(Vector3i const &c) const { return IsInside(c, 0); }
    int IndexOf(Vector3i const &c) const { return c.z() + m_Size.z() * (c.y() + m_Size.y() * c.x()); }
    Vector3d PositionOf(Vector3i const &c) const {
        return {m_Origin[0] + c[0] * m_Spacing, m_Origin[1] + c[1] * m_Spacing,
                m_Origin[2] + c[2] * m_Spacing};
    }
    Vector3i CoordOf(int index) const {
        return {index / m_Size.z() / m_Size.y(), index / m_Size.z() % m_Size.y(), index % m_Size.z()};
    }

    static constexpr int GetNumNeighbors() { return 6; }
    static constexpr int NeighborAxisOf(int ord) { return ord >> 1; }
    static constexpr int NeighborSideOf(int ord) { return ord & 1 ? 1 : -1; }
    static Vector3i NeighborOf(Vector3i const &c, int ord) {
        return c + Vector3i::Unit(NeighborAxisOf(ord)) * NeighborSideOf(ord);
    }

  private:
    double m_Spacing;
    double m_InvSpacing;
    Vector3i m_Size;
    Vector3d m_Origin;
};

template <class Func>
inline void ForEach(Grid const &grid, Func &&func) {
    for (int i = 0; i < grid.GetSize().x(); i++)
        for (int j = 0; j < grid.GetSize().y(); j++)
            for (int k = 0; k < grid.GetSize().z(); k++) func(Vector3i(i, j, k));
}

template <class Func>
inline void ForEach(std::array<Grid, 3> const &grids, Func &&func) {
    for (int axis = 0; axis < 3; axis++)
        ForEach(grids[axis], [&](Vector3i const &c) { func(axis, c); });
}

}  // namespace Pivot
