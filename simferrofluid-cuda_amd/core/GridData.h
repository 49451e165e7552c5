// GridData<T> (core/GridData.h:8-53): a std::vector in the grid's z-fastest
// order plus a reference to the owning Grid, which must outlive it.
#pragma once
#include "Grid.h"

#include <cstdio>
#include <cstdlib>
#include <istream>
#include <ostream>
#include <vector>

namespace Pivot {

template <class Type>
class GridData {
  public:
    explicit GridData(Grid const &grid, Type const &value = Type(0))
        : m_Grid{grid}, m_Data(grid.GetNumVertices(), value) {}
    GridData(GridData const &) = default;

    GridData &operator=(GridData const &rhs) {
        if (!(m_Grid == rhs.m_Grid)) {
            std::fprintf(stderr, "Failed to assign between GridData with different grids\n");
            std::exit(EXIT_FAILURE);
        }
        m_Data = rhs.m_Data;
        return *this;
    }

    Grid const &GetGrid() const { return m_Grid; }
    auto &GetData() { return m_Data; }
    auto const &GetData() const { return m_Data; }

    Type &operator[](Vector3i const &c) { return m_Data[m_Grid.IndexOf(c)]; }
    Type const &operator[](Vector3i const &c) const { return m_Data[m_Grid.IndexOf(c)]; }
    Type &operator[](int index) { return m_Data[index]; }
    Type const &operator[](int index) const { return m_Data[index]; }

    // raw values in index order, no header (core/GridData.h:47-48, IO::Write of a span)
    void Save(std::ostream &out) const {
        out.write(reinterpret_cast<char const *>(m_Data.data()), (std::streamsize)(m_Data.size() * sizeof(Type)));
    }
    void Load(std::istream &in) {
        in.read(reinterpret_cast<char *>(m_Data.data()), (std::streamsize)(m_Data.size() * sizeof(Type)));
    }

    void SetConstant(Type const &value) { std::fill(m_Data.begin(), m_Data.end(), value); }
    void SetZero() { SetConstant(Type(0)); }

  private:
    Grid const &m_Grid;
    std::vector<Type> m_Data;
};

}  // namespace Pivot
