// SGridData<T> (core/SGridData.h:7-41): one GridData per face axis.
#pragma once
#include "GridData.h"

namespace Pivot {

template <class Type>
class SGridData {
  public:
    explicit SGridData(std::array<Grid, 3> const &grids,
                       Vector3<Type> const &value = Vector3<Type>::Zero())
        : m_Grids{grids},
          m_Datas{GridData<Type>(m_Grids[0], value[0]), GridData<Type>(m_Grids[1], value[1]),
                  GridData<Type>(m_Grids[2], value[2])} {}
    SGridData(SGridData const &) = default;

    SGridData &operator=(SGridData const &rhs) {
        for (int a = 0; a < 3; a++) m_Datas[a] = rhs.m_Datas[a];
        return *this;
    }

    std::array<Grid, 3> const &GetGrids() const { return m_Grids; }
    // axis 0, 1, 2 in turn (core/SGridData.h:35-36)
    void Save(std::ostream &out) const {
        for (int i = 0; i < 3; i++) m_Datas[i].Save(out);
    }
    void Load(std::istream &in) {
        for (int i = 0; i < 3; i++) m_Datas[i].Load(in);
    }
    GridData<Type> &operator[](int axis) { return m_Datas[axis]; }
    GridData<Type> const &operator[](int axis) const { return m_Datas[axis]; }

  private:
    std::array<Grid, 3> const &m_Grids;
    std::array<GridData<Type>, 3> m_Datas;
};

}  // namespace Pivot
