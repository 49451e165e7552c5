// The parts of the reference's SurfaceMesh (core/SurfaceMesh.h) that
// MagneticField::CalculateVertexPressures reads: vertex positions and normals.
#pragma once
#include "Vector3.h"

#include <vector>

namespace Pivot {

struct SurfaceMesh {
    std::vector<Vector3d> Positions;
    std::vector<Vector3d> Normals;
    int size() const { return (int)Positions.size(); }
};

}  // namespace Pivot
