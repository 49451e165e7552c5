// The parts of the reference's SurfaceMesh (core/SurfaceMesh.h:6-38) that the
// magnetic and projection steps read: vertex positions, normals and mean
// curvatures, and the triangle indices.
#pragma once
#include "Vector3.h"

#include <cstdint>
#include <vector>

namespace Pivot {

struct SurfaceMesh {
    std::vector<Vector3d> Positions;
    std::vector<Vector3d> Normals;
    std::vector<std::uint32_t> Indices;
    std::vector<double> MeanCurvatures;
    int size() const { return (int)Positions.size(); }
};

}  // namespace Pivot
