// Extrapolation::Solve for staggered data (core/Extrapolation.h:27-37,
// core/Extrapolation.cpp:4-30) with the same signature; the validity
// predicate is host code and is evaluated on the host into one byte per
// face, the sweeps run on the GPU (sfp_extrapolate_faces).
#pragma once
#include "../../include/sfp.h"
#include "GpuContext.h"
#include "SGridData.h"

#include <cstdint>
#include <functional>
#include <type_traits>
#include <vector>

namespace Pivot {

class Extrapolation {
  public:
    template <typename Func>
        requires std::is_convertible_v<Func, std::function<bool(int, Vector3i const &)>>
    static void Solve(SGridData<double> &sgrData, double clearVal, int maxSteps, Func &&isValid) {
        auto const &grids = sgrData.GetGrids();
        std::vector<std::uint8_t> valid[3];
        for (int a = 0; a < 3; a++) {
            Grid const &g = grids[a];
            valid[a].assign(g.GetNumVertices(), 0);
            ForEach(g, [&](Vector3i const &face) { valid[a][g.IndexOf(face)] = isValid(a, face) ? 1 : 0; });
        }
        // the cell grid behind the face grids: face grid 0 has one more x layer
        Vector3i n = grids[0].GetSize();
        n[0] -= 1;
        double const h = grids[0].GetSpacing();
        Vector3d const o0 = grids[0].GetOrigin();
        Vector3d const origin{o0[0], o0[1] - h / 2, o0[2] - h / 2};
        sfp_context ctx = detail::SharedContext(n, h, origin);
        double *data[3];
        unsigned char const *v[3];
        for (int a = 0; a < 3; a++) {
            data[a] = sgrData[a].GetData().data();
            v[a] = valid[a].data();
        }
        sfp_status const st = sfp_extrapolate_faces(ctx, data, v, clearVal, maxSteps);
        if (st != SFP_OK) detail::Fail(ctx, st, "Extrapolation::Solve");
    }
};

}  // namespace Pivot
