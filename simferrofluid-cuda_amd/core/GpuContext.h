// A library context (sfp_context) per grid geometry for the drop-in pieces
// that have no object of their own to hold one (Extrapolation's static
// Solve, Collider::Enforce). Contexts live until the process ends.
#pragma once
#include "Vector3.h"

struct sfp_context_s;

namespace Pivot::detail {

// The context for a cell grid of `resolution` cells with the given spacing
// and lower corner (StaggeredGrid::GetOrigin()); created on first use.
sfp_context_s *SharedContext(Vector3i const &resolution, double spacing, Vector3d const &origin);

// Prints the library's error and exits, like the reference's fatal paths.
[[noreturn]] void Fail(sfp_context_s *ctx, int status, char const *what);

}  // namespace Pivot::detail
