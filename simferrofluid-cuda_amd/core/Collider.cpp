// Collider::Enforce (core/Collider.cpp:75-86) over the GPU library.
#include "Collider.h"

#include "../../include/sfp.h"
#include "GpuContext.h"

namespace Pivot {

void Collider::Enforce(SGridData<double> &fluidVelocity) const {
    sfp_context ctx = detail::SharedContext(m_SGrid.GetResolution(), m_SGrid.GetSpacing(), m_SGrid.GetOrigin());
    double *vel[3];
    double const *frac[3], *nrm[3], *cv[3];
    for (int a = 0; a < 3; a++) {
        vel[a] = fluidVelocity[a].GetData().data();
        frac[a] = m_Fraction[a].GetData().data();
        nrm[a] = m_Normal[a].GetData().data();
        cv[a] = Velocity[a].GetData().data();
    }
    sfp_status const st = sfp_collider_enforce(ctx, vel, frac, nrm, cv);
    if (st != SFP_OK) detail::Fail(ctx, st, "Collider::Enforce");
}

}  // namespace Pivot
