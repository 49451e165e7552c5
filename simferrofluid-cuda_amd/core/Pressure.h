// Drop-in replacement of the reference's Pressure class (core/Pressure.h:6-56)
// with the same public interface; the solve runs on the GPU through the C ABI
// in sfp.h instead of amgcl.
#pragma once
#include "Collider.h"
#include "GridData.h"
#include "SGridData.h"
#include "StaggeredGrid.h"

#include <functional>
#include <type_traits>

struct sfp_context_s;

namespace Pivot {

class Pressure {
  public:
    explicit Pressure(StaggeredGrid const &sgrid);
    ~Pressure();
    Pressure(Pressure const &) = delete;
    Pressure &operator=(Pressure const &) = delete;

    void Project(SGridData<double> &velocity, GridData<double> const &levelSet,
                 Collider const &collider, double volError = 0);

    void Reproject(SGridData<double> &velocity, GridData<double> const &levelSet,
                   Collider const &collider, double volError = 0);

    template <typename Func>
        requires(std::is_convertible_v<Func, std::function<double(int, Vector3i const &, double)>>)
    void SetPressureJump(Func &&pressureJump) {
        m_PressureJump = pressureJump;
    }

    // Simulation::ProjectVelocity's Project, then Extrapolation::Solve with
    // `steps` sweeps and its mask (fraction < 1 and a liquid cell on either
    // side), then collider.Enforce, with the velocity kept on the GPU between
    // them (one upload, one download). Not in the reference's interface.
    void ProjectExtrapolateEnforce(SGridData<double> &velocity, GridData<double> const &levelSet,
                                   Collider const &collider, double volError, int steps);

    // Iterations, FP64 relative residual and full-grid pressure (0 at air
    // cells) of the last solve; the reference keeps the pressure private (m_RdP).
    int LastIterations() const { return m_LastIters; }
    double LastResidual() const { return m_LastResidual; }
    std::vector<double> const &GetPressure() const { return m_Pressure; }

  private:
    void Solve(SGridData<double> &velocity, GridData<double> const &levelSet,
               Collider const &collider, double volError, bool reproject);
    // the jump callback on every ghost-fluid face into m_Jump; false: no callback
    bool EvaluateJumps(SGridData<double> const &velocity, GridData<double> const &levelSet,
                       Collider const &collider);

    StaggeredGrid const &m_SGrid;
    sfp_context_s *m_Ctx = nullptr;
    std::vector<double> m_Jump[3];
    std::vector<double> m_Pressure;
    int m_LastIters = 0;
    double m_LastResidual = 0;
    // Pressure jump: p_liquid - p_air
    std::function<double(int, Vector3i const &, double)> m_PressureJump = nullptr;
};

}  // namespace Pivot
