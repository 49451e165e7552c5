// Multigrid hierarchy and the FP64-outer flexible CG that replace amgcl's
// make_solver<amg<...>, cg/bicgstab> (core/Pressure.cpp:125-135,
// core/MagneticField.cpp:92-94,120-123).
#pragma once
#include "kernels.hpp"
#include "sfp_internal.hpp"
#include "unet.hpp"

#include <array>
#include <vector>

namespace sfp {

struct SolverConfig {
    double tol = 1e-8;
    int max_iters = 100;
    int precond = 0;  // sfp_precond
    int coarsest = 4;
    int cheby_degree = 3;
    double cheby_ratio = 4.0;
    double coarse_scale = 0.5;
    double stall_ratio = 0.5;
    int verbose = 0;
};

// Chebyshev recurrence coefficients for D^{-1}A on [lmax/ratio, lmax].
struct Cheby {
    int degree = 3;
    float cd[8] = {0}, cr[8] = {0};
    static Cheby make(int degree, double ratio);
};

// Timing of one tracked kernel: pairs of events recorded around each launch.
struct KernelTimer {
    std::vector<hipEvent_t> ev;
    std::vector<double> bytes;
    size_t used = 0;
    long long launches = 0;
    double total_ms = 0.0, total_bytes = 0.0;
    void begin(hipStream_t s, double b);
    void end(hipStream_t s);
    void harvest();  // after a stream sync: fold recorded pairs into the totals
    void reset();
    ~KernelTimer();
};

// Device workspace shared by the systems of one context.
struct Workspace {
    hipStream_t stream = nullptr;
    DevBuf<double> partials;
    DevBuf<KrylovScalars> scalars;
    DevBuf<int> err;
    KrylovScalars *host_sc = nullptr;  // pinned
    int *host_err = nullptr;           // pinned
    bool profiling = false;
    std::array<KernelTimer, K_COUNT> timers;
    // learned preconditioner: the context's network and the level set and
    // spacing of the current solve (for the operator features)
    UNet *net = nullptr;
    const double *phi = nullptr;
    double spacing = 1.0;
    void read_scalars();               // synchronous copy of *scalars to host_sc
};

struct SolveInfo {
    int iterations = 0;
    double rel = 0.0;
    bool converged = false;
    int fallback = 0;
    int precond_used = 0;
};

class System {
  public:
    void allocate(Dims d);
    Dims dims() const { return d_; }
    Level &fine() { return lv_[0]; }
    DevBuf<double> &rhs() { return rhs_; }
    DevBuf<double> &x() { return x_; }
    int num_levels() const { return (int)lv_.size(); }
    Level &level(int l) { return lv_[l]; }

    // Coarsen the fine operator into the hierarchy and invert the coarsest level.
    void build_hierarchy(Workspace &ws, const SolverConfig &cfg);
    // One preconditioner application z = B r (FP32 result, fine grid).
    float *precondition(Workspace &ws, const SolverConfig &cfg, const double *r64);
    // The learned preconditioner can run on this grid (each axis a multiple of 4).
    bool learned_ok(const Workspace &ws) const;
    // Network input features of the current operator (into the network input).
    void prepare_learned(Workspace &ws);
    float *precondition_learned(Workspace &ws, const double *r64);
    // Flexible CG on A x = rhs; x holds the start when warm.
    SolveInfo solve(Workspace &ws, const SolverConfig &cfg, bool warm);
    // Number of active rows (sets ws.scalars->pad[0]); synchronous.
    long long count_active(Workspace &ws);

  private:
    float *vlevel(Workspace &ws, const Cheby &cb, int l, const double *r64);
    Dims d_{};
    std::vector<Level> lv_;
    DevBuf<double> rhs_, x_, r_, p_, q_, w_;
    DevBuf<float> ainv_;
    int built_coarsest_ = -1;
};

void launch_count_active(hipStream_t s, Dims d, const float *diag32, double *partials,
                         KrylovScalars *sc);

}  // namespace sfp
