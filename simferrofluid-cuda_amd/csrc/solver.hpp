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

// FCG iterations enqueued between two reads of the residual on the host.
constexpr int kBatch = 4;

struct SolverConfig {
    double tol = 1e-8;
    int max_iters = 100;
    int precond = 0;  // sfp_precond: 0 MG, 1 learned, 2 Jacobi, 3 smoothing only
    int coarsest = 4;
    int cheby_degree = 3;
    double cheby_ratio = 4.0;
    int coarse_cheby_degree = 5;  // the smoother below the fine level
    double coarse_cheby_ratio = 6.0;
    double coarse_scale = 0.5;
    double stall_ratio = 0.5;
    int verbose = 0;
    // the learned cycle's smoother (0: the plain cycle's value)
    int net_cheby_degree = 0;
    double net_cheby_ratio = 0.0;
    int net_coarse_cheby_degree = 0;
    double net_coarse_cheby_ratio = 0.0;
    // magnetic box condition: 0 the reference's (Neumann box faces, cell 0
    // pinned), 1 phi = 0 on the box faces
    int magnetic_bc = 0;
    // the pressure system's coarse-level smoother (0: coarse_cheby_*): its
    // iteration counts do not gain from the magnetic system's stronger one
    int pressure_coarse_cheby_degree = 3;
    double pressure_coarse_cheby_ratio = 4.0;
    // this configuration as the pressure system runs it
    SolverConfig for_pressure() const {
        SolverConfig c = *this;
        if (pressure_coarse_cheby_degree > 0) c.coarse_cheby_degree = pressure_coarse_cheby_degree;
        if (pressure_coarse_cheby_ratio > 0) c.coarse_cheby_ratio = pressure_coarse_cheby_ratio;
        return c;
    }
};

// Chebyshev recurrence coefficients for D^{-1}A on [lmax/ratio, lmax].
struct Cheby {
    int degree = 3;
    float cd[8] = {0}, cr[8] = {0};
    static Cheby make(int degree, double ratio);
};
// The smoother of each level: `fine` on level 0, `coarse` below it.
struct Smoothing {
    Cheby fine, coarse;
    const Cheby &at(int l) const { return l == 0 ? fine : coarse; }
    static Smoothing make(const SolverConfig &cfg) {
        return Smoothing{Cheby::make(cfg.cheby_degree, cfg.cheby_ratio),
                         Cheby::make(cfg.coarse_cheby_degree, cfg.coarse_cheby_ratio)};
    }
    // the learned cycle's: the settings its network was trained under
    static Smoothing make_learned(const SolverConfig &cfg) {
        auto pick = [](double v, double plain) { return v > 0 ? v : plain; };
        return Smoothing{
            Cheby::make((int)pick(cfg.net_cheby_degree, cfg.cheby_degree),
                        pick(cfg.net_cheby_ratio, cfg.cheby_ratio)),
            Cheby::make((int)pick(cfg.net_coarse_cheby_degree, cfg.coarse_cheby_degree),
                        pick(cfg.net_coarse_cheby_ratio, cfg.coarse_cheby_ratio))};
    }
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
    // the end event of the open pair, for a launcher to record itself (after
    // its kernel, before a follow-up launch); then end_recorded()
    hipEvent_t pending_end() const { return ev[used + 1]; }
    void end_recorded() { used += 2; }
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
    // &scalars->stop while a batched solve runs (iteration kernels return at
    // once when it is nonzero), null otherwise
    const double *gate = nullptr;
    // run the V-cycle's small levels (<= kTailMaxCells cells) in one launch
    bool vcycle_tail = true;
    // run as one pass on large grids: bit 0, the last pre-smoothing step with
    // the restriction; bit 1, the last two post-smoothing steps (the flat step
    // and restriction beat the first pass on MI355X); bits 2 and 3: the same
    // on small grids, where launches cost more than the bytes they move; bit
    // 4: the first two pre-smoothing passes (degree 3, from the stored
    // first iterate) on large grids
    int fused_pairs = 2 | 16;
    // the stencil kernels rebuild a row-sum diagonal (Level::pinned_rowsum)
    // from the couplings instead of reading it
    bool rowsum_diag = true;
    // level whose x0 holds the first Chebyshev iterate of its b (stored by
    // the FCG update for the next cycle), or -1
    int x0_level = -1;
    // a second stream (default priority) on which the coarsest level's dense
    // inverse is formed while `stream` runs the rest of the set-up; side_in
    // orders it after the coarsening, side_out orders its first use after it
    hipStream_t side = nullptr;
    hipEvent_t side_in = nullptr, side_out = nullptr;
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
    // One preconditioner application z = B r (FP32 result, fine grid): the
    // V-cycle, Jacobi, or (cfg.precond 3) the learned cycle with a zero
    // correction in place of the network's.
    float *precondition(Workspace &ws, const SolverConfig &cfg, const double *r64);
    // The learned preconditioner can run on this hierarchy: a network is
    // loaded and the first coarse level has every axis a multiple of 4.
    bool learned_ok(const Workspace &ws) const;
    // Network input features of the first coarse level (into the network input).
    void prepare_learned(Workspace &ws);
    // The learned cycle: the V-cycle, then the residual it leaves restricted
    // to the first coarse level, the network's correction there (scaled to
    // be optimal in that level's energy norm), prolongation and fine
    // post-smoothing.
    float *precondition_learned(Workspace &ws, const SolverConfig &cfg, const double *r64);
    // The network's input of the learned cycle: the first coarse level's
    // right-hand side after the V-cycle of r64 (FP32).
    const float *restrict_smoothed(Workspace &ws, const SolverConfig &cfg, const double *r64);
    // Features of the first coarse level into f3 ([cells][3] FP32).
    void learned_features(Workspace &ws, float *f3);
    // Flexible CG on A x = rhs; x holds the start when warm.
    SolveInfo solve(Workspace &ws, const SolverConfig &cfg, bool warm) {
        return fcg(ws, cfg, 0, warm);
    }
    // Flexible CG with level `level`'s FP64 operator and the V-cycle from that
    // level down, on the first cells of rhs/x (from x = 0).
    SolveInfo solve_level(Workspace &ws, const SolverConfig &cfg, int level) {
        return fcg(ws, cfg, level, false);
    }
    // Number of active rows (sets ws.scalars->pad[0]); synchronous.
    long long count_active(Workspace &ws);
    // The operator is a Neumann problem made definite by one eliminated cell
    // (the magnetic potential's pin at cell 0). Its coarse levels then scale
    // only the couplings by coarse_scale and keep each aggregate's row sum
    // (the pin's eliminated couplings, in the aggregate holding cell 0, and
    // rounding elsewhere) unscaled: the pin's conductance is that of one
    // fine cell on every level. Scaled with the couplings it would halve per
    // level, the coarse problems would approach the singular Neumann one,
    // and the V-cycle would overshoot the near-constant mode.
    bool pinned_ = false;

  private:
    enum Cycle { kCycleLearned, kCycleSmooth, kCycleRestrictOnly };
    float *learned_cycle(Workspace &ws, const SolverConfig &cfg, const double *r64, Cycle c);
    float *post_smooth(Workspace &ws, const Smoothing &sm, int l, float *cur, const float *e);
    SolveInfo fcg(Workspace &ws, const SolverConfig &cfg, int level, bool warm);
    float *vlevel(Workspace &ws, const Smoothing &sm, int l, const double *r64, bool x0_ready = false);
    // levels l .. coarsest in one launch (vlevel with r64 null), or null
    // where the tail does not apply
    float *vcycle_tail(Workspace &ws, const Smoothing &sm, int l);
    void network_correction(Workspace &ws, float *x);
    void coarse_phi(Workspace &ws);
    DevBuf<double> phic_;  // level set restricted to the first coarse level
    Dims d_{};
    std::vector<Level> lv_;
    DevBuf<double> rhs_, x_, r_, p_, q_, w_;
    DevBuf<float> ainv_;
    // the dense inverse is being formed on ws.side: join before reading it
    // (or before the coarsest operator it reads is rewritten)
    bool ainv_pending_ = false;
    void join_ainv(Workspace &ws);
    int built_coarsest_ = -1;
};

void launch_count_active(hipStream_t s, Dims d, const float *diag32, double *partials,
                         KrylovScalars *sc);

}  // namespace sfp
