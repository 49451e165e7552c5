// Hierarchy setup, V-cycle orchestration and the FP64 flexible CG.
#include "solver.hpp"

#include "learned.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace sfp {

void hip_fail(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    throw HipError(buf);
}

// ---------------------------------------------------------------- Chebyshev
Cheby Cheby::make(int degree, double ratio) {
    // D^{-1}A of these weakly diagonally dominant M-matrices has spectrum in
    // (0, 2] (Gershgorin), so lambda_max = 2 is a safe bound needing no estimate.
    Cheby c;
    c.degree = std::max(1, std::min(degree, 8));
    const double lmax = 2.0, lmin = lmax / ratio;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    c.cd[0] = 0.f;
    c.cr[0] = (float)(1.0 / theta);
    for (int s = 1; s < c.degree; s++) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        c.cd[s] = (float)(rn * rho);
        c.cr[s] = (float)(2.0 * rn / delta);
        rho = rn;
    }
    return c;
}

// ---------------------------------------------------------------- timers
void KernelTimer::begin(hipStream_t s, double b) {
    if (used + 2 > ev.size()) {
        for (int t = 0; t < 64; t++) {
            hipEvent_t e;
            SFP_HIP(hipEventCreate(&e));
            ev.push_back(e);
        }
    }
    if (bytes.size() < ev.size() / 2) bytes.resize(ev.size() / 2);
    bytes[used / 2] = b;
    SFP_HIP(hipEventRecord(ev[used], s));
}
void KernelTimer::end(hipStream_t s) {
    SFP_HIP(hipEventRecord(ev[used + 1], s));
    used += 2;
}
void KernelTimer::harvest() {
    // launches of a batch past its stop return at once: a duration under a
    // fifth of the median is such a launch, not a measurement of the kernel
    std::vector<float> ms(used / 2);
    for (size_t t = 0; t + 1 < used; t += 2) SFP_HIP(hipEventElapsedTime(&ms[t / 2], ev[t], ev[t + 1]));
    std::vector<float> sorted = ms;
    std::sort(sorted.begin(), sorted.end());
    const float med = sorted.empty() ? 0.f : sorted[sorted.size() / 2];
    for (size_t q = 0; q < ms.size(); q++) {
        if (ms[q] < 0.2f * med) continue;
        total_ms += ms[q];
        total_bytes += bytes[q];
        launches++;
    }
    used = 0;
}
void KernelTimer::reset() {
    used = 0;
    launches = 0;
    total_ms = total_bytes = 0.0;
}
KernelTimer::~KernelTimer() {
    for (auto e : ev) (void)hipEventDestroy(e);
}

void Workspace::read_scalars() {
    SFP_HIP(hipMemcpyAsync(host_sc, scalars.p, sizeof(KrylovScalars), hipMemcpyDeviceToHost, stream));
    SFP_HIP(hipStreamSynchronize(stream));
}

// ---------------------------------------------------------------- System
// The stencil kernels skip wavefront rows whose cells are all inactive; the
// skipped cells of the level vectors then keep the zeros written here.
static void zero_vectors(Level &L) {
    const size_t n = L.d.cells();
    for (auto *v : {&L.x0, &L.x1, &L.dv, &L.b}) SFP_HIP(hipMemset(v->p, 0, n * sizeof(float)));
}

void System::allocate(Dims d) {
    if (d_.nx == d.nx && d_.ny == d.ny && d_.nz == d.nz && !lv_.empty()) return;
    d_ = d;
    lv_.clear();
    lv_.emplace_back();
    Level &L = lv_[0];
    L.d = d;
    const size_t n = d.cells();
    L.diag64.ensure(n);
    L.diag32.ensure(n);
    for (int a = 0; a < 3; a++) {
        L.cpl64[a].ensure(n);
        L.cpl32[a].ensure(n);
    }
    L.x0.ensure(n);
    L.x1.ensure(n);
    L.dv.ensure(n);
    L.b.ensure(n);
    zero_vectors(L);
    rhs_.ensure(n);
    x_.ensure(n);
    r_.ensure(n);
    p_.ensure(n);
    q_.ensure(n);
    w_.ensure(n);
    for (auto *v : {&r_, &p_, &q_, &w_}) SFP_HIP(hipMemset(v->p, 0, n * sizeof(double)));
    SFP_HIP(hipMemset(x_.p, 0, n * sizeof(double)));
    built_coarsest_ = -1;
}

void System::join_ainv(Workspace &ws) {
    if (!ainv_pending_) return;
    SFP_HIP(hipStreamWaitEvent(ws.stream, ws.side_out, 0));
    ainv_pending_ = false;
}

void System::build_hierarchy(Workspace &ws, const SolverConfig &cfg) {
    // an inverse still being formed reads the coarsest operator rebuilt below
    join_ainv(ws);
    const int coarsest = std::max(2, std::min(cfg.coarsest, 4));
    if (built_coarsest_ != coarsest) {
        lv_.resize(1);
        while (true) {
            Dims dl = lv_.back().d;
            if (std::max({dl.nx, dl.ny, dl.nz}) <= coarsest) break;
            Level L;
            L.d = dl.coarse();
            const size_t n = L.d.cells();
            L.diag64.ensure(n);
            L.diag32.ensure(n);
            for (int a = 0; a < 3; a++) {
                L.cpl64[a].ensure(n);
                L.cpl32[a].ensure(n);
            }
            L.x0.ensure(n);
            L.x1.ensure(n);
            L.dv.ensure(n);
            L.b.ensure(n);
            zero_vectors(L);
            lv_.push_back(std::move(L));
        }
        if ((int)lv_.back().d.cells() > kMaxCoarse) throw ArgError("coarsest level too large");
        ainv_.ensure((size_t)lv_.back().d.cells() * lv_.back().d.cells());
        built_coarsest_ = coarsest;
    }
    for (size_t l = 1; l < lv_.size(); l++) {
        Level &F = lv_[l - 1];
        Level &C = lv_[l];
        double *c64[3] = {C.cpl64[0].p, C.cpl64[1].p, C.cpl64[2].p};
        float *c32[3] = {C.cpl32[0].p, C.cpl32[1].p, C.cpl32[2].p};
        launch_coarsen(ws.stream, F.d, F.op64(), F.diag32.p, C.d, cfg.coarse_scale,
                       pinned_ ? 1.0 : cfg.coarse_scale, C.diag64.p, c64, C.diag32.p, c32);
    }
    // The dense inverse of the coarsest level (a one-block kernel bound by its
    // pivot steps) runs on the side stream, overlapping the clears below, the
    // initial residual and the host's read of it; its first reader joins it.
    if (ws.side) {
        SFP_HIP(hipEventRecord(ws.side_in, ws.stream));
        SFP_HIP(hipStreamWaitEvent(ws.side, ws.side_in, 0));
        launch_coarse_invert(ws.side, lv_.back().d, lv_.back().op64(), ainv_.p);
        SFP_HIP(hipEventRecord(ws.side_out, ws.side));
        ainv_pending_ = true;
    } else {
        launch_coarse_invert(ws.stream, lv_.back().d, lv_.back().op64(), ainv_.p);
    }
    // the row-tiled step pairs skip the loads of z rows without active cells
    for (Level &L : lv_) {
        if (row_pair(L.d) && !L.pinned_rowsum) {
            L.rowact.ensure((size_t)L.d.nx * L.d.ny);
            launch_row_activity(ws.stream, L.d, L.diag32.p, L.rowact.p);
        }
    }
    // Every kernel skips groups of inactive cells (rows, tiles, pairs of
    // rows) and writes 0 at the inactive cells of the groups it processes
    // only if their inputs are 0 there; the groups differ between kernels.
    // A new operator may have a new active set, so the vectors are cleared at
    // the inactive cells (values left on cells that were active before would
    // otherwise leak into z, w, q and the inner products). Active cells are
    // written before they are read in every cycle; p and q are written at
    // every cell of the update's groups by the first iteration and read in
    // those groups only.
    if (lv_.size() > (size_t)kClearMaxLevels) throw ArgError("too many multigrid levels");
    ClearArgs ca;
    ca.nlev = (int)lv_.size();
    size_t total = 0;
    for (size_t l = 0; l < lv_.size(); l++) {
        Level &L = lv_[l];
        total += L.d.cells();
        if (total >= ((size_t)1 << 32)) throw ArgError("grid too large for the level clear");
        ca.n[l] = (unsigned)L.d.cells();
        ca.start[l + 1] = (unsigned)total;
        ca.diag32[l] = L.diag32.p;
        float *const v[4] = {L.x0.p, L.x1.p, L.dv.p, L.b.p};
        for (int k = 0; k < 4; k++) ca.v[l][k] = v[k];
    }
    ca.w = w_.p;
    launch_clear_inactive(ws.stream, ca);
}

float *System::vcycle_tail(Workspace &ws, const Smoothing &sm, int l) {
    const int nlev = (int)lv_.size() - l;
    if (!ws.vcycle_tail || nlev < 2 || nlev > kTailMaxLevels || lv_[l].d.cells() > kTailMaxCells)
        return nullptr;
    TailArgs t;
    t.nlev = nlev;
    for (int q = 0; q < nlev; q++) {
        Level &L = lv_[l + q];
        t.lv[q] = TailLevel{L.d, L.op32(), L.b.p, L.x0.p, L.x1.p, L.dv.p};
    }
    for (int set = 0; set < 2; set++) {
        const Cheby &cb = set == 0 ? sm.at(l) : sm.coarse;  // the first level's, the others'
        t.deg[set] = cb.degree;
        for (int q = 0; q < 8; q++) {
            t.cd[set][q] = cb.cd[q];
            t.cr[set][q] = cb.cr[q];
        }
    }
    t.ainv = ainv_.p;
    if (tail_lds_bytes(t) > kTailLdsBytes) return nullptr;
    join_ainv(ws);
    launch_vcycle_tail(ws.stream, t, ws.gate);
    return tail_result(t);
}

// Grids whose kernels are bound by launch latency rather than bytes.
static bool small_grid(Dims d) { return d.cells() <= ((size_t)1 << 19); }

// Algorithmic bytes per cell of one single Chebyshev step: x, b, diag and 3
// couplings in and x out, plus the direction in and out by mode and the
// coarse correction (an eighth of a value per cell). rowsum: the flat
// kernels rebuild the diagonal instead of reading it.
static double step_bytes(int mode, bool rowsum) {
    double b = 4.0 * (rowsum ? 6 : 7);
    if (mode & SM_READ_D) b += 4.0;
    if (mode & SM_WRITE_D) b += 4.0;
    if (mode & SM_PROLONG) b += 0.5;
    return b;
}

// x0_ready: the level's x0 holds c0 D^-1 b, stored by the restriction that
// formed its b (coarse levels) or by the FCG kernels (ws.x0_level).
float *System::vlevel(Workspace &ws, const Smoothing &sm, int l, const double *r64, bool x0_ready) {
    if (!r64) {
        if (float *z = vcycle_tail(ws, sm, l)) return z;
    }
    const Cheby &cb = sm.at(l);
    Level &L = lv_[l];
    const hipStream_t st = ws.stream;
    const bool fine = (l == 0);
    const double cells = (double)L.d.cells();
    if (l + 1 == (int)lv_.size()) {
        if (r64) launch_smooth_first(st, L.d, L.op32(), r64, L.b.p, L.x1.p, 0.f, ws.gate);
        join_ainv(ws);
        launch_coarse_solve(st, (int)L.d.cells(), ainv_.p, L.b.p, L.x0.p, ws.gate);
        return L.x0.p;
    }
    const Op32 A = L.op32();
    // the flat kernels (nz a multiple of 4) and the pair kernels rebuild a
    // row-sum diagonal instead of reading it
    const bool diag_rebuilt = L.pinned_rowsum && flat_step(L.d);
    const double diag_bytes = diag_rebuilt ? 0.0 : 4.0, pair_diag_bytes = L.pinned_rowsum ? 0.0 : 4.0;
    float *cur = L.x0.p, *oth = L.x1.p;
    const int deg = cb.degree;
    auto smooth = [&](int mode, float cd, float cr, const double *r, const float *e) {
        // timed: the fine-level single steps but the fused first step
        const bool track = fine && ws.profiling && !(mode & SM_FIRST);
        if (track) ws.timers[K_SMOOTH].begin(st, step_bytes(mode, diag_rebuilt) * cells);
        launch_smooth(st, mode, L.d, A, r, L.b.p, cur, oth, L.dv.p, cb.cr[0], cd, cr, e, ws.gate);
        if (track) ws.timers[K_SMOOTH].end(st);
        std::swap(cur, oth);
    };
    // the last step and the restriction run as one pass where the grid pays
    // for the pass's tile overlap
    const bool fuse_restrict =
        deg >= 3 && (((ws.fused_pairs & 1) && pair_pays(L.d)) || ((ws.fused_pairs & 4) && small_grid(L.d)));
    if (deg == 1) {
        launch_smooth_first(st, L.d, A, r64, L.b.p, cur, cb.cr[0], ws.gate);
    } else {
        // steps 0 and 1 in one pass (step 0 read from x0 when the FCG update
        // stored it); then steps 2 .. deg-1
        const int first = (!r64 && (x0_ready || ws.x0_level == l)) ? SM_XFIRST : SM_FIRST;
        // at degree 3 the two passes from the stored first iterate can run as
        // one: the iterate is also step 1's previous direction
        const bool pre_pair = deg == 3 && !fuse_restrict && first == SM_XFIRST &&
                              (ws.fused_pairs & 16) && pair_pays(L.d);
        if (pre_pair) {
            // reads x0, b, diag, 3 couplings; writes x
            if (fine && ws.profiling) ws.timers[K_SMOOTH_PAIR].begin(st, (24.0 + pair_diag_bytes) * cells);
            launch_smooth_pair(st, L.d, A, L.b.p, cur, oth, cur, cb.cd[1], cb.cr[1], cb.cd[2], cb.cr[2],
                               ws.gate, L.rowact.p);
            if (fine && ws.profiling) ws.timers[K_SMOOTH_PAIR].end(st);
            std::swap(cur, oth);
        } else {
            smooth(first | (deg > 2 ? SM_WRITE_D : 0), cb.cd[1], cb.cr[1], r64, nullptr);
            for (int s = 2; s < deg - (fuse_restrict ? 1 : 0); s++)
                smooth(SM_READ_D | (s < deg - 1 ? SM_WRITE_D : 0), cb.cd[s], cb.cr[s], nullptr, nullptr);
        }
    }
    bool store_x0 = false;  // the restriction stores the coarse level's first iterate
    if (fuse_restrict) {
        // reads x, b, d, diag, 3 couplings; writes x and the coarse b
        if (fine && ws.profiling) ws.timers[K_SMOOTH_RESTRICT].begin(st, (28.0 + pair_diag_bytes + 0.5) * cells);
        launch_smooth_restrict(st, L.d, A, L.b.p, cur, oth, L.dv.p, cb.cd[deg - 1], cb.cr[deg - 1],
                               lv_[l + 1].b.p, ws.gate);
        if (fine && ws.profiling) ws.timers[K_SMOOTH_RESTRICT].end(st);
        std::swap(cur, oth);
    } else {
        // the coarse level's first iterate too, where its cycle's first step reads it
        Level &C = lv_[l + 1];
        store_x0 = l + 2 < (int)lv_.size() && sm.at(l + 1).degree >= 2;
        // reads x, b, diag, 3 couplings; writes the coarse b (and reads the
        // coarse diagonal, writes the coarse x0): per fine cell
        if (fine && ws.profiling) ws.timers[K_RESTRICT].begin(st, (20.0 + diag_bytes + 0.5 + (store_x0 ? 1.0 : 0.0)) * cells);
        launch_restrict(st, L.d, A, L.b.p, cur, C.b.p, ws.gate, store_x0 ? C.x0.p : nullptr,
                        C.diag32.p, sm.at(l + 1).cr[0]);
        if (fine && ws.profiling) ws.timers[K_RESTRICT].end(st);
    }
    const float *e = vlevel(ws, sm, l + 1, nullptr, store_x0);
    return post_smooth(ws, sm, l, cur, e);
}

float *System::post_smooth(Workspace &ws, const Smoothing &sm, int l, float *cur, const float *e) {
    const Cheby &cb = sm.at(l);
    Level &L = lv_[l];
    const hipStream_t st = ws.stream;
    const Op32 A = L.op32();
    float *oth = cur == L.x0.p ? L.x1.p : L.x0.p;
    const int deg = cb.degree;
    const bool track = l == 0 && ws.profiling;
    // the last two steps run as one pass where the grid is large enough to
    // pay for the pass's tile overlap (the direction they leave is not stored:
    // nothing reads it before the next cycle rewrites it)
    const int fused = (deg >= 3 && (((ws.fused_pairs & 2) && pair_pays(L.d)) ||
                                    ((ws.fused_pairs & 8) && small_grid(L.d)))) ? 2 : 0;
    // step 0 adds the coarse correction: in the flat step's gathers, else on
    // large grids by its own pointwise pass before a plain step (cheaper than
    // the x-plane tiles' gathers at every neighbour); equal values every way
    const bool add_first = !flat_step(L.d) && pair_pays(L.d);
    if (add_first) launch_prolong_add(st, L.d, L.diag32.p, e, cur, ws.gate);
    for (int s = 0; s < deg - fused; s++) {
        // step 0 adds the prolonged correction; the fine-level steps are timed
        const int mode = (s == 0 ? (add_first ? 0 : SM_PROLONG) : SM_READ_D) | (s < deg - 1 ? SM_WRITE_D : 0);
        const bool t = track;
        if (t) ws.timers[K_SMOOTH].begin(st, step_bytes(mode, L.pinned_rowsum && flat_step(L.d)) * (double)L.d.cells());
        launch_smooth(st, mode, L.d, A, nullptr, L.b.p, cur, oth, L.dv.p, cb.cr[0], s == 0 ? 0.f : cb.cd[s],
                      s == 0 ? cb.cr[0] : cb.cr[s], s == 0 ? e : nullptr, ws.gate);
        if (t) ws.timers[K_SMOOTH].end(st);
        std::swap(cur, oth);
    }
    if (fused) {
        // reads x, b, d, diag (unless rebuilt), 3 couplings and writes x
        if (track) ws.timers[K_SMOOTH_PAIR].begin(st, (L.pinned_rowsum ? 28.0 : 32.0) * (double)L.d.cells());
        launch_smooth_pair(st, L.d, A, L.b.p, cur, oth, L.dv.p, cb.cd[deg - 2], cb.cr[deg - 2],
                           cb.cd[deg - 1], cb.cr[deg - 1], ws.gate, L.rowact.p);
        if (track) ws.timers[K_SMOOTH_PAIR].end(st);
        std::swap(cur, oth);
    }
    return cur;
}

float *System::precondition(Workspace &ws, const SolverConfig &cfg, const double *r64) {
    if (cfg.precond == 2) {  // SFP_PRECOND_JACOBI
        launch_jacobi(ws.stream, d_, lv_[0].op32(), r64, lv_[0].x0.p, ws.gate);
        return lv_[0].x0.p;
    }
    if (cfg.precond == 3 && lv_.size() > 1) return learned_cycle(ws, cfg, r64, kCycleSmooth);
    return vlevel(ws, Smoothing::make(cfg), 0, r64);
}

float *System::learned_cycle(Workspace &ws, const SolverConfig &cfg, const double *r64, Cycle c) {
    // the smoother the network was trained under
    const Smoothing sm = Smoothing::make_learned(cfg);
    Level &L = lv_[0], &C = lv_[1];
    const hipStream_t st = ws.stream;
    float *z = vlevel(ws, sm, 0, r64);
    // residual the V-cycle leaves, restricted to the first coarse level
    launch_restrict(st, L.d, L.op32(), L.b.p, z, C.b.p, ws.gate);
    if (c == kCycleRestrictOnly) return C.b.p;
    if (c == kCycleLearned) {
        network_correction(ws, C.x0.p);
    } else {  // the same cycle without the network: a zero correction
        SFP_HIP(hipMemsetAsync(C.x0.p, 0, C.d.cells() * sizeof(float), st));
    }
    return post_smooth(ws, sm, 0, z, C.x0.p);
}

bool System::learned_ok(const Workspace &ws) const {
    if (!ws.net || lv_.size() < 2) return false;
    const Dims c = lv_[1].d;
    return c.nx % 4 == 0 && c.ny % 4 == 0 && c.nz % 4 == 0;
}

void System::coarse_phi(Workspace &ws) {
    if (!ws.phi) return;
    phic_.ensure(lv_[1].d.cells());
    launch_restrict_phi(ws.stream, d_, ws.phi, lv_[1].d, phic_.p);
}

void System::prepare_learned(Workspace &ws) {
    const Level &C = lv_[1];
    ws.net->ensure(UNetShape{1, C.d.nx, C.d.ny, C.d.nz}, false);
    coarse_phi(ws);
    launch_learned_features(ws.stream, C.d, C.op64(), C.diag32.p, ws.phi ? phic_.p : nullptr,
                            2.0 * ws.spacing, ws.net->input_buffer(), nullptr);
}

void System::learned_features(Workspace &ws, float *f3) {
    if (lv_.size() < 2) throw StateError("the hierarchy has a single level");
    const Level &C = lv_[1];
    coarse_phi(ws);
    launch_learned_features(ws.stream, C.d, C.op64(), C.diag32.p, ws.phi ? phic_.p : nullptr,
                            2.0 * ws.spacing, nullptr, f3);
}

void System::network_correction(Workspace &ws, float *x) {
    Level &C = lv_[1];
    const hipStream_t st = ws.stream;
    UNet &net = *ws.net;
    launch_pack_residual32(st, C.d, C.b.p, C.diag32.p, ws.partials.p, ws.scalars.p, net.input_buffer(),
                           ws.gate);
    net.forward(st, net.input_buffer(), net.output_buffer(), ws.gate);
    launch_scaled_correction(st, C.d, C.op32(), net.output_buffer(), C.b.p, ws.partials.p,
                             ws.scalars.p, x, ws.gate);
}

float *System::precondition_learned(Workspace &ws, const SolverConfig &cfg, const double *r64) {
    return learned_cycle(ws, cfg, r64, kCycleLearned);
}

const float *System::restrict_smoothed(Workspace &ws, const SolverConfig &cfg, const double *r64) {
    if (lv_.size() < 2) throw StateError("the hierarchy has a single level");
    return learned_cycle(ws, cfg, r64, kCycleRestrictOnly);
}

long long System::count_active(Workspace &ws) {
    launch_count_active(ws.stream, d_, lv_[0].diag32.p, ws.partials.p, ws.scalars.p);
    ws.read_scalars();
    return (long long)ws.host_sc->pad[0];
}

SolveInfo System::fcg(Workspace &ws, const SolverConfig &cfg, int level, bool warm) {
    SolveInfo info;
    info.precond_used = cfg.precond;
    const hipStream_t st = ws.stream;
    const Dims d = lv_[level].d;
    const size_t n = d.cells();
    const Op64 A = lv_[level].op64();
    const bool fine = level == 0;
    const bool profiling = fine && ws.profiling;
    if (!warm) SFP_HIP(hipMemsetAsync(x_.p, 0, n * sizeof(double), st));
    // The V-cycle preconditioners start from the residual rounded to FP32 in
    // the level's b, with the cycle's first Chebyshev iterate in the level's
    // x0: the residual kernels and the FCG update store both, so no cycle
    // reads the FP64 residual (the Jacobi preconditioner reads r).
    float *b32 = cfg.precond != 2 ? lv_[level].b.p : nullptr;
    float *x0 = b32 ? lv_[level].x0.p : nullptr;
    // learned preconditioner until an iteration reduces the residual by less
    // than stall_ratio; then the plain V-cycle with the plain smoother
    bool learned = fine && cfg.precond == 1 && learned_ok(ws);
    // the first Chebyshev iterate the residual and update kernels store is
    // that of the cycle which reads it next
    const Smoothing plain = Smoothing::make(cfg), net_sm = Smoothing::make_learned(cfg);
    auto first_coef = [&](bool l) {
        return ((l || (fine && cfg.precond == 3)) ? net_sm : plain).at(level).cr[0];
    };
    float c0 = first_coef(learned);
    launch_residual64(st, d, A, rhs_.p, warm ? x_.p : nullptr, r_.p, ws.partials.p,
                      ws.scalars.p, b32, x0, lv_[level].diag32.p, c0);
    ws.read_scalars();
    const double bb = ws.host_sc->bb;
    if (bb == 0.0) {  // zero right-hand side: x = 0
        SFP_HIP(hipMemsetAsync(x_.p, 0, n * sizeof(double), st));
        info.converged = true;
        info.rel = 0.0;
        return info;
    }
    double rel = std::sqrt(ws.host_sc->rr / bb);
    bool first = true;
    bool checked = true;  // rel is a true residual
    const double cells = (double)n;
    if (learned) prepare_learned(ws);
    bool b_ready = b32 != nullptr;  // lv_[level].b holds r rounded (and lv_[level].x0 its first iterate)
    // iterations to enqueue before the host reads the residual; the first
    // learned iteration runs alone: it decides whether the network pays,
    // and iterations enqueued behind a stall would still launch every
    // kernel of the learned cycle (each returning at once)
    int next_batch = learned ? 1 : kBatch;
    // x is updated every other iteration (X_LATER, then X_BOTH): `run` counts
    // the updates since it was last current; an odd count leaves alpha p of
    // the last one pending, added before anything reads x
    int run = 0;
    auto bring_x_up_to_date = [&]() {
        if (run % 2) launch_fcg_flush(st, d, lv_[level].diag32.p, p_.p, x_.p, ws.scalars.p);
        run = 0;
    };
    info.precond_used = learned ? 1 : (cfg.precond == 1 ? 0 : cfg.precond);
    while (true) {
        if (rel <= cfg.tol) {
            if (checked) break;
            // recursive residual reached tol: recompute its norm from x in
            // FP64 (the sums only), and restart from the true residual if the
            // two have drifted apart
            bring_x_up_to_date();
            launch_residual64(st, d, A, rhs_.p, x_.p, nullptr, ws.partials.p, ws.scalars.p);
            ws.read_scalars();
            rel = std::sqrt(ws.host_sc->rr / bb);
            checked = true;
            if (rel > cfg.tol) {
                launch_residual64(st, d, A, rhs_.p, x_.p, r_.p, ws.partials.p,
                                  ws.scalars.p, b32, x0, lv_[level].diag32.p, c0);
                first = true;
                b_ready = b32 != nullptr;
            }
            continue;
        }
        if (info.iterations >= cfg.max_iters) break;
        // Up to kBatch iterations are enqueued before the host looks at the
        // residual (after the first batch, as many as the residual's rate of
        // decrease predicts are left): the update's finalize raises `stop` on
        // the device once ||r|| <= tol ||b||, or (learned preconditioner)
        // once an iteration reduced ||r|| by less than stall_ratio, and every
        // kernel of the remaining iterations of the batch then returns at
        // once (profiled launches among them are dropped from the timings).
        const int batch = std::min(next_batch, cfg.max_iters - info.iterations);
        if (batch > 1) {
            launch_batch_begin(st, ws.scalars.p, cfg.tol * cfg.tol * bb,
                               learned ? cfg.stall_ratio * cfg.stall_ratio : 0.0);
            ws.gate = &ws.scalars.p->stop;
        }
        for (int it = 0; it < batch; it++) {
            const double *r64 = b_ready ? nullptr : r_.p;  // null: the cycle reads b
            ws.x0_level = b_ready ? level : -1;
            const float *z;
            if (learned)
                z = precondition_learned(ws, cfg, r64);
            else if (fine)
                z = precondition(ws, cfg, r64);
            else
                z = vlevel(ws, Smoothing::make(cfg), level, r64);
            // z, the couplings, r, q_prev (after the first iteration) in; w
            // out; and the diagonal and the FP32 one (skips inactive cells)
            // unless the operator has the row-sum structure, where every cell
            // but cell 0 is active and the diagonal is rebuilt
            // (the timings cover the stencil kernels, not the finalize after them)
            hipEvent_t done = nullptr;
            if (profiling) {
                ws.timers[K_MATVEC64].begin(st, (4.0 + (A.pinned_rowsum ? 24.0 : 4.0 + 32.0) + 8.0 +
                                                 (first ? 0.0 : 8.0) + 8.0) * cells);
                done = ws.timers[K_MATVEC64].pending_end();
            }
            launch_fcg_matvec(st, d, A, A.pinned_rowsum ? nullptr : lv_[level].diag32.p, z, r_.p, q_.p,
                              w_.p, first ? 1 : 0, ws.partials.p, ws.scalars.p, ws.gate, done);
            if (profiling) ws.timers[K_MATVEC64].end_recorded();
            // z, w, p, q, r in; p, q, r out; x in and out every other
            // iteration (and diag in, b and x0 out)
            const int xmode = (run + it) % 2 == 0 ? X_LATER : X_BOTH;
            if (profiling) {
                ws.timers[K_UPDATE64].begin(st, (4.0 + 8.0 * 4 + 8.0 * 3 + (xmode == X_LATER ? 0.0 : 16.0) +
                                                 (b32 ? 4.0 + 4.0 + 4.0 : 0.0)) * cells);
                done = ws.timers[K_UPDATE64].pending_end();
            }
            ws.x0_level = -1;
            launch_fcg_update(st, d, lv_[level].diag32.p, z, w_.p, p_.p, q_.p, x_.p, r_.p, b32,
                              b32 ? lv_[level].x0.p : nullptr, c0, first ? 1 : 0, xmode, ws.partials.p,
                              ws.scalars.p, ws.gate, done);
            if (profiling) ws.timers[K_UPDATE64].end_recorded();
            first = false;
            b_ready = b32 != nullptr;
        }
        ws.read_scalars();
        const bool stalled = batch > 1 && ws.host_sc->stalled != 0.0;
        ws.gate = nullptr;
        const int done = batch > 1 ? (int)ws.host_sc->iters : 1;
        run += done;
        info.iterations += done;
        if (!learned && cfg.precond == 1 && info.precond_used == 1) info.fallback += done;
        const double prev = rel;
        rel = std::sqrt(ws.host_sc->rr / bb);
        checked = false;
        // the next batch: as many iterations as the last ones' mean residual
        // reduction predicts are left (iterations enqueued past convergence
        // still launch every kernel, which returns at once)
        next_batch = kBatch;
        if (done > 0 && rel > 0.0 && rel < prev) {
            const double rate = std::pow(rel / prev, 1.0 / done);
            const double left = std::log(cfg.tol / rel) / std::log(rate);
            if (left < kBatch) next_batch = std::max(1, (int)std::ceil(left));
        }
        if (learned && (stalled || (batch == 1 && rel > cfg.stall_ratio * prev))) {
            // fall back to the plain V-cycle, restarting the recurrence from
            // the current iterate (a direction built with the other
            // preconditioner slows the V-cycle's iterations down); the first
            // iterate stored by the last update belongs to the learned
            // cycle's smoother
            learned = false;
            bring_x_up_to_date();
            first = true;
            const float c0_plain = first_coef(false);
            if (c0_plain != c0) b_ready = false;
            c0 = c0_plain;
            next_batch = kBatch;
        }
    }
    bring_x_up_to_date();
    if (!checked) {
        launch_residual64(st, d, A, rhs_.p, x_.p, nullptr, ws.partials.p, ws.scalars.p);
        ws.read_scalars();
        rel = std::sqrt(ws.host_sc->rr / bb);
    }
    info.rel = rel;
    info.converged = rel <= cfg.tol;
    if (info.fallback > 0) info.precond_used = 0;
    if (profiling)
        for (auto &t : ws.timers) t.harvest();
    return info;
}

}  // namespace sfp
