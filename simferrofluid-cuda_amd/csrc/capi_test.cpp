// Inspection hooks of include/sfp_test.h (operators, hierarchy, single kernels).
#include "../../include/sfp_test.h"
#include "context.hpp"

#include <vector>

using namespace sfp;

namespace {

template <class F>
sfp_status run(sfp_context ctx, F &&f) {
    if (!ctx) return SFP_ERR_ARGUMENT;
    try {
        SFP_HIP(hipSetDevice(ctx->device));
        f();
        return SFP_OK;
    } catch (const ArgError &e) {
        ctx->last_error = e.what();
        return SFP_ERR_ARGUMENT;
    } catch (const StateError &e) {
        ctx->last_error = e.what();
        return SFP_ERR_STATE;
    } catch (const HipError &e) {
        ctx->last_error = e.what();
        return SFP_ERR_HIP;
    } catch (const std::exception &e) {
        ctx->last_error = e.what();
        return SFP_ERR_ARGUMENT;
    }
}

// The configuration the system `which` solves with (the pressure system's
// own coarse smoother).
SolverConfig cfg_of(sfp_context ctx, int which) { return which == 0 ? ctx->cfg.for_pressure() : ctx->cfg; }

System &pick(sfp_context ctx, int which) {
    if (which == 0) {
        if (!ctx->pressure_ready || ctx->pressure_unknowns == 0)
            throw StateError("no pressure system assembled");
        return ctx->pressure;
    }
    if (which == 1) {
        if (!ctx->field_ready) throw StateError("no magnetic system assembled");
        return ctx->magnetic;
    }
    throw ArgError("which must be 0 (pressure) or 1 (magnetic)");
}

template <class T>
void d2h(sfp_context ctx, T *h, const T *d, size_t n) {
    if (!h) return;
    SFP_HIP(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, ctx->ws.stream));
}

}  // namespace

extern "C" {

sfp_status sfp_test_num_levels(sfp_context ctx, int which, int *levels) {
    return run(ctx, [&] {
        if (!levels) throw ArgError("null levels");
        *levels = pick(ctx, which).num_levels();
    });
}

sfp_status sfp_test_level_dims(sfp_context ctx, int which, int level, int dims[3]) {
    return run(ctx, [&] {
        System &s = pick(ctx, which);
        if (level < 0 || level >= s.num_levels() || !dims) throw ArgError("bad level");
        const Dims d = s.level(level).d;
        dims[0] = d.nx;
        dims[1] = d.ny;
        dims[2] = d.nz;
    });
}

sfp_status sfp_test_export_level(sfp_context ctx, int which, int level, double *diag,
                                 double *cpl_x, double *cpl_y, double *cpl_z, double *rhs) {
    return run(ctx, [&] {
        System &s = pick(ctx, which);
        if (level < 0 || level >= s.num_levels()) throw ArgError("bad level");
        Level &L = s.level(level);
        const size_t n = L.d.cells();
        d2h(ctx, diag, L.diag64.p, n);
        d2h(ctx, cpl_x, L.cpl64[0].p, n);
        d2h(ctx, cpl_y, L.cpl64[1].p, n);
        d2h(ctx, cpl_z, L.cpl64[2].p, n);
        if (rhs && level == 0) d2h(ctx, rhs, s.rhs().p, n);
        SFP_HIP(hipStreamSynchronize(ctx->ws.stream));
    });
}

sfp_status sfp_test_export_level32(sfp_context ctx, int which, int level, float *diag,
                                   float *cpl_x, float *cpl_y, float *cpl_z) {
    return run(ctx, [&] {
        System &s = pick(ctx, which);
        if (level < 0 || level >= s.num_levels()) throw ArgError("bad level");
        Level &L = s.level(level);
        const size_t n = L.d.cells();
        d2h(ctx, diag, L.diag32.p, n);
        d2h(ctx, cpl_x, L.cpl32[0].p, n);
        d2h(ctx, cpl_y, L.cpl32[1].p, n);
        d2h(ctx, cpl_z, L.cpl32[2].p, n);
        SFP_HIP(hipStreamSynchronize(ctx->ws.stream));
    });
}

sfp_status sfp_test_apply64(sfp_context ctx, int which, const float *z, double *w) {
    return run(ctx, [&] {
        System &s = pick(ctx, which);
        if (!z || !w) throw ArgError("null vector");
        const Dims d = s.dims();
        const size_t n = d.cells();
        DevBuf<float> dz;
        DevBuf<double> dr, dw;
        dz.ensure(n);
        dr.ensure(n);
        dw.ensure(n);
        SFP_HIP(hipMemcpyAsync(dz.p, z, n * sizeof(float), hipMemcpyHostToDevice, ctx->ws.stream));
        SFP_HIP(hipMemsetAsync(dr.p, 0, n * sizeof(double), ctx->ws.stream));
        launch_fcg_matvec(ctx->ws.stream, d, s.fine().op64(), nullptr, dz.p, dr.p, nullptr, dw.p, 1,
                          ctx->ws.partials.p, ctx->ws.scalars.p);
        d2h(ctx, w, dw.p, n);
        SFP_HIP(hipStreamSynchronize(ctx->ws.stream));
    });
}

sfp_status sfp_test_precondition(sfp_context ctx, int which, const double *r, float *z) {
    return run(ctx, [&] {
        System &s = pick(ctx, which);
        if (!r || !z) throw ArgError("null vector");
        const size_t n = s.dims().cells();
        DevBuf<double> dr;
        dr.ensure(n);
        SFP_HIP(hipMemcpyAsync(dr.p, r, n * sizeof(double), hipMemcpyHostToDevice, ctx->ws.stream));
        const float *dz = s.precondition(ctx->ws, cfg_of(ctx, which), dr.p);
        d2h(ctx, z, dz, n);
        SFP_HIP(hipStreamSynchronize(ctx->ws.stream));
    });
}

sfp_status sfp_test_residual(sfp_context ctx, int which, const double *x, double *r, double *rr,
                             double *bb) {
    return run(ctx, [&] {
        System &s = pick(ctx, which);
        if (!x) throw ArgError("null x");
        const Dims d = s.dims();
        const size_t n = d.cells();
        DevBuf<double> dx, dr;
        dx.ensure(n);
        dr.ensure(n);
        SFP_HIP(hipMemcpyAsync(dx.p, x, n * sizeof(double), hipMemcpyHostToDevice, ctx->ws.stream));
        launch_residual64(ctx->ws.stream, d, s.fine().op64(), s.rhs().p, dx.p, dr.p,
                          ctx->ws.partials.p, ctx->ws.scalars.p);
        d2h(ctx, r, dr.p, n);
        ctx->ws.read_scalars();
        if (rr) *rr = ctx->ws.host_sc->rr;
        if (bb) *bb = ctx->ws.host_sc->bb;
    });
}

sfp_status sfp_test_solve(sfp_context ctx, int which, int level, const double *rhs, double *x,
                          sfp_stats *stats) {
    return run(ctx, [&] {
        System &s = pick(ctx, which);
        if (!rhs || !x) throw ArgError("null vector");
        if (level < 0 || level >= s.num_levels()) throw ArgError("bad level");
        const size_t n = s.dims().cells(), nl = s.level(level).d.cells();
        const hipStream_t st = ctx->ws.stream;
        DevBuf<double> keep_b, keep_x;
        keep_b.ensure(n);
        keep_x.ensure(n);
        SFP_HIP(hipMemcpyAsync(keep_b.p, s.rhs().p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        SFP_HIP(hipMemcpyAsync(keep_x.p, s.x().p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        SFP_HIP(hipMemcpyAsync(s.rhs().p, rhs, nl * sizeof(double), hipMemcpyHostToDevice, st));
        SolveInfo info;
        try {
            info = level == 0 ? s.solve(ctx->ws, cfg_of(ctx, which), false)
                               : s.solve_level(ctx->ws, cfg_of(ctx, which), level);
        } catch (...) {
            (void)hipMemcpyAsync(s.rhs().p, keep_b.p, n * sizeof(double), hipMemcpyDeviceToDevice, st);
            (void)hipMemcpyAsync(s.x().p, keep_x.p, n * sizeof(double), hipMemcpyDeviceToDevice, st);
            (void)hipStreamSynchronize(st);
            throw;
        }
        d2h(ctx, x, s.x().p, nl);
        SFP_HIP(hipMemcpyAsync(s.rhs().p, keep_b.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        SFP_HIP(hipMemcpyAsync(s.x().p, keep_x.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        SFP_HIP(hipStreamSynchronize(st));
        if (stats) {
            *stats = sfp_stats{};
            stats->iterations = info.iterations;
            stats->converged = info.converged ? 1 : 0;
            stats->fallback = info.fallback;
            stats->precond_used = info.precond_used;
            stats->rel_residual = info.rel;
            stats->unknowns = (long long)nl;
        }
    });
}

sfp_status sfp_test_restrict_smoothed(sfp_context ctx, int which, const double *r, float *rc) {
    return run(ctx, [&] {
        System &s = pick(ctx, which);
        if (!r || !rc) throw ArgError("null vector");
        if (s.num_levels() < 2) throw StateError("the hierarchy has a single level");
        const size_t n = s.dims().cells();
        DevBuf<double> dr;
        dr.ensure(n);
        SFP_HIP(hipMemcpyAsync(dr.p, r, n * sizeof(double), hipMemcpyHostToDevice, ctx->ws.stream));
        const float *b = s.restrict_smoothed(ctx->ws, cfg_of(ctx, which), dr.p);
        d2h(ctx, rc, b, s.level(1).d.cells());
        SFP_HIP(hipStreamSynchronize(ctx->ws.stream));
    });
}

sfp_status sfp_test_smooth_pair(sfp_context ctx, int which, int level, const float *b,
                                const float *x1, const float *d1, float cdA, float crA,
                                float cdB, float crB, float *x3_pair, float *x3_steps) {
    return run(ctx, [&] {
        System &s = pick(ctx, which);
        if (level < 0 || level >= s.num_levels()) throw ArgError("bad level");
        if (!b || !x1) throw ArgError("null vector");
        const bool from_x1 = d1 == nullptr;  // d1 is x1 (stored first iterate)
        if (from_x1) d1 = x1;
        Level &L = s.level(level);
        const size_t n = L.d.cells();
        const hipStream_t st = ctx->ws.stream;
        DevBuf<float> db, dx1, dd, dx2, dx3;
        db.ensure(n);
        dx1.ensure(n);
        dd.ensure(n);
        dx2.ensure(n);
        dx3.ensure(n);
        SFP_HIP(hipMemcpyAsync(db.p, b, n * sizeof(float), hipMemcpyHostToDevice, st));
        SFP_HIP(hipMemcpyAsync(dx1.p, x1, n * sizeof(float), hipMemcpyHostToDevice, st));
        // the buffers the kernels skip on all-inactive rows hold zeros, as the
        // solver's level vectors do
        SFP_HIP(hipMemsetAsync(dx2.p, 0, n * sizeof(float), st));
        SFP_HIP(hipMemsetAsync(dx3.p, 0, n * sizeof(float), st));
        const Op32 A = L.op32();
        if (x3_pair) {
            SFP_HIP(hipMemcpyAsync(dd.p, d1, n * sizeof(float), hipMemcpyHostToDevice, st));
            launch_smooth_pair(st, L.d, A, db.p, dx1.p, dx3.p, from_x1 ? dx1.p : dd.p, cdA, crA, cdB, crB,
                               nullptr, L.rowact.p);
            d2h(ctx, x3_pair, dx3.p, n);
            SFP_HIP(hipStreamSynchronize(st));
        }
        if (x3_steps) {
            SFP_HIP(hipMemcpyAsync(dd.p, d1, n * sizeof(float), hipMemcpyHostToDevice, st));
            SFP_HIP(hipMemsetAsync(dx3.p, 0, n * sizeof(float), st));
            launch_smooth(st, (from_x1 ? SM_XFIRST : SM_READ_D) | SM_WRITE_D, L.d, A, nullptr, db.p, dx1.p,
                          dx2.p, dd.p, 0.f, cdA, crA, nullptr);
            launch_smooth(st, SM_READ_D, L.d, A, nullptr, db.p, dx2.p, dx3.p, dd.p, 0.f, cdB, crB,
                          nullptr);
            d2h(ctx, x3_steps, dx3.p, n);
            SFP_HIP(hipStreamSynchronize(st));
        }
    });
}

sfp_status sfp_test_smooth_restrict(sfp_context ctx, int which, int level, const float *b,
                                    const float *x1, const float *d1, float cd, float cr,
                                    float *x2_fused, float *bc_fused, float *x2_steps,
                                    float *bc_steps) {
    return run(ctx, [&] {
        System &s = pick(ctx, which);
        if (level < 0 || level + 1 >= s.num_levels()) throw ArgError("bad level");
        if (!b || !x1 || !d1) throw ArgError("null vector");
        Level &L = s.level(level);
        const Dims dc = s.level(level + 1).d;
        const size_t n = L.d.cells(), nc = dc.cells();
        const hipStream_t st = ctx->ws.stream;
        DevBuf<float> db, dx1, dd, dx2, dbc;
        db.ensure(n);
        dx1.ensure(n);
        dd.ensure(n);
        dx2.ensure(n);
        dbc.ensure(nc);
        SFP_HIP(hipMemcpyAsync(db.p, b, n * sizeof(float), hipMemcpyHostToDevice, st));
        SFP_HIP(hipMemcpyAsync(dx1.p, x1, n * sizeof(float), hipMemcpyHostToDevice, st));
        SFP_HIP(hipMemcpyAsync(dd.p, d1, n * sizeof(float), hipMemcpyHostToDevice, st));
        const Op32 A = L.op32();
        for (int fused = 1; fused >= 0; fused--) {
            float *xo = fused ? x2_fused : x2_steps;
            float *bo = fused ? bc_fused : bc_steps;
            if (!xo && !bo) continue;
            // zeros where the kernels skip all-inactive rows, as in the solver
            SFP_HIP(hipMemsetAsync(dx2.p, 0, n * sizeof(float), st));
            SFP_HIP(hipMemsetAsync(dbc.p, 0, nc * sizeof(float), st));
            if (fused) {
                launch_smooth_restrict(st, L.d, A, db.p, dx1.p, dx2.p, dd.p, cd, cr, dbc.p);
            } else {
                launch_smooth(st, SM_READ_D, L.d, A, nullptr, db.p, dx1.p, dx2.p, dd.p, 0.f, cd, cr,
                              nullptr);
                launch_restrict(st, L.d, A, db.p, dx2.p, dbc.p);
            }
            d2h(ctx, xo, dx2.p, n);
            d2h(ctx, bo, dbc.p, nc);
            SFP_HIP(hipStreamSynchronize(st));
        }
    });
}

sfp_status sfp_test_set_vcycle_tail(sfp_context ctx, int enable) {
    return run(ctx, [&] { ctx->ws.vcycle_tail = enable != 0; });
}

sfp_status sfp_test_set_fused_pairs(sfp_context ctx, int mask) {
    return run(ctx, [&] {
        if (mask < 0 || mask > 31) throw ArgError("fused pairs: mask is 0..31");
        ctx->ws.fused_pairs = mask;
    });
}

sfp_status sfp_test_set_rowsum_diagonal(sfp_context ctx, int enable) {
    return run(ctx, [&] { ctx->ws.rowsum_diag = enable != 0; });
}

}  // extern "C"
