// The C ABI (include/sfp.h): contexts, staging of host arrays, and the
// Pressure / MagneticField solve sequences on the device.
#include "context.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

using namespace sfp;

namespace {

thread_local std::string g_create_error;

size_t face_count(Dims d, int a) { return d.face(a).cells(); }

}  // namespace

// Error text of the calls without a context (sfp_last_error(NULL)).
void sfp::set_unbound_error(const char *msg) { g_create_error = msg; }

namespace {

sfp_params to_params(const SolverConfig &c) {
    sfp_params p;
    p.tol = c.tol;
    p.max_iters = c.max_iters;
    p.precond = c.precond;
    p.mg_coarsest = c.coarsest;
    p.cheby_degree = c.cheby_degree;
    p.cheby_ratio = c.cheby_ratio;
    p.coarse_scale = c.coarse_scale;
    p.stall_ratio = c.stall_ratio;
    p.verbose = c.verbose;
    p.coarse_cheby_degree = c.coarse_cheby_degree;
    p.coarse_cheby_ratio = c.coarse_cheby_ratio;
    p.net_cheby_degree = c.net_cheby_degree;
    p.net_cheby_ratio = c.net_cheby_ratio;
    p.net_coarse_cheby_degree = c.net_coarse_cheby_degree;
    p.net_coarse_cheby_ratio = c.net_coarse_cheby_ratio;
    p.magnetic_bc = c.magnetic_bc;
    p.pressure_coarse_cheby_degree = c.pressure_coarse_cheby_degree;
    p.pressure_coarse_cheby_ratio = c.pressure_coarse_cheby_ratio;
    return p;
}

template <class F>
sfp_status guarded(sfp_context ctx, F &&f) {
    if (!ctx) return SFP_ERR_ARGUMENT;
    try {
        SFP_HIP(hipSetDevice(ctx->device));
        f();
        ctx->last_error.clear();
        return SFP_OK;
    } catch (const ArgError &e) {
        ctx->last_error = e.what();
        return SFP_ERR_ARGUMENT;
    } catch (const InputError &e) {
        ctx->last_error = e.what();
        return SFP_ERR_INPUT;
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

void require3(const void *const p[3], const char *what) {
    if (!p || !p[0] || !p[1] || !p[2]) throw ArgError(std::string(what) + " must hold 3 non-null arrays");
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    SFP_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
}

void fill_stats(sfp_stats *st, const SolveInfo &info, long long unknowns, float setup, float solve,
                float apply) {
    if (!st) return;
    st->iterations = info.iterations;
    st->converged = info.converged ? 1 : 0;
    st->fallback = info.fallback;
    st->precond_used = info.precond_used;
    st->rel_residual = info.rel;
    st->setup_ms = setup;
    st->solve_ms = solve;
    st->apply_ms = apply;
    st->unknowns = unknowns;
}

void check_err(sfp_context ctx, const char *msg) {
    SFP_HIP(hipMemcpyAsync(ctx->ws.host_err, ctx->ws.err.p, sizeof(int), hipMemcpyDeviceToHost,
                           ctx->ws.stream));
    SFP_HIP(hipStreamSynchronize(ctx->ws.stream));
    if (*ctx->ws.host_err) {
        SFP_HIP(hipMemsetAsync(ctx->ws.err.p, 0, sizeof(int), ctx->ws.stream));
        SFP_HIP(hipStreamSynchronize(ctx->ws.stream));
        throw InputError(msg);
    }
}

// ---------------------------------------------------------------- pressure
void pressure_dev(sfp_context ctx, double *const vel[3], const double *phi,
                  const double *const frac[3], const double *const us[3],
                  const double *const jump[3], double vol_error, double *pressure, sfp_stats *st,
                  bool reproject) {
    require3((const void *const *)vel, "velocity");
    require3((const void *const *)frac, "fraction");
    require3((const void *const *)us, "solid_velocity");
    if (!phi) throw ArgError("level_set is null");
    if (jump) require3((const void *const *)jump, "jump");
    const hipStream_t s = ctx->ws.stream;
    const Dims d = ctx->d;
    System &sys = ctx->pressure;
    sys.allocate(d);
    ctx->liquid.ensure(d.cells());
    const Faces F{{frac[0], frac[1], frac[2]}}, V{{vel[0], vel[1], vel[2]}}, U{{us[0], us[1], us[2]}};
    Faces J{{nullptr, nullptr, nullptr}};
    if (jump) J = Faces{{jump[0], jump[1], jump[2]}};
    Level &L = sys.fine();
    SFP_HIP(hipEventRecord(ctx->ev[0], s));
    if (!reproject) {
        // the assembly overwrites the liquid mask and the fine operator: until
        // the new system is complete (checked, counted, coarsened) a
        // reprojection has nothing consistent to reuse
        ctx->pressure_ready = false;
        ctx->pressure_unknowns = 0;
        double *c64[3] = {L.cpl64[0].p, L.cpl64[1].p, L.cpl64[2].p};
        float *c32[3] = {L.cpl32[0].p, L.cpl32[1].p, L.cpl32[2].p};
        launch_pressure_assemble(s, d, phi, F, V, U, jump ? &J : nullptr, vol_error, L.diag64.p, c64,
                                 sys.rhs().p, L.diag32.p, c32, ctx->liquid.p, ctx->ws.err.p);
        check_err(ctx, "open (fraction < 1) face on the grid boundary next to a liquid cell");
        const long long unknowns = sys.count_active(ctx->ws);
        if (unknowns == 0) {  // Pressure::Project returns early
            ctx->pressure_ready = true;
            if (pressure) SFP_HIP(hipMemsetAsync(pressure, 0, d.cells() * sizeof(double), s));
            SFP_HIP(hipStreamSynchronize(s));
            fill_stats(st, SolveInfo{0, 0.0, true, 0, ctx->cfg.precond}, 0, 0.f, 0.f, 0.f);
            return;
        }
        sys.build_hierarchy(ctx->ws, ctx->cfg.for_pressure());
        ctx->pressure_unknowns = unknowns;
        ctx->pressure_ready = true;
    } else {
        if (!ctx->pressure_ready) throw StateError("reproject before project");
        if (ctx->pressure_unknowns == 0) {
            fill_stats(st, SolveInfo{0, 0.0, true, 0, ctx->cfg.precond}, 0, 0.f, 0.f, 0.f);
            return;
        }
        launch_pressure_rhs(s, d, ctx->liquid.p, F, V, U, vol_error, sys.rhs().p);
    }
    SFP_HIP(hipEventRecord(ctx->ev[1], s));
    ctx->ws.phi = phi;
    ctx->ws.spacing = ctx->spacing;
    SolveInfo info = sys.solve(ctx->ws, ctx->cfg.for_pressure(), reproject);
    SFP_HIP(hipEventRecord(ctx->ev[2], s));
    MutFaces Vm{{vel[0], vel[1], vel[2]}};
    launch_pressure_apply(s, d, phi, F, jump ? &J : nullptr, ctx->liquid.p, sys.x().p, Vm,
                          ctx->ws.err.p);
    if (pressure)
        SFP_HIP(hipMemcpyAsync(pressure, sys.x().p, d.cells() * sizeof(double),
                               hipMemcpyDeviceToDevice, s));
    SFP_HIP(hipEventRecord(ctx->ev[3], s));
    check_err(ctx, "open (fraction < 1) face on the grid boundary next to a liquid cell");
    fill_stats(st, info, ctx->pressure_unknowns, elapsed(ctx->ev[0], ctx->ev[1]),
               elapsed(ctx->ev[1], ctx->ev[2]), elapsed(ctx->ev[2], ctx->ev[3]));
}

// ---------------------------------------------------------------- magnetic
void magnetic_dev(sfp_context ctx, const double *phi, const double *const hext[3], double chi,
                  double band, double *potential, double *const htot_out[3], sfp_stats *st) {
    require3((const void *const *)hext, "external_field");
    if (!phi) throw ArgError("level_set is null");
    if (!(band > 0)) throw ArgError("band_width must be positive");
    const hipStream_t s = ctx->ws.stream;
    const Dims d = ctx->d;
    System &sys = ctx->magnetic;
    sys.allocate(d);
    for (int a = 0; a < 3; a++) {
        ctx->mu[a].ensure(face_count(d, a));
        ctx->htot[a].ensure(face_count(d, a));
    }
    Level &L = sys.fine();
    SFP_HIP(hipEventRecord(ctx->ev[0], s));
    MutFaces M{{ctx->mu[0].p, ctx->mu[1].p, ctx->mu[2].p}};
    launch_magnetic_mu(s, d, ctx->mg, phi, chi, band, M);
    const Faces Mf{{ctx->mu[0].p, ctx->mu[1].p, ctx->mu[2].p}};
    const Faces H{{hext[0], hext[1], hext[2]}};
    double *c64[3] = {L.cpl64[0].p, L.cpl64[1].p, L.cpl64[2].p};
    float *c32[3] = {L.cpl32[0].p, L.cpl32[1].p, L.cpl32[2].p};
    const int dirichlet = ctx->cfg.magnetic_bc == SFP_MAGNETIC_BC_DIRICHLET_BOX;
    launch_magnetic_assemble(s, d, Mf, H, ctx->spacing, L.diag64.p, c64, sys.rhs().p, L.diag32.p,
                             c32, dirichlet);
    // the row-sum diagonal holds for the reference's boundary only
    L.pinned_rowsum = ctx->ws.rowsum_diag && !dirichlet;
    sys.pinned_ = !dirichlet;
    sys.build_hierarchy(ctx->ws, ctx->cfg);
    SFP_HIP(hipEventRecord(ctx->ev[1], s));
    ctx->ws.phi = phi;
    ctx->ws.spacing = ctx->spacing;
    SolveInfo info = sys.solve(ctx->ws, ctx->cfg, ctx->magnetic_warm);
    ctx->magnetic_warm = true;
    SFP_HIP(hipEventRecord(ctx->ev[2], s));
    MutFaces T{{ctx->htot[0].p, ctx->htot[1].p, ctx->htot[2].p}};
    launch_magnetic_total_field(s, d, ctx->mg.inv, sys.x().p, H, T);
    ctx->field_ready = true;
    if (potential)
        SFP_HIP(hipMemcpyAsync(potential, sys.x().p, d.cells() * sizeof(double),
                               hipMemcpyDeviceToDevice, s));
    if (htot_out)
        for (int a = 0; a < 3; a++)
            if (htot_out[a])
                SFP_HIP(hipMemcpyAsync(htot_out[a], ctx->htot[a].p, face_count(d, a) * sizeof(double),
                                       hipMemcpyDeviceToDevice, s));
    SFP_HIP(hipEventRecord(ctx->ev[3], s));
    SFP_HIP(hipStreamSynchronize(s));
    fill_stats(st, info, (long long)d.cells() - (dirichlet ? 0 : 1), elapsed(ctx->ev[0], ctx->ev[1]),
               elapsed(ctx->ev[1], ctx->ev[2]), elapsed(ctx->ev[2], ctx->ev[3]));
}

// host staging helpers
const double *up(sfp_context ctx, DevBuf<double> &buf, const double *h, size_t n) {
    if (!h) return nullptr;
    buf.ensure(n);
    SFP_HIP(hipMemcpyAsync(buf.p, h, n * sizeof(double), hipMemcpyHostToDevice, ctx->ws.stream));
    return buf.p;
}
void down(sfp_context ctx, double *h, const double *dptr, size_t n) {
    SFP_HIP(hipMemcpyAsync(h, dptr, n * sizeof(double), hipMemcpyDeviceToHost, ctx->ws.stream));
}

sfp_status pressure_host(sfp_context ctx, double *const vel[3], const double *phi,
                         const double *const frac[3], const double *const us[3],
                         const double *const jump[3], double vol_error, double *pressure,
                         sfp_stats *st, bool reproject) {
    return guarded(ctx, [&] {
        require3((const void *const *)vel, "velocity");
        require3((const void *const *)frac, "fraction");
        require3((const void *const *)us, "solid_velocity");
        if (jump) require3((const void *const *)jump, "jump");
        if (!phi) throw ArgError("level_set is null");
        const Dims d = ctx->d;
        double *dv[3];
        const double *df[3], *du[3], *dj[3];
        for (int a = 0; a < 3; a++) {
            const size_t n = face_count(d, a);
            dv[a] = const_cast<double *>(up(ctx, ctx->s_face[0][a], vel[a], n));
            df[a] = up(ctx, ctx->s_face[1][a], frac[a], n);
            du[a] = up(ctx, ctx->s_face[2][a], us[a], n);
            dj[a] = jump ? up(ctx, ctx->s_face[3][a], jump[a], n) : nullptr;
        }
        const double *dphi = up(ctx, ctx->s_phi, phi, d.cells());
        double *dp = nullptr;
        if (pressure) {
            ctx->s_cell.ensure(d.cells());
            dp = ctx->s_cell.p;
        }
        pressure_dev(ctx, dv, dphi, df, du, jump ? dj : nullptr, vol_error, dp, st, reproject);
        for (int a = 0; a < 3; a++) down(ctx, vel[a], dv[a], face_count(d, a));
        if (pressure) down(ctx, pressure, dp, d.cells());
        SFP_HIP(hipStreamSynchronize(ctx->ws.stream));
    });
}

// ---------------------------------------------------------------- post-projection
void check_steps(int steps) {
    if (steps < 0) throw ArgError("steps must be >= 0");
}

void extrapolate_velocity_dev(sfp_context ctx, double *const vel[3], const double *phi,
                              const double *const frac[3], int steps) {
    require3((const void *const *)vel, "velocity");
    require3((const void *const *)frac, "fraction");
    if (!phi) throw ArgError("level_set is null");
    check_steps(steps);
    const hipStream_t s = ctx->ws.stream;
    const Dims d = ctx->d;
    for (int a = 0; a < 3; a++) {
        const Dims fd = d.face(a);
        ctx->post_valid[0][a].ensure(fd.cells());
        ctx->post_valid[1][a].ensure(fd.cells());
        launch_velocity_valid(s, d, a, phi, frac[a], ctx->post_valid[0][a].p);
        launch_extrapolate(s, fd, vel[a], ctx->post_valid[0][a].p, ctx->post_valid[1][a].p, 0.0, steps);
    }
    SFP_HIP(hipStreamSynchronize(s));
}

void collider_enforce_dev(sfp_context ctx, double *const vel[3], const double *const frac[3],
                          const double *const normal[3], const double *const colvel[3]) {
    require3((const void *const *)vel, "velocity");
    require3((const void *const *)frac, "fraction");
    require3((const void *const *)normal, "normal");
    require3((const void *const *)colvel, "collider_velocity");
    const hipStream_t s = ctx->ws.stream;
    const Dims d = ctx->d;
    for (int a = 0; a < 3; a++) {
        const size_t n = face_count(d, a);
        ctx->post_old[a].ensure(n);
        SFP_HIP(hipMemcpyAsync(ctx->post_old[a].p, vel[a], n * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    const Faces old{{ctx->post_old[0].p, ctx->post_old[1].p, ctx->post_old[2].p}};
    launch_collider_enforce(s, d, ctx->spacing, ctx->origin, old, Faces{{frac[0], frac[1], frac[2]}},
                            Faces{{normal[0], normal[1], normal[2]}},
                            Faces{{colvel[0], colvel[1], colvel[2]}}, MutFaces{{vel[0], vel[1], vel[2]}});
    SFP_HIP(hipStreamSynchronize(s));
}

}  // namespace

extern "C" {

sfp_status sfp_extrapolate_faces(sfp_context ctx, double *const data[3],
                                 const unsigned char *const valid[3], double clear_value,
                                 int steps) {
    return guarded(ctx, [&] {
        require3((const void *const *)data, "data");
        require3((const void *const *)valid, "valid");
        check_steps(steps);
        const hipStream_t s = ctx->ws.stream;
        const Dims d = ctx->d;
        for (int a = 0; a < 3; a++) {
            const Dims fd = d.face(a);
            const size_t n = fd.cells();
            double *dd = const_cast<double *>(up(ctx, ctx->s_face[0][a], data[a], n));
            ctx->post_valid[0][a].ensure(n);
            ctx->post_valid[1][a].ensure(n);
            SFP_HIP(hipMemcpyAsync(ctx->post_valid[0][a].p, valid[a], n, hipMemcpyHostToDevice, s));
            launch_extrapolate(s, fd, dd, ctx->post_valid[0][a].p, ctx->post_valid[1][a].p, clear_value,
                               steps);
            down(ctx, data[a], dd, n);
        }
        SFP_HIP(hipStreamSynchronize(s));
    });
}

sfp_status sfp_extrapolate_velocity(sfp_context ctx, double *const velocity[3],
                                    const double *level_set, const double *const fraction[3],
                                    int steps) {
    return guarded(ctx, [&] {
        require3((const void *const *)velocity, "velocity");
        require3((const void *const *)fraction, "fraction");
        if (!level_set) throw ArgError("level_set is null");
        const Dims d = ctx->d;
        double *dv[3];
        const double *df[3];
        for (int a = 0; a < 3; a++) {
            const size_t n = face_count(d, a);
            dv[a] = const_cast<double *>(up(ctx, ctx->s_face[0][a], velocity[a], n));
            df[a] = up(ctx, ctx->s_face[1][a], fraction[a], n);
        }
        const double *dphi = up(ctx, ctx->s_phi, level_set, d.cells());
        extrapolate_velocity_dev(ctx, dv, dphi, df, steps);
        for (int a = 0; a < 3; a++) down(ctx, velocity[a], dv[a], face_count(d, a));
        SFP_HIP(hipStreamSynchronize(ctx->ws.stream));
    });
}

sfp_status sfp_collider_enforce(sfp_context ctx, double *const velocity[3],
                                const double *const fraction[3], const double *const normal[3],
                                const double *const collider_velocity[3]) {
    return guarded(ctx, [&] {
        require3((const void *const *)velocity, "velocity");
        require3((const void *const *)fraction, "fraction");
        require3((const void *const *)normal, "normal");
        require3((const void *const *)collider_velocity, "collider_velocity");
        const Dims d = ctx->d;
        double *dv[3];
        const double *df[3], *dn[3], *dc[3];
        for (int a = 0; a < 3; a++) {
            const size_t n = face_count(d, a);
            dv[a] = const_cast<double *>(up(ctx, ctx->s_face[0][a], velocity[a], n));
            df[a] = up(ctx, ctx->s_face[1][a], fraction[a], n);
            dn[a] = up(ctx, ctx->s_face[2][a], normal[a], n);
            dc[a] = up(ctx, ctx->s_face[3][a], collider_velocity[a], n);
        }
        collider_enforce_dev(ctx, dv, df, dn, dc);
        for (int a = 0; a < 3; a++) down(ctx, velocity[a], dv[a], face_count(d, a));
        SFP_HIP(hipStreamSynchronize(ctx->ws.stream));
    });
}

sfp_status sfp_projection_substep(sfp_context ctx, double *const velocity[3], const double *level_set,
                                  const double *const fraction[3],
                                  const double *const solid_velocity[3],
                                  const double *const jump[3], double vol_error, int steps,
                                  const double *const normal[3],
                                  const double *const collider_velocity[3], double *pressure,
                                  sfp_stats *stats) {
    return guarded(ctx, [&] {
        require3((const void *const *)velocity, "velocity");
        require3((const void *const *)fraction, "fraction");
        require3((const void *const *)solid_velocity, "solid_velocity");
        require3((const void *const *)normal, "normal");
        require3((const void *const *)collider_velocity, "collider_velocity");
        if (jump) require3((const void *const *)jump, "jump");
        if (!level_set) throw ArgError("level_set is null");
        check_steps(steps);
        const Dims d = ctx->d;
        // one upload of every input, the three passes on the device, one
        // download of the velocity (and of the pressure when asked)
        double *dv[3];
        const double *df[3], *du[3], *dj[3], *dn[3], *dc[3];
        for (int a = 0; a < 3; a++) {
            const size_t n = face_count(d, a);
            dv[a] = const_cast<double *>(up(ctx, ctx->s_face[0][a], velocity[a], n));
            df[a] = up(ctx, ctx->s_face[1][a], fraction[a], n);
            du[a] = up(ctx, ctx->s_face[2][a], solid_velocity[a], n);
            dj[a] = jump ? up(ctx, ctx->s_face[3][a], jump[a], n) : nullptr;
            dn[a] = up(ctx, ctx->s_face[4][a], normal[a], n);
            dc[a] = up(ctx, ctx->s_face[5][a], collider_velocity[a], n);
        }
        const double *dphi = up(ctx, ctx->s_phi, level_set, d.cells());
        double *dp = nullptr;
        if (pressure) {
            ctx->s_cell.ensure(d.cells());
            dp = ctx->s_cell.p;
        }
        pressure_dev(ctx, dv, dphi, df, du, jump ? dj : nullptr, vol_error, dp, stats, false);
        extrapolate_velocity_dev(ctx, dv, dphi, df, steps);
        collider_enforce_dev(ctx, dv, df, dn, dc);
        for (int a = 0; a < 3; a++) down(ctx, velocity[a], dv[a], face_count(d, a));
        if (pressure) down(ctx, pressure, dp, d.cells());
        SFP_HIP(hipStreamSynchronize(ctx->ws.stream));
    });
}

sfp_status sfp_extrapolate_velocity_device(sfp_context ctx, double *const velocity[3],
                                           const double *level_set,
                                           const double *const fraction[3], int steps) {
    return guarded(ctx, [&] { extrapolate_velocity_dev(ctx, velocity, level_set, fraction, steps); });
}

sfp_status sfp_collider_enforce_device(sfp_context ctx, double *const velocity[3],
                                       const double *const fraction[3],
                                       const double *const normal[3],
                                       const double *const collider_velocity[3]) {
    return guarded(ctx, [&] { collider_enforce_dev(ctx, velocity, fraction, normal, collider_velocity); });
}

const char *sfp_status_string(sfp_status s) {
    switch (s) {
        case SFP_OK: return "ok";
        case SFP_ERR_ARGUMENT: return "invalid argument";
        case SFP_ERR_INPUT: return "input outside what the reference handles";
        case SFP_ERR_HIP: return "HIP runtime error";
        case SFP_ERR_NO_DEVICE: return "no usable GPU";
        case SFP_ERR_STATE: return "call out of order";
        case SFP_ERR_UNSUPPORTED: return "unsupported";
    }
    return "unknown status";
}

void sfp_default_params(sfp_params *p) {
    if (p) *p = to_params(SolverConfig{});
}

sfp_status sfp_create(sfp_context *out, int nx, int ny, int nz, double spacing,
                      const double origin[3], int device) {
    if (!out) return SFP_ERR_ARGUMENT;
    *out = nullptr;
    if (nx <= 0 || ny <= 0 || nz <= 0 || !(spacing > 0) || !origin) {
        g_create_error = "grid extents and spacing must be positive, origin non-null";
        return SFP_ERR_ARGUMENT;
    }
    if ((long long)nx * ny * nz >= (1LL << 31)) {
        g_create_error = "grid too large: the reference indexes cells with int";
        return SFP_ERR_ARGUMENT;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        g_create_error = "no HIP device";
        return SFP_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= count) {
        g_create_error = "device index out of range";
        return SFP_ERR_NO_DEVICE;
    }
    sfp_context ctx = new sfp_context_s;
    ctx->device = device;
    ctx->d = Dims{nx, ny, nz};
    ctx->spacing = spacing;
    for (int a = 0; a < 3; a++) ctx->origin[a] = origin[a];
    ctx->mg = mag_geometry(ctx->d, spacing, origin);
    try {
        SFP_HIP(hipSetDevice(device));
        SFP_HIP(hipStreamCreateWithFlags(&ctx->ws.stream, hipStreamNonBlocking));
        SFP_HIP(hipStreamCreateWithFlags(&ctx->ws.side, hipStreamNonBlocking));
        SFP_HIP(hipEventCreateWithFlags(&ctx->ws.side_in, hipEventDisableTiming));
        SFP_HIP(hipEventCreateWithFlags(&ctx->ws.side_out, hipEventDisableTiming));
        for (auto &e : ctx->ev) SFP_HIP(hipEventCreate(&e));
        ctx->ws.partials.ensure(reduction_partials(ctx->d));
        ctx->ws.scalars.ensure(1);
        ctx->ws.err.ensure(1);
        SFP_HIP(hipMemset(ctx->ws.scalars.p, 0, sizeof(KrylovScalars)));
        SFP_HIP(hipMemset(ctx->ws.err.p, 0, sizeof(int)));
        SFP_HIP(hipHostMalloc((void **)&ctx->ws.host_sc, sizeof(KrylovScalars), hipHostMallocDefault));
        SFP_HIP(hipHostMalloc((void **)&ctx->ws.host_err, sizeof(int), hipHostMallocDefault));
    } catch (const std::exception &e) {
        g_create_error = e.what();
        sfp_destroy(ctx);
        return SFP_ERR_HIP;
    }
    *out = ctx;
    return SFP_OK;
}

void sfp_destroy(sfp_context ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->ws.stream) (void)hipStreamSynchronize(ctx->ws.stream);
    if (ctx->ws.side) (void)hipStreamSynchronize(ctx->ws.side);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {ctx->ws.side_in, ctx->ws.side_out})
        if (e) (void)hipEventDestroy(e);
    if (ctx->ws.host_sc) (void)hipHostFree(ctx->ws.host_sc);
    if (ctx->ws.host_err) (void)hipHostFree(ctx->ws.host_err);
    hipStream_t s = ctx->ws.stream, side = ctx->ws.side;
    ctx->ws.stream = nullptr;
    ctx->ws.side = nullptr;
    delete ctx;  // frees device buffers
    if (s) (void)hipStreamDestroy(s);
    if (side) (void)hipStreamDestroy(side);
}

sfp_status sfp_set_params(sfp_context ctx, const sfp_params *p) {
    if (!ctx || !p) return SFP_ERR_ARGUMENT;
    if (!(p->tol > 0) || p->max_iters < 0 || p->precond < 0 || p->precond > 3 ||
        p->mg_coarsest < 2 || p->mg_coarsest > 4 || p->cheby_degree < 1 || p->cheby_degree > 8 ||
        !(p->cheby_ratio > 1) || !(p->coarse_scale > 0) || !(p->stall_ratio > 0) ||
        p->coarse_cheby_degree < 1 || p->coarse_cheby_degree > 8 || !(p->coarse_cheby_ratio > 1) ||
        p->net_cheby_degree < 0 || p->net_cheby_degree > 8 || p->net_coarse_cheby_degree < 0 ||
        p->net_coarse_cheby_degree > 8 || !(p->net_cheby_ratio == 0 || p->net_cheby_ratio > 1) ||
        !(p->net_coarse_cheby_ratio == 0 || p->net_coarse_cheby_ratio > 1) ||
        (p->magnetic_bc != SFP_MAGNETIC_BC_REFERENCE && p->magnetic_bc != SFP_MAGNETIC_BC_DIRICHLET_BOX) ||
        p->pressure_coarse_cheby_degree < 0 || p->pressure_coarse_cheby_degree > 8 ||
        !(p->pressure_coarse_cheby_ratio == 0 || p->pressure_coarse_cheby_ratio > 1)) {
        ctx->last_error = "parameter out of range";
        return SFP_ERR_ARGUMENT;
    }
    if (p->precond == SFP_PRECOND_LEARNED && !ctx->net) {
        ctx->last_error = "learned preconditioner needs a network (sfp_set_network)";
        return SFP_ERR_UNSUPPORTED;
    }
    SolverConfig &c = ctx->cfg;
    c.tol = p->tol;
    c.max_iters = p->max_iters;
    c.precond = p->precond;
    c.coarsest = p->mg_coarsest;
    c.cheby_degree = p->cheby_degree;
    c.cheby_ratio = p->cheby_ratio;
    c.coarse_scale = p->coarse_scale;
    c.stall_ratio = p->stall_ratio;
    c.verbose = p->verbose;
    c.coarse_cheby_degree = p->coarse_cheby_degree;
    c.coarse_cheby_ratio = p->coarse_cheby_ratio;
    c.net_cheby_degree = p->net_cheby_degree;
    c.net_cheby_ratio = p->net_cheby_ratio;
    c.net_coarse_cheby_degree = p->net_coarse_cheby_degree;
    c.net_coarse_cheby_ratio = p->net_coarse_cheby_ratio;
    c.magnetic_bc = p->magnetic_bc;
    c.pressure_coarse_cheby_degree = p->pressure_coarse_cheby_degree;
    c.pressure_coarse_cheby_ratio = p->pressure_coarse_cheby_ratio;
    return SFP_OK;
}

sfp_status sfp_get_params(sfp_context ctx, sfp_params *p) {
    if (!ctx || !p) return SFP_ERR_ARGUMENT;
    *p = to_params(ctx->cfg);
    return SFP_OK;
}

const char *sfp_last_error(sfp_context ctx) {
    return ctx ? ctx->last_error.c_str() : g_create_error.c_str();
}

void *sfp_stream(sfp_context ctx) { return ctx ? (void *)ctx->ws.stream : nullptr; }

sfp_status sfp_reset_warm_start(sfp_context ctx) {
    return guarded(ctx, [&] { ctx->magnetic_warm = false; });
}

sfp_status sfp_pressure_project(sfp_context ctx, double *const velocity[3],
                                const double *level_set, const double *const fraction[3],
                                const double *const solid_velocity[3],
                                const double *const jump[3], double vol_error, double *pressure,
                                sfp_stats *stats) {
    return pressure_host(ctx, velocity, level_set, fraction, solid_velocity, jump, vol_error,
                         pressure, stats, false);
}

sfp_status sfp_pressure_reproject(sfp_context ctx, double *const velocity[3],
                                  const double *level_set, const double *const fraction[3],
                                  const double *const solid_velocity[3],
                                  const double *const jump[3], double vol_error,
                                  double *pressure, sfp_stats *stats) {
    return pressure_host(ctx, velocity, level_set, fraction, solid_velocity, jump, vol_error,
                         pressure, stats, true);
}

sfp_status sfp_pressure_project_device(sfp_context ctx, double *const velocity[3],
                                       const double *level_set, const double *const fraction[3],
                                       const double *const solid_velocity[3],
                                       const double *const jump[3], double vol_error,
                                       double *pressure, sfp_stats *stats) {
    return guarded(ctx, [&] {
        pressure_dev(ctx, velocity, level_set, fraction, solid_velocity, jump, vol_error, pressure,
                     stats, false);
    });
}

sfp_status sfp_pressure_reproject_device(sfp_context ctx, double *const velocity[3],
                                         const double *level_set,
                                         const double *const fraction[3],
                                         const double *const solid_velocity[3],
                                         const double *const jump[3], double vol_error,
                                         double *pressure, sfp_stats *stats) {
    return guarded(ctx, [&] {
        pressure_dev(ctx, velocity, level_set, fraction, solid_velocity, jump, vol_error, pressure,
                     stats, true);
    });
}

sfp_status sfp_magnetic_solve(sfp_context ctx, const double *level_set,
                              const double *const external_field[3], double susceptibility,
                              double band_width, double *potential, double *const total_field[3],
                              sfp_stats *stats) {
    return guarded(ctx, [&] {
        require3((const void *const *)external_field, "external_field");
        if (!level_set) throw ArgError("level_set is null");
        const Dims d = ctx->d;
        const double *dh[3];
        for (int a = 0; a < 3; a++) dh[a] = up(ctx, ctx->s_face[4][a], external_field[a], face_count(d, a));
        const double *dphi = up(ctx, ctx->s_phi, level_set, d.cells());
        double *dpot = nullptr;
        if (potential) {
            ctx->s_cell.ensure(d.cells());
            dpot = ctx->s_cell.p;
        }
        magnetic_dev(ctx, dphi, dh, susceptibility, band_width, dpot, nullptr, stats);
        if (potential) down(ctx, potential, dpot, d.cells());
        if (total_field)
            for (int a = 0; a < 3; a++)
                if (total_field[a]) down(ctx, total_field[a], ctx->htot[a].p, face_count(d, a));
        SFP_HIP(hipStreamSynchronize(ctx->ws.stream));
    });
}

sfp_status sfp_magnetic_solve_device(sfp_context ctx, const double *level_set,
                                     const double *const external_field[3],
                                     double susceptibility, double band_width, double *potential,
                                     double *const total_field[3], sfp_stats *stats) {
    return guarded(ctx, [&] {
        magnetic_dev(ctx, level_set, external_field, susceptibility, band_width, potential,
                     total_field, stats);
    });
}

sfp_status sfp_magnetic_vertex_pressures(sfp_context ctx, int nverts, const double *positions,
                                         const double *normals, double susceptibility,
                                         double *pressures) {
    return guarded(ctx, [&] {
        if (nverts < 0) throw ArgError("nverts must be >= 0");
        if (!ctx->field_ready) throw StateError("vertex pressures before a magnetic solve");
        if (nverts == 0) return;
        if (!positions || !normals || !pressures) throw ArgError("null vertex array");
        const size_t n3 = 3 * (size_t)nverts;
        const double *dp = up(ctx, ctx->s_vert[0], positions, n3);
        const double *dn = up(ctx, ctx->s_vert[1], normals, n3);
        ctx->s_vert[2].ensure(nverts);
        const Faces T{{ctx->htot[0].p, ctx->htot[1].p, ctx->htot[2].p}};
        launch_vertex_pressures(ctx->ws.stream, ctx->d, ctx->mg, T, susceptibility, nverts, dp, dn,
                                ctx->s_vert[2].p);
        down(ctx, pressures, ctx->s_vert[2].p, nverts);
        SFP_HIP(hipStreamSynchronize(ctx->ws.stream));
    });
}

// ---------------------------------------------------------------- profiling
sfp_status sfp_profile_enable(sfp_context ctx, int enable) {
    if (!ctx) return SFP_ERR_ARGUMENT;
    ctx->ws.profiling = enable != 0;
    for (auto &t : ctx->ws.timers) t.reset();
    return SFP_OK;
}

sfp_status sfp_profile_query(sfp_context ctx, int kernel, sfp_kernel_profile *out) {
    if (!ctx || !out || kernel < 0 || kernel >= K_COUNT) return SFP_ERR_ARGUMENT;
    const KernelTimer &t = ctx->ws.timers[kernel];
    out->launches = t.launches;
    out->total_ms = t.total_ms;
    out->bytes = t.total_bytes;
    return SFP_OK;
}

}  // extern "C"
