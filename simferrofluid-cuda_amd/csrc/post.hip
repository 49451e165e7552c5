// The steps that follow the pressure projection in Simulation::ProjectVelocity
// (core/Simulation.cpp:236-241): velocity extrapolation into the air and the
// solid-wall boundary condition, on gfx950 in FP64.
//
// Built with -ffp-contract=off and written in the reference's evaluation
// order (neighbour ordinal 0..5 = -x,+x,-y,+y,-z,+z; TriLerp corner order and
// weight products as core/TriLerp.h:27-36; dot products summed x, y, z).
#include "../../include/sfp.h"
#include "device_util.hpp"
#include "kernels.hpp"

#include <mutex>

namespace sfp {
namespace {

__device__ __forceinline__ void coord_of(Dims g, size_t q, int c[3]) {
    c[2] = (int)(q % (size_t)g.nz);
    const size_t r = q / (size_t)g.nz;
    c[1] = (int)(r % (size_t)g.ny);
    c[0] = (int)(r / (size_t)g.ny);
}

// valid[f] = fraction < 1 and a liquid cell (phi <= 0) on either side of the
// face (core/Simulation.cpp:236-240); a side outside the cell grid counts as
// not liquid.
__global__ __launch_bounds__(kBlock) void k_velocity_valid(Dims d, int axis, const double *__restrict__ phi,
                                                           const double *__restrict__ frac,
                                                           unsigned char *__restrict__ valid) {
    const Dims fd = d.face(axis);
    const size_t q = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= fd.cells()) return;
    int f[3];
    coord_of(fd, q, f);
    bool v = false;
    if (frac[q] < 1) {
        // cell0 = face - e_axis, cell1 = face (StaggeredGrid::AdjCellOfFace)
        int c0[3] = {f[0], f[1], f[2]};
        c0[axis] -= 1;
        const bool in0 = c0[axis] >= 0;
        const bool in1 = f[axis] < d.n(axis);
        if (in0 && phi[d.idx(c0[0], c0[1], c0[2])] <= 0)
            v = true;
        else if (in1 && phi[d.idx(f[0], f[1], f[2])] <= 0)
            v = true;
    }
    valid[q] = v ? 1 : 0;
}

// One sweep of Extrapolation::Solve (core/Extrapolation.cpp:6-29): an invalid
// point with valid neighbours takes their mean and becomes valid, one
// without takes clear. Only invalid points are written and only valid ones
// read, so the sweep is free of races. Every entry of valid1 is written (a
// point that stays invalid gets 0, what the reference's copy of the mask holds
// there, since validity only grows), so valid1 needs no initial copy.
__global__ __launch_bounds__(kBlock) void k_extrapolate(Dims g, double *data,
                                                        const unsigned char *__restrict__ valid,
                                                        unsigned char *__restrict__ valid1,
                                                        double clear) {
    const size_t q = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= g.cells()) return;
    if (valid[q]) {
        valid1[q] = 1;
        return;
    }
    int c[3];
    coord_of(g, q, c);
    int cnt = 0;
    double sum = 0;
    for (int ord = 0; ord < 6; ord++) {
        int nb[3] = {c[0], c[1], c[2]};
        nb[ord >> 1] += (ord & 1) ? 1 : -1;
        if (nb[0] < 0 || nb[1] < 0 || nb[2] < 0 || nb[0] >= g.nx || nb[1] >= g.ny || nb[2] >= g.nz)
            continue;
        const size_t p = g.idx(nb[0], nb[1], nb[2]);
        if (valid[p]) {
            sum += data[p];
            cnt++;
        }
    }
    if (cnt > 0) {
        data[q] = sum / cnt;
        valid1[q] = 1;
    } else {
        data[q] = clear;
        valid1[q] = 0;
    }
}

struct FaceGeom {
    double origin[3][3];  // origin of face grid a
    double spacing, inv;
};

// TriLerp::Interpolate of one face grid at pos, with clamped reads
// (core/TriLerp.h:22-36,54-61, GridData::At).
__device__ double lerp_face(Dims g, const double org[3], double spacing, double inv,
                            const double *__restrict__ data, const double pos[3]) {
    int lo[3];
    double w0[3], w1[3];
    for (int a = 0; a < 3; a++) lo[a] = (int)floor((pos[a] - org[a]) * inv - 0.0);
    for (int a = 0; a < 3; a++) {
        const double fr = ((pos[a] - org[a]) - (double)lo[a] * spacing) * inv;
        w0[a] = 1. - fr;
        w1[a] = fr;
    }
    double val = 0;
    for (int p = 0; p < 8; p++) {
        const int ox = p & 1, oy = (p >> 1) & 1, oz = (p >> 2) & 1;
        int c[3] = {lo[0] + ox, lo[1] + oy, lo[2] + oz};
        for (int a = 0; a < 3; a++) c[a] = min(max(c[a], 0), g.n(a) - 1);
        const double wt = (ox ? w1[0] : w0[0]) * (oy ? w1[1] : w0[1]) * (oz ? w1[2] : w0[2]);
        val += data[g.idx(c[0], c[1], c[2])] * wt;
    }
    return val;
}

// Collider::Enforce on face grid `axis` (core/Collider.cpp:75-86): at every
// face with solid fraction exactly 1 the fluid velocity loses the normal
// component of its difference to the collider velocity; `old` is the
// velocity before the call, read through TriLerp.
__global__ __launch_bounds__(kBlock) void k_enforce(Dims d, int axis, FaceGeom geo, Faces old, Faces frac,
                                                    Faces normal, Faces colvel,
                                                    double *__restrict__ out) {
    const Dims fd = d.face(axis);
    const size_t q = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (q >= fd.cells()) return;
    if (!(frac.a[axis][q] == 1.)) return;
    int f[3];
    coord_of(fd, q, f);
    double pos[3];
    for (int a = 0; a < 3; a++) pos[a] = geo.origin[axis][a] + (double)f[a] * geo.spacing;
    double vel[3], vel0[3], n[3];
    for (int b = 0; b < 3; b++) {
        const Dims gb = d.face(b);
        vel[b] = lerp_face(gb, geo.origin[b], geo.spacing, geo.inv, old.a[b], pos);
        vel0[b] = lerp_face(gb, geo.origin[b], geo.spacing, geo.inv, colvel.a[b], pos);
        n[b] = lerp_face(gb, geo.origin[b], geo.spacing, geo.inv, normal.a[b], pos);
    }
    // Eigen's normalized(): unchanged when the squared norm is not positive
    const double z = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (z > 0) {
        const double r = sqrt(z);
        for (int b = 0; b < 3; b++) n[b] = n[b] / r;
    }
    const double dot = (vel[0] - vel0[0]) * n[0] + (vel[1] - vel0[1]) * n[1] + (vel[2] - vel0[2]) * n[2];
    out[q] = vel[axis] - dot * n[axis];
}

unsigned blocks_of(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

void launch_velocity_valid(hipStream_t s, Dims d, int axis, const double *phi, const double *frac,
                           unsigned char *valid) {
    hipLaunchKernelGGL(k_velocity_valid, dim3(blocks_of(d.face(axis).cells())), dim3(kBlock), 0, s, d,
                       axis, phi, frac, valid);
    SFP_LAUNCH_CHECK();
}

void launch_extrapolate(hipStream_t s, Dims g, double *data, unsigned char *valid,
                        unsigned char *scratch, double clear, int steps) {
    unsigned char *a = valid, *b = scratch;
    for (int it = 0; it < steps; it++) {
        hipLaunchKernelGGL(k_extrapolate, dim3(blocks_of(g.cells())), dim3(kBlock), 0, s, g, data, a, b,
                           clear);
        SFP_LAUNCH_CHECK();
        std::swap(a, b);
    }
}

void launch_collider_enforce(hipStream_t s, Dims d, double spacing, const double origin[3], Faces old,
                             Faces frac, Faces normal, Faces colvel, MutFaces vel) {
    FaceGeom geo;
    // StaggeredGrid face grid a: origin + (1 - e_a) * spacing / 2 (core/StaggeredGrid.h:15-19)
    for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) geo.origin[a][c] = origin[c] + ((c == a ? 0.0 : 1.0) * spacing) / 2;
    geo.spacing = spacing;
    geo.inv = 1 / spacing;
    for (int a = 0; a < 3; a++) {
        hipLaunchKernelGGL(k_enforce, dim3(blocks_of(d.face(a).cells())), dim3(kBlock), 0, s, d, a, geo, old,
                           frac, normal, colvel, vel.a[a]);
        SFP_LAUNCH_CHECK();
    }
}

}  // namespace sfp

// ---------------------------------------------------------------- without allocation
// sfp_extrapolate_velocity_noalloc: the mask and its sweep buffer live in one
// file-scope device array. Its host-side name is not a device address: each
// device's copy is found with hipGetSymbolAddress (with that device current)
// on the first call there, and only that address is passed to kernels.
namespace {

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;  // two masks of up to 32 Mi faces
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// This device's copy of `scratch` (looked up once per device), or null.
unsigned char *device_scratch(int dev, hipError_t *err) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (*err != hipSuccess) return nullptr;
        g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" sfp_status sfp_extrapolate_velocity_noalloc(int nx, int ny, int nz, double *const velocity[3],
                                                        const double *level_set,
                                                        const double *const fraction[3], int steps,
                                                        void *stream) {
    using namespace sfp;
    if (nx <= 0 || ny <= 0 || nz <= 0 || steps < 0) {
        set_unbound_error("grid extents must be positive and steps >= 0");
        return SFP_ERR_ARGUMENT;
    }
    const Dims d{nx, ny, nz};
    for (int a = 0; a < 3; a++) {
        const Dims f = d.face(a);
        if (2.0 * (double)f.nx * (double)f.ny * (double)f.nz > (double)kNoAllocScratchBytes) {
            set_unbound_error("grid too large for the fixed scratch of the no-allocation entry point "
                              "(two bytes per face of each face grid, 64 MiB)");
            return SFP_ERR_ARGUMENT;
        }
    }
    if (!velocity || !fraction || !level_set || !velocity[0] || !velocity[1] || !velocity[2] ||
        !fraction[0] || !fraction[1] || !fraction[2]) {
        set_unbound_error("velocity, level_set and fraction must be non-null device arrays");
        return SFP_ERR_ARGUMENT;
    }
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_unbound_error("no current HIP device");
        return SFP_ERR_NO_DEVICE;
    }
    unsigned char *base = device_scratch(dev, &e);
    if (!base) {
        set_unbound_error(hipGetErrorString(e));
        return SFP_ERR_HIP;
    }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    try {
        for (int a = 0; a < 3; a++) {
            const Dims fd = d.face(a);
            unsigned char *mask = base, *sweep = base + fd.cells();
            launch_velocity_valid(s, d, a, level_set, fraction[a], mask);
            launch_extrapolate(s, fd, velocity[a], mask, sweep, 0.0, steps);
        }
    } catch (const HipError &err) {
        set_unbound_error(err.what());
        return SFP_ERR_HIP;
    }
    return SFP_OK;
}
