// Assembly of the two systems, the ghost-fluid velocity update and the
// magnetic field update, on gfx950 in FP64.
//
// Built with -ffp-contract=off and written in the reference's evaluation
// order (face ordinal 0..5 = -x,+x,-y,+y,-z,+z; left-to-right products), so
// diagonals, couplings and right-hand sides are bit-identical to the
// reference's CSR assembly; only sin() (device libm vs host libm) can differ
// in the last bits inside the magnetic smoothing band.
#include "device_util.hpp"
#include "kernels.hpp"

namespace sfp {

MagGeom mag_geometry(Dims d, double spacing, const double origin[3]) {
    // MagneticField's StaggeredGrid is built from the simulation grid's centre
    // (core/MagneticField.cpp:11-16); its origin is that centre minus half the
    // extent (core/StaggeredGrid.h:10), which can differ from `origin` by an ulp.
    MagGeom g;
    const int n[3] = {d.nx, d.ny, d.nz};
    double ext_origin[3];
    for (int a = 0; a < 3; a++) {
        const double center = origin[a] + (double)n[a] * spacing / 2.;
        ext_origin[a] = center - (double)n[a] * spacing / 2;
    }
    for (int f = 0; f < 3; f++)
        for (int a = 0; a < 3; a++)
            g.face_origin[f][a] = ext_origin[a] + (a == f ? 0.0 : 1.0) * spacing / 2;
    for (int a = 0; a < 3; a++) g.ls_origin[a] = origin[a] + spacing / 2;
    g.spacing = spacing;
    g.inv = 1 / spacing;
    return g;
}

// x planes per thread of the assembly, field and update kernels (4 measured
// best on MI355X at 256^3: pressure assembly 383 -> 364 us, total field 87 ->
// 78 us per face grid against 16 planes)
constexpr int kAssemblyPlanes = 4;

namespace {

// std::max(theta, .001) including its NaN behaviour (core/Pressure.cpp:62,161)
__device__ __forceinline__ double max_001(double t) { return (t < .001) ? .001 : t; }

struct Coord {
    int v[3];
};

__device__ __forceinline__ bool valid(Dims d, const int c[3]) {
    return c[0] >= 0 && c[1] >= 0 && c[2] >= 0 && c[0] < d.nx && c[1] < d.ny && c[2] < d.nz;
}

// TriLerp::Interpolate with clamped reads (core/TriLerp.h:22-36,54-61).
template <bool kSkipZero = false>
__device__ double trilerp(Dims g, const double org[3], double spacing, double inv,
                          const double *__restrict__ data, const double pos[3]) {
    int lo[3];
    double w0[3], w1[3];
    for (int a = 0; a < 3; a++) {
        const double t = (pos[a] - org[a]) * inv - 0.0;
        lo[a] = (int)floor(t);
    }
    for (int a = 0; a < 3; a++) {
        const double fr = ((pos[a] - org[a]) - (double)lo[a] * spacing) * inv;
        w0[a] = 1. - fr;
        w1[a] = fr;
    }
    // kSkipZero: a corner of weight exactly 0 adds a zero (the data are
    // finite), which leaves the sum unchanged but for the sign of a zero sum;
    // it is not loaded (at face centres of the cell grid, 6 of the 8 corners).
    double val = 0;
#pragma unroll
    for (int p = 0; p < 8; p++) {
        const int ox = p & 1, oy = (p >> 1) & 1, oz = (p >> 2) & 1;
        int c[3] = {lo[0] + ox, lo[1] + oy, lo[2] + oz};
        for (int a = 0; a < 3; a++) c[a] = min(max(c[a], 0), g.n(a) - 1);
        const double wt = (ox ? w1[0] : w0[0]) * (oy ? w1[1] : w0[1]) * (oz ? w1[2] : w0[2]);
        if (!kSkipZero || wt != 0.) val += data[g.idx(c[0], c[1], c[2])] * wt;
    }
    return val;
}

// MagneticField::Heaviside (core/MagneticField.h:44-54)
__device__ __forceinline__ double heaviside(double phi, double band) {
    const double pi = 3.141592653589793238462643383279502884;
    if (phi < -band) return 0.;
    if (phi > band) return 1.;
    return .5 * (1. + phi / band + sin(pi * phi / band) / pi);
}

__global__ __launch_bounds__(kBlock) void k_pressure_assemble(CellGrid cg, 
    Dims d, const double *__restrict__ phi, Faces frac, Faces vel, Faces us, Faces jump,
    int has_jump, double vol_error, double *__restrict__ diag64, MutFaces cpl64,
    double *__restrict__ rhs, float *__restrict__ diag32, float *cpl32x, float *cpl32y,
    float *cpl32z, unsigned char *__restrict__ liquid, int *err) {
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    if (k >= d.nz || j >= d.ny) return;
    float *cpl32[3] = {cpl32x, cpl32y, cpl32z};
    for (int i = i0, ie = min(i0 + cg.planes, d.nx); i < ie; i++) {
        const size_t c = d.idx(i, j, k);
        const double pc = phi[c];
        const bool liq = pc <= 0;
        liquid[c] = liq ? 1 : 0;
        double low[3] = {0.0, 0.0, 0.0};
        if (!liq) {
            diag64[c] = 1.0;
            rhs[c] = 0.0;
            diag32[c] = 0.f;
        } else {
            const int cell[3] = {i, j, k};
            // every operand of the six faces is loaded before any is used:
            // face by face, each face's loads would wait on the one before
            double fr[6], pn[6], uv[6], usv[6], jv[6];
            bool nbv[6];
#pragma unroll
            for (int ord = 0; ord < 6; ord++) {
                const int axis = ord >> 1, side = (ord & 1) ? 1 : -1;
                int nb[3] = {cell[0], cell[1], cell[2]};
                nb[axis] += side;
                int face[3] = {cell[0], cell[1], cell[2]};
                face[axis] += (ord & 1);
                const size_t fidx = d.face(axis).idx(face[0], face[1], face[2]);
                nbv[ord] = valid(d, nb);
                fr[ord] = frac.a[axis][fidx];
                pn[ord] = nbv[ord] ? phi[d.idx(nb[0], nb[1], nb[2])] : 0.0;
                uv[ord] = vel.a[axis][fidx];
                usv[ord] = us.a[axis][fidx];
                jv[ord] = has_jump ? jump.a[axis][fidx] : 0.0;
            }
            double diag = 0;
            double div = 0;
#pragma unroll
            for (int ord = 0; ord < 6; ord++) {
                const int axis = ord >> 1, side = (ord & 1) ? 1 : -1;
                const double weight = 1 - fr[ord];
                if (weight > 0) {
                    if (!nbv[ord]) {
                        atomicExch(err, 1);
                    } else {
                        if (pn[ord] <= 0) {
                            diag += weight;
                            if (side < 0) low[axis] = weight;
                        } else {
                            const double theta = pc / (pc - pn[ord]);
                            const double ic = 1. / max_001(theta);
                            diag += weight * ic;
                            if (has_jump) div += jv[ord] * weight * ic;
                        }
                    }
                    div -= (double)side * weight * uv[ord];
                }
                if (weight < 1) div -= (double)side * (1 - weight) * usv[ord];
            }
            div += vol_error;
            rhs[c] = div;
            const double dg = diag != 0 ? diag : 1.;
            diag64[c] = dg;
            diag32[c] = (float)dg;
        }
        for (int a = 0; a < 3; a++) {
            cpl64.a[a][c] = low[a];
            cpl32[a][c] = (float)low[a];
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_pressure_rhs(CellGrid cg, Dims d,
                                                         const unsigned char *__restrict__ liquid,
                                                         Faces frac, Faces vel, Faces us,
                                                         double vol_error,
                                                         double *__restrict__ rhs) {
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    if (k >= d.nz || j >= d.ny) return;
    for (int i = i0, ie = min(i0 + cg.planes, d.nx); i < ie; i++) {
        const size_t c = d.idx(i, j, k);
        if (!liquid[c]) {
            rhs[c] = 0.0;
            continue;
        }
        const int cell[3] = {i, j, k};
        double div = 0;
        for (int ord = 0; ord < 6; ord++) {
            const int axis = ord >> 1, side = (ord & 1) ? 1 : -1;
            int face[3] = {cell[0], cell[1], cell[2]};
            face[axis] += (ord & 1);
            const size_t fidx = d.face(axis).idx(face[0], face[1], face[2]);
            const double weight = 1 - frac.a[axis][fidx];
            if (weight > 0) div -= (double)side * weight * vel.a[axis][fidx];
            if (weight < 1) div -= (double)side * (1 - weight) * us.a[axis][fidx];
        }
        div += vol_error;
        rhs[c] = div;
    }
}

__global__ __launch_bounds__(kBlock) void k_pressure_apply(CellGrid cg, Dims d, int axis,
                                                           const double *__restrict__ phi,
                                                           const double *__restrict__ frac,
                                                           const double *__restrict__ jump,
                                                           const unsigned char *__restrict__ liquid,
                                                           const double *__restrict__ p,
                                                           double *__restrict__ vel, int *err) {
    const Dims fd = d.face(axis);
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    if (k >= fd.nz || j >= fd.ny) return;
    for (int i = i0, ie = min(i0 + cg.planes, fd.nx); i < ie; i++) {
        const size_t fidx = fd.idx(i, j, k);
        const double weight = 1. - frac[fidx];
        if (weight == 0.) continue;
        int c1[3] = {i, j, k};
        int c0[3] = {i, j, k};
        c0[axis] -= 1;
        const bool v0 = valid(d, c0), v1 = valid(d, c1);
        if (!v0 || !v1) {
            // open face on the grid boundary: the reference reads outside the grid
            const int *in = v0 ? c0 : c1;
            if (liquid[d.idx(in[0], in[1], in[2])]) atomicExch(err, 1);
            continue;
        }
        const size_t i0 = d.idx(c0[0], c0[1], c0[2]), i1 = d.idx(c1[0], c1[1], c1[2]);
        const bool l0 = liquid[i0] != 0, l1 = liquid[i1] != 0;
        if (!l0 && !l1) continue;
        double u = vel[fidx];
        if (l0 && l1) {
            u -= (p[i1] - p[i0]) * weight;
        } else {
            const double phi0 = phi[i0], phi1 = phi[i1];
            const double theta = phi0 / (phi0 - phi1);
            const double pin = l0 ? p[i0] : p[i1];
            const double pintf = jump ? jump[fidx] : 0.;
            const double ic = 1. / max_001(l0 ? theta : 1 - theta);
            u -= (double)(l0 ? +1 : -1) * (pintf - pin) * ic * weight;
        }
        vel[fidx] = u;
    }
}

__global__ __launch_bounds__(kBlock) void k_magnetic_mu(CellGrid cg, Dims d, int axis, MagGeom g,
                                                        const double *__restrict__ phi, double chi,
                                                        double band, double *__restrict__ mu) {
    const Dims fd = d.face(axis);
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    if (k >= fd.nz || j >= fd.ny) return;
    for (int i = i0, ie = min(i0 + cg.planes, fd.nx); i < ie; i++) {
        const int f[3] = {i, j, k};
        const size_t fidx = fd.idx(i, j, k);
        if (f[axis] <= 0 || f[axis] >= d.n(axis)) {  // IsBoundaryFace, width 0
            mu[fidx] = 1.;
            continue;
        }
        double pos[3];
        for (int a = 0; a < 3; a++) pos[a] = g.face_origin[axis][a] + (double)f[a] * g.spacing;
        // (the Heaviside of +0 and -0 is the same, so trilerp's zero sign
        // does not matter)
        const double ph = trilerp<true>(d, g.ls_origin, g.spacing, g.inv, phi, pos);
        mu[fidx] = 1. + (1. - heaviside(ph, band)) * chi;
    }
}

// dirichlet = 0: MagneticField::BuildMatrix / AssembleRHS (box faces skipped,
// cell 0 pinned by elimination). dirichlet = 1, an opt-in variant with no
// reference counterpart: phi = 0 on the box faces (a mirrored ghost value, so
// a box face adds 2 mu_face to the diagonal and its (mu_face - 1) H_ext flux
// to the right-hand side, like an interior face), every cell an unknown.
__global__ __launch_bounds__(kBlock) void k_magnetic_assemble(CellGrid cg, Dims d, Faces mu, Faces hext,
                                                              double spacing, int dirichlet,
                                                              double *__restrict__ diag64,
                                                              MutFaces cpl64,
                                                              double *__restrict__ rhs,
                                                              float *__restrict__ diag32,
                                                              float *cpl32x, float *cpl32y,
                                                              float *cpl32z) {
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    if (k >= d.nz || j >= d.ny) return;
    float *cpl32[3] = {cpl32x, cpl32y, cpl32z};
    for (int i = i0, ie = min(i0 + cg.planes, d.nx); i < ie; i++) {
        const size_t c = d.idx(i, j, k);
        double low[3] = {0.0, 0.0, 0.0};
        if (c == 0 && !dirichlet) {  // pinned cell: identity row, zero right-hand side
            diag64[c] = 1.0;
            rhs[c] = 0.0;
            diag32[c] = 0.f;
        } else {
            const int cell[3] = {i, j, k};
            // the six faces' operands are loaded before any is used (every
            // face, boundary faces included, has an entry in the face grids)
            double mv[6], hv[6];
            bool nbv[6], pin[6];
#pragma unroll
            for (int ord = 0; ord < 6; ord++) {
                const int axis = ord >> 1, side = (ord & 1) ? 1 : -1;
                int nb[3] = {cell[0], cell[1], cell[2]};
                nb[axis] += side;
                int face[3] = {cell[0], cell[1], cell[2]};
                face[axis] += (ord & 1);
                const size_t fidx = d.face(axis).idx(face[0], face[1], face[2]);
                nbv[ord] = valid(d, nb);
                pin[ord] = !dirichlet && nb[0] == 0 && nb[1] == 0 && nb[2] == 0;
                mv[ord] = mu.a[axis][fidx];
                hv[ord] = hext.a[axis][fidx];
            }
            double diag = 0.;
            double b = 0;
            // FP32 diagonal: the sum of the FP32 couplings in the same order
            // (in FP64, rounded once), which the stencil kernels can rebuild
            // from the couplings (next to the pinned cell the eliminated
            // coupling stays in the diagonal, which is then the FP64 one
            // rounded)
            double diag_f = 0.;
            bool next_to_pin = false;
#pragma unroll
            for (int ord = 0; ord < 6; ord++) {
                const int axis = ord >> 1, side = (ord & 1) ? 1 : -1;
                if (!nbv[ord]) {
                    if (dirichlet) {  // box face: ghost value -phi
                        diag += 2. * mv[ord];
                        b += (mv[ord] - 1.) * hv[ord] * (double)(-side) * spacing;
                    }
                    continue;
                }
                const double weight = mv[ord];
                // column 0 is dropped from every other row (pin by elimination)
                if (side < 0 && !pin[ord]) low[axis] = weight;
                diag += weight;
                diag_f += (double)(float)weight;
                next_to_pin = next_to_pin || pin[ord];
                b += (weight - 1.) * hv[ord] * (double)(-side) * spacing;
            }
            diag64[c] = diag;
            rhs[c] = b;
            // (with box faces in the diagonal the FP32 one is not a row sum:
            // the FP64 one rounded)
            diag32[c] = (float)(next_to_pin || dirichlet ? diag : diag_f);
        }
        for (int a = 0; a < 3; a++) {
            cpl64.a[a][c] = low[a];
            cpl32[a][c] = (float)low[a];
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_magnetic_total_field(CellGrid cg, Dims d, int axis, double inv,
                                                                 const double *__restrict__ pot,
                                                                 const double *__restrict__ hext,
                                                                 double *__restrict__ htot) {
    const Dims fd = d.face(axis);
    int k, j, i0;
    cell_tile(cg, k, j, i0);
    if (k >= fd.nz || j >= fd.ny) return;
    for (int i = i0, ie = min(i0 + cg.planes, fd.nx); i < ie; i++) {
        const size_t fidx = fd.idx(i, j, k);
        int c0[3] = {i, j, k};
        c0[axis] -= 1;
        const int c1[3] = {i, j, k};
        double h;
        if (valid(d, c0) && valid(d, c1))
            h = (pot[d.idx(c0[0], c0[1], c0[2])] - pot[d.idx(c1[0], c1[1], c1[2])]) * inv;
        else
            h = 0.;
        h += hext[fidx];
        htot[fidx] = h;
    }
}

__global__ __launch_bounds__(kBlock) void k_vertex_pressures(Dims d, MagGeom g, Faces htot,
                                                             double chi, int nverts,
                                                             const double *__restrict__ pos,
                                                             const double *__restrict__ nrm,
                                                             double *__restrict__ out) {
    const double pi = 3.141592653589793238462643383279502884;
    const double mu0 = 4e-7 * pi;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nverts) return;
    const double P[3] = {pos[3 * (size_t)v], pos[3 * (size_t)v + 1], pos[3 * (size_t)v + 2]};
    double H[3];
    for (int a = 0; a < 3; a++)
        H[a] = trilerp(d.face(a), g.face_origin[a], g.spacing, g.inv, htot.a[a], P);
    const double n0 = nrm[3 * (size_t)v], n1 = nrm[3 * (size_t)v + 1], n2 = nrm[3 * (size_t)v + 2];
    double coeff = H[0] * H[0] + H[1] * H[1] + H[2] * H[2];
    const double hn = H[0] * n0 + H[1] * n1 + H[2] * n2;
    coeff += chi * chi / (4 * chi + 4) * hn * hn;
    out[v] = mu0 * 0.5 * chi * coeff;
}

}  // namespace

void launch_pressure_assemble(hipStream_t st, Dims d, const double *phi, Faces frac, Faces vel,
                              Faces us, const Faces *jump, double vol_error, double *diag64,
                              double *const cpl64[3], double *rhs, float *diag32,
                              float *const cpl32[3], unsigned char *liquid, int *err) {
    Launch L = cell_launch(d, 0, kAssemblyPlanes);
    Faces j = jump ? *jump : Faces{{nullptr, nullptr, nullptr}};
    MutFaces c{{cpl64[0], cpl64[1], cpl64[2]}};
    hipLaunchKernelGGL(k_pressure_assemble, L.grid, L.block, 0, st, L.cg, d, phi, frac, vel, us, j,
                       jump ? 1 : 0, vol_error, diag64, c, rhs, diag32, cpl32[0], cpl32[1],
                       cpl32[2], liquid, err);
    SFP_LAUNCH_CHECK();
}

void launch_pressure_rhs(hipStream_t st, Dims d, const unsigned char *liquid, Faces frac,
                         Faces vel, Faces us, double vol_error, double *rhs) {
    Launch L = cell_launch(d, 0, kAssemblyPlanes);
    hipLaunchKernelGGL(k_pressure_rhs, L.grid, L.block, 0, st, L.cg, d, liquid, frac, vel, us,
                       vol_error, rhs);
    SFP_LAUNCH_CHECK();
}

void launch_pressure_apply(hipStream_t st, Dims d, const double *phi, Faces frac,
                           const Faces *jump, const unsigned char *liquid, const double *p,
                           MutFaces vel, int *err) {
    for (int a = 0; a < 3; a++) {
        Launch L = cell_launch(d.face(a), 0, kAssemblyPlanes);
        hipLaunchKernelGGL(k_pressure_apply, L.grid, L.block, 0, st, L.cg, d, a, phi, frac.a[a],
                           jump ? jump->a[a] : nullptr, liquid, p, vel.a[a], err);
        SFP_LAUNCH_CHECK();
    }
}

void launch_magnetic_mu(hipStream_t st, Dims d, const MagGeom &g, const double *phi, double chi,
                        double band, MutFaces mu) {
    for (int a = 0; a < 3; a++) {
        Launch L = cell_launch(d.face(a), 0, kAssemblyPlanes);
        hipLaunchKernelGGL(k_magnetic_mu, L.grid, L.block, 0, st, L.cg, d, a, g, phi, chi, band, mu.a[a]);
        SFP_LAUNCH_CHECK();
    }
}

void launch_magnetic_assemble(hipStream_t st, Dims d, Faces mu, Faces hext, double spacing,
                              double *diag64, double *const cpl64[3], double *rhs,
                              float *diag32, float *const cpl32[3], int dirichlet) {
    Launch L = cell_launch(d, 0, kAssemblyPlanes);
    MutFaces c{{cpl64[0], cpl64[1], cpl64[2]}};
    hipLaunchKernelGGL(k_magnetic_assemble, L.grid, L.block, 0, st, L.cg, d, mu, hext, spacing,
                       dirichlet, diag64, c, rhs, diag32, cpl32[0], cpl32[1], cpl32[2]);
    SFP_LAUNCH_CHECK();
}

void launch_magnetic_total_field(hipStream_t st, Dims d, double inv, const double *pot,
                                 Faces hext, MutFaces htot) {
    for (int a = 0; a < 3; a++) {
        Launch L = cell_launch(d.face(a), 0, kAssemblyPlanes);
        hipLaunchKernelGGL(k_magnetic_total_field, L.grid, L.block, 0, st, L.cg, d, a, inv, pot,
                           hext.a[a], htot.a[a]);
        SFP_LAUNCH_CHECK();
    }
}

void launch_vertex_pressures(hipStream_t st, Dims d, const MagGeom &g, Faces htot, double chi,
                             int nverts, const double *pos, const double *nrm, double *out) {
    if (nverts <= 0) return;
    const unsigned blocks = (nverts + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_vertex_pressures, dim3(blocks), dim3(kBlock), 0, st, d, g, htot, chi,
                       nverts, pos, nrm, out);
    SFP_LAUNCH_CHECK();
}

}  // namespace sfp
