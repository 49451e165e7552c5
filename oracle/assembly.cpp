// FP64 restatement of the reference assembly, projection and field update.
// Test infrastructure only (see oracle.h). Built with -ffp-contract=off so
// every a*b+c below rounds twice, as the reference's MSVC /fp:precise build does.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <utility>
#include <vector>

// Planted mistakes for the oracle's own closed-form tests (orc_set_mutation):
// each one changes a single expression of the restatement so that a
// closed-form check must fail. 0, the default, is the faithful restatement.
enum Mutation {
    MUT_NONE = 0,
    MUT_HEAVISIDE_NO_SIN = 1,   // band interior without the sin term
    MUT_HEAVISIDE_HALF_BAND = 2,  // band of half the width
    MUT_JUMP_SIGN = 3,          // jump subtracted in the rhs and in the update (consistently)
    MUT_JUMP_NO_COEF = 4,       // rhs jump term without 1/max(theta, .001)
    MUT_NO_CLAMP = 5,           // 1/theta at both ghost-fluid sites
    MUT_TRILERP_SWAP = 6,       // trilinear weights (1 - f, f) swapped
    MUT_FACE_ORIGIN = 7,        // face grids of the field placed at the cell centres
    MUT_BOX_FACE_WEIGHT = 8,    // Dirichlet box: mu instead of 2 mu per box face
};
static int g_mutation = MUT_NONE;

namespace {

struct Dims {
    int n[3];
    int64_t index(int i, int j, int k) const {  // core/Grid.h:28
        return k + (int64_t)n[2] * (j + (int64_t)n[1] * i);
    }
    bool valid(int i, int j, int k) const {  // core/Grid.h:24-25
        return i >= 0 && j >= 0 && k >= 0 && i < n[0] && j < n[1] && k < n[2];
    }
};

Dims face_dims(Dims const &c, int axis) {  // core/StaggeredGrid.h:15-19
    Dims f = c;
    f.n[axis] += 1;
    return f;
}

// core/Grid.h:37-40 and core/StaggeredGrid.h:55-58
inline int axis_of(int ord) { return ord >> 1; }
inline int side_of(int ord) { return ord & 1 ? 1 : -1; }

// std::max(theta, .001) with std::max's exact semantics (NaN passes through)
inline double max_001(double t) {
    if (g_mutation == MUT_NO_CLAMP) return t;
    return (t < .001) ? .001 : t;
}

void coord_of(Dims const &d, int64_t idx, int c[3]) {  // core/Grid.h:31
    c[0] = (int)(idx / d.n[2] / d.n[1]);
    c[1] = (int)(idx / d.n[2] % d.n[1]);
    c[2] = (int)(idx % d.n[2]);
}

// Trilinear interpolation on a grid with (origin, spacing, inv), clamped reads
// (core/TriLerp.h:22-36,54-61; core/Grid.h:27,29,33,35; core/GridData.h:25).
double trilerp(Dims const &g, const double org[3], double spacing, double inv,
               const double *data, const double pos[3]) {
    int lo[3];
    double w[2][3];
    for (int a = 0; a < 3; a++) {
        double t = (pos[a] - org[a]) * inv - 0.0;
        lo[a] = (int)std::floor(t);
    }
    for (int a = 0; a < 3; a++) {
        double fr = ((pos[a] - org[a]) - (double)lo[a] * spacing) * inv;
        w[0][a] = 1. - fr;
        w[1][a] = fr;
        if (g_mutation == MUT_TRILERP_SWAP) std::swap(w[0][a], w[1][a]);
    }
    static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    double val = 0;
    for (int p = 0; p < 8; p++) {
        int c[3];
        for (int a = 0; a < 3; a++) {
            c[a] = lo[a] + off[p][a];
            c[a] = std::min(std::max(c[a], 0), g.n[a] - 1);
        }
        double wt = w[off[p][0]][0] * w[off[p][1]][1] * w[off[p][2]][2];
        val += data[g.index(c[0], c[1], c[2])] * wt;
    }
    return val;
}

// MagneticField::Heaviside (core/MagneticField.h:44-54)
double heaviside(double phi, double band) {
    const double pi = 3.141592653589793238462643383279502884;
    if (g_mutation == MUT_HEAVISIDE_HALF_BAND) band *= .5;
    if (phi < -band) return 0.;
    if (phi > band) return 1.;
    if (g_mutation == MUT_HEAVISIDE_NO_SIN) return .5 * (1. + phi / band);
    return .5 * (1. + phi / band + std::sin(pi * phi / band) / pi);
}

// Geometry of the magnetic solver's own StaggeredGrid (boundary width 0, built
// from the simulation grid's centre) -- core/MagneticField.cpp:11-16 and the
// StaggeredGrid ctor core/StaggeredGrid.h:8-21.
struct ExtGeometry {
    double origin[3];           // m_SGridExt origin
    double face_origin[3][3];   // face grid origins
};

ExtGeometry ext_geometry(Dims const &c, double spacing, const double sim_origin[3]) {
    ExtGeometry g;
    for (int a = 0; a < 3; a++) {
        double center = sim_origin[a] + (double)c.n[a] * spacing / 2.;
        g.origin[a] = center - (double)c.n[a] * spacing / 2;
    }
    for (int f = 0; f < 3; f++)
        for (int a = 0; a < 3; a++)
            g.face_origin[f][a] =
                g.origin[a] + (a == f && g_mutation != MUT_FACE_ORIGIN ? 0.0 : 1.0) * spacing / 2;
    return g;
}

}  // namespace

extern "C" {

void orc_set_mutation(int m) { g_mutation = m; }

int orc_pressure_unknowns(int nx, int ny, int nz, const double *phi, int *grid2mat,
                          int *mat2grid) {
    int64_t total = (int64_t)nx * ny * nz;
    int n = 0;
    for (int64_t idx = 0; idx < total; idx++) {  // ForEach order == linear order
        if (phi[idx] <= 0) {
            grid2mat[idx] = n;
            mat2grid[n] = (int)idx;
            n++;
        } else {
            grid2mat[idx] = -1;
        }
    }
    return n;
}

int64_t orc_pressure_matrix(int nx, int ny, int nz, const double *phi,
                            const double *const frac[3], const double *const vel[3],
                            const double *const usolid[3], const double *const jump[3],
                            double vol_error, const int *grid2mat, const int *mat2grid, int n,
                            int *rowptr, int *col, double *val, double *rhs) {
    Dims d{{nx, ny, nz}};
    Dims fd[3] = {face_dims(d, 0), face_dims(d, 1), face_dims(d, 2)};
    int64_t nnz = 0;
    rowptr[0] = 0;
    for (int r = 0; r < n; r++) {
        int cell[3];
        coord_of(d, mat2grid[r], cell);
        double diag = 0;
        double div = 0;
        std::vector<std::pair<int, double>> row;
        for (int i = 0; i < 6; i++) {
            int axis = axis_of(i), side = side_of(i);
            int nb[3] = {cell[0], cell[1], cell[2]};
            nb[axis] += side;
            int face[3] = {cell[0], cell[1], cell[2]};
            face[axis] += (i & 1);
            int64_t fidx = fd[axis].index(face[0], face[1], face[2]);
            double weight = 1 - frac[axis][fidx];
            if (weight > 0) {
                if (!d.valid(nb[0], nb[1], nb[2])) return -1;
                int64_t nidx = d.index(nb[0], nb[1], nb[2]);
                int c = grid2mat[nidx];
                if (c >= 0) {
                    diag += weight;
                    row.emplace_back(c, -weight);
                } else {
                    double pc = phi[mat2grid[r]];
                    double theta = pc / (pc - phi[nidx]);
                    double ic = 1. / max_001(theta);
                    diag += weight * ic;
                    if (jump) {
                        if (g_mutation == MUT_JUMP_SIGN)
                            div -= jump[axis][fidx] * weight * ic;
                        else if (g_mutation == MUT_JUMP_NO_COEF)
                            div += jump[axis][fidx] * weight;
                        else
                            div += jump[axis][fidx] * weight * ic;
                    }
                }
                div -= side * weight * vel[axis][fidx];
            }
            if (weight < 1) div -= side * (1 - weight) * usolid[axis][fidx];
        }
        div += vol_error;
        rhs[r] = div;
        row.emplace_back(r, diag ? diag : 1.);
        std::sort(row.begin(), row.end(),
                  [](auto const &a, auto const &b) { return a.first < b.first; });
        for (auto const &e : row) {
            col[nnz] = e.first;
            val[nnz] = e.second;
            nnz++;
        }
        rowptr[r + 1] = (int)nnz;
    }
    return nnz;
}

void orc_pressure_rhs(int nx, int ny, int nz, const double *const frac[3],
                      const double *const vel[3], const double *const usolid[3],
                      double vol_error, const int *mat2grid, int n, double *rhs) {
    Dims d{{nx, ny, nz}};
    Dims fd[3] = {face_dims(d, 0), face_dims(d, 1), face_dims(d, 2)};
    for (int r = 0; r < n; r++) {
        int cell[3];
        coord_of(d, mat2grid[r], cell);
        double div = 0;
        for (int i = 0; i < 6; i++) {
            int axis = axis_of(i), side = side_of(i);
            int face[3] = {cell[0], cell[1], cell[2]};
            face[axis] += (i & 1);
            int64_t fidx = fd[axis].index(face[0], face[1], face[2]);
            double weight = 1 - frac[axis][fidx];
            if (weight > 0) div -= side * weight * vel[axis][fidx];
            if (weight < 1) div -= side * (1 - weight) * usolid[axis][fidx];
        }
        div += vol_error;
        rhs[r] = div;
    }
}

int orc_pressure_apply(int nx, int ny, int nz, const double *phi, const double *const frac[3],
                       const double *const jump[3], const int *grid2mat, const double *p,
                       double *const vel[3]) {
    Dims d{{nx, ny, nz}};
    int bad = 0;
    for (int axis = 0; axis < 3; axis++) {
        Dims fd = face_dims(d, axis);
        for (int i = 0; i < fd.n[0]; i++)
            for (int j = 0; j < fd.n[1]; j++)
                for (int k = 0; k < fd.n[2]; k++) {
                    int64_t fidx = fd.index(i, j, k);
                    int c1[3] = {i, j, k};
                    int c0[3] = {i, j, k};
                    c0[axis] -= 1;
                    double weight = 1. - frac[axis][fidx];
                    if (weight == 0.) continue;
                    if (!d.valid(c0[0], c0[1], c0[2]) || !d.valid(c1[0], c1[1], c1[2])) {
                        bad = 1;
                        continue;
                    }
                    int64_t i0 = d.index(c0[0], c0[1], c0[2]);
                    int64_t i1 = d.index(c1[0], c1[1], c1[2]);
                    int id0 = grid2mat[i0], id1 = grid2mat[i1];
                    if (id0 < 0 && id1 < 0) continue;
                    if (id0 >= 0 && id1 >= 0) {
                        vel[axis][fidx] -= (p[id1] - p[id0]) * weight;
                    } else {
                        double phi0 = phi[i0], phi1 = phi[i1];
                        double theta = phi0 / (phi0 - phi1);
                        double pin = id0 >= 0 ? p[id0] : p[id1];
                        double pintf = jump ? jump[axis][fidx] : 0.;
                        if (g_mutation == MUT_JUMP_SIGN) pintf = -pintf;
                        double ic = 1. / max_001(id0 >= 0 ? theta : 1 - theta);
                        vel[axis][fidx] -= (id0 >= 0 ? +1 : -1) * (pintf - pin) * ic * weight;
                    }
                }
    }
    return bad ? -1 : 0;
}

void orc_magnetic_coeff(int nx, int ny, int nz, double spacing, const double origin[3],
                        const double *phi, double chi, double band, double *const mu[3]) {
    Dims d{{nx, ny, nz}};
    ExtGeometry g = ext_geometry(d, spacing, origin);
    // level-set grid: the simulation's cell grid
    double ls_org[3];
    for (int a = 0; a < 3; a++) ls_org[a] = origin[a] + spacing / 2;
    double inv = 1 / spacing;
    for (int axis = 0; axis < 3; axis++) {
        Dims fd = face_dims(d, axis);
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < fd.n[0]; i++)
            for (int j = 0; j < fd.n[1]; j++)
                for (int k = 0; k < fd.n[2]; k++) {
                    int f[3] = {i, j, k};
                    int64_t fidx = fd.index(i, j, k);
                    // IsBoundaryFace with boundary width 0 (core/StaggeredGrid.h:48)
                    if (f[axis] <= 0 || f[axis] >= d.n[axis]) {
                        mu[axis][fidx] = 1.;
                        continue;
                    }
                    double pos[3];
                    for (int a = 0; a < 3; a++)
                        pos[a] = g.face_origin[axis][a] + (double)f[a] * spacing;
                    double ph = trilerp(d, ls_org, spacing, inv, phi, pos);
                    mu[axis][fidx] = 1. + (1. - heaviside(ph, band)) * chi;
                }
    }
}

int64_t orc_magnetic_matrix(int nx, int ny, int nz, const double *const mu[3], int bc, int *rowptr,
                            int *col, double *val) {
    Dims d{{nx, ny, nz}};
    Dims fd[3] = {face_dims(d, 0), face_dims(d, 1), face_dims(d, 2)};
    int64_t total = (int64_t)nx * ny * nz;
    int64_t nnz = 0;
    rowptr[0] = 0;
    std::vector<std::pair<int, double>> row;
    for (int64_t r = 0; r < total; r++) {
        row.clear();
        if (r == 0 && bc == 0) {
            row.emplace_back(0, 1.);
        } else {
            int cell[3];
            coord_of(d, r, cell);
            double diag = 0.;
            for (int i = 0; i < 6; i++) {
                int axis = axis_of(i), side = side_of(i);
                int nb[3] = {cell[0], cell[1], cell[2]};
                nb[axis] += side;
                int face[3] = {cell[0], cell[1], cell[2]};
                face[axis] += (i & 1);
                double weight = mu[axis][fd[axis].index(face[0], face[1], face[2])];
                if (!d.valid(nb[0], nb[1], nb[2])) {
                    // Dirichlet box: phi = 0 on the face, ghost value -phi
                    if (bc == 1) diag += (g_mutation == MUT_BOX_FACE_WEIGHT ? 1. : 2.) * weight;
                    continue;
                }
                int64_t c = d.index(nb[0], nb[1], nb[2]);
                if (c != 0 || bc == 1) row.emplace_back((int)c, -weight);
                diag += weight;
            }
            row.emplace_back((int)r, diag);
        }
        std::sort(row.begin(), row.end(),
                  [](auto const &a, auto const &b) { return a.first < b.first; });
        for (auto const &e : row) {
            col[nnz] = e.first;
            val[nnz] = e.second;
            nnz++;
        }
        rowptr[r + 1] = (int)nnz;
    }
    return nnz;
}

void orc_magnetic_rhs(int nx, int ny, int nz, double spacing, const double *const mu[3],
                      const double *const hext[3], int bc, double *rhs) {
    Dims d{{nx, ny, nz}};
    Dims fd[3] = {face_dims(d, 0), face_dims(d, 1), face_dims(d, 2)};
    int64_t total = (int64_t)nx * ny * nz;
    for (int64_t r = 0; r < total; r++) rhs[r] = 0;
#pragma omp parallel for schedule(static)
    for (int64_t r = bc == 1 ? 0 : 1; r < total; r++) {
        int cell[3];
        coord_of(d, r, cell);
        for (int i = 0; i < 6; i++) {
            int axis = axis_of(i), side = side_of(i);
            int nb[3] = {cell[0], cell[1], cell[2]};
            nb[axis] += side;
            if (!d.valid(nb[0], nb[1], nb[2]) && bc == 0) continue;
            int face[3] = {cell[0], cell[1], cell[2]};
            face[axis] += (i & 1);
            int64_t fidx = fd[axis].index(face[0], face[1], face[2]);
            double weight = mu[axis][fidx];
            rhs[r] += (weight - 1.) * hext[axis][fidx] * (-side) * spacing;
        }
    }
}

// ||rhs - A x||^2 of MagneticField::BuildMatrix's system without storing A
// (rows as in orc_magnetic_matrix: row 0 the identity, column 0 dropped),
// for grids whose CSR would not fit; *bb receives ||rhs||^2.
double orc_magnetic_residual(int nx, int ny, int nz, const double *const mu[3], int bc,
                             const double *rhs, const double *x, double *bb) {
    Dims d{{nx, ny, nz}};
    Dims fd[3] = {face_dims(d, 0), face_dims(d, 1), face_dims(d, 2)};
    int64_t total = (int64_t)nx * ny * nz;
    double rr = 0.0, b2 = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rr, b2)
    for (int64_t r = 0; r < total; r++) {
        double ax;
        if (r == 0 && bc == 0) {
            ax = x[0];
        } else {
            int cell[3];
            coord_of(d, r, cell);
            double diag = 0., off = 0.;
            for (int i = 0; i < 6; i++) {
                int axis = axis_of(i), side = side_of(i);
                int nb[3] = {cell[0], cell[1], cell[2]};
                nb[axis] += side;
                int face[3] = {cell[0], cell[1], cell[2]};
                face[axis] += (i & 1);
                double weight = mu[axis][fd[axis].index(face[0], face[1], face[2])];
                if (!d.valid(nb[0], nb[1], nb[2])) {
                    if (bc == 1) diag += 2. * weight;
                    continue;
                }
                int64_t c = d.index(nb[0], nb[1], nb[2]);
                if (c != 0 || bc == 1) off += weight * x[c];
                diag += weight;
            }
            ax = diag * x[r] - off;
        }
        double e = rhs[r] - ax;
        rr += e * e;
        b2 += rhs[r] * rhs[r];
    }
    *bb = b2;
    return rr;
}

void orc_magnetic_total_field(int nx, int ny, int nz, double spacing, const double *potential,
                              const double *const hext[3], double *const htotal[3]) {
    Dims d{{nx, ny, nz}};
    double inv = 1 / spacing;
    for (int axis = 0; axis < 3; axis++) {
        Dims fd = face_dims(d, axis);
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < fd.n[0]; i++)
            for (int j = 0; j < fd.n[1]; j++)
                for (int k = 0; k < fd.n[2]; k++) {
                    int64_t fidx = fd.index(i, j, k);
                    int c0[3] = {i, j, k};
                    c0[axis] -= 1;
                    double h;
                    if (d.valid(c0[0], c0[1], c0[2]) && d.valid(i, j, k))
                        h = (potential[d.index(c0[0], c0[1], c0[2])] - potential[d.index(i, j, k)]) *
                            inv;
                    else
                        h = 0.;
                    h += hext[axis][fidx];
                    htotal[axis][fidx] = h;
                }
    }
}

void orc_magnetic_vertex_pressures(int nx, int ny, int nz, double spacing, const double origin[3],
                                   const double *const htotal[3], double chi, int nverts,
                                   const double *positions, const double *normals,
                                   double *pressures) {
    const double pi = 3.141592653589793238462643383279502884;
    const double mu0 = 4e-7 * pi;
    Dims d{{nx, ny, nz}};
    ExtGeometry g = ext_geometry(d, spacing, origin);
    double inv = 1 / spacing;
#pragma omp parallel for schedule(static)
    for (int v = 0; v < nverts; v++) {
        const double *pos = positions + 3 * (int64_t)v;
        const double *nrm = normals + 3 * (int64_t)v;
        double H[3];
        for (int a = 0; a < 3; a++)
            H[a] = trilerp(face_dims(d, a), g.face_origin[a], spacing, inv, htotal[a], pos);
        double coeff = H[0] * H[0] + H[1] * H[1] + H[2] * H[2];
        double hn = H[0] * nrm[0] + H[1] * nrm[1] + H[2] * nrm[2];
        coeff += chi * chi / (4 * chi + 4) * hn * hn;
        pressures[v] = mu0 * 0.5 * chi * coeff;
    }
}

}  // extern "C"

// ---------------------------------------------------------------- after the projection

// Simulation::ProjectVelocity's extrapolation mask (core/Simulation.cpp:236-240):
// fraction < 1 and a liquid cell on either side (outside the grid: not liquid).
void orc_velocity_valid(int nx, int ny, int nz, const double *phi, const double *const frac[3],
                        unsigned char *const valid[3]) {
    const Dims d{{nx, ny, nz}};
    for (int axis = 0; axis < 3; axis++) {
        const Dims fd = face_dims(d, axis);
        const int64_t nf = (int64_t)fd.n[0] * fd.n[1] * fd.n[2];
        for (int64_t q = 0; q < nf; q++) {
            int f[3];
            coord_of(fd, q, f);
            bool v = false;
            if (frac[axis][q] < 1) {
                int c0[3] = {f[0], f[1], f[2]};
                c0[axis] -= 1;  // StaggeredGrid::AdjCellOfFace(axis, face, 0)
                const bool liq0 = d.valid(c0[0], c0[1], c0[2]) && phi[d.index(c0[0], c0[1], c0[2])] <= 0;
                const bool liq1 = d.valid(f[0], f[1], f[2]) && phi[d.index(f[0], f[1], f[2])] <= 0;
                v = liq0 || liq1;
            }
            valid[axis][q] = v ? 1 : 0;
        }
    }
}

// Extrapolation::Solve on the three face grids (core/Extrapolation.h:27-37,
// core/Extrapolation.cpp:4-30); valid is copied, not modified.
void orc_extrapolate_faces(int nx, int ny, int nz, double *const data[3],
                           const unsigned char *const valid_in[3], double clear, int steps) {
    const Dims d{{nx, ny, nz}};
    for (int axis = 0; axis < 3; axis++) {
        const Dims g = face_dims(d, axis);
        const int64_t n = (int64_t)g.n[0] * g.n[1] * g.n[2];
        std::vector<unsigned char> valid(valid_in[axis], valid_in[axis] + n), valid1 = valid;
        double *gr = data[axis];
        for (int iter = 0; iter < steps; iter++) {
            for (int64_t q = 0; q < n; q++) {
                if (!valid[q]) {
                    int c[3];
                    coord_of(g, q, c);
                    int cnt = 0;
                    double sum = 0;
                    for (int ord = 0; ord < 6; ord++) {
                        int nb[3] = {c[0], c[1], c[2]};
                        nb[axis_of(ord)] += side_of(ord);
                        if (g.valid(nb[0], nb[1], nb[2]) && valid[g.index(nb[0], nb[1], nb[2])]) {
                            sum += gr[g.index(nb[0], nb[1], nb[2])];
                            cnt++;
                        }
                    }
                    if (cnt > 0) {
                        gr[q] = sum / cnt;
                        valid1[q] = 1;
                    } else {
                        gr[q] = clear;
                    }
                } else {
                    valid1[q] = 1;
                }
            }
            valid.swap(valid1);
        }
    }
}

// Collider::Enforce (core/Collider.cpp:75-86) on the simulation's staggered
// grid: origin is StaggeredGrid::GetOrigin() (core/StaggeredGrid.h:13),
// face grid a starts at origin + (1 - e_a) * spacing / 2.
void orc_collider_enforce(int nx, int ny, int nz, double spacing, const double origin[3],
                          double *const vel[3], const double *const frac[3],
                          const double *const normal[3], const double *const colvel[3]) {
    const Dims d{{nx, ny, nz}};
    const double inv = 1 / spacing;
    double forg[3][3];
    for (int a = 0; a < 3; a++)
        for (int c = 0; c < 3; c++) forg[a][c] = origin[c] + ((c == a ? 0.0 : 1.0) * spacing) / 2;
    std::vector<double> old[3];
    for (int a = 0; a < 3; a++) {
        const Dims fd = face_dims(d, a);
        old[a].assign(vel[a], vel[a] + (int64_t)fd.n[0] * fd.n[1] * fd.n[2]);
    }
    for (int axis = 0; axis < 3; axis++) {
        const Dims fd = face_dims(d, axis);
        const int64_t nf = (int64_t)fd.n[0] * fd.n[1] * fd.n[2];
        for (int64_t q = 0; q < nf; q++) {
            if (!(frac[axis][q] == 1.)) continue;
            int f[3];
            coord_of(fd, q, f);
            double pos[3];
            for (int a = 0; a < 3; a++) pos[a] = forg[axis][a] + (double)f[a] * spacing;
            double v[3], v0[3], n[3];
            for (int b = 0; b < 3; b++) {
                const Dims gb = face_dims(d, b);
                v[b] = trilerp(gb, forg[b], spacing, inv, old[b].data(), pos);
                v0[b] = trilerp(gb, forg[b], spacing, inv, colvel[b], pos);
                n[b] = trilerp(gb, forg[b], spacing, inv, normal[b], pos);
            }
            const double z = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];  // Eigen normalized()
            if (z > 0) {
                const double r = std::sqrt(z);
                for (int b = 0; b < 3; b++) n[b] = n[b] / r;
            }
            const double dot = (v[0] - v0[0]) * n[0] + (v[1] - v0[1]) * n[1] + (v[2] - v0[2]) * n[2];
            vel[axis][q] = v[axis] - dot * n[axis];
        }
    }
}
