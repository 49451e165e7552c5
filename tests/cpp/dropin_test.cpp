// Drives the drop-in Pressure and MagneticField classes the way the
// reference's Simulation does (core/Simulation.cpp:193-243), then the whole
// ProjectionStep (magnetic pressure, jump, projection, extrapolation, wall
// condition), and checks every result against the FP64 oracle.
// Exit status 0 = all checks passed.
#include "MagneticField.h"
#include "Pressure.h"
#include "Simulation.h"
#include "oracle.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

using namespace Pivot;

namespace {

int g_fail = 0;
void check(bool ok, char const *what, double value = 0) {
    std::printf("  [%s] %s (%.3e)\n", ok ? "ok" : "FAIL", what, value);
    if (!ok) g_fail++;
}

// deterministic uniform [0,1) (SplitMix64)
double urand(uint64_t &state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

struct Csr {
    std::vector<int> ptr, col;
    std::vector<double> val;
};

double solve_ref(int n, Csr const &A, std::vector<double> const &b, std::vector<double> &x, bool cg) {
    double rel, ts, tv;
    x.assign(n, 0.0);
    if (cg)
        orc_cg(n, A.ptr.data(), A.col.data(), A.val.data(), b.data(), x.data(), 1e-12, 2000, 2, &rel, &ts, &tv);
    else
        orc_bicgstab(n, A.ptr.data(), A.col.data(), A.val.data(), b.data(), x.data(), 1e-12, 2000, 2, &rel,
                     &ts, &tv);
    return rel;
}

}  // namespace

int main() {
    // ragged grid with an off-centre box: origin rounding is exercised
    StaggeredGrid const sgrid(2, 1.0 / 22, Vector3i(20, 18, 22), Vector3d(0.41, 0.52, 0.47));
    Vector3i const n = sgrid.GetResolution();
    int const ncell = sgrid.GetNumCells();
    Grid const &cg = sgrid.GetCellGrid();
    auto const &fg = sgrid.GetFaceGrids();
    uint64_t rng = 12345;

    GridData<double> levelSet(cg);
    ForEach(cg, [&](Vector3i const &c) {
        Vector3d const p = cg.PositionOf(c);
        double const sphere = (p - Vector3d(0.42, 0.62, 0.46)).norm() - 0.14;
        double const pool = p.y() - 0.4 + 0.03 * std::cos(7 * p.x());
        levelSet[c] = std::min(sphere, pool) + 0.01 * (urand(rng) - 0.5);
    });
    Collider collider(sgrid);
    collider.SetBoxWalls();
    SGridData<double> velocity(fg);
    for (int a = 0; a < 3; a++)
        ForEach(fg[a], [&](Vector3i const &f) {
            velocity[a][f] = 2 * urand(rng) - 1;
            collider.Velocity[a][f] = 0.3 * (urand(rng) - 0.5);
            double const fr = urand(rng);
            if (f[1] < 5 && collider.GetFraction()[a][f] == 0) collider.Fraction()[a][f] = fr < 0.4 ? 0.0 : fr;
        });
    auto jump = [](int axis, Vector3i const &face, double theta) {
        return 0.01 * (axis + 1) + 0.002 * face.x() - 0.001 * face.z() + 0.0 * theta;
    };

    // ---------------------------------------------------------------- Pressure::Project
    std::printf("Pressure::Project %dx%dx%d\n", n[0], n[1], n[2]);
    SGridData<double> vel = velocity;
    Pressure pressure(sgrid);
    pressure.SetPressureJump(jump);
    double const volError = 0.0031;
    pressure.Project(vel, levelSet, collider, volError);
    std::printf("\n");

    // oracle system with the same per-face jump values
    std::vector<double> jumpFace[3];
    for (int a = 0; a < 3; a++) {
        jumpFace[a].assign(fg[a].GetNumVertices(), 0.0);
        ForEach(fg[a], [&](Vector3i const &f) { jumpFace[a][fg[a].IndexOf(f)] = jump(a, f, 0.0); });
    }
    double const *frac[3], *us[3], *vin[3], *jp[3];
    for (int a = 0; a < 3; a++) {
        frac[a] = collider.GetFraction()[a].GetData().data();
        us[a] = collider.Velocity[a].GetData().data();
        vin[a] = velocity[a].GetData().data();
        jp[a] = jumpFace[a].data();
    }
    std::vector<int> g2m(ncell), m2g(ncell);
    int const nu = orc_pressure_unknowns(n[0], n[1], n[2], levelSet.GetData().data(), g2m.data(), m2g.data());
    Csr A;
    A.ptr.resize(nu + 1);
    A.col.resize(7 * (size_t)nu);
    A.val.resize(7 * (size_t)nu);
    std::vector<double> rhs(nu);
    int64_t const nnz = orc_pressure_matrix(n[0], n[1], n[2], levelSet.GetData().data(), frac, vin, us, jp,
                                            volError, g2m.data(), m2g.data(), nu, A.ptr.data(), A.col.data(),
                                            A.val.data(), rhs.data());
    check(nnz > 0, "oracle assembly", (double)nnz);
    std::vector<double> pc(nu);
    for (int r = 0; r < nu; r++) pc[r] = pressure.GetPressure()[m2g[r]];
    double const rel = orc_rel_residual(nu, A.ptr.data(), A.col.data(), A.val.data(), rhs.data(), pc.data());
    check(rel <= 1e-8, "oracle relative residual of the GPU pressure <= 1e-8", rel);
    check(std::fabs(rel - pressure.LastResidual()) < 1e-12, "GPU-reported residual matches", pressure.LastResidual());
    std::vector<double> xref;
    solve_ref(nu, A, rhs, xref, false);
    double dmax = 0, smax = 0;
    for (int r = 0; r < nu; r++) {
        dmax = std::max(dmax, std::fabs(pc[r] - xref[r]));
        smax = std::max(smax, std::fabs(xref[r]));
    }
    check(dmax <= 1e-6 * smax, "L-inf distance to the 1e-12 oracle solution (relative)", dmax / smax);
    // velocity update bit for bit given the GPU pressure
    std::vector<double> vref[3];
    double *vr[3];
    for (int a = 0; a < 3; a++) {
        vref[a] = velocity[a].GetData();
        vr[a] = vref[a].data();
    }
    orc_pressure_apply(n[0], n[1], n[2], levelSet.GetData().data(), frac, jp, g2m.data(), pc.data(), vr);
    bool same = true;
    for (int a = 0; a < 3; a++) same &= vref[a] == vel[a].GetData();
    check(same, "projected velocities bit-identical to the oracle's update");

    // ---------------------------------------------------------------- Pressure::Reproject
    std::printf("Pressure::Reproject\n");
    SGridData<double> vel2 = velocity;
    for (int a = 0; a < 3; a++)
        for (auto &v : vel2[a].GetData()) v = 0.5 * v + 0.05;
    pressure.SetPressureJump([](int, Vector3i const &, double) { return 0.0; });
    pressure.Reproject(vel2, levelSet, collider, volError);
    std::printf("\n");
    double const *vin2[3] = {vel2[0].GetData().data(), vel2[1].GetData().data(), vel2[2].GetData().data()};
    std::vector<double> rhs2(nu);
    {
        // SetRightHandSide reads the velocities before the update: rebuild them
        SGridData<double> v2 = velocity;
        for (int a = 0; a < 3; a++)
            for (auto &v : v2[a].GetData()) v = 0.5 * v + 0.05;
        double const *v2p[3] = {v2[0].GetData().data(), v2[1].GetData().data(), v2[2].GetData().data()};
        orc_pressure_rhs(n[0], n[1], n[2], frac, v2p, us, volError, m2g.data(), nu, rhs2.data());
    }
    (void)vin2;
    for (int r = 0; r < nu; r++) pc[r] = pressure.GetPressure()[m2g[r]];
    double const rel2 = orc_rel_residual(nu, A.ptr.data(), A.col.data(), A.val.data(), rhs2.data(), pc.data());
    check(rel2 <= 1e-8, "reproject: oracle relative residual <= 1e-8", rel2);

    // ---------------------------------------------------------------- MagneticField::Solve
    std::printf("MagneticField::Solve\n");
    double const chi = 2.0, band = 2 * sgrid.GetSpacing();
    MagneticField field(sgrid, 0.33, band);
    field.SetSusceptibility(chi);
    field.SetBandWidth(band);
    SGridData<double> hext(fg);
    for (int a = 0; a < 3; a++)
        ForEach(fg[a], [&](Vector3i const &f) { hext[a][f] = a == 1 ? 1.0 : (a == 0 ? 0.2 : -0.1); });
    field.Solve(levelSet, hext);
    std::printf("\n");
    Vector3d const o = sgrid.GetOrigin();
    double const origin[3] = {o[0], o[1], o[2]};
    std::vector<double> mu[3];
    double *mup[3];
    double const *muc[3], *hp[3];
    for (int a = 0; a < 3; a++) {
        mu[a].assign(fg[a].GetNumVertices(), 0.0);
        mup[a] = mu[a].data();
        muc[a] = mu[a].data();
        hp[a] = hext[a].GetData().data();
    }
    orc_magnetic_coeff(n[0], n[1], n[2], sgrid.GetSpacing(), origin, levelSet.GetData().data(), chi, band, mup);
    Csr M;
    M.ptr.resize(ncell + 1);
    M.col.resize(7 * (size_t)ncell);
    M.val.resize(7 * (size_t)ncell);
    orc_magnetic_matrix(n[0], n[1], n[2], muc, 0, M.ptr.data(), M.col.data(), M.val.data());
    std::vector<double> mrhs(ncell);
    orc_magnetic_rhs(n[0], n[1], n[2], sgrid.GetSpacing(), muc, hp, 0, mrhs.data());
    double const mrel = orc_rel_residual(ncell, M.ptr.data(), M.col.data(), M.val.data(), mrhs.data(),
                                         field.GetPotential().data());
    check(mrel <= 1e-6, "oracle relative residual of the GPU potential <= 1e-6 (shipped tol)", mrel);
    check(field.GetPotential()[0] == 0.0, "pinned cell 0 holds 0");
    std::vector<double> href[3];
    double *hr[3];
    for (int a = 0; a < 3; a++) {
        href[a].assign(fg[a].GetNumVertices(), 0.0);
        hr[a] = href[a].data();
    }
    orc_magnetic_total_field(n[0], n[1], n[2], sgrid.GetSpacing(), field.GetPotential().data(), hp, hr);
    same = true;
    for (int a = 0; a < 3; a++) same &= href[a] == field.GetTotalField(a);
    check(same, "total field bit-identical to the oracle's AssignTotalField");

    // ---------------------------------------------------------------- CalculateVertexPressures
    SurfaceMesh mesh;
    for (int v = 0; v < 300; v++) {
        Vector3d p(o[0] + urand(rng) * n[0] * sgrid.GetSpacing(), o[1] + urand(rng) * n[1] * sgrid.GetSpacing(),
                   o[2] + urand(rng) * n[2] * sgrid.GetSpacing());
        Vector3d nn(urand(rng) - 0.5, urand(rng) - 0.5, urand(rng) - 0.5);
        mesh.Positions.push_back(p);
        mesh.Normals.push_back(nn / nn.norm());
    }
    std::vector<double> const vp = field.CalculateVertexPressures(mesh);
    std::vector<double> pos, nrm, vpr(mesh.size());
    for (int v = 0; v < mesh.size(); v++)
        for (int a = 0; a < 3; a++) {
            pos.push_back(mesh.Positions[v][a]);
            nrm.push_back(mesh.Normals[v][a]);
        }
    double const *hc[3] = {field.GetTotalField(0).data(), field.GetTotalField(1).data(), field.GetTotalField(2).data()};
    orc_magnetic_vertex_pressures(n[0], n[1], n[2], sgrid.GetSpacing(), origin, hc, chi, mesh.size(), pos.data(),
                                  nrm.data(), vpr.data());
    double vmax = 0, verr = 0;
    for (int v = 0; v < mesh.size(); v++) {
        vmax = std::max(vmax, std::fabs(vpr[v]));
        verr = std::max(verr, std::fabs(vpr[v] - vp[v]));
    }
    check(verr <= 8 * 2.2e-16 * vmax, "vertex pressures match the oracle to 8 ulp of the max", verr / vmax);

    // ---------------------------------------------------------------- ProjectionStep
    std::printf("ProjectionStep::ProjectVelocity (Simulation::ProjectVelocity)\n");
    // a contour with one vertex on every cell-centre edge the interface crosses
    Contour contour(cg);
    SurfaceMesh &cm = contour.GetMesh();
    for (int a = 0; a < 3; a++) {
        Grid const &eg = contour.GetEdgeGrid(a);
        ForEach(eg, [&](Vector3i const &e) {
            Vector3i const e1 = e + Vector3i::Unit(a);
            double const p0 = levelSet[e], p1 = levelSet[e1];
            if ((p0 <= 0) == (p1 <= 0)) return;
            double const t = p0 / (p0 - p1);
            Vector3d const x0 = cg.PositionOf(e);
            Vector3d pos = x0;
            pos[a] += t * sgrid.GetSpacing();
            contour.EdgeMark(a)[e] = cm.size();
            cm.Positions.push_back(pos);
            cm.Normals.push_back(Vector3d::Unit(a) * (p0 <= 0 ? 1.0 : -1.0));
            cm.MeanCurvatures.push_back(5.0 + 2.0 * urand(rng));
        });
    }
    for (int a = 0; a < 3; a++)
        ForEach(fg[a], [&](Vector3i const &f) {
            bool const lo = f[a] == 0, hi = f[a] == fg[a].GetSize()[a] - 1;
            collider.Normal()[a][f] = lo ? 1.0 : (hi ? -1.0 : 0.0);
        });
    ProjectionStep::Settings st;
    st.magneticEnabled = true;
    st.surfaceTensionEnabled = true;
    st.surfaceTensionCoeff = 0.07;
    st.liquidDensity = 1000.0;
    st.chi = 2.0;
    st.fieldApplied = [](Vector3d const &p, double time) { return Vector3d(0.1 * p.z(), 1.0 + 0.2 * time, 0.0); };
    SGridData<double> vstep = velocity;
    ProjectionStep step(sgrid, vstep, levelSet, collider, contour, st);
    double const dt = 1e-3;
    step.ProjectVelocity(dt, 0.25, volError);
    std::printf("\n");
    std::vector<double> const &mp = step.GetMagneticPressures();
    check((int)mp.size() == cm.size() && cm.size() > 100, "magnetic pressure per contour vertex", (double)cm.size());
    // the jump the step handed to the projection, on every interior face
    double const inv = sgrid.GetInvSpacing();
    std::vector<double> jumpStep[3];
    double const *js[3];
    for (int a = 0; a < 3; a++) {
        jumpStep[a].assign(fg[a].GetNumVertices(), 0.0);
        ForEach(fg[a], [&](Vector3i const &f) {
            if (f[a] == 0 || f[a] == n[a]) return;
            int const idx = contour.VertexIndexOf(a, f - Vector3i::Unit(a));
            if (idx < 0) return;
            double pj = 0;
            pj += cm.MeanCurvatures[idx] * st.surfaceTensionCoeff / st.liquidDensity * inv * dt;
            pj -= mp[idx] / st.liquidDensity * inv * dt;
            jumpStep[a][fg[a].IndexOf(f)] = pj;
        });
        js[a] = jumpStep[a].data();
    }
    std::vector<double> rhs3(nu);
    Csr A3;
    A3.ptr.resize(nu + 1);
    A3.col.resize(7 * (size_t)nu);
    A3.val.resize(7 * (size_t)nu);
    orc_pressure_matrix(n[0], n[1], n[2], levelSet.GetData().data(), frac, vin, us, js, volError, g2m.data(),
                        m2g.data(), nu, A3.ptr.data(), A3.col.data(), A3.val.data(), rhs3.data());
    for (int r = 0; r < nu; r++) pc[r] = step.GetPressure().GetPressure()[m2g[r]];
    double const rel3 = orc_rel_residual(nu, A3.ptr.data(), A3.col.data(), A3.val.data(), rhs3.data(), pc.data());
    check(rel3 <= 1e-8, "step: oracle relative residual of the GPU pressure (curvature + magnetic jump)", rel3);
    // given that pressure, update + extrapolation + wall condition match the oracle bit for bit
    std::vector<double> vchain[3];
    double *vc[3];
    for (int a = 0; a < 3; a++) {
        vchain[a] = velocity[a].GetData();
        vc[a] = vchain[a].data();
    }
    orc_pressure_apply(n[0], n[1], n[2], levelSet.GetData().data(), frac, js, g2m.data(), pc.data(), vc);
    std::vector<unsigned char> vmask[3];
    unsigned char *vm[3];
    unsigned char const *vmc[3];
    for (int a = 0; a < 3; a++) {
        vmask[a].assign(fg[a].GetNumVertices(), 0);
        vm[a] = vmask[a].data();
        vmc[a] = vm[a];
    }
    orc_velocity_valid(n[0], n[1], n[2], levelSet.GetData().data(), frac, vm);
    orc_extrapolate_faces(n[0], n[1], n[2], vc, vmc, 0.0, 6);
    double const *nrmp[3] = {collider.GetNormal()[0].GetData().data(), collider.GetNormal()[1].GetData().data(),
                             collider.GetNormal()[2].GetData().data()};
    orc_collider_enforce(n[0], n[1], n[2], sgrid.GetSpacing(), origin, vc, frac, nrmp, us);
    bool chain = true;
    for (int a = 0; a < 3; a++) chain &= vchain[a] == vstep[a].GetData();
    check(chain, "step: projected, extrapolated, wall-enforced velocity bit-identical to the oracle chain");
    // closed form: on a wall face away from edges the normal velocity equals the collider's
    double werr = 0;
    ForEach(fg[0], [&](Vector3i const &f) {
        if (f[0] != 0 || f[1] < 1 || f[2] < 1 || f[1] >= n[1] - 1 || f[2] >= n[2] - 1) return;
        werr = std::max(werr, std::fabs(vstep[0][f] - collider.Velocity[0][f]));
    });
    check(werr <= 1e-15, "step: wall-normal velocity equals the collider velocity on the x = min wall", werr);

    std::printf("%s (%d failed)\n", g_fail ? "FAILED" : "PASSED", g_fail);
    return g_fail ? 1 : 0;
}
