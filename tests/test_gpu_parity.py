"""GPU parity of the two Poisson solves against the FP64 oracle.

Assembly (matrix and right-hand side) is compared bit for bit with the
oracle's restatement of core/Pressure.cpp:38-78 and core/MagneticField.cpp:50-118;
the magnetic permeability goes through sin() (device vs host libm), so those
entries are compared to 4 ulp. Solves are compared through the FP64 relative
residual on the oracle's own CSR matrix (<= tol) and the L-infinity distance
to the oracle's 1e-13 solution. The velocity and field updates are compared
bit for bit given the GPU's solution.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import sfp
from sfp import scenes
from helpers import aggregation_prolongation, csr_from_full_grid, full_grid_operator, ulp_diff

pytestmark = pytest.mark.gpu


def _pressure_scenes():
    return [("pool_drop_32", scenes.pressure_scene(32, seed=3)),
            ("random_16", scenes.random_pressure_scene(16, seed=5)),
            ("random_ragged", scenes.random_pressure_scene((13, 17, 11), seed=7))]


@pytest.mark.parametrize("name,sc", _pressure_scenes())
def test_pressure_assembly_bit_exact(oracle, name, sc):
    print(f"[case] pressure assembly {name} shape={sc['geo'].n}", flush=True)
    ref = oracle.pressure_system(sc)
    with sfp.Context.for_scene(sc) as ctx:
        vel, p, st = ctx.pressure_project(sc)
        got = ctx.export_level(0, 0)
    diag, cpl, rhs = full_grid_operator(ref["A"], ref["rhs"], ref["mat2grid"], sc["geo"].n)
    assert st["unknowns"] == ref["n"]
    assert np.array_equal(got["diag"], diag)
    for a in range(3):
        assert np.array_equal(got["cpl"][a], cpl[a]), f"coupling axis {a}"
    assert np.array_equal(got["rhs"], rhs)


@pytest.mark.parametrize("name,sc", _pressure_scenes())
def test_pressure_project_parity(oracle, name, sc):
    print(f"[case] pressure project {name} shape={sc['geo'].n}", flush=True)
    ref = oracle.pressure_system(sc)
    A, b, m2g = ref["A"], ref["rhs"], ref["mat2grid"]
    x_ref = spla.spsolve(A.tocsc(), b)
    with sfp.Context.for_scene(sc, tol=1e-10) as ctx:
        vel, p, st = ctx.pressure_project(sc)
    x = p.ravel()[m2g]
    rel = oracle.rel_residual(A, b, x)
    print(f"  iters={st['iterations']} rel={rel:.3e} gpu_rel={st['rel_residual']:.3e}", flush=True)
    assert st["converged"] == 1
    assert rel <= 1e-10 * 1.01
    assert abs(rel - st["rel_residual"]) <= 1e-12
    scale = max(np.abs(x_ref).max(), 1e-300)
    assert np.abs(x - x_ref).max() / scale < 1e-7
    # air cells carry no pressure
    air = np.ones(p.size, bool)
    air[m2g] = False
    assert np.all(p.ravel()[air] == 0)
    # velocity update, bit for bit given the GPU pressure
    ref_vel = oracle.pressure_apply(sc, ref["grid2mat"], x)
    for a in range(3):
        assert np.array_equal(vel[a], ref_vel[a]), f"velocity axis {a}"


def test_pressure_default_tol_and_reproject(oracle):
    sc = scenes.random_pressure_scene(16, seed=11)
    print("[case] pressure reproject random_16", flush=True)
    ref = oracle.pressure_system(sc)
    with sfp.Context.for_scene(sc) as ctx:
        vel, p, st = ctx.pressure_project(sc)
        assert st["converged"] == 1 and st["rel_residual"] <= 1e-8
        # Reproject: same matrix, right-hand side without jump terms, warm start
        sc2 = dict(sc)
        sc2["velocity"] = [v * 0.5 + 0.1 for v in sc["velocity"]]
        vel2, p2, st2 = ctx.pressure_project(sc2, reproject=True)
    rhs2 = oracle.pressure_rhs(sc2, ref["mat2grid"])
    x2 = p2.ravel()[ref["mat2grid"]]
    assert oracle.rel_residual(ref["A"], rhs2, x2) <= 1e-8
    ref_vel = oracle.pressure_apply(sc2, ref["grid2mat"], x2)
    for a in range(3):
        assert np.array_equal(vel2[a], ref_vel[a])


@pytest.mark.parametrize("shape", [(12, 14, 16), (12, 14, 15), (9, 6, 130)])
def test_pressure_apply64_matches_csr(oracle, shape):
    """The FP64 matvec of the FCG loop (over cell pairs where nz is even, over
    x-plane tiles where it is odd) against the oracle's CSR matrix."""
    sc = scenes.random_pressure_scene(shape, seed=2)
    print(f"[case] apply64 {shape}", flush=True)
    ref = oracle.pressure_system(sc)
    with sfp.Context.for_scene(sc) as ctx:
        ctx.pressure_project(sc)
        rng = np.random.default_rng(0)
        z = rng.standard_normal(sc["geo"].n).astype(np.float32)
        liquid = (sc["phi"] <= 0)
        z = np.where(liquid, z, 0).astype(np.float32)
        w = ctx.apply64(0, z)
    m2g = ref["mat2grid"]
    w_ref = oracle.spmv(ref["A"], z.ravel()[m2g].astype(np.float64))
    err = np.abs(w.ravel()[m2g] - w_ref).max() / np.abs(w_ref).max()
    assert err < 1e-14


def _mag(name):
    if name == "spike_ragged":
        geo = scenes.Geometry((18, 20, 22), 1.0 / 22, center=(0.41, 0.5, 0.55))
        X, Y, Z = geo.cell_centers()
        phi = np.ascontiguousarray(np.broadcast_to(Y - 0.45 - 0.04 * np.cos(9 * X + 2 * Z), geo.n)).copy()
        return dict(geo=geo, phi=phi, hext=[np.full(geo.face_shape(a), (0.3, 1.5, -0.2)[a]) for a in range(3)],
                    chi=4.0, band=2.0 * geo.spacing)
    n, chi = (32, 2.0) if name == "wavy_32_chi2" else (24, 8.0)
    return scenes.magnetic_scene(n, seed=1, chi=chi)


@pytest.mark.parametrize("name", ["wavy_32_chi2", "wavy_24_chi8", "spike_ragged"])
def test_magnetic_assembly(oracle, name):
    sc = _mag(name)
    print(f"[case] magnetic assembly {name} shape={sc['geo'].n}", flush=True)
    ref = oracle.magnetic_system(sc)
    with sfp.Context.for_scene(sc) as ctx:
        ctx.magnetic_solve(sc)
        got = ctx.export_level(1, 0)
    n = sc["geo"].ncells
    diag, cpl, rhs = full_grid_operator(ref["A"], ref["rhs"], np.arange(n), sc["geo"].n)
    assert ulp_diff(got["diag"], diag) <= 4
    for a in range(3):
        assert ulp_diff(got["cpl"][a], cpl[a]) <= 4
    assert np.abs(got["rhs"] - rhs).max() <= 8 * np.finfo(float).eps * np.abs(rhs).max()


@pytest.mark.parametrize("name", ["wavy_32_chi2", "wavy_24_chi8", "spike_ragged"])
def test_magnetic_solve_parity(oracle, name):
    sc = _mag(name)
    print(f"[case] magnetic solve {name} shape={sc['geo'].n}", flush=True)
    ref = oracle.magnetic_system(sc)
    x_ref = spla.spsolve(ref["A"].tocsc(), ref["rhs"])
    with sfp.Context.for_scene(sc, tol=1e-10) as ctx:
        pot, H, st = ctx.magnetic_solve(sc)
    x = pot.ravel()
    rel = oracle.rel_residual(ref["A"], ref["rhs"], x)
    print(f"  iters={st['iterations']} rel={rel:.3e}", flush=True)
    assert st["converged"] == 1
    assert rel <= 1.01e-10
    assert x[0] == 0.0
    assert np.abs(x - x_ref).max() / np.abs(x_ref).max() < 1e-6
    H_ref = oracle.magnetic_total_field(sc, pot)
    for a in range(3):
        assert np.array_equal(H[a], H_ref[a])


# ------------------------------------------------------------------ opt-in Dirichlet box
@pytest.mark.parametrize("name", ["wavy_32_chi2", "spike_ragged"])
def test_magnetic_dirichlet_box(oracle, name):
    """sfp_params.magnetic_bc = SFP_MAGNETIC_BC_DIRICHLET_BOX (no reference
    counterpart): the assembled operator and right-hand side against the
    oracle's bc = 1 system (within 4 ulp like the reference path's: the
    permeability goes through the device's sin in the band), the solve to
    tolerance on that system, and the total field bit for bit given the
    potential."""
    sc = _mag(name)
    print(f"[case] magnetic Dirichlet box {name} shape={sc['geo'].n}", flush=True)
    ref = oracle.magnetic_system(sc, bc=1)
    x_ref = spla.spsolve(ref["A"].tocsc(), ref["rhs"])
    with sfp.Context.for_scene(sc, tol=1e-10, magnetic_bc=sfp.MAGNETIC_BC_DIRICHLET_BOX) as ctx:
        pot, H, st = ctx.magnetic_solve(sc)
        got = ctx.export_level(1, 0)
    n = sc["geo"].ncells
    diag, cpl, rhs = full_grid_operator(ref["A"], ref["rhs"], np.arange(n), sc["geo"].n)
    assert ulp_diff(got["diag"], diag) <= 4
    for a in range(3):
        assert ulp_diff(got["cpl"][a], cpl[a]) <= 4
    assert np.abs(got["rhs"] - rhs).max() <= 8 * np.finfo(float).eps * np.abs(rhs).max()
    x = pot.ravel()
    rel = oracle.rel_residual(ref["A"], ref["rhs"], x)
    print(f"  iters={st['iterations']} rel={rel:.3e} unknowns={st['unknowns']}", flush=True)
    assert st["converged"] == 1 and rel <= 1.01e-10
    assert st["unknowns"] == n  # cell 0 is an unknown too
    assert np.abs(x - x_ref).max() / np.abs(x_ref).max() < 1e-6
    H_ref = oracle.magnetic_total_field(sc, pot)
    for a in range(3):
        assert np.array_equal(H[a], H_ref[a])
    # the reference boundary on the same context afterwards: pinned again
    with sfp.Context.for_scene(sc, tol=1e-10) as ctx:
        ctx.set_params(magnetic_bc=sfp.MAGNETIC_BC_DIRICHLET_BOX)
        ctx.magnetic_solve(sc, want_field=False)
        ctx.set_params(magnetic_bc=sfp.MAGNETIC_BC_REFERENCE)
        ctx.reset_warm_start()
        pot0, _, st0 = ctx.magnetic_solve(sc, want_field=False)
    ref0 = oracle.magnetic_system(sc)
    assert pot0.ravel()[0] == 0.0 and oracle.rel_residual(ref0["A"], ref0["rhs"], pot0.ravel()) <= 1.01e-10


@pytest.mark.parametrize("n", [(96, 40, 6), (7, 130, 3), (9, 11, 200)])
def test_anisotropic_grids_solve_to_tolerance(oracle, n):
    """Flat, thin and tall grids: hierarchies whose axes stop halving at
    different levels (single-cell axes, levels of every size class) still
    converge to the tolerance on the oracle's systems."""
    print(f"[case] anisotropic {n}", flush=True)
    geo = scenes.Geometry(n, 1.0 / max(n))
    X, Y, Z = geo.cell_centers()
    phi = np.ascontiguousarray(np.broadcast_to(
        Y - 0.47 * n[1] / max(n) - 0.03 * np.cos(7 * X + 3 * Z), geo.n)).copy()
    msc = dict(geo=geo, phi=phi, hext=[np.full(geo.face_shape(a), (0.2, 1.0, -0.1)[a]) for a in range(3)],
               chi=2.0, band=1.5 * geo.spacing)
    ref = oracle.magnetic_system(msc)
    with sfp.Context.for_scene(msc, tol=1e-9) as ctx:
        pot, _, st = ctx.magnetic_solve(msc)
    assert st["converged"] == 1
    assert oracle.rel_residual(ref["A"], ref["rhs"], pot.ravel()) <= 1.01e-9
    psc = scenes.random_pressure_scene(n, seed=3)
    pref = oracle.pressure_system(psc)
    with sfp.Context.for_scene(psc, tol=1e-9) as ctx:
        _, p, st = ctx.pressure_project(psc)
    assert st["converged"] == 1
    assert oracle.rel_residual(pref["A"], pref["rhs"], p.ravel()[pref["mat2grid"]]) <= 1.01e-9


@pytest.mark.parametrize("n", [(64, 64, 64), (21, 13, 50), (40, 18, 34)])
def test_magnetic_rowsum_diagonal_matches_stored(n):
    """The magnetic FCG matvec rebuilds each diagonal entry from the couplings
    it loads; solves with that and with the stored diagonal give the same
    potential and iteration count, bit for bit (the pinned cell's neighbours
    included)."""
    print(f"[case] row-sum diagonal {n}", flush=True)
    geo = scenes.Geometry(n, 1.0 / max(n))
    X, Y, Z = geo.cell_centers()
    phi = np.ascontiguousarray(np.broadcast_to(
        Y - 0.45 * n[1] / max(n) - 0.04 * np.cos(5 * X + 4 * Z), geo.n)).copy()
    msc = dict(geo=geo, phi=phi, hext=[np.full(geo.face_shape(a), (0.3, 1.0, -0.2)[a]) for a in range(3)],
               chi=3.0, band=1.5 * geo.spacing)
    out = []
    with sfp.Context.for_scene(msc) as ctx:
        for on in (True, False, True):
            ctx.set_rowsum_diagonal(on)
            ctx.reset_warm_start()
            pot, _, st = ctx.magnetic_solve(msc)
            assert st["converged"] == 1
            out.append((pot, st["iterations"]))
    for pot, its in out[1:]:
        assert its == out[0][1]
        np.testing.assert_array_equal(pot, out[0][0])


def test_magnetic_warm_start_and_vertex_pressures(oracle):
    sc = _mag("wavy_32_chi2")
    print("[case] magnetic warm start + vertex pressures", flush=True)
    rng = np.random.default_rng(3)
    pos = rng.uniform(0.05, 0.95, size=(500, 3))
    nrm = rng.standard_normal((500, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    with sfp.Context.for_scene(sc, tol=1e-9) as ctx:
        pot1, H1, st1 = ctx.magnetic_solve(sc)
        pot2, H2, st2 = ctx.magnetic_solve(sc)  # warm: starts at the solution
        vp = ctx.vertex_pressures(pos, nrm, sc["chi"])
    assert st2["iterations"] <= 1
    vp_ref = oracle.magnetic_vertex_pressures(sc, H2, pos, nrm)
    assert np.abs(vp - vp_ref).max() <= 4 * np.finfo(float).eps * np.abs(vp_ref).max()


def test_magnetic_zero_susceptibility_gives_applied_field(oracle):
    sc = _mag("wavy_32_chi2")
    print("[case] chi = 0", flush=True)
    with sfp.Context.for_scene(sc) as ctx:
        pot, H, st = ctx.magnetic_solve(sc, chi=0.0)
    assert np.all(pot == 0)
    for a in range(3):
        assert np.array_equal(H[a], sc["hext"][a])


@pytest.mark.parametrize("which,case", [(0, "ragged"), (1, "ragged"), (0, "z4"), (1, "z4")])
def test_coarsening_is_scaled_galerkin(oracle, which, case):
    """The coarse operators are ½ P^T A P of the aggregation prolongation P,
    on grids whose z extent is not a multiple of 4 (the x-plane tiled
    coarsening) and on grids where it is (the flat one). The magnetic
    operator (pinned cell 0) keeps its row sums unscaled: ½ G + ½ diag(G 1)
    with G = P^T A P, which differs from ½ G in the aggregate of cell 0."""
    print(f"[case] coarsening which={which} {case}", flush=True)
    if which == 0:
        sc = scenes.random_pressure_scene((12, 10, 14) if case == "ragged" else (16, 12, 20), seed=9)
    else:
        sc = _mag("spike_ragged") if case == "ragged" else scenes.magnetic_scene(32, seed=3)
    with sfp.Context.for_scene(sc) as ctx:
        if which == 0:
            ctx.pressure_project(sc)
        else:
            ctx.magnetic_solve(sc)
        scale = ctx.params().coarse_scale
        L0 = ctx.export_level(which, 0)
        L1 = ctx.export_level(which, 1)
        f32 = ctx.export_level32(which, 0)
    active = f32["diag"] > 0
    A0 = csr_from_full_grid(L0["diag"], L0["cpl"])
    P, cs = aggregation_prolongation(L0["diag"].shape, active)
    G = (P.T @ A0 @ P).tocsr()
    rs_scale = 1.0 if which == 1 else scale
    Ac = (scale * G + (rs_scale - scale) * sp.diags(G @ np.ones(G.shape[0]))).tocsr()
    A1 = csr_from_full_grid(L1["diag"], L1["cpl"])
    act_c = np.asarray(P.sum(axis=0)).ravel() > 0
    sel = np.flatnonzero(act_c)
    D = (A1[sel][:, sel] - Ac[sel][:, sel])
    assert abs(D).max() <= 1e-12 * abs(Ac).max()
    # inactive aggregates are identity rows
    ina = np.flatnonzero(~act_c)
    assert np.all(L1["diag"].ravel()[ina] == 1.0)


@pytest.mark.parametrize("shape,chi", [((33, 27, 45), 4.0), ((40, 36, 1), 2.0), ((24, 1, 30), 8.0)])
def test_magnetic_ragged_grids_vs_oracle(oracle, shape, chi):
    """The magnetic potential on grids with an odd row length (the x-plane
    tiled matvec and smoother) and one cell thick in y or z: the solve meets
    the tolerance on the oracle's system."""
    geo = scenes.Geometry(shape, 1.0 / max(shape))
    X, Y, Z = geo.cell_centers()
    phi = np.ascontiguousarray(np.broadcast_to(X + Y + Z - 0.5 - 0.05 * np.cos(9 * X + 4 * Z),
                                               geo.n)).copy()
    sc = dict(geo=geo, phi=phi, chi=chi, band=1.5 * geo.spacing,
              hext=[np.full(geo.face_shape(a), (0.2, 1.0, -0.3)[a]) for a in range(3)])
    print(f"[case] magnetic ragged grid {shape} chi={chi}", flush=True)
    ref = oracle.magnetic_system(sc)
    with sfp.Context.for_scene(sc, tol=1e-9, max_iters=200) as ctx:
        pot, _, st = ctx.magnetic_solve(sc, want_field=False)
    rel = oracle.rel_residual(ref["A"], ref["rhs"], pot.ravel())
    print(f"  {st['iterations']} its, oracle rel {rel:.2e}", flush=True)
    assert st["converged"] == 1 and rel <= 1e-9 * 1.01
