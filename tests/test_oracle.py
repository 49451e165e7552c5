"""Pins the FP64 CPU oracle (oracle/) on the CPU.

The reference ships no tests, fixtures or known answers and cannot be built
here, so the oracle is pinned by:
  * agreement, bit for bit, with an independent numpy transcription of the
    same reference code (tests/numpy_ref.py);
  * closed forms: K1 planar magnetic interface (B_y = H_ext exactly),
    K2 hydrostatic column (projected velocities vanish), K3 divergence after
    projection equals volError on every liquid cell, K5 manufactured
    solutions, chi = 0 gives H = H_ext;
  * invariants: symmetry, diagonal dominance, zero row sums of the pinned
    Neumann operator, the sum of the magnetic right-hand side;
  * planted mistakes: each mutation of the numpy transcription must fail at
    least one of the checks above.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from sfp import scenes
import numpy_ref as nr
from helpers import full_grid_operator, ulp_diff


def _pscenes():
    return {
        "pool_drop_16": scenes.pressure_scene(16, seed=2),
        "random_12": scenes.random_pressure_scene(12, seed=3),
        "random_ragged": scenes.random_pressure_scene((9, 13, 7), seed=4),
        "random_nojump": dict(scenes.random_pressure_scene(10, seed=5), jump=None),
    }


PSCENES = _pscenes()


# ------------------------------------------------------------------ pressure
@pytest.mark.parametrize("name", list(PSCENES))
def test_pressure_assembly_matches_numpy(oracle, name):
    sc = PSCENES[name]
    ref = oracle.pressure_system(sc)
    diag, cpl, rhs = full_grid_operator(ref["A"], ref["rhs"], ref["mat2grid"], sc["geo"].n)
    liq, d2, c2, r2 = nr.pressure_assembly(sc)
    assert ref["n"] == int(liq.sum())
    assert np.array_equal(ref["mat2grid"], np.flatnonzero(liq.ravel()))
    assert np.array_equal(diag, d2)
    for a in range(3):
        assert np.array_equal(cpl[a], c2[a])
    assert np.array_equal(rhs, r2)


@pytest.mark.parametrize("name", list(PSCENES))
def test_pressure_matrix_symmetric_and_dominant(oracle, name):
    A = oracle.pressure_system(PSCENES[name])["A"]
    assert abs(A - A.T).max() == 0
    d = A.diagonal()
    off = np.asarray(abs(A).sum(axis=1)).ravel() - np.abs(d)
    assert np.all(d >= off * (1 - 1e-15))
    assert np.all(A.data[A.indices != np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))] < 0)


@pytest.mark.parametrize("name", list(PSCENES))
def test_pressure_apply_matches_numpy(oracle, name):
    sc = PSCENES[name]
    ref = oracle.pressure_system(sc)
    rng = np.random.default_rng(7)
    p = rng.standard_normal(ref["n"])
    got = oracle.pressure_apply(sc, ref["grid2mat"], p)
    pf = np.zeros(sc["geo"].ncells)
    pf[ref["mat2grid"]] = p
    liq = sc["phi"] <= 0
    exp = nr.pressure_apply(sc, liq, pf.reshape(sc["geo"].n))
    for a in range(3):
        assert np.array_equal(got[a], exp[a])


def test_pressure_reproject_rhs_has_no_jump_terms(oracle):
    sc = PSCENES["random_12"]
    ref = oracle.pressure_system(sc)
    r2 = oracle.pressure_rhs(sc, ref["mat2grid"])
    r3 = oracle.pressure_system(dict(sc, jump=None))["rhs"]
    assert np.array_equal(r2, r3)


def _singular_components(A):
    """Connected groups of unknowns with no Dirichlet (ghost-fluid) contact:
    their block of A is a singular pure-Neumann operator."""
    from scipy.sparse.csgraph import connected_components
    ncomp, lab = connected_components(A, directed=False)
    d = A.diagonal()
    extra = d - (np.asarray(abs(A).sum(axis=1)).ravel() - np.abs(d))
    bad = []
    for c in range(ncomp):
        m = lab == c
        if m.sum() > 1 and extra[m].max() <= 1e-12:
            bad.append(c)
    return bad


def _k3_scene(n=12, seed=0, jump=True, vol_error=0.0123):
    """Walls with w in {0,1}: closed box plus random closed interior faces,
    redrawn until every liquid region touches air (a nonsingular system)."""
    import oracle as o
    for attempt in range(50):
        sc = scenes.random_pressure_scene(n, seed=seed, vol_error=vol_error)
        rng = np.random.default_rng(1000 * seed + attempt)
        for a in range(3):
            f = np.where(rng.random(sc["fraction"][a].shape) < 0.04, 1.0, 0.0)
            sc["fraction"][a] = np.maximum(f, scenes.wall_fractions(sc["geo"])[a])
        if not jump:
            sc["jump"] = None
        if not _singular_components(o.pressure_system(sc)["A"]):
            return sc
    raise RuntimeError("no nonsingular K3 scene found")


def _outflow(sc, vel, liq):
    """sum over faces of side * (w u + (1 - w) u_solid) for every cell."""
    n = sc["phi"].shape
    out = np.zeros(n)
    for axis in range(3):
        w = 1 - sc["fraction"][axis]
        flux = w * vel[axis] + (1 - w) * sc["solid_velocity"][axis]
        lo = [slice(0, n[b]) for b in range(3)]
        hi = [slice(0, n[b]) for b in range(3)]
        hi[axis] = slice(1, n[axis] + 1)
        out += flux[tuple(hi)] - flux[tuple(lo)]
    return out


def _k3_residual(sc, assemble, apply_):
    """max |outflow - volError| over non-isolated liquid cells, relative to max |u|."""
    liq, diag, cpl, rhs = assemble(sc)
    from helpers import csr_from_full_grid
    A = csr_from_full_grid(diag, cpl)
    idx = np.flatnonzero(liq.ravel())
    Al = A[idx][:, idx].tocsc()
    p = np.zeros(liq.size)
    p[idx] = spla.spsolve(Al, rhs.ravel()[idx])
    vel = apply_(sc, liq, p.reshape(liq.shape))
    out = _outflow(sc, vel, liq)
    open_faces = np.zeros(liq.shape, bool)
    n = liq.shape
    for axis in range(3):
        w = 1 - sc["fraction"][axis]
        lo = [slice(0, n[b]) for b in range(3)]
        hi = [slice(0, n[b]) for b in range(3)]
        hi[axis] = slice(1, n[axis] + 1)
        open_faces |= (w[tuple(lo)] > 0) | (w[tuple(hi)] > 0)
    cells = liq & open_faces
    scale = max(np.abs(v).max() for v in sc["velocity"])
    return np.abs(out[cells] - sc["vol_error"]).max() / scale


def _oracle_assemble(oracle):
    def f(sc):
        ref = oracle.pressure_system(sc)
        d, c, r = full_grid_operator(ref["A"], ref["rhs"], ref["mat2grid"], sc["geo"].n)
        return sc["phi"] <= 0, d, c, r
    return f


def _oracle_apply(oracle):
    def f(sc, liq, p):
        g2m, m2g = oracle.pressure_unknowns(sc["phi"])
        return oracle.pressure_apply(sc, g2m, p.ravel()[m2g])
    return f


@pytest.mark.parametrize("jump", [True, False])
def test_k3_divergence_equals_vol_error(oracle, jump):
    sc = _k3_scene(12, seed=1, jump=jump)
    assert _k3_residual(sc, _oracle_assemble(oracle), _oracle_apply(oracle)) < 1e-10
    assert _k3_residual(sc, nr.pressure_assembly, nr.pressure_apply) < 1e-10


def test_k2_hydrostatic_column(oracle):
    n = 14
    geo = scenes.Geometry(n, 1.0 / n)
    X, Y, Z = geo.cell_centers()
    phi = np.ascontiguousarray(np.broadcast_to(Y - 0.537 + 0 * X + 0 * Z, geo.n)).copy()
    g_dt = 9.8 * 1e-3
    vel = [np.zeros(geo.face_shape(a)) for a in range(3)]
    vel[1][:] = -g_dt
    sc = dict(geo=geo, phi=phi, velocity=vel, fraction=scenes.wall_fractions(geo),
              solid_velocity=[np.zeros(geo.face_shape(a)) for a in range(3)], jump=None,
              vol_error=0.0)
    ref = oracle.pressure_system(sc)
    p = spla.spsolve(ref["A"].tocsc(), ref["rhs"])
    new = oracle.pressure_apply(sc, ref["grid2mat"], p)
    liq = phi <= 0
    for a in range(3):
        w = 1 - sc["fraction"][a]
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a] = slice(0, n)
        hi[a] = slice(1, n + 1)
        touched = np.zeros(geo.face_shape(a), bool)
        pad = np.zeros(geo.face_shape(a), bool)
        pad[tuple(lo)] |= liq
        pad[tuple(hi)] |= liq
        touched = pad & (w > 0)
        assert np.abs(new[a][touched]).max() < 1e-12 * g_dt
    # the liquid pressure rises linearly with depth (hydrostatic)
    pf = np.zeros(geo.ncells)
    pf[ref["mat2grid"]] = p
    col = pf.reshape(geo.n)[3, :, 5]
    depth = col[liq[3, :, 5]]
    assert np.all(np.diff(depth) < 0)


def test_k5_manufactured_solution_all_solvers(oracle):
    sc = PSCENES["random_12"]
    A = oracle.pressure_system(sc)["A"]
    x_true = np.random.default_rng(3).standard_normal(A.shape[0])
    b = oracle.spmv(A, x_true)
    assert np.allclose(b, A @ x_true, rtol=1e-14, atol=1e-14)
    for method, pc in [("cg", "amg"), ("bicgstab", "amg"), ("cg", "jacobi"), ("bicgstab", "jacobi")]:
        x, it, rel, _, _ = oracle.solve(A, b, method, pc, tol=1e-12, maxit=2000)
        assert rel <= 1e-12, (method, pc, rel)
        assert np.abs(x - x_true).max() < 1e-9 * np.abs(x_true).max(), (method, pc)


def test_amg_iteration_counts_are_mesh_independent(oracle):
    its = []
    for n in (16, 32):
        ps = oracle.pressure_system(scenes.pressure_scene(n, seed=0))
        _, it, rel, _, _ = oracle.solve(ps["A"], ps["rhs"], "bicgstab", "amg", tol=1e-8)
        assert rel <= 1e-8
        its.append(it)
    assert max(its) <= 15 and its[1] <= its[0] + 5


# ------------------------------------------------------------------ magnetic
def _mscenes():
    ragged = scenes.Geometry((9, 11, 7), 1.0 / 11, center=(0.43, 0.5, 0.52))
    X, Y, Z = ragged.cell_centers()
    phi = np.ascontiguousarray(np.broadcast_to(Y - 0.45 - 0.05 * np.cos(7 * X + 3 * Z), ragged.n)).copy()
    return {
        "wavy_12_chi2": scenes.magnetic_scene(12, seed=1, chi=2.0),
        "wavy_10_chi8": scenes.magnetic_scene(10, seed=2, chi=8.0),
        "ragged_chi4": dict(geo=ragged, phi=phi, chi=4.0, band=2 * ragged.spacing,
                            hext=[np.full(ragged.face_shape(a), (0.3, 1.2, -0.4)[a]) for a in range(3)]),
    }


MSCENES = _mscenes()


@pytest.mark.parametrize("name", list(MSCENES))
def test_magnetic_mu_matches_numpy(oracle, name):
    sc = MSCENES[name]
    mu = oracle.magnetic_coeff(sc)
    mu2 = nr.magnetic_mu(sc)
    for a in range(3):
        assert np.abs(mu[a] - mu2[a]).max() <= 1e-13 * sc["chi"]
        n = sc["geo"].n
        lo = [slice(None)] * 3
        lo[a] = 0
        assert np.all(mu[a][tuple(lo)] == 1.0)
        assert mu[a].min() >= 1.0 and mu[a].max() <= 1.0 + sc["chi"]


@pytest.mark.parametrize("name", list(MSCENES))
def test_magnetic_system_matches_numpy(oracle, name):
    sc = MSCENES[name]
    ref = oracle.magnetic_system(sc)
    A2, b2 = nr.magnetic_system(sc, ref["mu"])
    D = ref["A"] - A2
    assert D.nnz == 0 or abs(D).max() == 0
    assert np.array_equal(ref["A"].indptr, A2.indptr) and np.array_equal(ref["A"].indices, A2.indices)
    assert np.array_equal(ref["rhs"], b2)


@pytest.mark.parametrize("name", list(MSCENES))
def test_magnetic_invariants(oracle, name):
    sc = MSCENES[name]
    ref = oracle.magnetic_system(sc)
    A, b, mu = ref["A"], ref["rhs"], ref["mu"]
    geo = sc["geo"]
    assert abs(A - A.T).max() == 0
    # A 1: rows sum to zero except row 0 (identity) and the neighbours of cell 0,
    # whose dropped column leaves mu of the shared face
    rs = A @ np.ones(A.shape[0])
    exp = np.zeros(A.shape[0])
    exp[0] = 1.0
    strides = (geo.n[1] * geo.n[2], geo.n[2], 1)
    for a in range(3):
        f = [0, 0, 0]
        f[a] = 1
        exp[strides[a]] = mu[a][tuple(f)]
    assert np.abs(rs - exp).max() <= 1e-13 * (1 + sc["chi"])
    # sum of the right-hand side: interior face terms telescope; what is left
    # is minus cell 0's own (dropped) row
    h = sc["hext"]
    row0 = 0.0
    for a in range(3):
        f = [0, 0, 0]
        f[a] = 1
        row0 += (mu[a][tuple(f)] - 1.0) * h[a][tuple(f)] * (-1) * geo.spacing
    assert abs(b.sum() + row0) <= 1e-12 * max(1.0, np.abs(b).sum())


@pytest.mark.parametrize("name", list(MSCENES))
def test_magnetic_diagonal_is_ordered_coupling_sum(oracle, name):
    """Away from cell 0 and its upper neighbours, each diagonal entry equals,
    bit for bit, the row's off-diagonal magnitudes added in the order -x, +x,
    -y, +y, -z, +z (what the FCG matvec rebuilds instead of reading the
    diagonal); at those four cells it does not."""
    sc = MSCENES[name]
    A = oracle.magnetic_system(sc)["A"].tocoo()
    nx, ny, nz = sc["geo"].n
    n = nx * ny * nz
    off = {}
    for r, c, v in zip(A.row, A.col, A.data):
        if r != c:
            off[(int(r), int(c) - int(r))] = -float(v)
    diag = A.diagonal()
    strides = (ny * nz, nz, 1)
    special = {0, 1, nz, ny * nz}
    for r in range(n):
        acc = 0.0
        for s in strides:
            acc = (acc + off.get((r, -s), 0.0)) + off.get((r, s), 0.0)
        if r in special:
            assert acc != diag[r]
        else:
            assert acc == diag[r], r


@pytest.mark.parametrize("name", list(MSCENES))
def test_matrix_free_magnetic_residual_matches_csr(oracle, name):
    """The row-by-row residual used at sizes whose CSR does not fit agrees
    with the CSR system's, and with its right-hand side."""
    sc = MSCENES[name]
    ref = oracle.magnetic_system(sc)
    assert np.array_equal(oracle.magnetic_rhs(sc, ref["mu"]), ref["rhs"])
    rng = np.random.default_rng(5)
    for x in (rng.standard_normal(ref["n"]), spla.spsolve(ref["A"].tocsc(), ref["rhs"])):
        r_csr = oracle.rel_residual(ref["A"], ref["rhs"], x)
        r_mf = oracle.magnetic_rel_residual(sc, ref["mu"], ref["rhs"], x)
        assert abs(r_mf - r_csr) <= 1e-12 * max(r_csr, 1e-16) + 1e-15


def test_k1_planar_interface_gives_exact_flux(oracle):
    n = 12
    geo = scenes.Geometry(n, 1.0 / n)
    X, Y, Z = geo.cell_centers()
    y0 = 0.45 + 0.3 / n
    phi = np.ascontiguousarray(np.broadcast_to(Y - y0 + 0 * X + 0 * Z, geo.n)).copy()
    H = 1.7
    sc = dict(geo=geo, phi=phi, chi=3.0, band=1.5 / n,
              hext=[np.full(geo.face_shape(a), (0.0, H, 0.0)[a]) for a in range(3)])
    ref = oracle.magnetic_system(sc)
    pot = spla.spsolve(ref["A"].tocsc(), ref["rhs"])
    Ht = oracle.magnetic_total_field(sc, pot)
    mu = ref["mu"]
    inner = (slice(None), slice(1, n), slice(None))
    assert np.abs(mu[1][inner] * Ht[1][inner] - H).max() < 1e-10 * H
    for a in (0, 2):
        s = [slice(None)] * 3
        s[a] = slice(1, n)
        assert np.abs(Ht[a][tuple(s)]).max() < 1e-10 * H
    P = pot.reshape(geo.n)
    assert np.abs(P - P[:1, :, :1]).max() < 1e-10 * np.abs(P).max()


def test_chi_zero_gives_applied_field(oracle):
    sc = dict(MSCENES["wavy_12_chi2"], chi=0.0)
    ref = oracle.magnetic_system(sc)
    assert np.all(ref["rhs"] == 0)
    x, it, rel, _, _ = oracle.solve(ref["A"], ref["rhs"], "cg", "amg")
    assert it == 0 and np.all(x == 0)
    Ht = oracle.magnetic_total_field(sc, x)
    for a in range(3):
        assert np.array_equal(Ht[a], sc["hext"][a])


def test_total_field_matches_numpy(oracle):
    sc = MSCENES["ragged_chi4"]
    pot = np.random.default_rng(1).standard_normal(sc["geo"].ncells)
    a = oracle.magnetic_total_field(sc, pot)
    b = nr.magnetic_total_field(sc, pot)
    for i in range(3):
        assert np.array_equal(a[i], b[i])


def test_vertex_pressures_closed_form(oracle):
    sc = MSCENES["wavy_12_chi2"]
    geo = sc["geo"]
    # uniform total field: trilinear interpolation returns it exactly
    Hc = np.array([0.3, 1.1, -0.2])
    Ht = [np.full(geo.face_shape(a), Hc[a]) for a in range(3)]
    rng = np.random.default_rng(0)
    pos = rng.uniform(0.0, 1.0, (50, 3))
    nrm = rng.standard_normal((50, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    got = oracle.magnetic_vertex_pressures(sc, Ht, pos, nrm)
    chi = sc["chi"]
    hn = nrm @ Hc
    exp = 4e-7 * np.pi * 0.5 * chi * (Hc @ Hc + chi * chi / (4 * chi + 4) * hn * hn)
    assert np.abs(got - exp).max() <= 1e-14 * exp.max()


# ------------------------------------------------------------------ planted mistakes
def _pressure_checks_fail(oracle, mutation):
    """True when at least one pressure check rejects the mutated transcription."""
    sc = _k3_scene(12, seed=2, jump=True)
    if _k3_residual(sc, lambda s: nr.pressure_assembly(s, mutation),
                    lambda s, l, p: nr.pressure_apply(s, l, p, mutation)) > 1e-8:
        return True
    for name, s in list(PSCENES.items()) + [("near_interface", _near_interface_scene())]:
        ref = oracle.pressure_system(s)
        d, c, r = full_grid_operator(ref["A"], ref["rhs"], ref["mat2grid"], s["geo"].n)
        liq, d2, c2, r2 = nr.pressure_assembly(s, mutation)
        if not (np.array_equal(d, d2) and np.array_equal(r, r2)):
            return True
    return False


def _near_interface_scene():
    sc = scenes.random_pressure_scene(10, seed=8)
    phi = sc["phi"].copy()
    phi[np.abs(phi) < 0.02] = -1e-6   # liquid cells almost on the interface: theta < .001
    sc["phi"] = phi
    return sc


@pytest.mark.parametrize("mutation", ["no_jump", "solid_sign", "theta_side", "no_clamp"])
def test_planted_pressure_mistakes_are_caught(oracle, mutation):
    assert not _pressure_checks_fail(oracle, None)
    assert _pressure_checks_fail(oracle, mutation)


def _magnetic_checks_fail(oracle, mutation):
    for name, sc in MSCENES.items():
        mu = oracle.magnetic_coeff(sc)
        mu2 = nr.magnetic_mu(sc, mutation)
        if any(np.abs(mu[a] - mu2[a]).max() > 1e-13 * sc["chi"] for a in range(3)):
            return True
        A, b = nr.magnetic_system(sc, mu, mutation)
        if abs(A - A.T).max() != 0:
            return True
    # closed form K1 on the mutated transcription
    n = 10
    geo = scenes.Geometry(n, 1.0 / n)
    X, Y, Z = geo.cell_centers()
    phi = np.ascontiguousarray(np.broadcast_to(Y - 0.43 + 0 * X + 0 * Z, geo.n)).copy()
    sc = dict(geo=geo, phi=phi, chi=3.0, band=1.5 / n,
              hext=[np.full(geo.face_shape(a), (0.0, 1.0, 0.0)[a]) for a in range(3)])
    mu = nr.magnetic_mu(sc, mutation)
    A, b = nr.magnetic_system(sc, mu, mutation)
    pot = spla.spsolve(A.tocsc(), b)
    Ht = nr.magnetic_total_field(sc, pot)
    inner = (slice(None), slice(1, n), slice(None))
    return np.abs(mu[1][inner] * Ht[1][inner] - 1.0).max() > 1e-8


@pytest.mark.parametrize("mutation", ["band_half", "mu_side", "rhs_sign", "pin_keep"])
def test_planted_magnetic_mistakes_are_caught(oracle, mutation):
    assert not _magnetic_checks_fail(oracle, None)
    assert _magnetic_checks_fail(oracle, mutation)
