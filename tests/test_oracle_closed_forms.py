"""Closed forms that pin four branches of the FP64 oracle on their own, with a
planted mistake in the oracle (oracle.mutated) that each one must catch:

  * the Heaviside band interior of the magnetic permeability
    (core/MagneticField.h:44-54, core/MagneticField.cpp:50-60): on a planar
    interface the faces sit at phi = 0, +-eps/4, +-eps/2, +-3eps/4, +-eps, where
    H takes values fixed by sin(pi/4) and sin(pi/2);
  * the ghost-fluid jump's sign and scale at both of its sites
    (core/Pressure.cpp:64 and :159-163): with u = 0, closed walls and a
    uniform jump J, p = J on every liquid cell solves the system and the
    update leaves u unchanged; J -> -J flips p;
  * the max(theta, .001) clamp (core/Pressure.cpp:62,161): a liquid cell at
    phi = -1e-9 next to air gets exactly 1000 w on its diagonal, and the
    update moves its face by 1000 w p;
  * the trilinear weights and face-grid origins behind the vertex pressures
    (core/TriLerp.h:22-36,54-61, core/MagneticField.cpp:37-48): a linear total
    field is interpolated exactly at any interior point;
  * the library's opt-in Dirichlet box for the magnetic potential (no
    reference counterpart): with mu = 1 its operator is the 7-point Laplacian
    with phi = 0 on the box faces (mirrored ghosts), whose eigenvectors are
    products of sin(k pi (i + 1/2) / n) with eigenvalues sum 2 - 2 cos(k pi / n).
"""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as orc
from sfp import scenes


# ------------------------------------------------------------------ Heaviside band
def _planar_mu_errors(oracle, chi=3.0):
    """Largest |mu - closed form| over every face of a planar interface."""
    n = 16
    geo = scenes.Geometry(n, 1.0 / n)          # origin 0, every position exact
    h = geo.spacing
    eps = 2 * h
    y0 = 8 * h                                  # y-face 8 lies on the interface
    X, Y, Z = geo.cell_centers()
    phi = np.ascontiguousarray(np.broadcast_to(Y - y0 + 0 * X + 0 * Z, geo.n)).copy()
    sc = dict(geo=geo, phi=phi, chi=chi, band=eps)
    mu = oracle.magnetic_coeff(sc)
    s = math.sqrt(0.5)                          # sin(pi/4)
    # H(t eps) for the t the faces meet (t = +-1 and beyond: 1 or 0)
    H = {0.0: 0.5, 0.25: 0.5 * (1.25 + s / math.pi), -0.25: 0.5 * (0.75 - s / math.pi),
         0.5: 0.5 * (1.5 + 1 / math.pi), -0.5: 0.5 * (0.5 - 1 / math.pi),
         0.75: 0.5 * (1.75 + s / math.pi), -0.75: 0.5 * (0.25 - s / math.pi)}

    def expect(phi_f):
        t = round(phi_f / eps * 4) / 4
        if t >= 1.0:
            return 1.0
        if t <= -1.0:
            return 1.0 + chi
        return 1.0 + (1.0 - H[t]) * chi

    err = 0.0
    for a in range(3):
        shp = geo.face_shape(a)
        ys = np.arange(shp[1]) * h + (0.0 if a == 1 else h / 2)   # face y positions
        for j in range(shp[1]):
            m = mu[a][:, j, :]
            # faces on the box boundary of their own axis keep mu = 1
            inside = np.ones(m.shape, bool)
            if a == 0:
                inside[0, :] = inside[-1, :] = False
            elif a == 2:
                inside[:, 0] = inside[:, -1] = False
            elif j == 0 or j == n:
                inside[:] = False
            if inside.any():
                err = max(err, np.abs(m[inside] - expect(ys[j] - y0)).max())
            if not inside.all():
                err = max(err, np.abs(m[~inside] - 1.0).max())
    return err


def test_heaviside_band_interior_closed_form(oracle):
    assert _planar_mu_errors(oracle) <= 1e-13


@pytest.mark.parametrize("mutation", ["heaviside_no_sin", "heaviside_half_band"])
def test_heaviside_planted_mistakes_fail_the_closed_form(oracle, mutation):
    with orc.mutated(mutation):
        assert _planar_mu_errors(oracle) > 1e-3


# ------------------------------------------------------------------ jump sign and scale
def _jump_scenes(J):
    """u = 0, no solid velocity, no volume error, a uniform jump J on every
    face: a flat pool and the sphere-over-pool scene (walls closed)."""
    out = []
    n = 12
    geo = scenes.Geometry(n, 1.0 / n)
    X, Y, Z = geo.cell_centers()
    flat = np.ascontiguousarray(np.broadcast_to(Y - 0.4 - 0.3 / n + 0 * X + 0 * Z, geo.n)).copy()
    drop = scenes.pressure_scene(n, seed=3)["phi"]
    for phi in (flat, drop):
        out.append(dict(geo=geo, phi=phi, fraction=scenes.wall_fractions(geo),
                        velocity=[np.zeros(geo.face_shape(a)) for a in range(3)],
                        solid_velocity=[np.zeros(geo.face_shape(a)) for a in range(3)],
                        jump=[np.full(geo.face_shape(a), J) for a in range(3)], vol_error=0.0))
    return out


def _jump_errors(oracle, J):
    """(max |p - J| / |J|, max |u_after|) over the jump scenes."""
    ep, eu = 0.0, 0.0
    for sc in _jump_scenes(J):
        ref = oracle.pressure_system(sc)
        p = spla.spsolve(ref["A"].tocsc(), ref["rhs"])
        ep = max(ep, np.abs(p - J).max() / abs(J))
        vel = oracle.pressure_apply(sc, ref["grid2mat"], p)
        eu = max(eu, max(np.abs(v).max() for v in vel))
    return ep, eu


@pytest.mark.parametrize("J", [0.37, -0.37])
def test_uniform_jump_gives_uniform_liquid_pressure(oracle, J):
    ep, eu = _jump_errors(oracle, J)
    assert ep <= 1e-12
    assert eu <= 1e-9 * abs(J)


@pytest.mark.parametrize("mutation", ["jump_sign", "jump_no_coef"])
def test_jump_planted_mistakes_fail_the_closed_form(oracle, mutation):
    with orc.mutated(mutation):
        ep, _ = _jump_errors(oracle, 0.37)
    assert ep > 1e-3


# ------------------------------------------------------------------ theta clamp
def _clamp_scene():
    """Liquid below y = 4 cells (phi = -1) and air above (phi = 1), a bubble
    cell at (2,1,2); the liquid cells (4,3,4) (air at +y) and (2,2,2) (air at
    -y) sit at phi = -1e-9, so theta ~ 1e-9 there, clamped to .001."""
    n = 8
    geo = scenes.Geometry(n, 1.0 / n)
    phi = np.where(np.arange(n)[None, :, None] >= 4, 1.0, -1.0) * np.ones(geo.n)
    phi[2, 1, 2] = 1.0
    phi[4, 3, 4] = -1e-9
    phi[2, 2, 2] = -1e-9
    rng = np.random.default_rng(7)
    return dict(geo=geo, phi=phi, fraction=scenes.wall_fractions(geo),
                velocity=[rng.uniform(-1, 1, geo.face_shape(a)) for a in range(3)],
                solid_velocity=[np.zeros(geo.face_shape(a)) for a in range(3)],
                jump=None, vol_error=0.0)


def _clamp_errors(oracle):
    """Deviations from the clamped closed form: diagonal 5 + 1000 at both
    cells, and the y-face next to each moved by exactly 1000 p (the update
    subtracts (+-1)(0 - p) 1000 w with w = 1)."""
    sc = _clamp_scene()
    ref = oracle.pressure_system(sc)
    g2m = ref["grid2mat"]
    diag = ref["A"].diagonal()
    rng = np.random.default_rng(1)
    p = rng.uniform(0.5, 1.5, ref["n"])
    vel = oracle.pressure_apply(sc, g2m, p)
    errs = []
    for cell, face, sign in (((4, 3, 4), (4, 4, 4), +1), ((2, 2, 2), (2, 2, 2), -1)):
        r = g2m[cell]
        errs.append(abs(diag[r] - 1005.0))
        du = vel[1][face] - sc["velocity"][1][face]
        errs.append(abs(du / (sign * p[r]) - 1000.0) / 1000.0)
    return max(errs)


def test_theta_clamp_closed_form(oracle):
    assert _clamp_errors(oracle) <= 1e-13


def test_clamp_planted_mistake_fails_the_closed_form(oracle):
    with orc.mutated("no_clamp"):
        assert _clamp_errors(oracle) > 1.0


# ------------------------------------------------------------------ trilinear vertex pressures
def _linear_field_errors(oracle):
    """Relative error of the vertex pressures of a linear total field
    H(x) = c + G x at random interior points, against the closed form."""
    geo = scenes.Geometry((10, 12, 14), 1.0 / 16, center=(0.3, 0.55, 0.45))
    rng = np.random.default_rng(11)
    c = rng.uniform(-1, 1, 3)
    G = rng.uniform(-2, 2, (3, 3))
    Ht = []
    for a in range(3):
        Xf, Yf, Zf = geo.face_centers(a)
        Ht.append(np.ascontiguousarray(c[a] + G[a, 0] * Xf + G[a, 1] * Yf + G[a, 2] * Zf
                                       + 0 * (Xf + Yf + Zf)))
    lo = np.array(geo.origin) + 1.5 * geo.spacing
    hi = np.array(geo.origin) + (np.array(geo.n) - 1.5) * geo.spacing
    pos = rng.uniform(lo, hi, (200, 3))
    nrm = rng.standard_normal((200, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    chi = 1.7
    sc = dict(geo=geo, chi=chi)
    got = oracle.magnetic_vertex_pressures(sc, Ht, pos, nrm)
    H = c[None, :] + pos @ G.T
    hn = np.sum(H * nrm, axis=1)
    # magnetic pressure mu0 chi / 2 (|H|^2 + chi^2 / (4 chi + 4) (H.n)^2)
    exp = 4e-7 * np.pi * 0.5 * chi * (np.sum(H * H, axis=1) + chi * chi / (4 * chi + 4) * hn * hn)
    return np.abs(got - exp).max() / np.abs(exp).max()


def test_vertex_pressures_of_a_linear_field(oracle):
    assert _linear_field_errors(oracle) <= 1e-12


@pytest.mark.parametrize("mutation", ["trilerp_swap", "face_origin"])
def test_trilerp_planted_mistakes_fail_the_closed_form(oracle, mutation):
    with orc.mutated(mutation):
        assert _linear_field_errors(oracle) > 1e-4


# ------------------------------------------------------------------ Dirichlet box
def _box_errors(oracle, shape=(7, 5, 6)):
    geo = scenes.Geometry(shape, 1.0 / max(shape))
    X, Y, Z = geo.cell_centers()
    phi = np.ascontiguousarray(np.broadcast_to(Y - 0.5 + 0 * X + 0 * Z, geo.n)).copy()
    sc = dict(geo=geo, phi=phi, chi=0.0, band=1.5 * geo.spacing,
              hext=[np.zeros(geo.face_shape(a)) for a in range(3)])
    A = orc.magnetic_system(sc, bc=1)["A"]
    idx = [np.arange(n) + 0.5 for n in shape]
    worst = 0.0
    for k in [(1, 1, 1), (2, 3, 1), (7, 5, 6), (4, 1, 5)]:
        f = [np.sin(k[a] * np.pi * idx[a] / shape[a]) for a in range(3)]
        u = (f[0][:, None, None] * f[1][None, :, None] * f[2][None, None, :]).ravel()
        lam = sum(2.0 - 2.0 * np.cos(k[a] * np.pi / shape[a]) for a in range(3))
        worst = max(worst, np.linalg.norm(A @ u - lam * u) / np.linalg.norm(u))
    return worst, A


def test_dirichlet_box_eigenvectors(oracle):
    err, A = _box_errors(oracle)
    assert err <= 1e-12
    # no pinned row: row 0 has its neighbours; the operator is symmetric
    assert A[0, 1] == -1.0 and A[0, 0] == 3.0 + 3 * 2.0  # three neighbours, three box faces
    assert abs(A - A.T).max() == 0.0


def test_box_planted_mistake_fails_the_closed_form(oracle):
    with orc.mutated("box_face_weight"):
        err, _ = _box_errors(oracle)
    assert err > 1e-2
