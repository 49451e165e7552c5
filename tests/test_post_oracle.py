"""Pins the oracle's restatement of the steps after the projection
(Simulation::ProjectVelocity, core/Simulation.cpp:236-241): the extrapolation
mask, Extrapolation::Solve (core/Extrapolation.cpp:4-30) and Collider::Enforce
(core/Collider.cpp:75-86), against an independent numpy transcription,
closed forms and planted mistakes."""
import numpy as np
import pytest

from sfp import scenes

SHIFTS = [(0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1)]  # ord 0..5


def np_velocity_valid(phi, frac):
    out = []
    liq = phi <= 0
    for a in range(3):
        pad = [(0, 0)] * 3
        pad[a] = (1, 1)
        lp = np.pad(liq, pad, constant_values=False)  # outside the grid: air
        sl0 = [slice(None)] * 3
        sl1 = [slice(None)] * 3
        sl0[a] = slice(0, -1)   # cell0 = face - e_a
        sl1[a] = slice(1, None)  # cell1 = face
        out.append(((frac[a] < 1) & (lp[tuple(sl0)] | lp[tuple(sl1)])).astype(np.uint8))
    return out


def np_extrapolate(data, valid, clear, steps):
    out = []
    for g, v0 in zip(data, valid):
        g = g.copy()
        v = v0.astype(bool)
        for _ in range(steps):
            s = np.zeros_like(g)
            cnt = np.zeros(g.shape, np.int64)
            for ax, side in SHIFTS:
                pad = [(0, 0)] * 3
                pad[ax] = (1, 1)
                gp = np.pad(g, pad)
                vp = np.pad(v, pad, constant_values=False)
                sl = [slice(None)] * 3
                sl[ax] = slice(1 + side, 1 + side + g.shape[ax])
                nv, ng = vp[tuple(sl)], gp[tuple(sl)]
                s = np.where(nv, s + ng, s)
                cnt += nv
            fill = (~v) & (cnt > 0)
            newg = np.where(fill, s / np.maximum(cnt, 1), g)
            newg = np.where((~v) & (cnt == 0), clear, newg)
            g = newg
            v = v | fill
        out.append(g)
    return out


def np_trilerp(shape, org, h, data, pos):
    inv = 1 / h
    lo = np.floor((pos - org) * inv - 0.0).astype(np.int64)
    fr = ((pos - org) - lo * h) * inv
    w = [1.0 - fr, fr]
    val = np.zeros(pos.shape[:-1])
    for p in range(8):
        o = (p & 1, (p >> 1) & 1, (p >> 2) & 1)
        c = [np.clip(lo[..., a] + o[a], 0, shape[a] - 1) for a in range(3)]
        wt = w[o[0]][..., 0] * w[o[1]][..., 1] * w[o[2]][..., 2]
        val = val + data[c[0], c[1], c[2]] * wt
    return val


def np_enforce(geo, vel, frac, normal, colvel, normalize=True):
    h = geo.spacing
    org = np.array(geo.origin)
    forg = [org + np.array([(0.0 if c == a else 1.0) * h / 2 for c in range(3)]) for a in range(3)]
    out = [v.copy() for v in vel]
    for a in range(3):
        idx = np.argwhere(frac[a] == 1.0)
        if idx.size == 0:
            continue
        pos = forg[a][None, :] + idx.astype(np.float64) * h
        v = [np_trilerp(vel[b].shape, forg[b], h, vel[b], pos) for b in range(3)]
        v0 = [np_trilerp(colvel[b].shape, forg[b], h, colvel[b], pos) for b in range(3)]
        n = [np_trilerp(normal[b].shape, forg[b], h, normal[b], pos) for b in range(3)]
        if normalize:
            z = n[0] * n[0] + n[1] * n[1] + n[2] * n[2]
            r = np.sqrt(np.where(z > 0, z, 1.0))
            n = [np.where(z > 0, nb / r, nb) for nb in n]
        dot = (v[0] - v0[0]) * n[0] + (v[1] - v0[1]) * n[1] + (v[2] - v0[2]) * n[2]
        out[a][tuple(idx.T)] = v[a] - dot * n[a]
    return out


def random_faces(geo, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-scale, scale, geo.face_shape(a)) for a in range(3)]


def fractions_with_walls(geo, seed):
    """Box walls (fraction 1) plus an interior band of exact 1s, exact 0s and
    fractional values."""
    rng = np.random.default_rng(seed)
    fr = scenes.wall_fractions(geo)
    out = []
    for a in range(3):
        r = rng.uniform(size=geo.face_shape(a))
        f = np.where(r < 0.15, 1.0, np.where(r < 0.3, 0.4 * r, 0.0))
        out.append(np.maximum(f, fr[a]))
    return out


@pytest.mark.parametrize("n", [(12, 12, 12), (9, 14, 11)])
def test_velocity_valid_matches_numpy(oracle, n):
    sc = scenes.random_pressure_scene(n, 3)
    got = oracle.velocity_valid(sc)
    want = np_velocity_valid(sc["phi"], sc["fraction"])
    for a in range(3):
        np.testing.assert_array_equal(got[a], want[a])
        assert got[a].sum() > 0


@pytest.mark.parametrize("steps", [0, 1, 6])
def test_extrapolate_matches_numpy(oracle, steps):
    geo = scenes.Geometry((10, 13, 9), 0.1)
    rng = np.random.default_rng(4)
    data = random_faces(geo, 5)
    valid = [(rng.uniform(size=geo.face_shape(a)) < 0.1).astype(np.uint8) for a in range(3)]
    got = oracle.extrapolate_faces(geo, data, valid, clear=-7.5, steps=steps)
    want = np_extrapolate(data, valid, -7.5, steps)
    for a in range(3):
        np.testing.assert_array_equal(got[a], want[a])
        v = valid[a].astype(bool)
        np.testing.assert_array_equal(got[a][v], data[a][v])  # valid faces keep their values


def test_extrapolate_single_source_closed_form(oracle):
    """One valid face of value 3: after s sweeps every face within L1
    distance s holds 3 (a mean of 3s), every other face the clear value."""
    geo = scenes.Geometry((9, 9, 9), 1.0)
    data = [np.zeros(geo.face_shape(a)) for a in range(3)]
    valid = [np.zeros(geo.face_shape(a), np.uint8) for a in range(3)]
    src = (4, 4, 4)
    for a in range(3):
        data[a][src] = 3.0
        valid[a][src] = 1
    for steps in (1, 2, 5):
        out = oracle.extrapolate_faces(geo, data, valid, clear=-1.0, steps=steps)
        for a in range(3):
            I, J, K = np.indices(geo.face_shape(a))
            dist = np.abs(I - 4) + np.abs(J - 4) + np.abs(K - 4)
            np.testing.assert_array_equal(out[a], np.where(dist <= steps, 3.0, -1.0))


def test_enforce_matches_numpy_and_keeps_open_faces(oracle):
    geo = scenes.Geometry((8, 11, 7), 1.0 / 11, center=(0.1, -0.3, 0.5))
    vel = random_faces(geo, 6)
    frac = fractions_with_walls(geo, 7)
    normal = random_faces(geo, 8)
    normal[0][0, 0, 0] = normal[1][0, 0, 0] = normal[2][0, 0, 0] = 0.0  # zero normal somewhere
    colvel = random_faces(geo, 9, 0.2)
    got = oracle.collider_enforce(geo, vel, frac, normal, colvel)
    want = np_enforce(geo, vel, frac, normal, colvel)
    for a in range(3):
        np.testing.assert_array_equal(got[a], want[a])
        keep = frac[a] != 1.0
        np.testing.assert_array_equal(got[a][keep], vel[a][keep])
        assert np.any(got[a][~keep] != vel[a][~keep])
    # a planted mistake (the normal left unnormalized) is caught
    bad = np_enforce(geo, vel, frac, normal, colvel, normalize=False)
    assert max(np.abs(bad[a] - got[a]).max() for a in range(3)) > 1e-3


def test_enforce_uniform_fields_closed_form(oracle):
    """Uniform fluid velocity u, collider velocity u0 and normal m: at solid
    faces the result is u - ((u - u0) . m/|m|) m/|m| (trilinear weights sum to
    1 up to rounding)."""
    geo = scenes.Geometry((6, 7, 5), 0.125)
    u, u0, m = np.array([0.3, -1.2, 0.7]), np.array([0.1, 0.05, -0.2]), np.array([1.0, 2.0, -2.0])
    vel = [np.full(geo.face_shape(a), u[a]) for a in range(3)]
    colvel = [np.full(geo.face_shape(a), u0[a]) for a in range(3)]
    normal = [np.full(geo.face_shape(a), m[a]) for a in range(3)]
    frac = scenes.wall_fractions(geo)
    out = oracle.collider_enforce(geo, vel, frac, normal, colvel)
    nh = m / np.linalg.norm(m)
    want = u - np.dot(u - u0, nh) * nh
    for a in range(3):
        solid = frac[a] == 1.0
        np.testing.assert_allclose(out[a][solid], want[a], rtol=0, atol=1e-14)
        np.testing.assert_array_equal(out[a][~solid], u[a])
