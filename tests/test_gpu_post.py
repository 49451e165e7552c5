"""GPU tests of the steps after the projection through the C ABI: the velocity
extrapolation of Simulation::ProjectVelocity, the general face-grid
Extrapolation::Solve and Collider::Enforce, each bit-exact against the oracle
(same FP64 operations in the same order), host and device entry points."""
import numpy as np
import pytest

import sfp
from sfp import scenes
from test_post_oracle import fractions_with_walls, random_faces

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [(16, 16, 16), (17, 20, 23), (64, 64, 64)])
def test_extrapolate_velocity_bit_exact(oracle, n):
    print(f"[case] extrapolate velocity {n}", flush=True)
    sc = scenes.random_pressure_scene(n, 5)
    geo = sc["geo"]
    want = oracle.extrapolate_faces(geo, sc["velocity"], oracle.velocity_valid(sc), 0.0, 6)
    with sfp.Context.for_scene(sc) as ctx:
        got = ctx.extrapolate_velocity(sc["velocity"], sc["phi"], sc["fraction"], 6)
    for a in range(3):
        np.testing.assert_array_equal(got[a], want[a])


@pytest.mark.parametrize("steps", [0, 1, 6, 11])
def test_extrapolate_faces_bit_exact(oracle, steps):
    geo = scenes.Geometry((13, 9, 18), 0.05)
    print(f"[case] extrapolate faces steps={steps}", flush=True)
    rng = np.random.default_rng(steps)
    data = random_faces(geo, 11)
    valid = [(rng.uniform(size=geo.face_shape(a)) < 0.07).astype(np.uint8) for a in range(3)]
    want = oracle.extrapolate_faces(geo, data, valid, -2.0, steps)
    with sfp.Context(geo.n, geo.spacing, geo.origin) as ctx:
        got = ctx.extrapolate_faces(data, valid, -2.0, steps)
    for a in range(3):
        np.testing.assert_array_equal(got[a], want[a])


@pytest.mark.parametrize("n", [(8, 11, 7), (32, 32, 32)])
def test_collider_enforce_bit_exact(oracle, n):
    geo = scenes.Geometry(n, 1.0 / max(n), center=(0.1, -0.3, 0.5))
    print(f"[case] collider enforce {n}", flush=True)
    vel = random_faces(geo, 6)
    frac = fractions_with_walls(geo, 7)
    normal = random_faces(geo, 8)
    normal[1][2, 3, 1] = normal[0][2, 3, 1] = normal[2][2, 3, 1] = 0.0
    colvel = random_faces(geo, 9, 0.2)
    want = oracle.collider_enforce(geo, vel, frac, normal, colvel)
    with sfp.Context(geo.n, geo.spacing, geo.origin) as ctx:
        got = ctx.collider_enforce(vel, frac, normal, colvel)
    for a in range(3):
        np.testing.assert_array_equal(got[a], want[a])


def test_device_entry_points_match_host(oracle):
    import torch
    sc = scenes.random_pressure_scene((24, 20, 28), 8)
    geo = sc["geo"]
    print("[case] post-projection device entry points", flush=True)
    normal = random_faces(geo, 12)
    colvel = random_faces(geo, 13, 0.1)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)  # noqa: E731
    vel_d = [t(a) for a in sc["velocity"]]
    phi_d = t(sc["phi"])
    frac_d = [t(a) for a in sc["fraction"]]
    nrm_d = [t(a) for a in normal]
    col_d = [t(a) for a in colvel]
    torch.cuda.synchronize()
    with sfp.Context.for_scene(sc) as ctx:
        p = lambda ts: [x.data_ptr() for x in ts]  # noqa: E731
        ctx.extrapolate_velocity_device(p(vel_d), phi_d.data_ptr(), p(frac_d), 6)
        ctx.collider_enforce_device(p(vel_d), p(frac_d), p(nrm_d), p(col_d))
        got = [v.cpu().numpy() for v in vel_d]
        h = ctx.extrapolate_velocity(sc["velocity"], sc["phi"], sc["fraction"], 6)
        h = ctx.collider_enforce(h, sc["fraction"], normal, colvel)
    want = oracle.extrapolate_faces(geo, sc["velocity"], oracle.velocity_valid(sc), 0.0, 6)
    want = oracle.collider_enforce(geo, want, sc["fraction"], normal, colvel)
    for a in range(3):
        np.testing.assert_array_equal(got[a], want[a])
        np.testing.assert_array_equal(h[a], want[a])


def test_post_errors():
    geo = scenes.Geometry((8, 8, 8), 0.125)
    with sfp.Context(geo.n, geo.spacing, geo.origin) as ctx:
        z = [np.zeros(geo.face_shape(a)) for a in range(3)]
        with pytest.raises(sfp.SfpError) as e:
            ctx.extrapolate_velocity(z, np.zeros(geo.n), z, -1)
        assert e.value.status == sfp.ERR_ARGUMENT
        st = sfp.lib().sfp_collider_enforce(ctx.handle, None, None, None, None)
        assert st == sfp.ERR_ARGUMENT


def test_checkpoint_round_trip_into_the_solver(oracle, tmp_path):
    """A scene written as a reference checkpoint and read back drives the GPU
    projection to the same velocities as the original arrays, and the
    solution meets the tolerance on the oracle's system."""
    from sfp import checkpoint
    sc = scenes.pressure_scene(48, 3)
    n = sc["geo"].n
    path = str(tmp_path / "5.sav")
    print("[case] checkpoint round trip 48^3 pressure", flush=True)
    checkpoint.write(path, n, sc["velocity"], sc["phi"], init_volume=0.2, vol_error=0.0)
    ck = checkpoint.read(path, n)
    sc2 = dict(sc, velocity=ck["velocity"], phi=ck["phi"])
    with sfp.Context.for_scene(sc) as ctx:
        v1, p1, st1 = ctx.pressure_project(sc)
        v2, p2, st2 = ctx.pressure_project(sc2)
    for a in range(3):
        np.testing.assert_array_equal(v1[a], v2[a])
    ps = oracle.pressure_system(sc2)
    assert oracle.rel_residual(ps["A"], ps["rhs"], p2.ravel()[ps["mat2grid"]]) <= 1e-8
