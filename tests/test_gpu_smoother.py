"""GPU tests of the one-pass kernels of the V-cycle through the C ABI: the
pair of post-smoothing steps, the last pre-smoothing step fused with the
restriction, and the small levels run in one launch. Each is bit-identical to
the separate launches on pressure and magnetic levels (ragged and odd extents,
inactive cells, one and several x planes per block, grids narrower than a
tile); the first two are also within FP32 rounding of an FP64 numpy
statement on the same FP32 operator."""
import numpy as np
import pytest

import sfp
from sfp import scenes

pytestmark = pytest.mark.gpu

COEF = (0.31, 0.72, 0.27, 0.66)  # cdA, crA, cdB, crB


def np_apply(diag, cpl, x):
    """A x for the full-grid operator: diagonal plus A[c, c - e_a] = -cpl_a[c]
    (the coupling stored on the lower face of c; unused on the first layer)."""
    y = diag * x
    for a in range(3):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a] = slice(0, -1)
        hi[a] = slice(1, None)
        lo, hi = tuple(lo), tuple(hi)
        c = cpl[a][hi]
        y[hi] -= c * x[lo]
        y[lo] -= c * x[hi]
    return y


def np_two_steps(diag, cpl, b, x1, d1, cd_a, cr_a, cd_b, cr_b):
    inv = np.where(diag > 0, 1.0 / np.where(diag > 0, diag, 1.0), 0.0)
    d2 = cd_a * d1 + cr_a * inv * (b - np_apply(diag, cpl, x1))
    x2 = x1 + d2
    return x2 + cd_b * d2 + cr_b * inv * (b - np_apply(diag, cpl, x2))


def np_restrict(r):
    """Sum over each coarse cell's (up to 8) children."""
    pad = [(0, n % 2) for n in r.shape]
    rp = np.pad(r, pad)
    nx, ny, nz = (n // 2 for n in rp.shape)
    return rp.reshape(nx, 2, ny, 2, nz, 2).sum(axis=(1, 3, 5))


def _system(kind, n, seed):
    if kind == "pressure":
        sc = scenes.random_pressure_scene(n, seed=seed)
        ctx = sfp.Context.for_scene(sc)
        ctx.pressure_project(sc)
        return ctx, 0
    sc = scenes.magnetic_scene(n, seed=seed)
    ctx = sfp.Context.for_scene(sc)
    ctx.magnetic_solve(sc)
    return ctx, 1


@pytest.mark.parametrize("kind,n,level,pre", [
    ("pressure", (72, 66, 130), 0, False),   # one plane per block, 3 x 5 tiles, air cells
    ("pressure", (128, 128, 128), 0, False),  # several planes per block
    ("magnetic", 64, 0, False),               # every cell active
    ("pressure", (96, 80, 112), 1, False),    # coarse level narrower than a tile in z
    # z rows of 193..256 cells: the row-tiled kernel (whole rows per wavefront)
    ("pressure", (40, 50, 256), 0, False),    # one plane per block, air rows skipped
    ("pressure", (40, 50, 256), 0, True),     # the first pre-smoothing passes (d1 = x1)
    ("pressure", (300, 61, 200), 0, False),   # runs of several planes, lanes past the row's end
    ("magnetic", (70, 33, 256), 0, False),    # row-sum diagonal rebuilt, every row active
    ("magnetic", (70, 33, 256), 0, True),
])
def test_pair_matches_two_steps(kind, n, level, pre):
    print(f"[case] smooth pair {kind} {n} level {level} pre {pre}", flush=True)
    ctx, which = _system(kind, n, 3)
    with ctx:
        op = ctx.export_level32(which, level)
        diag = op["diag"]
        act = diag > 0
        assert act.any()
        if kind == "pressure" and level == 0:
            assert not act.all()
        rng = np.random.default_rng(11)
        vec = lambda s: (rng.uniform(-s, s, diag.shape) * act).astype(np.float32)  # noqa: E731
        b, x1, d1 = vec(1.0), vec(1.0), vec(0.1)
        if pre:
            d1 = x1  # the stored first iterate is also the previous direction
        x3p, x3s = ctx.smooth_pair(which, level, b, x1, None if pre else d1, *COEF)
        # repeated launches: the same bits each time (no race between the
        # block's wavefronts over the shared planes)
        for _ in range(4):
            again, _ = ctx.smooth_pair(which, level, b, x1, None if pre else d1, *COEF)
            np.testing.assert_array_equal(again, x3p)
    np.testing.assert_array_equal(x3p, x3s)
    assert np.all(x3p[~act] == 0)
    f64 = lambda a: a.astype(np.float64)  # noqa: E731
    ref = np_two_steps(f64(diag), [f64(c) for c in op["cpl"]], f64(b), f64(x1), f64(d1), *COEF)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(x3p, ref, rtol=0, atol=2e-5 * scale)
    # a planted mistake in the statement (step B's direction term dropped) is caught
    bad = np_two_steps(f64(diag), [f64(c) for c in op["cpl"]], f64(b), f64(x1), f64(d1),
                       COEF[0], COEF[1], 0.0, COEF[3])
    assert np.abs(bad - x3p).max() > 100 * 2e-5 * scale


@pytest.mark.parametrize("kind,n,level", [
    ("pressure", (72, 66, 130), 0),   # two planes per block, air cells
    ("pressure", (129, 67, 65), 0),   # odd extents: single-child coarse cells
    ("magnetic", 64, 0),
    ("pressure", (96, 80, 112), 1),
])
def test_smooth_restrict_matches_separate_launches(kind, n, level):
    print(f"[case] smooth+restrict {kind} {n} level {level}", flush=True)
    ctx, which = _system(kind, n, 4)
    with ctx:
        op = ctx.export_level32(which, level)
        diag = op["diag"]
        act = diag > 0
        rng = np.random.default_rng(12)
        vec = lambda s: (rng.uniform(-s, s, diag.shape) * act).astype(np.float32)  # noqa: E731
        b, x1, d1 = vec(1.0), vec(1.0), vec(0.1)
        x2f, bcf, x2s, bcs = ctx.smooth_restrict(which, level, b, x1, d1, COEF[0], COEF[1])
        for _ in range(4):
            x2a, bca, _, _ = ctx.smooth_restrict(which, level, b, x1, d1, COEF[0], COEF[1])
            np.testing.assert_array_equal(x2a, x2f)
            np.testing.assert_array_equal(bca, bcf)
    np.testing.assert_array_equal(x2f, x2s)
    np.testing.assert_array_equal(bcf, bcs)
    f64 = lambda a: a.astype(np.float64)  # noqa: E731
    dg, cpl = f64(diag), [f64(c) for c in op["cpl"]]
    inv = np.where(dg > 0, 1.0 / np.where(dg > 0, dg, 1.0), 0.0)
    x2 = f64(x1) + COEF[0] * f64(d1) + COEF[1] * inv * (f64(b) - np_apply(dg, cpl, f64(x1)))
    np.testing.assert_allclose(x2f, x2, rtol=0, atol=2e-5 * np.abs(x2).max())
    res = np.where(act, f64(b) - np_apply(dg, cpl, f64(x2f)), 0.0)
    want = np_restrict(res)
    assert bcf.shape == want.shape
    np.testing.assert_allclose(bcf, want, rtol=0, atol=1e-4 * np.abs(want).max())
    # a planted mistake (restricting with one child per coarse cell) is caught
    assert np.abs(res[::2, ::2, ::2] - bcf).max() > 1e-2 * np.abs(want).max()


@pytest.mark.parametrize("kind", ["pressure", "magnetic"])
def test_repeated_solves_are_bitwise_identical(kind):
    """Solves of one system on a grid that takes the one-pass kernels repeat
    bit for bit (iteration count and solution)."""
    print(f"[case] repeated {kind} solves at 160^3", flush=True)
    sc = scenes.pressure_scene(160, 2) if kind == "pressure" else scenes.magnetic_scene(160, seed=2)
    with sfp.Context.for_scene(sc) as ctx:
        runs = []
        for _ in range(3):
            ctx.reset_warm_start()
            runs.append(ctx.pressure_project(sc) if kind == "pressure" else ctx.magnetic_solve(sc))
    for r in runs[1:]:
        assert r[-1]["iterations"] == runs[0][-1]["iterations"]
        np.testing.assert_array_equal(r[1] if kind == "pressure" else r[0],
                                      runs[0][1] if kind == "pressure" else runs[0][0])


def test_changed_active_set_matches_fresh_context():
    """Projections on one context whose liquid region shrinks and grows
    between calls (cells that were unknowns become air and back) give the
    bits of the same projections on fresh contexts, on a grid that takes the
    one-pass kernels."""
    base = scenes.pressure_scene(144, 4)
    wet = dict(base, phi=np.ascontiguousarray(base["phi"] - 0.04))
    dry = dict(base, phi=np.ascontiguousarray(base["phi"] + 0.04))
    print("[case] 144^3 projections with changing liquid regions", flush=True)
    with sfp.Context.for_scene(base) as ctx:
        seq = [ctx.pressure_project(sc) for sc in (wet, dry, wet, dry)]
    for sc, got in zip((wet, dry), seq[:2]):
        with sfp.Context.for_scene(base) as fresh:
            want = fresh.pressure_project(sc)
        assert got[2]["iterations"] == want[2]["iterations"] <= 20
        np.testing.assert_array_equal(got[1], want[1])
        for a in range(3):
            np.testing.assert_array_equal(got[0][a], want[0][a])
    for later, first in zip(seq[2:], seq[:2]):
        assert later[2]["iterations"] == first[2]["iterations"]
        np.testing.assert_array_equal(later[1], first[1])


@pytest.mark.parametrize("n", [(144, 144, 144), (131, 129, 133)])
def test_pair_path_solves_to_tolerance(oracle, n):
    """Grids where the V-cycle takes the one-pass kernels (nz >= 64, above
    128^3 cells; even and odd extents) converge to the tolerance on the
    oracle's system."""
    sc = scenes.random_pressure_scene(n, 1) if n[0] % 2 else scenes.pressure_scene(n[0], 1)
    print(f"[case] pressure {n} through the one-pass kernels", flush=True)
    with sfp.Context.for_scene(sc) as ctx:
        v, p, st = ctx.pressure_project(sc)
    assert st["converged"] and st["iterations"] <= 20
    ps = oracle.pressure_system(sc)
    assert oracle.rel_residual(ps["A"], ps["rhs"], p.ravel()[ps["mat2grid"]]) <= 1e-8


@pytest.mark.parametrize("kind,n", [("pressure", (14, 14, 14)), ("pressure", (40, 37, 45)),
                                    ("magnetic", 33), ("pressure", (96, 80, 112)),
                                    ("magnetic", 64), ("pressure", (64, 48, 32))])
@pytest.mark.parametrize("degree,coarse", [(1, 1), (2, 5), (3, 4), (4, 3), (3, 1)])
def test_vcycle_tail_matches_level_by_level(kind, n, degree, coarse):
    """The V-cycle with its small levels in one launch gives the same
    preconditioned vector, bit for bit, as level-by-level launches, and the
    whole solve the same solution and iteration count (fine and coarse
    levels with their own smoother degrees). The 64^3 and 64 x 48 x 32 grids
    put z extents divisible by four into the tail (its 16-byte sweeps), the
    others odd or ragged ones (one cell per thread)."""
    print(f"[case] vcycle tail {kind} {n} degree {degree}/{coarse}", flush=True)
    sc = scenes.random_pressure_scene(n, seed=5) if kind == "pressure" else \
        scenes.magnetic_scene(n, seed=5)
    which = 0 if kind == "pressure" else 1
    with sfp.Context.for_scene(sc, cheby_degree=degree, coarse_cheby_degree=coarse) as ctx:
        solve = ctx.pressure_project if which == 0 else ctx.magnetic_solve
        out = {}
        for tail in (False, True):
            ctx.set_vcycle_tail(tail)
            ctx.reset_warm_start()
            res = solve(sc)
            rng = np.random.default_rng(3)
            diag = ctx.export_level32(which, 0)["diag"]
            r = rng.standard_normal(diag.shape) * (diag > 0)
            out[tail] = (res, ctx.precondition(which, r))
    (a, za), (b, zb) = out[False], out[True]
    np.testing.assert_array_equal(za, zb)
    assert a[-1]["iterations"] == b[-1]["iterations"]
    for u, v in zip(a[:-1], b[:-1]):
        if isinstance(u, list):
            for p, q in zip(u, v):
                np.testing.assert_array_equal(p, q)
        else:
            np.testing.assert_array_equal(u, v)



@pytest.mark.parametrize("kind", ["pressure", "magnetic"])
def test_fused_pairs_match_separate_launches(kind):
    """Solves with the step pairs as one pass and as separate launches (flat
    single steps and the restriction), on the large and on the small levels,
    give the same solution and iteration count, bit for bit."""
    print(f"[case] fused pairs on/off {kind} 160^3", flush=True)
    sc = scenes.pressure_scene(160, 3) if kind == "pressure" else scenes.magnetic_scene(160, seed=3)
    with sfp.Context.for_scene(sc) as ctx:
        out = {}
        for mask in (3, 0, 1, 2, 14, 18, 31):
            ctx.set_fused_pairs(mask)
            ctx.reset_warm_start()
            out[mask] = ctx.pressure_project(sc) if kind == "pressure" else ctx.magnetic_solve(sc)
    a = out[3]
    for m in (0, 1, 2, 14, 18, 31):
        b = out[m]
        assert a[-1]["iterations"] == b[-1]["iterations"]
        for u, v in zip(a[:-1], b[:-1]):
            if isinstance(u, list):
                for p, q in zip(u, v):
                    np.testing.assert_array_equal(p, q)
            elif u is not None:
                np.testing.assert_array_equal(u, v)
