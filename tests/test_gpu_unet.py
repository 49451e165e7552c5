"""GPU tests of the learned preconditioner: the U-Net kernels (BF16 MFMA
implicit-GEMM convolutions, FP32 accumulation) against closed forms and the
FP64 oracle, training on the GPU, the operator features, and FCG with the
network inside the solver (convergence, fall-back, and the ablation against
the method without it).

Tolerances: activations and weights are rounded to BF16 (8 significant
bits, relative rounding error 2^-9) between layers, so outputs are compared
with the oracle by relative L2 error; hand-built weight sets whose values are
exact in BF16 are compared exactly."""
import os

import numpy as np
import pytest

import sfp
import unet_params as U
from sfp import scenes

pytestmark = pytest.mark.gpu


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("d", [(1, 0, 0), (0, -1, 0), (0, 0, 1), (-1, 1, -1)])
def test_shift_taps_exact(d):
    shape = (2, 8, 12, 20)
    print(f"[case] shift taps d={d} shape={shape}", flush=True)
    x = np.random.default_rng(1).normal(size=shape + (4,)).astype(np.float32)
    with sfp.UNet(*shape, training=False) as net:
        for ch in (0, 2):
            net.set_params(U.shift_params(*d, channel=ch))
            out = net.forward(x)
            want = np.maximum(U.shifted(bf16(x[..., ch]), *d), 0.0)
            assert np.array_equal(out, want), (ch, np.abs(out - want).max())


def test_coarse_path_block_means():
    shape = (1, 16, 8, 24)
    print(f"[case] pool/up path shape={shape}", flush=True)
    x = np.random.default_rng(2).uniform(0.0, 1.0, size=shape + (4,)).astype(np.float32)
    with sfp.UNet(*shape, training=False) as net:
        net.set_params(U.coarse_path_params())
        out = net.forward(x)
    want = U.block_mean_up(bf16(x[..., 0]).astype(np.float64), 4)
    # two BF16 roundings of means in [0, 1]: error <= 2 * 2^-9
    assert np.abs(out - want).max() <= 4e-3


@pytest.mark.parametrize("shape", [(1, 16, 16, 16), (2, 8, 12, 20), (1, 4, 4, 4)])
def test_forward_matches_oracle(oracle, shape):
    print(f"[case] forward vs oracle shape={shape}", flush=True)
    p = U.random_params(21).astype(np.float32)
    x = np.random.default_rng(22).normal(size=shape + (4,)).astype(np.float32)
    with sfp.UNet(*shape, training=False) as net:
        net.set_params(p)
        out = net.forward(x)
    # the oracle sees the BF16-rounded weights and inputs the GPU starts from
    want, _ = oracle.unet_forward(bf16(p).astype(np.float64), bf16(x).astype(np.float64))
    err = rel_l2(out, want)
    print(f"forward rel L2 {err:.3e}")
    assert err <= 2e-2


def test_inference_and_training_forward_agree():
    """Inference reads the half-resolution tensors through the convolution
    (up-sampling fused into the staging), training up-samples explicitly:
    nearest up-sampling is a copy, so both give the same bits."""
    shape = (2, 8, 12, 16)
    print(f"[case] fused vs explicit up-sampling shape={shape}", flush=True)
    p = U.random_params(41).astype(np.float32)
    x = np.random.default_rng(42).normal(size=shape + (4,)).astype(np.float32)
    outs = []
    for training in (False, True):
        with sfp.UNet(*shape, training=training) as net:
            net.set_params(p)
            outs.append(net.forward(x))
    assert np.array_equal(outs[0], outs[1])


def test_gradients_match_oracle(oracle):
    shape = (1, 8, 8, 12)
    print(f"[case] gradients vs oracle shape={shape}", flush=True)
    rng = np.random.default_rng(23)
    p = U.random_params(24).astype(np.float32)
    x = rng.normal(size=shape + (4,)).astype(np.float32)
    t = rng.normal(size=shape).astype(np.float32)
    m = (rng.uniform(size=shape) < 0.8).astype(np.float32)
    with sfp.UNet(*shape, training=True) as net:
        net.set_params(p)
        loss, g = net.gradients(x, t, m)
    loss_o, go = oracle.unet_gradients(bf16(p).astype(np.float64), bf16(x).astype(np.float64), t, m)
    assert loss == pytest.approx(loss_o, rel=2e-2)
    blocks = []
    for k in range(10):
        s = slice(U.OFF_W[k], U.OFF_W[k] + 27 * U.CIN[k] * U.COUT[k])
        if k == 0:  # only input channels 0..3 carry data
            gw = U.conv_weights(g, 0)[:, :, :4].ravel()
            gow = U.conv_weights(go, 0)[:, :, :4].ravel()
            assert np.all(U.conv_weights(g, 0)[:, :, 4:] == 0.0)
            blocks.append(("w0", gw, gow))
        else:
            blocks.append((f"w{k}", g[s], go[s]))
        blocks.append((f"b{k}", U.conv_bias(g, k), U.conv_bias(go, k)))
    blocks.append(("head", g[U.OFF_HEAD:], go[U.OFF_HEAD:]))
    stats = []
    for name, a, b in blocks:
        e = rel_l2(a, b)
        cos = float(np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300))
        print(f"{name:5s} rel L2 {e:.3e} cos {cos:.6f}")
        stats.append((name, e, cos))
    bad = [s for s in stats if not (s[1] <= 0.1 and s[2] >= 0.995)]
    assert not bad, bad


def test_training_reduces_loss_and_is_deterministic():
    import torch
    shape = (1, 16, 16, 16)
    print(f"[case] training shape={shape}", flush=True)
    rng = np.random.default_rng(25)
    x = rng.normal(size=shape + (4,)).astype(np.float32)
    t = U.block_mean_up(x[..., 0].astype(np.float64), 2).astype(np.float32)
    dev = torch.device("cuda:0")
    xd, td = torch.from_numpy(x).to(dev), torch.from_numpy(t).to(dev)
    md = torch.ones(shape, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    runs = []
    for _ in range(2):
        with sfp.UNet(*shape, training=True) as net:
            net.init_params(7)
            losses = [net.train_step_device(xd.data_ptr(), td.data_ptr(), md.data_ptr(), 2e-3)
                      for _ in range(80)]
            runs.append((losses, net.get_params()))
    losses = runs[0][0]
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.25 * losses[0]
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1], runs[1][1])


def test_unet_errors():
    with pytest.raises(sfp.SfpError) as e:
        sfp.UNet(1, 6, 8, 8)
    assert e.value.status == sfp.ERR_ARGUMENT
    with sfp.UNet(1, 8, 8, 8, training=False) as net:
        with pytest.raises(sfp.SfpError) as e:
            net.gradients(np.zeros((1, 8, 8, 8, 4), np.float32), np.zeros((1, 8, 8, 8), np.float32))
        assert e.value.status == sfp.ERR_STATE
        with pytest.raises(sfp.SfpError) as e:
            net.set_params(np.zeros(10, np.float32))
        assert e.value.status == sfp.ERR_ARGUMENT


# ------------------------------------------------------------------ inside the solver
def _coarse_phi(phi):
    nx, ny, nz = phi.shape
    out = np.zeros(((nx + 1) // 2, (ny + 1) // 2, (nz + 1) // 2))
    cnt = np.zeros_like(out)
    for a in range(2):
        for b in range(2):
            for c in range(2):
                ch = phi[a::2, b::2, c::2]
                out[:ch.shape[0], :ch.shape[1], :ch.shape[2]] += ch
                cnt[:ch.shape[0], :ch.shape[1], :ch.shape[2]] += 1
    return out / cnt


def test_features_match_coarse_operator():
    sc = scenes.pressure_scene(32, 3)
    print("[case] learned features 32^3 pressure (coarse level 16^3)", flush=True)
    with sfp.Context.for_scene(sc) as ctx:
        ctx.pressure_project(sc, want_pressure=False)
        f = ctx.learned_features(0, sc["phi"])
        op = ctx.export_level(0, 1)
        d32 = ctx.export_level32(0, 1)["diag"]
    assert f.shape == (16, 16, 16, 3)
    act = d32 > 0
    diag, cpl = op["diag"], op["cpl"]
    off = cpl[0] + cpl[1] + cpl[2]
    off[:-1] += cpl[0][1:]
    off[:, :-1] += cpl[1][:, 1:]
    off[:, :, :-1] += cpl[2][:, :, 1:]
    h = 2 * sc["geo"].spacing
    want = np.zeros(f.shape)
    safe = np.where(act, diag, 6.0)
    want[..., 0] = np.where(act, np.log(safe / 6.0), 0.0)
    want[..., 1] = np.where(act, (diag - off) / safe, 0.0)
    want[..., 2] = np.where(act, np.clip(_coarse_phi(sc["phi"]) / h, -4, 4) / 4, 0.0)
    np.testing.assert_allclose(f, want, rtol=1e-6, atol=1e-6)
    assert act.sum() > 100 and np.all(f[~act] == 0)


def _apply(op, x):
    from sfp.learn import apply_full
    return apply_full(op["diag"], op["cpl"], x)


def test_coarse_level_hooks():
    """restrict_smoothed gives the restricted residual a V-cycle leaves and
    solve_rhs(level=1) solves the coarse operator to the tolerance."""
    sc = scenes.magnetic_scene(32, 4, 2.0)
    print("[case] coarse-level hooks 32^3 magnetic", flush=True)
    with sfp.Context.for_scene(sc, tol=1e-10) as ctx:
        ctx.magnetic_solve(sc, want_field=False)
        op0 = ctx.export_level(1, 0)
        op1 = ctx.export_level(1, 1)
        act1 = ctx.export_level32(1, 1)["diag"] > 0
        r = np.random.default_rng(5).normal(size=op0["diag"].shape)
        r *= ctx.export_level32(1, 0)["diag"] > 0
        rc = ctx.restrict_smoothed(1, r).astype(np.float64)
        ec, st = ctx.solve_rhs(1, rc, level=1)
        # the system's own solve is untouched by the hook
        x_before, _, _ = ctx.residual(1, np.zeros(op0["diag"].shape))
    assert st["converged"] and st["unknowns"] == act1.size
    res = (rc - _apply(op1, ec)) * act1
    assert np.linalg.norm(res) <= 1e-9 * np.linalg.norm(rc * act1)
    assert np.linalg.norm(rc) > 0
    np.testing.assert_array_equal(x_before, op0["rhs"])


def _solve(sc, precond, params=None, kind="pressure", **kw):
    with sfp.Context.for_scene(sc) as ctx:
        if params is not None:
            ctx.set_network(params)
        ctx.set_params(precond=precond, **kw)
        if kind == "pressure":
            vel, p, st = ctx.pressure_project(sc)
            return vel, p, st
        pot, _, st = ctx.magnetic_solve(sc, want_field=False)
        return None, pot, st


def test_learned_requires_network():
    sc = scenes.pressure_scene(16, 0)
    with sfp.Context.for_scene(sc) as ctx:
        with pytest.raises(sfp.SfpError) as e:
            ctx.set_params(precond=sfp.PRECOND_LEARNED)
        assert e.value.status == sfp.ERR_UNSUPPORTED
        with pytest.raises(sfp.SfpError) as e:
            ctx.set_network(np.zeros(5, np.float32))
        assert e.value.status == sfp.ERR_ARGUMENT


def test_learned_random_network_converges_and_falls_back():
    """With an untrained network the learned cycle still converges (its
    correction is energy-optimally scaled, so it cannot undo the V-cycle) to
    the MG-PCG answer; with a stall ratio no iteration can meet, the solve
    falls back to the plain V-cycle after the first iteration and converges."""
    sc = scenes.pressure_scene(64, 0)
    print("[case] learned (random weights) 64^3 pressure", flush=True)
    p = U.random_params(31).astype(np.float32)
    vel_l, pres_l, st_l = _solve(sc, sfp.PRECOND_LEARNED, p, stall_ratio=10.0)
    vel_m, pres_m, st_m = _solve(sc, sfp.PRECOND_MG)
    vel_f, pres_f, st_f = _solve(sc, sfp.PRECOND_LEARNED, p, stall_ratio=1e-6)
    print(st_l, st_m, st_f)
    assert st_l["converged"] and st_l["rel_residual"] <= 1e-8
    assert st_l["fallback"] == 0 and st_l["precond_used"] == sfp.PRECOND_LEARNED
    assert st_f["converged"] and st_f["rel_residual"] <= 1e-8
    assert st_f["fallback"] == st_f["iterations"] - 1 and st_f["precond_used"] == sfp.PRECOND_MG
    for pres, vel in ((pres_l, vel_l), (pres_f, vel_f)):
        assert rel_l2(pres, pres_m) <= 1e-6
        for a in range(3):
            assert np.abs(vel[a] - vel_m[a]).max() <= 1e-6 * np.abs(vel_m[a]).max()


def test_learned_grid_not_multiple_of_4_uses_multigrid():
    sc = scenes.pressure_scene(30, 0)
    p = U.random_params(32).astype(np.float32)
    _, _, st = _solve(sc, sfp.PRECOND_LEARNED, p)
    assert st["converged"] and st["precond_used"] == sfp.PRECOND_MG and st["fallback"] == 0


def _trained():
    from sfp import learn
    if not os.path.exists(learn.WEIGHTS):
        pytest.skip("no trained weights in the package")
    params, meta = learn.load_weights()
    return params, learn.smoother(meta)


@pytest.mark.parametrize("kind", ["pressure", "magnetic"])
def test_ablation_trained_network(kind):
    """What the trained network gains over the same cycle without it (a zero
    correction: the V-cycle plus the same post-smoothing, both with the
    smoother settings the weights were trained under) and over plain
    MG-PCG (the library's default smoother), at a fixed budget of 5
    iterations, on 128^3 scenes it was not
    trained on; and the learned solve with fall-back meets the tolerance and
    agrees with plain MG-PCG. Measured with the committed weights: the
    pressure residual after 5 iterations is several times lower than both
    baselines; on the magnetic system the network gains nothing (its error is
    dominated by global modes the coarse levels already handle), so there the
    test only bounds the loss."""
    params, sm = _trained()
    sc = scenes.pressure_scene(128, 11) if kind == "pressure" else scenes.magnetic_scene(128, 12, 2.0)
    print(f"[case] ablation {kind} 128^3", flush=True)
    budget = 5
    kw = dict(max_iters=budget, tol=1e-14)
    _, _, st_l = _solve(sc, sfp.PRECOND_LEARNED, params, kind, stall_ratio=10.0, **kw, **sm)
    _, _, st_s = _solve(sc, sfp.PRECOND_SMOOTH, None, kind, **kw, **sm)
    _, _, st_m = _solve(sc, sfp.PRECOND_MG, None, kind, **kw)
    _, _, st_j = _solve(sc, sfp.PRECOND_JACOBI, None, kind, **kw)
    print(f"after {budget} its: learned {st_l['rel_residual']:.3e}  without network "
          f"{st_s['rel_residual']:.3e}  MG {st_m['rel_residual']:.3e}  "
          f"Jacobi {st_j['rel_residual']:.3e}")
    assert st_l["fallback"] == 0 and st_l["precond_used"] == sfp.PRECOND_LEARNED
    assert st_l["rel_residual"] < 1e-3 * st_j["rel_residual"]
    if kind == "pressure":
        assert st_l["rel_residual"] < 0.5 * st_s["rel_residual"]
        assert st_l["rel_residual"] < 0.5 * st_m["rel_residual"]
    else:
        assert st_l["rel_residual"] < 3.0 * st_s["rel_residual"]
    vel_l, pres_l, st = _solve(sc, sfp.PRECOND_LEARNED, params, kind, **sm)
    _, pres_m, _ = _solve(sc, sfp.PRECOND_MG, None, kind)
    print("learned with fall-back:", st)
    assert st["converged"] and st["rel_residual"] <= 1e-8
    assert rel_l2(pres_l, pres_m) <= 1e-6
