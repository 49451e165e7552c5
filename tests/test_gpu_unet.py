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
def test_features_match_operator():
    sc = scenes.pressure_scene(32, 3)
    print("[case] learned features 32^3 pressure", flush=True)
    with sfp.Context.for_scene(sc) as ctx:
        ctx.pressure_project(sc, want_pressure=False)
        f = ctx.learned_features(0, sc["phi"])
        op = ctx.export_level(0, 0)
        d32 = ctx.export_level32(0, 0)["diag"]
    act = d32 > 0
    diag, cpl = op["diag"], op["cpl"]
    off = cpl[0] + cpl[1] + cpl[2]
    off[:-1] += cpl[0][1:]
    off[:, :-1] += cpl[1][:, 1:]
    off[:, :, :-1] += cpl[2][:, :, 1:]
    dx = sc["geo"].spacing
    want = np.zeros(f.shape)
    want[..., 0] = np.where(act, np.log(np.where(act, diag, 6.0) / 6.0), 0.0)
    want[..., 1] = np.where(act, (diag - off) / np.where(act, diag, 1.0), 0.0)
    want[..., 2] = np.where(act, np.clip(sc["phi"] / dx, -4, 4) / 4, 0.0)
    np.testing.assert_allclose(f, want, rtol=1e-6, atol=1e-6)
    assert act.sum() > 1000 and np.all(f[~act] == 0)


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


def test_learned_random_network_falls_back_and_converges():
    """An untrained network is a poor preconditioner: FCG falls back to the
    multigrid V-cycle and still meets the tolerance with the same answer."""
    sc = scenes.pressure_scene(64, 0)
    print("[case] learned (random weights) 64^3 pressure", flush=True)
    p = U.random_params(31).astype(np.float32)
    vel_l, pres_l, st_l = _solve(sc, sfp.PRECOND_LEARNED, p)
    vel_m, pres_m, st_m = _solve(sc, sfp.PRECOND_MG)
    print(st_l, st_m)
    assert st_l["converged"] and st_l["rel_residual"] <= 1e-8
    assert st_l["fallback"] > 0 and st_l["precond_used"] == sfp.PRECOND_MG
    assert rel_l2(pres_l, pres_m) <= 1e-6
    for a in range(3):
        assert np.abs(vel_l[a] - vel_m[a]).max() <= 1e-6 * np.abs(vel_m[a]).max()


def test_learned_grid_not_multiple_of_4_uses_multigrid():
    sc = scenes.pressure_scene(30, 0)
    p = U.random_params(32).astype(np.float32)
    _, _, st = _solve(sc, sfp.PRECOND_LEARNED, p)
    assert st["converged"] and st["precond_used"] == sfp.PRECOND_MG and st["fallback"] == 0


def _trained():
    from sfp import learn
    if not os.path.exists(learn.WEIGHTS):
        pytest.skip("no trained weights in the package")
    return learn.load_weights()[0]


@pytest.mark.parametrize("kind", ["pressure", "magnetic"])
def test_ablation_trained_network(kind):
    """What the trained network gains over FCG without it (Jacobi scaling
    only), at a fixed iteration budget, on a scene it was not trained on;
    and the learned solve with fall-back meets the tolerance."""
    params = _trained()
    sc = scenes.pressure_scene(64, 11) if kind == "pressure" else scenes.magnetic_scene(64, 12, 2.0)
    print(f"[case] ablation {kind} 64^3", flush=True)
    budget = 12
    _, _, st_l = _solve(sc, sfp.PRECOND_LEARNED, params, kind, max_iters=budget, tol=1e-14,
                        stall_ratio=10.0)
    _, _, st_j = _solve(sc, sfp.PRECOND_JACOBI, None, kind, max_iters=budget, tol=1e-14)
    print(f"after {budget} its: learned {st_l['rel_residual']:.3e}  jacobi {st_j['rel_residual']:.3e}")
    assert st_l["fallback"] == 0 and st_l["precond_used"] == sfp.PRECOND_LEARNED
    assert st_l["rel_residual"] < 0.5 * st_j["rel_residual"]
    _, _, st = _solve(sc, sfp.PRECOND_LEARNED, params, kind)
    print("learned with fall-back:", st)
    assert st["converged"] and st["rel_residual"] <= 1e-8
