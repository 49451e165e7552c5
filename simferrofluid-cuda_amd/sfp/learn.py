"""Training of the U-Net preconditioner (include/sfp_learned.h).

The learned preconditioner runs the multigrid V-cycle and then corrects the
residual it leaves: restricted to the first coarse level, that residual r_c
(scaled to unit RMS over the active coarse cells) and three features of the
coarse operator go through the network, whose output approximates the exact
coarse correction e_c = A_c^-1 r_c; it is prolonged and post-smoothed on the
fine grid. The cycle rescales the network output to be optimal in A_c's
energy norm, so the network is trained on directions: the target of a sample
is e_c scaled to unit RMS, with a masked MSE over the active cells.

Samples come from systems the library assembles itself (pressure projections
of drops over pools and magnetostatic solves of wavy interfaces). Fine
residuals are

* the system's right-hand side;
* residuals of unpreconditioned CG after a few iterations;
* white and box-smoothed noise on the active cells;

each goes through one V-cycle of the library's own kernels and the residual it
leaves is restricted (sfp_test_restrict_smoothed); e_c is the library's
multigrid-preconditioned solve on the coarse level (sfp_test_solve). Training runs on
the GPU through sfp_unet_train_step_device (the library's forward, backward
and Adam kernels); PyTorch only holds the device arrays, stacks batches and
flips them for augmentation.

Usage (from the package directory):
    python -m sfp.learn [--n 128 --steps 6000 --batch 8 --out weights/unet_precond.f32]
    python -m sfp.learn --calibrate 256 --calibration-out calib.json
The calibration measures the fall-back threshold (stall_ratio) the weights'
metadata records; see calibrate().
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

from . import Context, UNet, PRECOND_LEARNED, PRECOND_MG, unet_param_count
from . import scenes

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "weights",
                       "unet_precond.f32")
# The V-cycle smoother settings a network is trained under: its samples are the
# residuals that cycle leaves, so the learned preconditioner runs with them.
SMOOTHER_KEYS = ("cheby_degree", "cheby_ratio", "coarse_cheby_degree", "coarse_cheby_ratio")


def smoother(meta):
    """The smoother settings recorded with trained weights, as the set_params
    keywords of the learned cycle's smoother (net_cheby_degree, ...): the
    learned cycle runs with them, the plain V-cycle after a fall-back with the
    context's own."""
    return {"net_" + k: meta["smoother"][k] for k in SMOOTHER_KEYS}


def learned_params(meta):
    """set_params keywords of the learned preconditioner with these weights:
    the trained-under smoother and the fall-back threshold recorded with them
    (stall_ratio, from the plain V-cycle's measured per-iteration residual
    reduction; 0.5 for weights recorded without one)."""
    return dict(precond=PRECOND_LEARNED, stall_ratio=float(meta.get("stall_ratio", 0.5)),
                **smoother(meta))


# ------------------------------------------------------------------ scenes
def drops_scene(n: int, seed: int):
    """A pool of random depth with 1-3 random spheres above it in a closed box."""
    geo = scenes.Geometry(n, 1.0 / n)
    X, Y, Z = geo.cell_centers()
    u = scenes.uniform(seed, scenes.ID_MISC, 16)
    phi = Y - (0.12 + 0.25 * u[0]) + 0 * X + 0 * Z
    for k in range(1 + int(u[1] * 3)):
        c = (0.25 + 0.5 * u[2 + 4 * k], 0.45 + 0.3 * u[3 + 4 * k], 0.25 + 0.5 * u[4 + 4 * k])
        r = 0.06 + 0.12 * u[5 + 4 * k]
        phi = np.minimum(phi, np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - r)
    sc = scenes.pressure_scene(n, seed)
    sc["phi"] = np.ascontiguousarray(phi)
    return sc


# ------------------------------------------------------------------ operator helpers
def apply_full(diag, cpl, x):
    """A x for the full-grid 7-point operator (A[c, c-e_a] = -cpl_a[c])."""
    w = diag * x
    for a in range(3):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a] = slice(0, -1)
        hi[a] = slice(1, None)
        lo, hi = tuple(lo), tuple(hi)
        w[hi] -= cpl[a][hi] * x[lo]
        w[lo] -= cpl[a][hi] * x[hi]
    return w


def cg_residuals(diag, cpl, b, active, keep):
    """Residuals of unpreconditioned CG from x = 0 at the iterations in keep."""
    x = np.zeros_like(b)
    r = b * active
    p = r.copy()
    rr = float(np.vdot(r, r))
    out = {}
    for it in range(1, max(keep) + 1):
        q = apply_full(diag, cpl, p) * active
        alpha = rr / float(np.vdot(p, q))
        x += alpha * p
        r -= alpha * q
        rr_new = float(np.vdot(r, r))
        p = r + (rr_new / rr) * p
        rr = rr_new
        if it in keep:
            out[it] = r.copy()
    return out


def _rms(v, active):
    cnt = max(int(active.sum()), 1)
    return float(np.sqrt(np.sum((v * active) ** 2) / cnt))


# ------------------------------------------------------------------ harvest
def harvest_system(ctx: Context, which: int, phi, rng, cg_keep=(2, 6, 16, 40), noise=2):
    """(input [n,n,n,4], target [n,n,n], mask [n,n,n]) samples on the first
    coarse level of one assembled system."""
    op = ctx.export_level(which, 0)
    active = (ctx.export_level32(which, 0)["diag"] > 0).astype(np.float64)
    active_c = ctx.export_level32(which, 1)["diag"] > 0
    feats = ctx.learned_features(which, phi)
    diag, cpl, b = op["diag"], op["cpl"], op["rhs"]
    res = [b * active]
    res += list(cg_residuals(diag, cpl, b, active, set(cg_keep)).values())
    for k in range(noise):
        # white noise, then box-smoothed by 2, 1 and 4 passes in turn
        w = rng.standard_normal(b.shape) * active
        passes = (0, 2, 1, 4)[k % 4]
        if passes:
            w = scenes._box_smooth(w, passes) * active
        res.append(w)
    out = []
    for r in res:
        if _rms(r, active) == 0.0:
            continue
        rc = ctx.restrict_smoothed(which, r).astype(np.float64) * active_c
        sr = _rms(rc, active_c)
        if sr == 0.0:
            continue
        ec, st = ctx.solve_rhs(which, rc, level=1)
        if not st["converged"]:
            continue
        se = _rms(ec, active_c)
        if se == 0.0:
            continue
        inp = np.empty(rc.shape + (4,), np.float32)
        inp[..., 0] = (rc / sr).astype(np.float32)
        inp[..., 1:] = feats
        out.append((inp, (ec * active_c / se).astype(np.float32), active_c.astype(np.float32)))
    return out


CHIS = (0.5, 1.0, 2.0, 4.0, 8.0)


def scene_seeds(count):
    """Seeds of the training systems: half pressure scenes (100, 101, ...),
    half magnetic scenes with chi cycling through CHIS; none is used by the
    tests or the benchmark (seeds 0-9)."""
    half = count // 2
    pressure = tuple(range(100, 100 + half))
    magnetic = tuple((100 + half + i, CHIS[i % len(CHIS)]) for i in range(count - half))
    return pressure, magnetic


def harvest(n=128, pressure_seeds=(100, 101, 102, 103, 104, 105, 106, 107),
            magnetic=((108, 0.5), (109, 1.0), (110, 2.0), (111, 4.0), (112, 8.0), (113, 2.0)),
            seed=0, log=print, smoothing=None, cg_keep=(2, 6, 16, 40), noise=2):
    samples = []
    rng = np.random.default_rng(seed)
    with Context(n, 1.0 / n, scenes.Geometry(n, 1.0 / n).origin, tol=1e-7, max_iters=200,
                 precond=PRECOND_MG, **(smoothing or {})) as ctx:
        for s in pressure_seeds:
            sc = drops_scene(n, s)
            ctx.pressure_project(sc, want_pressure=False)
            got = harvest_system(ctx, 0, sc["phi"], rng, cg_keep, noise)
            samples += got
            log(f"harvest pressure seed {s}: {len(got)} samples")
        for s, chi in magnetic:
            sc = scenes.magnetic_scene(n, s, chi)
            ctx.magnetic_solve(sc, want_field=False)
            got = harvest_system(ctx, 1, sc["phi"], rng, cg_keep, noise)
            samples += got
            log(f"harvest magnetic seed {s} chi {chi}: {len(got)} samples")
    return samples


# ------------------------------------------------------------------ training
def train(samples, steps=6000, lr=1e-3, seed=0, batch=8, init=None, log=print, log_every=200,
          crop=0):
    """Adam on the masked MSE over batches of randomly flipped samples (crop >
    0: random crop^3 windows of them, the network being fully convolutional);
    returns (params, losses)."""
    import torch  # device arrays only

    shape = samples[0][0].shape[:3]
    if crop:
        if crop % 4 or any(crop > e for e in shape):
            raise ValueError(f"crop {crop} must be a multiple of 4 within the samples' {shape}")
        out_shape = (crop, crop, crop)
    else:
        out_shape = shape
    dev = torch.device("cuda:0")
    inp = [torch.from_numpy(s[0]).to(dev) for s in samples]
    tgt = [torch.from_numpy(s[1]).to(dev) for s in samples]
    msk = [torch.from_numpy(s[2]).to(dev) for s in samples]
    torch.cuda.synchronize()
    rng = np.random.default_rng(seed)
    losses = []
    with UNet(batch, *out_shape, training=True) as net:
        if init is None:
            net.init_params(seed + 1)
        else:
            net.set_params(init)
        t0 = time.time()
        for step in range(steps):
            xs, ts, ms = [], [], []
            for i in rng.integers(len(samples), size=batch):
                axes = [a for a in range(3) if rng.uniform() < 0.5]
                x, t, m = inp[i], tgt[i], msk[i]
                if crop:
                    o = [int(rng.integers(0, e - crop + 1)) for e in shape]
                    win = tuple(slice(o[a], o[a] + crop) for a in range(3))
                    x, t, m = x[win], t[win], m[win]
                if axes:
                    x, t, m = torch.flip(x, axes), torch.flip(t, axes), torch.flip(m, axes)
                xs.append(x)
                ts.append(t)
                ms.append(m)
            x = torch.stack(xs).contiguous()
            t = torch.stack(ts).contiguous()
            m = torch.stack(ms).contiguous()
            torch.cuda.synchronize()
            frac = step / max(steps - 1, 1)
            rate = lr * (0.05 + 0.95 * 0.5 * (1.0 + np.cos(np.pi * frac)))  # cosine decay
            loss = net.train_step_device(x.data_ptr(), t.data_ptr(), m.data_ptr(), rate)
            losses.append(loss)
            if log and (step % log_every == 0 or step == steps - 1):
                log(f"step {step:5d} loss {loss:.4f} mean100 {np.mean(losses[-100:]):.4f} "
                    f"lr {rate:.2e} {time.time() - t0:.1f}s")
        params = net.get_params()
    return params, losses


def save_weights(params, path, meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    params = np.ascontiguousarray(params, dtype="<f4")
    params.tofile(path)
    meta = dict(meta)
    meta["count"] = int(params.size)
    meta["sha256"] = hashlib.sha256(params.tobytes()).hexdigest()
    with open(path + ".json", "w") as f:
        json.dump(meta, f, indent=1)


def load_weights(path=WEIGHTS):
    """Weights and their metadata; checks the count and the checksum."""
    params = np.fromfile(path, dtype="<f4")
    with open(path + ".json") as f:
        meta = json.load(f)
    if params.size != unet_param_count() or params.size != meta["count"]:
        raise ValueError(f"{path}: {params.size} parameters, expected {unet_param_count()}")
    if hashlib.sha256(params.tobytes()).hexdigest() != meta["sha256"]:
        raise ValueError(f"{path}: checksum mismatch")
    return params.astype(np.float32), meta


# ------------------------------------------------------------------ calibration
def calibrate(params, meta, n=256, log=print):
    """Measures, on this GPU and at n^3, what a learned iteration must achieve
    to pay for itself, and returns (stall_ratio, record).

    For the pressure projection and the magnetic systems (chi 0.5, 2, 8) of
    the benchmark scenes: the plain MG-PCG solve's mean per-iteration
    residual reduction f and time per iteration c_mg, the learned
    preconditioner's time per iteration c_l (fall-back off) and its reduction
    in each of its first three iterations. A learned iteration costs
    rho = c_l / c_mg plain iterations, which reduce the residual by f^rho: the
    break-even reduction. stall_ratio is the geometric mean of the systems'
    break-evens, so the solver keeps the network only while an iteration
    reduces ||r|| by more than rho plain iterations would."""
    from . import scenes as sc_mod

    systems = [("pressure", sc_mod.pressure_scene(n, 0))]
    systems += [(f"magnetic_chi{c:g}", sc_mod.magnetic_scene(n, 0, c)) for c in (0.5, 2.0, 8.0)]
    lp = learned_params(meta)

    def solve(ctx, name, sc):
        if name == "pressure":
            return ctx.pressure_project(sc, want_pressure=False)[2]
        ctx.reset_warm_start()
        return ctx.magnetic_solve(sc, want_field=False)[2]

    rec, evens = {}, []
    for name, sc in systems:
        r = {}
        with Context.for_scene(sc, tol=1e-8, max_iters=200) as ctx:
            solve(ctx, name, sc)
            st = solve(ctx, name, sc)
            r["mg_iterations"] = st["iterations"]
            r["mg_factor"] = st["rel_residual"] ** (1.0 / st["iterations"])
            r["mg_ms_per_it"] = st["solve_ms"] / st["iterations"]
        with Context.for_scene(sc, tol=1e-8, max_iters=200) as ctx:
            ctx.set_network(params)
            ctx.set_params(**dict(lp, stall_ratio=100.0))  # never falls back
            solve(ctx, name, sc)
            st = solve(ctx, name, sc)
            r["learned_iterations"] = st["iterations"]
            r["learned_ms_per_it"] = st["solve_ms"] / st["iterations"]
            rels = []
            for k in (1, 2, 3):
                ctx.set_params(max_iters=k)
                rels.append(solve(ctx, name, sc)["rel_residual"])
            r["learned_factors"] = [rels[0], rels[1] / rels[0], rels[2] / rels[1]]
        r["cost_ratio"] = r["learned_ms_per_it"] / r["mg_ms_per_it"]
        r["break_even"] = r["mg_factor"] ** r["cost_ratio"]
        evens.append(r["break_even"])
        log(f"[calibrate] {name}: MG {r['mg_iterations']} its, factor {r['mg_factor']:.3f}, "
            f"{r['mg_ms_per_it']:.3f} ms/it; learned {r['learned_ms_per_it']:.3f} ms/it "
            f"(x{r['cost_ratio']:.2f}), first factors "
            + ", ".join(f"{v:.3f}" for v in r["learned_factors"]) + f"; break-even {r['break_even']:.2e}")
        rec[name] = r
    stall = float(np.exp(np.mean(np.log(evens))))
    return stall, dict(n=n, systems=rec, stall_ratio=stall)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--systems", type=int, default=14,
                    help="training systems (half pressure, half magnetic scenes)")
    ap.add_argument("--crop", type=int, default=0,
                    help="train on random crop^3 windows of the coarse-level samples (0: whole)")
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--cheby", default=None, metavar="DEG:RATIO:COARSE_DEG:COARSE_RATIO",
                    help="V-cycle smoother to train under (default: the library's)")
    ap.add_argument("--cg-keep", default="2,6,16,40", metavar="I,J,...",
                    help="CG iterations whose residuals become samples (with the right-hand side)")
    ap.add_argument("--noise", type=int, default=2,
                    help="noise residuals per system (white, then box-smoothed by 2, 1, 4 passes)")
    ap.add_argument("--calibrate", type=int, default=0, metavar="N",
                    help="measure the fall-back threshold of the weights at --out on N^3 scenes "
                         "instead of training; the record goes to --calibration-out")
    ap.add_argument("--calibration-out", default=None,
                    help="JSON file for the calibration record (default: print it)")
    a = ap.parse_args(argv)
    import torch  # its HIP runtime must initialise before libsfp's
    torch.cuda.init()
    t0 = time.time()
    log = lambda s: print(s, flush=True)  # noqa: E731
    if a.calibrate:
        params, meta = load_weights(a.out)
        stall, rec = calibrate(params, meta, a.calibrate, log=log)
        rec["sha256"] = meta["sha256"]
        if a.calibration_out:
            with open(a.calibration_out, "w") as f:
                json.dump(rec, f, indent=1)
        log(json.dumps(dict(stall_ratio=stall)))
        return 0
    from . import default_params
    dp = default_params()
    sm = {k: getattr(dp, k) for k in SMOOTHER_KEYS}
    if a.cheby:
        v = a.cheby.split(":")
        sm = dict(cheby_degree=int(v[0]), cheby_ratio=float(v[1]), coarse_cheby_degree=int(v[2]),
                  coarse_cheby_ratio=float(v[3]))
    ps, ms = scene_seeds(a.systems)
    cg_keep = tuple(int(v) for v in a.cg_keep.split(","))
    samples = harvest(a.n, ps, ms, seed=a.seed, log=log, smoothing=sm, cg_keep=cg_keep,
                      noise=a.noise)
    th = time.time() - t0
    log(f"{len(samples)} samples in {th:.1f}s")
    params, losses = train(samples, a.steps, a.lr, a.seed, a.batch, log=log, crop=a.crop)
    save_weights(params, a.out, dict(n=a.n, steps=a.steps, lr=a.lr, seed=a.seed, batch=a.batch,
                                     samples=len(samples), systems=a.systems, crop=a.crop,
                                     cg_keep=list(cg_keep), noise=a.noise,
                                     crops_seen=a.steps * a.batch if a.crop else 0,
                                     harvest_s=round(th, 1),
                                     train_s=round(time.time() - t0 - th, 1),
                                     loss_first100=float(np.mean(losses[:100])),
                                     loss_last100=float(np.mean(losses[-100:])), smoother=sm))
    log(f"saved {a.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
