"""Training of the U-Net preconditioner (include/sfp_learned.h).

The network maps a residual r (scaled to unit RMS over the active cells) and
three operator features to a search direction z for the FP64 flexible CG.
FCG is invariant to the scale of z, so the network is trained on directions:
the target of a sample is e = A^-1 r scaled to unit RMS, with a masked MSE
over the active cells.

Samples come from systems the library assembles itself (pressure projections
of drops over pools and magnetostatic solves of wavy interfaces):

* r = the system's right-hand side;
* r = residuals of unpreconditioned CG after a few iterations (what is left
  after the easy, high-frequency part is gone);
* r = white and box-smoothed noise on the active cells;

and e is computed by the library's own multigrid-preconditioned solve
(sfp_test_solve). Training runs on the GPU through sfp_unet_train_step_device
(the library's forward, backward and Adam kernels); PyTorch only holds the
device arrays and flips them for augmentation.

Usage: python -m sfp.learn --out weights/unet_precond.f32 [--n 128 --steps 1500]
(from the package directory), or train() from Python.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

from . import Context, UNet, PRECOND_MG, unet_param_count
from . import scenes

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "weights",
                       "unet_precond.f32")


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
    """(input [N,N,N,4], target [N,N,N], mask [N,N,N]) samples of one system."""
    op = ctx.export_level(which, 0)
    d32 = ctx.export_level32(which, 0)["diag"]
    active = (d32 > 0).astype(np.float64)
    feats = ctx.learned_features(which, phi)
    diag, cpl, b = op["diag"], op["cpl"], op["rhs"]
    res = [b * active]
    res += list(cg_residuals(diag, cpl, b, active, set(cg_keep)).values())
    for k in range(noise):
        w = rng.standard_normal(b.shape) * active
        if k % 2 == 1:
            w = scenes._box_smooth(w, 2) * active
        res.append(w)
    out = []
    for r in res:
        sr = _rms(r, active)
        if sr == 0.0:
            continue
        e, st = ctx.solve_rhs(which, r)
        if not st["converged"]:
            continue
        se = _rms(e, active)
        if se == 0.0:
            continue
        inp = np.empty(b.shape + (4,), np.float32)
        inp[..., 0] = (r * active / sr).astype(np.float32)
        inp[..., 1:] = feats
        out.append((inp, (e * active / se).astype(np.float32), active.astype(np.float32)))
    return out


def harvest(n=128, pressure_seeds=(1, 2, 3, 4), magnetic=((5, 0.5), (6, 2.0), (7, 8.0)), seed=0,
            log=print):
    samples = []
    rng = np.random.default_rng(seed)
    with Context(n, 1.0 / n, scenes.Geometry(n, 1.0 / n).origin, tol=1e-7, max_iters=200,
                 precond=PRECOND_MG) as ctx:
        for s in pressure_seeds:
            sc = drops_scene(n, s)
            ctx.pressure_project(sc, want_pressure=False)
            got = harvest_system(ctx, 0, sc["phi"], rng)
            samples += got
            log(f"harvest pressure seed {s}: {len(got)} samples")
        for s, chi in magnetic:
            sc = scenes.magnetic_scene(n, s, chi)
            ctx.magnetic_solve(sc, want_field=False)
            got = harvest_system(ctx, 1, sc["phi"], rng)
            samples += got
            log(f"harvest magnetic seed {s} chi {chi}: {len(got)} samples")
    return samples


# ------------------------------------------------------------------ training
def train(samples, steps=1500, lr=1e-3, seed=0, init=None, log=print, log_every=100):
    """Adam on the masked MSE with random axis flips; returns (params, losses)."""
    import torch  # device arrays only

    shape = samples[0][0].shape[:3]
    dev = torch.device("cuda:0")
    inp = [torch.from_numpy(s[0]).to(dev) for s in samples]
    tgt = [torch.from_numpy(s[1]).to(dev) for s in samples]
    msk = [torch.from_numpy(s[2]).to(dev) for s in samples]
    torch.cuda.synchronize()
    rng = np.random.default_rng(seed)
    losses = []
    with UNet(1, *shape, training=True) as net:
        if init is None:
            net.init_params(seed + 1)
        else:
            net.set_params(init)
        t0 = time.time()
        for step in range(steps):
            i = int(rng.integers(len(samples)))
            axes = [a for a in range(3) if rng.uniform() < 0.5]
            x, t, m = inp[i], tgt[i], msk[i]
            if axes:
                x = torch.flip(x, axes).contiguous()
                t = torch.flip(t, axes).contiguous()
                m = torch.flip(m, axes).contiguous()
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


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=WEIGHTS)
    a = ap.parse_args(argv)
    import torch  # its HIP runtime must initialise before libsfp's
    torch.cuda.init()
    t0 = time.time()
    log = lambda s: print(s, flush=True)  # noqa: E731
    samples = harvest(a.n, seed=a.seed, log=log)
    th = time.time() - t0
    log(f"{len(samples)} samples in {th:.1f}s")
    params, losses = train(samples, a.steps, a.lr, a.seed, log=log)
    save_weights(params, a.out, dict(n=a.n, steps=a.steps, lr=a.lr, seed=a.seed,
                                     samples=len(samples), harvest_s=round(th, 1),
                                     train_s=round(time.time() - t0 - th, 1),
                                     loss_first100=float(np.mean(losses[:100])),
                                     loss_last100=float(np.mean(losses[-100:]))))
    log(f"saved {a.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
