"""An independent, vectorized numpy transcription of the reference assembly
(core/Pressure.cpp:38-78,80-102,137-166; core/MagneticField.cpp:50-143),
used to cross-check the C++ oracle on the CPU.

`mutation` plants one plausible mistake, so the tests can show that the
checks they rely on (agreement with the oracle, closed forms, invariants)
catch it:
  "no_jump"       drop the ghost-fluid jump term from the right-hand side
  "solid_sign"    wrong sign of the solid-velocity term
  "theta_side"    use theta (not 1 - theta) for the minus-side faces in the update
  "no_clamp"      forget the max(theta, .001) clamp
  "band_half"     use half the band width in the smoothed Heaviside
  "mu_side"       put mu of the cell's own centre on the face instead of the face average
  "rhs_sign"      flip the magnetic right-hand side
  "pin_keep"      keep column 0 in the neighbours' rows (no elimination)
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def _shift(a, axis, d, fill):
    """b[c] = a[c + d e_axis] (fill outside)."""
    b = np.full_like(a, fill)
    src = [slice(None)] * 3
    dst = [slice(None)] * 3
    n = a.shape[axis]
    if d > 0:
        src[axis] = slice(d, n)
        dst[axis] = slice(0, n - d)
    else:
        src[axis] = slice(0, n + d)
        dst[axis] = slice(-d, n)
    b[tuple(dst)] = a[tuple(src)]
    return b


def _face(a_face, axis, upper, n):
    """Face values of the lower (upper=False) or upper face of every cell."""
    s = [slice(0, n[b]) for b in range(3)]
    if upper:
        s[axis] = slice(1, n[axis] + 1)
    return a_face[tuple(s)]


def pressure_assembly(sc, mutation=None):
    """Returns (liquid mask, diag, [cpl_x, cpl_y, cpl_z], rhs) on the full grid."""
    phi = sc["phi"]
    n = phi.shape
    liq = phi <= 0
    diag = np.zeros(n)
    div = np.zeros(n)
    low = [np.zeros(n) for _ in range(3)]
    jump = sc.get("jump")
    for ord_ in range(6):
        axis, side = ord_ >> 1, (1 if ord_ & 1 else -1)
        w = 1 - _face(sc["fraction"][axis], axis, side > 0, n)
        u = _face(sc["velocity"][axis], axis, side > 0, n)
        us = _face(sc["solid_velocity"][axis], axis, side > 0, n)
        nb_phi = _shift(phi, axis, side, np.nan)
        nb_in = ~np.isnan(_shift(np.zeros(n), axis, side, np.nan))
        nb_liq = nb_in & (nb_phi <= 0)
        open_ = w > 0
        if np.any(liq & open_ & ~nb_in):
            raise ValueError("open boundary face next to liquid")
        m_l = liq & open_ & nb_liq
        m_a = liq & open_ & nb_in & ~nb_liq
        diag = np.where(m_l, diag + w, diag)
        if side < 0:
            low[axis] = np.where(m_l, w, 0.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            theta = phi / (phi - nb_phi)
        clamp = theta if mutation == "no_clamp" else np.where(theta < .001, .001, theta)
        with np.errstate(divide="ignore"):
            ic = 1. / clamp
        diag = np.where(m_a, diag + w * ic, diag)
        if jump is not None and mutation != "no_jump":
            J = _face(jump[axis], axis, side > 0, n)
            div = np.where(m_a, div + J * w * ic, div)
        div = np.where(liq & open_, div - side * w * u, div)
        ssign = -side if mutation == "solid_sign" else side
        div = np.where(liq & (w < 1), div - ssign * (1 - w) * us, div)
    div = div + sc.get("vol_error", 0.0)
    rhs = np.where(liq, div, 0.0)
    diag = np.where(liq, np.where(diag != 0, diag, 1.0), 1.0)
    return liq, diag, low, rhs


def pressure_apply(sc, liq, p, mutation=None):
    """ApplyProjection on copies of the velocities; p on the full grid."""
    phi = sc["phi"]
    n = phi.shape
    out = []
    jump = sc.get("jump")
    for axis in range(3):
        u = np.array(sc["velocity"][axis], copy=True)
        w = 1. - sc["fraction"][axis]
        # interior faces only (boundary faces must be closed)
        s_in = [slice(None)] * 3
        s_in[axis] = slice(1, n[axis])
        lo = [slice(None)] * 3
        lo[axis] = slice(0, n[axis] - 1)
        hi = [slice(None)] * 3
        hi[axis] = slice(1, n[axis])
        l0, l1 = liq[tuple(lo)], liq[tuple(hi)]
        p0, p1 = p[tuple(lo)], p[tuple(hi)]
        f0, f1 = phi[tuple(lo)], phi[tuple(hi)]
        wi = w[tuple(s_in)]
        ui = u[tuple(s_in)]
        with np.errstate(divide="ignore", invalid="ignore"):
            theta = f0 / (f0 - f1)
        both = l0 & l1 & (wi != 0)
        one = (l0 ^ l1) & (wi != 0)
        pin = np.where(l0, p0, p1)
        pintf = jump[axis][tuple(s_in)] if jump is not None else 0.
        tt = theta if mutation == "theta_side" else np.where(l0, theta, 1 - theta)
        with np.errstate(divide="ignore"):
            ic = 1. / np.where(tt < .001, .001, tt)
        sgn = np.where(l0, 1.0, -1.0)
        new = np.where(both, ui - (p1 - p0) * wi, ui)
        new = np.where(one, ui - sgn * (pintf - pin) * ic * wi, new)
        u[tuple(s_in)] = new
        out.append(u)
    return out


def heaviside(phi, band):
    pi = np.pi
    mid = .5 * (1. + phi / band + np.sin(pi * phi / band) / pi)
    return np.where(phi < -band, 0., np.where(phi > band, 1., mid))


def magnetic_mu(sc, mutation=None):
    """Face permeability; TriLerp of phi at an interior face centre reduces to
    the two-cell average (weights 1/2, 1/2 up to rounding)."""
    geo = sc["geo"]
    phi = sc["phi"]
    n = phi.shape
    band = sc["band"] * (0.5 if mutation == "band_half" else 1.0)
    chi = sc["chi"]
    mu = []
    for axis in range(3):
        m = np.ones(geo.face_shape(axis))
        lo = [slice(None)] * 3
        lo[axis] = slice(0, n[axis] - 1)
        hi = [slice(None)] * 3
        hi[axis] = slice(1, n[axis])
        face_phi = phi[tuple(hi)] if mutation == "mu_side" else 0.5 * phi[tuple(lo)] + 0.5 * phi[tuple(hi)]
        s_in = [slice(None)] * 3
        s_in[axis] = slice(1, n[axis])
        m[tuple(s_in)] = 1. + (1. - heaviside(face_phi, band)) * chi
        mu.append(m)
    return mu


def magnetic_system(sc, mu, mutation=None):
    """CSR matrix (pinned cell 0) and right-hand side, z-fastest order."""
    geo = sc["geo"]
    n = geo.n
    N = geo.ncells
    idx = np.arange(N).reshape(n)
    strides = (n[1] * n[2], n[2], 1)
    rows, cols, vals = [], [], []
    diag = np.zeros(n)
    rhs = np.zeros(n)
    dx = geo.spacing
    for ord_ in range(6):
        axis, side = ord_ >> 1, (1 if ord_ & 1 else -1)
        m = _face(mu[axis], axis, side > 0, n)
        h = _face(sc["hext"][axis], axis, side > 0, n)
        valid = ~np.isnan(_shift(np.zeros(n), axis, side, np.nan))
        diag = np.where(valid, diag + m, diag)
        sgn = side if mutation == "rhs_sign" else -side
        rhs = np.where(valid, rhs + (m - 1.) * h * sgn * dx, rhs)
        r = idx[valid]
        c = r + side * strides[axis]
        keep = (c != 0) | (mutation == "pin_keep")
        keep = keep & (r != 0)
        rows.append(r[keep])
        cols.append(c[keep])
        vals.append(-m[valid][keep])
    d = diag.ravel().copy()
    d[0] = 1.0
    rows.append(idx.ravel())
    cols.append(idx.ravel())
    vals.append(d)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    A.sort_indices()
    b = rhs.ravel().copy()
    b[0] = 0.0
    return A, b


def magnetic_total_field(sc, pot):
    geo = sc["geo"]
    n = geo.n
    pot = pot.reshape(n)
    out = []
    for axis in range(3):
        h = np.zeros(geo.face_shape(axis))
        lo = [slice(None)] * 3
        lo[axis] = slice(0, n[axis] - 1)
        hi = [slice(None)] * 3
        hi[axis] = slice(1, n[axis])
        s_in = [slice(None)] * 3
        s_in[axis] = slice(1, n[axis])
        h[tuple(s_in)] = (pot[tuple(lo)] - pot[tuple(hi)]) * (1 / geo.spacing)
        out.append(h + sc["hext"][axis])
    return out
