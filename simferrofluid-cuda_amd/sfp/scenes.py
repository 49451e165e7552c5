"""Seeded synthetic scenes for the two grid Poisson solves.

Every random number comes from a counter-based generator: SplitMix64 of a key
built from (seed, array id, linear index), so any array can be regenerated
independently, in any order, on any machine, bit for bit.

Layout follows the reference containers (core/Grid.h:28): linear index
``z + nz*(y + ny*x)``; numpy arrays are therefore shaped ``(nx, ny, nz)`` in C
order. Face grid ``a`` has extent ``N + e_a`` (core/StaggeredGrid.h:15-19).

Scenes:

* ``pressure_scene``  -- sphere drop above a pool, box walls, smoothed random
  face velocities (the 64^3 ghost-fluid projection case, rescaled to any N).
* ``magnetic_scene``  -- cosine-perturbed flat interface, uniform applied field
  (the 128^3 FDM magnetostatic case, rescaled to any N).
* ``spike_scene``     -- 64-mode spike-pattern interface (the 512^3 case).
* ``random_pressure_scene`` -- fractional solid fractions, solid velocities,
  jumps and a volume error, for exercising every branch of the assembly.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

# array ids keep the random streams of different arrays disjoint
ID_VEL = 0          # +axis (0..2)
ID_FRAC = 4         # +axis
ID_USOLID = 8       # +axis
ID_JUMP = 12        # +axis
ID_MODE_A = 16
ID_MODE_M = 17
ID_MODE_N = 18
ID_MODE_T = 19
ID_MISC = 24


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014) on uint64 arrays."""
    with np.errstate(over="ignore"):
        z = x.astype(np.uint64) + _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, array_id: int, n: int, start: int = 0) -> np.ndarray:
    """n doubles in [0,1) for stream (seed, array_id), entries start..start+n-1."""
    idx = np.arange(start, start + n, dtype=np.uint64)
    key = (np.uint64(seed & 0xFFFF) << np.uint64(48)) ^ (np.uint64(array_id & 0xFF) << np.uint64(40)) ^ idx
    return (splitmix64(key) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


class Geometry:
    """Cell/face geometry of a StaggeredGrid (core/StaggeredGrid.h:8-21).

    ``origin`` is the grid's lower corner, computed like the reference ctor:
    ``center - resolution*spacing/2``.
    """

    def __init__(self, n, spacing: float, center=(0.5, 0.5, 0.5)):
        if np.isscalar(n):
            n = (int(n),) * 3
        self.n = tuple(int(v) for v in n)
        self.spacing = float(spacing)
        self.origin = tuple(float(c) - (float(r) * self.spacing) / 2 for c, r in zip(center, self.n))

    @property
    def ncells(self) -> int:
        return self.n[0] * self.n[1] * self.n[2]

    def face_shape(self, axis: int):
        s = list(self.n)
        s[axis] += 1
        return tuple(s)

    def cell_centers(self):
        """(X, Y, Z) broadcastable arrays of cell-centre coordinates."""
        h = self.spacing
        o = [self.origin[a] + h / 2 for a in range(3)]
        xs = [o[a] + np.arange(self.n[a], dtype=np.float64) * h for a in range(3)]
        return (xs[0][:, None, None], xs[1][None, :, None], xs[2][None, None, :])

    def face_centers(self, axis: int):
        h = self.spacing
        o = [self.origin[a] + (0.0 if a == axis else h / 2) for a in range(3)]
        shp = self.face_shape(axis)
        xs = [o[a] + np.arange(shp[a], dtype=np.float64) * h for a in range(3)]
        return (xs[0][:, None, None], xs[1][None, :, None], xs[2][None, None, :])


def _box_smooth(u: np.ndarray, passes: int) -> np.ndarray:
    """``passes`` rounds of the 3x3x3 box mean with clamped (edge) indexing,
    applied as three 1-D 3-point means (the clamped box is separable)."""
    for _ in range(passes):
        for a in range(3):
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[a] = slice(0, -1)
            hi[a] = slice(1, None)
            m = u.copy()
            m[tuple(hi)] += u[tuple(lo)]   # left neighbour
            m[tuple(lo)] += u[tuple(hi)]   # right neighbour
            e0 = [slice(None)] * 3
            e1 = [slice(None)] * 3
            e0[a] = slice(0, 1)
            e1[a] = slice(-1, None)
            m[tuple(e0)] += u[tuple(e0)]   # clamped left edge
            m[tuple(e1)] += u[tuple(e1)]   # clamped right edge
            u = m / 3.0
    return u


def wall_fractions(geo: Geometry):
    """Solid fraction 1 on the outermost faces of every face grid, else 0."""
    fr = []
    for a in range(3):
        f = np.zeros(geo.face_shape(a))
        idx = [slice(None)] * 3
        idx[a] = 0
        f[tuple(idx)] = 1.0
        idx[a] = geo.n[a]
        f[tuple(idx)] = 1.0
        fr.append(f)
    return fr


def pressure_scene(n: int = 64, seed: int = 0):
    """Sphere (r=0.2 at (.5,.6,.5)) above a pool (y<0.25) in a closed unit box.

    Returns a dict with geometry, phi (cells), velocity[3], fraction[3],
    solid_velocity[3] (zeros), jump=None, vol_error=0.
    """
    geo = Geometry(n, 1.0 / n)
    X, Y, Z = geo.cell_centers()
    sphere = np.sqrt((X - 0.5) ** 2 + (Y - 0.6) ** 2 + (Z - 0.5) ** 2) - 0.2
    phi = np.minimum(sphere, Y - 0.25 + 0 * X + 0 * Z)
    vel = []
    for a in range(3):
        shp = geo.face_shape(a)
        u = uniform(seed, ID_VEL + a, int(np.prod(shp))).reshape(shp) * 2.0 - 1.0
        vel.append(_box_smooth(u, 3))
    return dict(geo=geo, phi=np.ascontiguousarray(phi), velocity=vel,
                fraction=wall_fractions(geo),
                solid_velocity=[np.zeros(geo.face_shape(a)) for a in range(3)],
                jump=None, vol_error=0.0)


def _modes(seed: int, count: int, amp: float, mmax: int = 6):
    a = uniform(seed, ID_MODE_A, count) * amp
    m = 1 + np.floor(uniform(seed, ID_MODE_M, count) * mmax).astype(np.int64)
    nn = 1 + np.floor(uniform(seed, ID_MODE_N, count) * mmax).astype(np.int64)
    th = uniform(seed, ID_MODE_T, count) * (2.0 * np.pi)
    return a, m, nn, th


def _wavy_phi(geo: Geometry, y0: float, modes):
    X, Y, Z = geo.cell_centers()
    a, m, nn, th = modes
    s = np.zeros((geo.n[0], 1, geo.n[2]))
    for k in range(len(a)):
        s = s + a[k] * np.cos(2.0 * np.pi * (m[k] * X + nn[k] * Z) + th[k])
    return np.ascontiguousarray(np.broadcast_to(Y - y0 - s, geo.n)).copy()


def uniform_field(geo: Geometry, h):
    return [np.full(geo.face_shape(a), float(h[a])) for a in range(3)]


def magnetic_scene(n=128, seed: int = 0, chi: float = 2.0, band_cells: float = 1.5):
    """Cosine-perturbed interface y = 0.3 + sum of 8 modes, H_ext = (0,1,0);
    n is the cells per axis (or (nx, ny, nz), spacing 1 / max)."""
    geo = Geometry(n, 1.0 / (n if np.isscalar(n) else max(n)))
    phi = _wavy_phi(geo, 0.3, _modes(seed, 8, 0.01))
    return dict(geo=geo, phi=phi, hext=uniform_field(geo, (0.0, 1.0, 0.0)),
                chi=chi, band=band_cells * geo.spacing)


def spike_scene(n: int = 512, seed: int = 0, chi: float = 4.0):
    """Spike-pattern interface: y = 0.2 + 64 random cosine modes (|a| <= 0.005)."""
    geo = Geometry(n, 1.0 / n)
    phi = _wavy_phi(geo, 0.2, _modes(seed, 64, 0.005, mmax=12))
    return dict(geo=geo, phi=phi, hext=uniform_field(geo, (0.0, 1.5, 0.0)),
                chi=chi, band=1.5 * geo.spacing)


def random_pressure_scene(n=16, seed: int = 0, vol_error: float = 0.0137):
    """Pressure inputs that reach every branch of the ghost-fluid assembly.

    Solid fractions are 1 on the box faces and fractional (including values
    > 0.9 and exact 0) on a band of interior faces; solid velocities, jumps and
    a volume error are non-zero.
    """
    if np.isscalar(n):
        n = (int(n),) * 3
    geo = Geometry(n, 1.0 / max(n))
    X, Y, Z = geo.cell_centers()
    r = uniform(seed, ID_MISC, geo.ncells).reshape(geo.n)
    phi = np.minimum(np.sqrt((X - 0.5) ** 2 + (Y - 0.62) ** 2 + (Z - 0.48) ** 2) - 0.21,
                     Y - 0.33 + 0.05 * np.cos(6.0 * X) + 0 * Z) + 0.02 * (r - 0.5)
    phi = np.ascontiguousarray(phi)
    vel, frac, us, jump = [], [], [], []
    walls = wall_fractions(geo)
    for a in range(3):
        shp = geo.face_shape(a)
        cnt = int(np.prod(shp))
        vel.append(uniform(seed, ID_VEL + a, cnt).reshape(shp) * 2.0 - 1.0)
        f = uniform(seed, ID_FRAC + a, cnt).reshape(shp)
        Xf, Yf, Zf = geo.face_centers(a)
        band = (Yf + 0 * Xf + 0 * Zf) < 0.2
        f = np.where(band, np.where(f < 0.3, 0.0, np.where(f > 0.85, 1.0, f)), 0.0)
        frac.append(np.maximum(f, walls[a]))
        us.append(uniform(seed, ID_USOLID + a, cnt).reshape(shp) - 0.5)
        jump.append((uniform(seed, ID_JUMP + a, cnt).reshape(shp) - 0.5) * 0.3)
    return dict(geo=geo, phi=phi, velocity=vel, fraction=frac, solid_velocity=us,
                jump=jump, vol_error=vol_error)
