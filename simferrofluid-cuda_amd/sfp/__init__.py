"""Python binding of libsfp.so (the C ABI in include/sfp.h), used by the tests,
the smoke check and bench.py.

There is no fallback: if libsfp.so is missing or no GPU is usable, the calls
raise. Arrays follow the reference layout (z-fastest, numpy C order on shape
(nx, ny, nz); face grid a has extent n + e_a).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "libsfp.so"))
_LIB = None

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)

OK, ERR_ARGUMENT, ERR_INPUT, ERR_HIP, ERR_NO_DEVICE, ERR_STATE, ERR_UNSUPPORTED = range(7)
PRECOND_MG, PRECOND_LEARNED, PRECOND_JACOBI, PRECOND_SMOOTH = 0, 1, 2, 3
MAGNETIC_BC_REFERENCE, MAGNETIC_BC_DIRICHLET_BOX = 0, 1
KERNEL_SMOOTH, KERNEL_MATVEC64, KERNEL_UPDATE64, KERNEL_RESTRICT = 0, 1, 2, 3
KERNEL_SMOOTH_PAIR, KERNEL_SMOOTH_RESTRICT = 4, 5
KERNEL_NAMES = {0: "smooth", 1: "matvec64", 2: "update64", 3: "restrict", 4: "smooth_pair",
                5: "smooth_restrict"}

# exported symbols, per header (checked by the CPU tests)
EXPORTS = {
    "sfp.h": ["sfp_create", "sfp_destroy", "sfp_default_params", "sfp_set_params",
              "sfp_get_params", "sfp_last_error", "sfp_status_string", "sfp_stream",
              "sfp_reset_warm_start", "sfp_pressure_project", "sfp_pressure_reproject",
              "sfp_magnetic_solve", "sfp_magnetic_vertex_pressures",
              "sfp_pressure_project_device", "sfp_pressure_reproject_device",
              "sfp_magnetic_solve_device", "sfp_extrapolate_faces", "sfp_extrapolate_velocity",
              "sfp_collider_enforce", "sfp_extrapolate_velocity_device",
              "sfp_collider_enforce_device", "sfp_extrapolate_velocity_noalloc",
              "sfp_projection_substep"],
    "sfp_profile.h": ["sfp_profile_enable", "sfp_profile_query"],
    "sfp_test.h": ["sfp_test_num_levels", "sfp_test_level_dims", "sfp_test_export_level",
                   "sfp_test_export_level32", "sfp_test_apply64", "sfp_test_precondition",
                   "sfp_test_residual", "sfp_test_solve",
                   "sfp_test_restrict_smoothed", "sfp_test_smooth_pair",
                   "sfp_test_smooth_restrict", "sfp_test_set_vcycle_tail",
                   "sfp_test_set_fused_pairs", "sfp_test_set_rowsum_diagonal"],
    "sfp_learned.h": ["sfp_unet_param_count", "sfp_set_network", "sfp_learned_features",
                      "sfp_unet_create", "sfp_unet_destroy", "sfp_unet_set_params",
                      "sfp_unet_get_params", "sfp_unet_init_params", "sfp_unet_forward",
                      "sfp_unet_gradients", "sfp_unet_train_step_device", "sfp_unet_last_error",
                      "sfp_unet_profile_forward"],
}


class Params(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iters", C.c_int), ("precond", C.c_int),
                ("mg_coarsest", C.c_int), ("cheby_degree", C.c_int), ("cheby_ratio", C.c_double),
                ("coarse_scale", C.c_double), ("stall_ratio", C.c_double), ("verbose", C.c_int),
                ("coarse_cheby_degree", C.c_int), ("coarse_cheby_ratio", C.c_double),
                ("net_cheby_degree", C.c_int), ("net_cheby_ratio", C.c_double),
                ("net_coarse_cheby_degree", C.c_int), ("net_coarse_cheby_ratio", C.c_double),
                ("magnetic_bc", C.c_int), ("pressure_coarse_cheby_degree", C.c_int),
                ("pressure_coarse_cheby_ratio", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("fallback", C.c_int),
                ("precond_used", C.c_int), ("rel_residual", C.c_double), ("setup_ms", C.c_double),
                ("solve_ms", C.c_double), ("apply_ms", C.c_double), ("unknowns", C.c_longlong)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class KernelProfile(C.Structure):
    _fields_ = [("launches", C.c_longlong), ("total_ms", C.c_double), ("bytes", C.c_double)]


class SfpError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"sfp status {status}: {message}")
        self.status = status


P3 = _dp * 3
PP = C.POINTER(_dp)  # parameter type of a `double *const a[3]` argument


def lib():
    """Load libsfp.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libsfp.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.sfp_create.restype = C.c_int
        L.sfp_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_double, _dp, C.c_int]
        L.sfp_destroy.restype = None
        L.sfp_destroy.argtypes = [vp]
        L.sfp_default_params.restype = None
        L.sfp_default_params.argtypes = [C.POINTER(Params)]
        L.sfp_set_params.restype = C.c_int
        L.sfp_set_params.argtypes = [vp, C.POINTER(Params)]
        L.sfp_get_params.restype = C.c_int
        L.sfp_get_params.argtypes = [vp, C.POINTER(Params)]
        L.sfp_last_error.restype = C.c_char_p
        L.sfp_last_error.argtypes = [vp]
        L.sfp_status_string.restype = C.c_char_p
        L.sfp_status_string.argtypes = [C.c_int]
        L.sfp_stream.restype = vp
        L.sfp_stream.argtypes = [vp]
        L.sfp_reset_warm_start.restype = C.c_int
        L.sfp_reset_warm_start.argtypes = [vp]
        for name in ("sfp_pressure_project", "sfp_pressure_reproject",
                     "sfp_pressure_project_device", "sfp_pressure_reproject_device"):
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = [vp, PP, _dp, PP, PP, vp, C.c_double, _dp, C.POINTER(Stats)]
        for name in ("sfp_magnetic_solve", "sfp_magnetic_solve_device"):
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = [vp, _dp, PP, C.c_double, C.c_double, _dp, vp, C.POINTER(Stats)]
        U3 = C.POINTER(C.POINTER(C.c_ubyte))
        L.sfp_extrapolate_faces.restype = C.c_int
        L.sfp_extrapolate_faces.argtypes = [vp, PP, U3, C.c_double, C.c_int]
        for name in ("sfp_extrapolate_velocity", "sfp_extrapolate_velocity_device"):
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = [vp, PP, _dp, PP, C.c_int]
        L.sfp_projection_substep.restype = C.c_int
        L.sfp_projection_substep.argtypes = [vp, PP, _dp, PP, PP, vp, C.c_double, C.c_int, PP, PP, _dp,
                                             C.POINTER(Stats)]
        L.sfp_extrapolate_velocity_noalloc.restype = C.c_int
        L.sfp_extrapolate_velocity_noalloc.argtypes = [C.c_int] * 3 + [PP, _dp, PP, C.c_int, vp]
        for name in ("sfp_collider_enforce", "sfp_collider_enforce_device"):
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = [vp, PP, PP, PP, PP]
        L.sfp_magnetic_vertex_pressures.restype = C.c_int
        L.sfp_magnetic_vertex_pressures.argtypes = [vp, C.c_int, _dp, _dp, C.c_double, _dp]
        L.sfp_profile_enable.restype = C.c_int
        L.sfp_profile_enable.argtypes = [vp, C.c_int]
        L.sfp_profile_query.restype = C.c_int
        L.sfp_profile_query.argtypes = [vp, C.c_int, C.POINTER(KernelProfile)]
        L.sfp_test_num_levels.restype = C.c_int
        L.sfp_test_num_levels.argtypes = [vp, C.c_int, _ip]
        L.sfp_test_level_dims.restype = C.c_int
        L.sfp_test_level_dims.argtypes = [vp, C.c_int, C.c_int, _ip]
        L.sfp_test_export_level.restype = C.c_int
        L.sfp_test_export_level.argtypes = [vp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.sfp_test_export_level32.restype = C.c_int
        L.sfp_test_export_level32.argtypes = [vp, C.c_int, C.c_int, _fp, _fp, _fp, _fp]
        L.sfp_test_apply64.restype = C.c_int
        L.sfp_test_apply64.argtypes = [vp, C.c_int, _fp, _dp]
        L.sfp_test_precondition.restype = C.c_int
        L.sfp_test_precondition.argtypes = [vp, C.c_int, _dp, _fp]
        L.sfp_test_residual.restype = C.c_int
        L.sfp_test_residual.argtypes = [vp, C.c_int, _dp, _dp, _dp, _dp]
        L.sfp_test_solve.restype = C.c_int
        L.sfp_test_solve.argtypes = [vp, C.c_int, C.c_int, _dp, _dp, C.POINTER(Stats)]
        L.sfp_test_restrict_smoothed.restype = C.c_int
        L.sfp_test_restrict_smoothed.argtypes = [vp, C.c_int, _dp, _fp]
        L.sfp_test_set_vcycle_tail.restype = C.c_int
        L.sfp_test_set_vcycle_tail.argtypes = [vp, C.c_int]
        L.sfp_test_set_fused_pairs.restype = C.c_int
        L.sfp_test_set_fused_pairs.argtypes = [vp, C.c_int]
        L.sfp_test_set_rowsum_diagonal.restype = C.c_int
        L.sfp_test_set_rowsum_diagonal.argtypes = [vp, C.c_int]
        L.sfp_test_smooth_restrict.restype = C.c_int
        L.sfp_test_smooth_restrict.argtypes = [vp, C.c_int, C.c_int, _fp, _fp, _fp, C.c_float,
                                               C.c_float, _fp, _fp, _fp, _fp]
        L.sfp_test_smooth_pair.restype = C.c_int
        L.sfp_test_smooth_pair.argtypes = [vp, C.c_int, C.c_int, _fp, _fp, _fp, C.c_float, C.c_float,
                                           C.c_float, C.c_float, _fp, _fp]
        L.sfp_unet_param_count.restype = C.c_longlong
        L.sfp_unet_param_count.argtypes = []
        L.sfp_set_network.restype = C.c_int
        L.sfp_set_network.argtypes = [vp, _fp, C.c_longlong]
        L.sfp_learned_features.restype = C.c_int
        L.sfp_learned_features.argtypes = [vp, C.c_int, _dp, _fp]
        L.sfp_unet_create.restype = C.c_int
        L.sfp_unet_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_int]
        L.sfp_unet_destroy.restype = None
        L.sfp_unet_destroy.argtypes = [vp]
        for name in ("sfp_unet_set_params", "sfp_unet_get_params"):
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = [vp, _fp, C.c_longlong]
        L.sfp_unet_init_params.restype = C.c_int
        L.sfp_unet_init_params.argtypes = [vp, C.c_ulonglong]
        L.sfp_unet_forward.restype = C.c_int
        L.sfp_unet_forward.argtypes = [vp, _fp, _fp]
        L.sfp_unet_gradients.restype = C.c_int
        L.sfp_unet_gradients.argtypes = [vp, _fp, _fp, _fp, _dp, _fp]
        L.sfp_unet_train_step_device.restype = C.c_int
        L.sfp_unet_train_step_device.argtypes = [vp, _fp, _fp, _fp, C.c_float, _dp]
        L.sfp_unet_profile_forward.restype = C.c_int
        L.sfp_unet_profile_forward.argtypes = [vp, C.c_int, _dp, _dp]
        L.sfp_unet_last_error.restype = C.c_char_p
        L.sfp_unet_last_error.argtypes = [vp]
        _LIB = L
    return _LIB


def _d(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.dtype == np.float64 and a.flags.c_contiguous, "need C-contiguous float64"
        return a.ctypes.data_as(_dp)
    return C.cast(C.c_void_p(int(a)), _dp)  # device pointer (int)


def _f(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.dtype == np.float32 and a.flags.c_contiguous, "need C-contiguous float32"
        return a.ctypes.data_as(_fp)
    return C.cast(C.c_void_p(int(a)), _fp)  # device pointer (int)


def _p3(arrs):
    if arrs is None:
        return None
    return P3(*[_d(a) for a in arrs])


def extrapolate_velocity_noalloc(n, vel, phi, fraction, steps=6, stream=None):
    """sfp_extrapolate_velocity_noalloc on device arrays (int pointers) of an
    n = (nx, ny, nz) grid, enqueued on `stream` (a hipStream_t as an int;
    None: the default stream); returns without waiting."""
    st = lib().sfp_extrapolate_velocity_noalloc(int(n[0]), int(n[1]), int(n[2]), _p3(vel), _d(phi),
                                                _p3(fraction), int(steps),
                                                None if stream is None else C.c_void_p(int(stream)))
    if st != OK:
        raise SfpError(st, lib().sfp_last_error(None).decode())


def default_params() -> Params:
    p = Params()
    lib().sfp_default_params(C.byref(p))
    return p


class Context:
    """A solver context bound to one grid and one GPU (sfp_create)."""

    def __init__(self, n, spacing, origin=(0.0, 0.0, 0.0), device=0, **params):
        if np.isscalar(n):
            n = (int(n),) * 3
        self.n = tuple(int(v) for v in n)
        self.spacing = float(spacing)
        self.origin = tuple(float(v) for v in origin)
        self._h = C.c_void_p()
        org = (C.c_double * 3)(*self.origin)
        st = lib().sfp_create(C.byref(self._h), *self.n, self.spacing,
                              C.cast(org, _dp), int(device))
        if st != OK:
            raise SfpError(st, lib().sfp_last_error(None).decode())
        if params:
            self.set_params(**params)

    @classmethod
    def for_scene(cls, sc, device=0, **params):
        g = sc["geo"]
        return cls(g.n, g.spacing, g.origin, device, **params)

    def close(self):
        if self._h:
            lib().sfp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        if st != OK:
            raise SfpError(st, lib().sfp_last_error(self._h).decode())

    # The library copies cells() doubles per cell array and (n + e_a) n^2 per
    # face array of axis a, whatever the caller's buffer holds: host arrays
    # are checked against the context's grid before any library call.
    def _face_shape(self, a):
        return tuple(v + (1 if i == a else 0) for i, v in enumerate(self.n))

    def _cells(self, a, what):
        if a is not None and tuple(np.shape(a)) != self.n:
            raise ValueError(f"{what}: shape {tuple(np.shape(a))}, the grid is {self.n}")
        return a

    def _faces(self, arrs, what):
        if arrs is None:
            return arrs
        if len(arrs) != 3:
            raise ValueError(f"{what}: need 3 face arrays, got {len(arrs)}")
        for a, v in enumerate(arrs):
            if tuple(np.shape(v)) != self._face_shape(a):
                raise ValueError(f"{what}[{a}]: shape {tuple(np.shape(v))}, face grid {a} is "
                                 f"{self._face_shape(a)}")
        return arrs

    def _vertices(self, a, what):
        if np.ndim(a) != 2 or np.shape(a)[1] != 3:
            raise ValueError(f"{what}: need an (nverts, 3) array, got shape {tuple(np.shape(a))}")
        return a

    @property
    def handle(self):
        return self._h

    def params(self) -> Params:
        p = Params()
        self._check(lib().sfp_get_params(self._h, C.byref(p)))
        return p

    def set_params(self, **kw):
        p = self.params()
        for k, v in kw.items():
            if not hasattr(p, k):
                raise KeyError(k)
            setattr(p, k, v)
        self._check(lib().sfp_set_params(self._h, C.byref(p)))

    def reset_warm_start(self):
        self._check(lib().sfp_reset_warm_start(self._h))

    # ------------------------------------------------------------ host API
    def pressure_project(self, sc, reproject=False, want_pressure=True):
        """Pressure::Project on copies of sc's velocities. Returns
        (velocity[3], pressure N^3 or None, stats dict)."""
        self._faces(sc["velocity"], "velocity")
        self._cells(sc["phi"], "phi")
        self._faces(sc["fraction"], "fraction")
        self._faces(sc["solid_velocity"], "solid_velocity")
        self._faces(sc.get("jump"), "jump")
        vel = [np.array(v, dtype=np.float64, copy=True, order="C") for v in sc["velocity"]]
        phi = np.ascontiguousarray(sc["phi"], dtype=np.float64)
        frac = [np.ascontiguousarray(a, dtype=np.float64) for a in sc["fraction"]]
        us = [np.ascontiguousarray(a, dtype=np.float64) for a in sc["solid_velocity"]]
        jump = sc.get("jump")
        jump = None if jump is None else [np.ascontiguousarray(a, dtype=np.float64) for a in jump]
        p = np.zeros(phi.shape) if want_pressure else None
        st = Stats()
        f = lib().sfp_pressure_reproject if reproject else lib().sfp_pressure_project
        self._check(f(self._h, _p3(vel), _d(phi), _p3(frac), _p3(us),
                      C.cast(_p3(jump), C.c_void_p) if jump is not None else None,
                      float(sc.get("vol_error", 0.0)), _d(p), C.byref(st)))
        return vel, p, st.as_dict()

    def magnetic_solve(self, sc, chi=None, band=None, want_field=True):
        """MagneticField::Solve. Returns (potential N^3, total_field[3] or None, stats)."""
        self._cells(sc["phi"], "phi")
        self._faces(sc["hext"], "hext")
        phi = np.ascontiguousarray(sc["phi"], dtype=np.float64)
        hext = [np.ascontiguousarray(a, dtype=np.float64) for a in sc["hext"]]
        pot = np.zeros(phi.shape)
        H = [np.zeros(self._face_shape(a)) for a in range(3)] if want_field else None
        st = Stats()
        self._check(lib().sfp_magnetic_solve(
            self._h, _d(phi), _p3(hext), float(sc["chi"] if chi is None else chi),
            float(sc["band"] if band is None else band), _d(pot),
            C.cast(_p3(H), C.c_void_p) if H is not None else None, C.byref(st)))
        return pot, H, st.as_dict()

    def vertex_pressures(self, positions, normals, chi):
        self._vertices(positions, "positions")
        self._vertices(normals, "normals")
        if np.shape(normals) != np.shape(positions):
            raise ValueError(f"normals: shape {np.shape(normals)}, positions {np.shape(positions)}")
        positions = np.ascontiguousarray(positions, dtype=np.float64)
        normals = np.ascontiguousarray(normals, dtype=np.float64)
        nv = positions.shape[0]
        out = np.zeros(max(nv, 1))
        self._check(lib().sfp_magnetic_vertex_pressures(self._h, nv, _d(positions), _d(normals),
                                                        float(chi), _d(out)))
        return out[:nv]

    # ------------------------------------------------------------ after the projection
    def extrapolate_faces(self, data, valid, clear_value=0.0, steps=6):
        """Extrapolation::Solve on the three face grids; returns new arrays."""
        self._faces(data, "data")
        self._faces(valid, "valid")
        out = [np.array(a, dtype=np.float64, copy=True, order="C") for a in data]
        masks = [np.ascontiguousarray(m, dtype=np.uint8) for m in valid]
        mp = (C.POINTER(C.c_ubyte) * 3)(*[m.ctypes.data_as(C.POINTER(C.c_ubyte)) for m in masks])
        self._check(lib().sfp_extrapolate_faces(self._h, _p3(out), mp, float(clear_value), int(steps)))
        return out

    def extrapolate_velocity(self, vel, phi, fraction, steps=6):
        """The velocity extrapolation of Simulation::ProjectVelocity; returns new arrays."""
        self._faces(vel, "velocity")
        self._cells(phi, "phi")
        self._faces(fraction, "fraction")
        out = [np.array(a, dtype=np.float64, copy=True, order="C") for a in vel]
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        frac = [np.ascontiguousarray(a, dtype=np.float64) for a in fraction]
        self._check(lib().sfp_extrapolate_velocity(self._h, _p3(out), _d(phi), _p3(frac), int(steps)))
        return out

    def collider_enforce(self, vel, fraction, normal, collider_velocity):
        """Collider::Enforce; returns new arrays."""
        self._faces(vel, "velocity")
        self._faces(fraction, "fraction")
        self._faces(normal, "normal")
        self._faces(collider_velocity, "collider_velocity")
        out = [np.array(a, dtype=np.float64, copy=True, order="C") for a in vel]
        arrs = [[np.ascontiguousarray(a, dtype=np.float64) for a in g]
                for g in (fraction, normal, collider_velocity)]
        self._check(lib().sfp_collider_enforce(self._h, _p3(out), *[_p3(g) for g in arrs]))
        return out

    def projection_substep(self, sc, normal, collider_velocity, steps=6, want_pressure=False):
        """sfp_projection_substep: the projection, the velocity extrapolation
        and Collider::Enforce with the velocity kept on the device; returns
        (velocity[3], pressure or None, stats) on copies of sc's arrays."""
        for name, arrs in (("velocity", sc["velocity"]), ("fraction", sc["fraction"]),
                           ("solid_velocity", sc["solid_velocity"]), ("jump", sc.get("jump")),
                           ("normal", normal), ("collider_velocity", collider_velocity)):
            self._faces(arrs, name)
        self._cells(sc["phi"], "phi")
        vel = [np.array(v, dtype=np.float64, copy=True, order="C") for v in sc["velocity"]]
        c64 = lambda arrs: [np.ascontiguousarray(a, dtype=np.float64) for a in arrs]  # noqa: E731
        frac, us, nrm, cv = c64(sc["fraction"]), c64(sc["solid_velocity"]), c64(normal), c64(collider_velocity)
        jump = sc.get("jump")
        jump = None if jump is None else c64(jump)
        phi = np.ascontiguousarray(sc["phi"], dtype=np.float64)
        p = np.zeros(phi.shape) if want_pressure else None
        st = Stats()
        self._check(lib().sfp_projection_substep(
            self._h, _p3(vel), _d(phi), _p3(frac), _p3(us),
            C.cast(_p3(jump), C.c_void_p) if jump is not None else None,
            float(sc.get("vol_error", 0.0)), int(steps), _p3(nrm), _p3(cv), _d(p), C.byref(st)))
        return vel, p, st.as_dict()

    def extrapolate_velocity_device(self, vel, phi, fraction, steps=6):
        self._check(lib().sfp_extrapolate_velocity_device(self._h, _p3(vel), _d(phi), _p3(fraction),
                                                          int(steps)))

    def collider_enforce_device(self, vel, fraction, normal, collider_velocity):
        self._check(lib().sfp_collider_enforce_device(self._h, _p3(vel), _p3(fraction), _p3(normal),
                                                      _p3(collider_velocity)))

    # ------------------------------------------------------------ device API (int pointers)
    def pressure_project_device(self, vel, phi, frac, us, jump, vol_error, pressure=None,
                                reproject=False):
        st = Stats()
        f = lib().sfp_pressure_reproject_device if reproject else lib().sfp_pressure_project_device
        self._check(f(self._h, _p3(vel), _d(phi), _p3(frac), _p3(us),
                      C.cast(_p3(jump), C.c_void_p) if jump is not None else None,
                      float(vol_error), _d(pressure), C.byref(st)))
        return st.as_dict()

    def magnetic_solve_device(self, phi, hext, chi, band, potential=None, total_field=None):
        st = Stats()
        self._check(lib().sfp_magnetic_solve_device(
            self._h, _d(phi), _p3(hext), float(chi), float(band), _d(potential),
            C.cast(_p3(total_field), C.c_void_p) if total_field is not None else None,
            C.byref(st)))
        return st.as_dict()

    # ------------------------------------------------------------ learned preconditioner
    def set_network(self, params):
        """Load U-Net weights (float32, unet_param_count() values); afterwards
        precond=PRECOND_LEARNED may be selected."""
        params = np.ascontiguousarray(params, dtype=np.float32)
        self._check(lib().sfp_set_network(self._h, _f(params), params.size))

    def learned_features(self, which, phi=None):
        """Network input channels 1..3 on the first coarse level of the last
        assembled system, shape level_dims(which, 1) + (3,)."""
        out = np.zeros(self.level_dims(which, 1) + (3,), np.float32)
        self._cells(phi, "phi")
        ph = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
        self._check(lib().sfp_learned_features(self._h, int(which), _d(ph), _f(out)))
        return out

    # ------------------------------------------------------------ profiling
    def profile(self, enable=True):
        self._check(lib().sfp_profile_enable(self._h, 1 if enable else 0))

    def kernel_profile(self, kernel):
        k = KernelProfile()
        self._check(lib().sfp_profile_query(self._h, int(kernel), C.byref(k)))
        return {"launches": k.launches, "total_ms": k.total_ms, "bytes": k.bytes}

    # ------------------------------------------------------------ test hooks
    def num_levels(self, which):
        v = C.c_int(0)
        self._check(lib().sfp_test_num_levels(self._h, which, C.byref(v)))
        return v.value

    def level_dims(self, which, level):
        d = (C.c_int * 3)()
        self._check(lib().sfp_test_level_dims(self._h, which, level, C.cast(d, _ip)))
        return tuple(d)

    def export_level(self, which, level):
        dims = self.level_dims(which, level)
        arrs = [np.zeros(dims) for _ in range(5)]
        self._check(lib().sfp_test_export_level(self._h, which, level, *[_d(a) for a in arrs]))
        diag, cx, cy, cz, rhs = arrs
        return dict(diag=diag, cpl=[cx, cy, cz], rhs=rhs if level == 0 else None)

    def export_level32(self, which, level):
        dims = self.level_dims(which, level)
        arrs = [np.zeros(dims, np.float32) for _ in range(4)]
        self._check(lib().sfp_test_export_level32(self._h, which, level, *[_f(a) for a in arrs]))
        return dict(diag=arrs[0], cpl=arrs[1:])

    def apply64(self, which, z):
        self._cells(z, "z")
        z = np.ascontiguousarray(z, dtype=np.float32)
        w = np.zeros(z.shape)
        self._check(lib().sfp_test_apply64(self._h, which, _f(z), _d(w)))
        return w

    def precondition(self, which, r):
        self._cells(r, "r")
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros(r.shape, np.float32)
        self._check(lib().sfp_test_precondition(self._h, which, _d(r), _f(z)))
        return z

    def residual(self, which, x):
        self._cells(x, "x")
        x = np.ascontiguousarray(x, dtype=np.float64)
        r = np.zeros(x.shape)
        rr = C.c_double(0)
        bb = C.c_double(0)
        self._check(lib().sfp_test_residual(self._h, which, _d(x), _d(r), C.byref(rr), C.byref(bb)))
        return r, rr.value, bb.value

    def solve_rhs(self, which, rhs, level=0):
        """x = A_level^-1 rhs (level 0: the assembled system; from x = 0);
        returns (x, stats)."""
        want = self.n if level == 0 else self.level_dims(which, level)
        if tuple(np.shape(rhs)) != tuple(want) and np.size(rhs) != int(np.prod(want)):
            raise ValueError(f"rhs: shape {tuple(np.shape(rhs))}, level {level} is {tuple(want)}")
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros(rhs.shape)
        st = Stats()
        self._check(lib().sfp_test_solve(self._h, int(which), int(level), _d(rhs), _d(x),
                                         C.byref(st)))
        return x, st.as_dict()

    def restrict_smoothed(self, which, r):
        """The residual one V-cycle leaves for right-hand side r, restricted
        to the first coarse level (FP32, coarse dims)."""
        self._cells(r, "r")
        r = np.ascontiguousarray(r, dtype=np.float64)
        rc = np.zeros(self.level_dims(which, 1), np.float32)
        self._check(lib().sfp_test_restrict_smoothed(self._h, int(which), _d(r), _f(rc)))
        return rc

    def smooth_pair(self, which, level, b, x1, d1, cd_a, cr_a, cd_b, cr_b):
        """The last two post-smoothing steps on a level, from the one-pass
        kernel and from two single steps: (x3_pair, x3_steps), FP32."""
        shape = self.level_dims(which, level)
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32).reshape(shape)  # noqa: E731
        b, x1 = f32(b), f32(x1)
        d1 = None if d1 is None else f32(d1)  # None: d1 is x1 (first pre-smoothing passes)
        x3p = np.zeros(shape, np.float32)
        x3s = np.zeros(shape, np.float32)
        self._check(lib().sfp_test_smooth_pair(self._h, int(which), int(level), _f(b), _f(x1), _f(d1),
                                               cd_a, cr_a, cd_b, cr_b, _f(x3p), _f(x3s)))
        return x3p, x3s

    def set_vcycle_tail(self, enable):
        """Small V-cycle levels in one launch (True, the default) or level by level."""
        self._check(lib().sfp_test_set_vcycle_tail(self._h, 1 if enable else 0))

    def set_fused_pairs(self, mask):
        """Step pairs as one pass: bit 0, the last pre-smoothing step with the
        restriction, and bit 1, the last two post-smoothing steps, on large
        grids; bits 2 and 3, the same on small grids; bit 4, the first two
        pre-smoothing passes of a degree-3 cycle on large grids. The default
        is 18 (bits 1 and 4); True is 31, False 0."""
        mask = 31 if mask is True else 0 if mask is False else int(mask)
        self._check(lib().sfp_test_set_fused_pairs(self._h, mask))

    def set_rowsum_diagonal(self, enable):
        """Let the stencil kernels of the magnetic system rebuild the FP64 and
        FP32 diagonals from the couplings (True, the default) or read them
        (False), from the next magnetic solve on; the results are the same bit
        for bit."""
        self._check(lib().sfp_test_set_rowsum_diagonal(self._h, 1 if enable else 0))

    def smooth_restrict(self, which, level, b, x1, d1, cd, cr):
        """The last pre-smoothing step and the restriction on a level, from
        the one-pass kernel and from the separate launches:
        (x2_fused, bc_fused, x2_steps, bc_steps), FP32."""
        shape = self.level_dims(which, level)
        cshape = self.level_dims(which, level + 1)
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32).reshape(shape)  # noqa: E731
        b, x1, d1 = f32(b), f32(x1), f32(d1)
        out = [np.zeros(shape, np.float32), np.zeros(cshape, np.float32),
               np.zeros(shape, np.float32), np.zeros(cshape, np.float32)]
        self._check(lib().sfp_test_smooth_restrict(self._h, int(which), int(level), _f(b), _f(x1),
                                                   _f(d1), cd, cr, *[_f(a) for a in out]))
        return tuple(out)


def unet_param_count() -> int:
    return int(lib().sfp_unet_param_count())


class UNet:
    """The stand-alone U-Net (sfp_unet_*): forward, gradients and Adam steps on
    the GPU for volumes of shape (batch, nx, ny, nz), each axis a multiple of 4."""

    def __init__(self, batch, nx, ny, nz, training=True, device=0):
        self.shape = (int(batch), int(nx), int(ny), int(nz))
        self._h = C.c_void_p()
        st = lib().sfp_unet_create(C.byref(self._h), int(device), *self.shape, 1 if training else 0)
        if st != OK:
            raise SfpError(st, lib().sfp_unet_last_error(None).decode())

    def close(self):
        if self._h:
            lib().sfp_unet_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        if st != OK:
            raise SfpError(st, lib().sfp_unet_last_error(self._h).decode())

    def set_params(self, params):
        params = np.ascontiguousarray(params, dtype=np.float32)
        self._check(lib().sfp_unet_set_params(self._h, _f(params), params.size))

    def get_params(self):
        p = np.zeros(unet_param_count(), np.float32)
        self._check(lib().sfp_unet_get_params(self._h, _f(p), p.size))
        return p

    def init_params(self, seed):
        self._check(lib().sfp_unet_init_params(self._h, int(seed)))

    def forward(self, inp):
        inp = np.ascontiguousarray(inp, dtype=np.float32)
        assert inp.shape == self.shape + (4,)
        out = np.zeros(self.shape, np.float32)
        self._check(lib().sfp_unet_forward(self._h, _f(inp), _f(out)))
        return out

    def gradients(self, inp, target, mask=None):
        inp = np.ascontiguousarray(inp, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
        g = np.zeros(unet_param_count(), np.float32)
        loss = C.c_double(0)
        self._check(lib().sfp_unet_gradients(self._h, _f(inp), _f(target), _f(mask), C.byref(loss),
                                             _f(g)))
        return loss.value, g

    def profile_forward(self, reps=10):
        """(ms per forward pass, [ms per convolution] * 10) on the network's
        device-resident input, timed with HIP events."""
        fwd = C.c_double(0)
        conv = np.zeros(10)
        self._check(lib().sfp_unet_profile_forward(self._h, int(reps), C.byref(fwd), _d(conv)))
        return fwd.value, conv

    def train_step_device(self, inp, target, mask, lr):
        """One Adam step on device arrays (int pointers); returns the loss before it."""
        loss = C.c_double(0)
        self._check(lib().sfp_unet_train_step_device(self._h, _f(inp), _f(target),
                                                     None if mask is None else _f(mask),
                                                     float(lr), C.byref(loss)))
        return loss.value
