"""ctypes front end of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, as the checker. The shipped library never uses it.

Each function mirrors one reference routine (see oracle.h for file:line) and
takes the scene dictionaries produced by ``sfp.scenes``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def lib():
    global _LIB
    if _LIB is None:
        # SFP_ORACLE_LIB: another build of the same sources (the sanitizer
        # build, oracle/_san/liboracle.so)
        path = os.environ.get("SFP_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        P3 = _dp * 3
        L.orc_pressure_unknowns.restype = C.c_int
        L.orc_pressure_unknowns.argtypes = [C.c_int] * 3 + [_dp, _ip, _ip]
        L.orc_pressure_matrix.restype = C.c_int64
        L.orc_pressure_matrix.argtypes = [C.c_int] * 3 + [_dp, P3, P3, P3, C.c_void_p, C.c_double,
                                                           _ip, _ip, C.c_int, _ip, _ip, _dp, _dp]
        L.orc_pressure_rhs.restype = None
        L.orc_pressure_rhs.argtypes = [C.c_int] * 3 + [P3, P3, P3, C.c_double, _ip, C.c_int, _dp]
        L.orc_pressure_apply.restype = C.c_int
        L.orc_pressure_apply.argtypes = [C.c_int] * 3 + [_dp, P3, C.c_void_p, _ip, _dp, P3]
        L.orc_magnetic_coeff.restype = None
        L.orc_magnetic_coeff.argtypes = [C.c_int] * 3 + [C.c_double, _dp, _dp, C.c_double,
                                                          C.c_double, P3]
        L.orc_magnetic_matrix.restype = C.c_int64
        L.orc_magnetic_matrix.argtypes = [C.c_int] * 3 + [P3, C.c_int, _ip, _ip, _dp]
        L.orc_magnetic_rhs.restype = None
        L.orc_magnetic_rhs.argtypes = [C.c_int] * 3 + [C.c_double, P3, P3, C.c_int, _dp]
        L.orc_magnetic_residual.restype = C.c_double
        L.orc_magnetic_residual.argtypes = [C.c_int] * 3 + [P3, C.c_int, _dp, _dp, _dp]
        L.orc_magnetic_total_field.restype = None
        L.orc_magnetic_total_field.argtypes = [C.c_int] * 3 + [C.c_double, _dp, P3, P3]
        L.orc_magnetic_vertex_pressures.restype = None
        L.orc_magnetic_vertex_pressures.argtypes = [C.c_int] * 3 + [C.c_double, _dp, P3, C.c_double,
                                                                     C.c_int, _dp, _dp, _dp]
        L.orc_spmv.restype = None
        L.orc_spmv.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp]
        L.orc_rel_residual.restype = C.c_double
        L.orc_rel_residual.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp]
        for name in ("orc_cg", "orc_bicgstab"):
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = [C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_double, C.c_int, C.c_int,
                          _dp, _dp, _dp]
        L.orc_unet_param_count.restype = C.c_longlong
        L.orc_unet_param_count.argtypes = []
        for name in ("orc_unet_forward", "orc_unet_gradients"):
            f = getattr(L, name)
            f.restype = C.c_double
            f.argtypes = [C.c_int] * 4 + [_dp, _dp, C.c_void_p, C.c_void_p, _dp]
        U3 = C.POINTER(C.c_ubyte) * 3
        L.orc_velocity_valid.restype = None
        L.orc_velocity_valid.argtypes = [C.c_int] * 3 + [_dp, P3, U3]
        L.orc_extrapolate_faces.restype = None
        L.orc_extrapolate_faces.argtypes = [C.c_int] * 3 + [P3, U3, C.c_double, C.c_int]
        L.orc_collider_enforce.restype = None
        L.orc_collider_enforce.argtypes = [C.c_int] * 3 + [C.c_double, _dp, P3, P3, P3, P3]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_mutation.restype = None
        L.orc_set_mutation.argtypes = [C.c_int]
        _LIB = L
    return _LIB


# Planted mistakes (orc_set_mutation in assembly.cpp): each changes one
# expression of the restatement so that a closed-form test must fail.
MUTATIONS = {"heaviside_no_sin": 1, "heaviside_half_band": 2, "jump_sign": 3,
             "jump_no_coef": 4, "no_clamp": 5, "trilerp_swap": 6, "face_origin": 7,
             "box_face_weight": 8}


class mutated:
    """``with oracle.mutated("jump_sign"): ...`` runs the block on an oracle
    with that planted mistake (None: the faithful one)."""

    def __init__(self, name):
        self.code = 0 if name is None else MUTATIONS[name]

    def __enter__(self):
        lib().orc_set_mutation(self.code)
        return self

    def __exit__(self, *exc):
        lib().orc_set_mutation(0)


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def _p3(arrs):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in arrs]
    return (_dp * 3)(*[_d(a) for a in arrs]), arrs


def num_threads() -> int:
    return lib().orc_num_threads()


# ---------------------------------------------------------------- pressure
def pressure_unknowns(phi: np.ndarray):
    nx, ny, nz = phi.shape
    g2m = np.empty(phi.size, np.int32)
    m2g = np.empty(phi.size, np.int32)
    n = lib().orc_pressure_unknowns(nx, ny, nz, _d(np.ascontiguousarray(phi)), _i(g2m), _i(m2g))
    return g2m.reshape(phi.shape), m2g[:n].copy()


def pressure_system(sc):
    """Pressure::SetUnKnowns + BuildProjectionMatrix. Returns dict with
    grid2mat, mat2grid, A (scipy CSR, n x n), rhs (n)."""
    phi = np.ascontiguousarray(sc["phi"], dtype=np.float64)
    nx, ny, nz = phi.shape
    g2m, m2g = pressure_unknowns(phi)
    n = m2g.size
    rowptr = np.zeros(n + 1, np.int32)
    col = np.zeros(7 * max(n, 1), np.int32)
    val = np.zeros(7 * max(n, 1), np.float64)
    rhs = np.zeros(max(n, 1), np.float64)
    fr, k1 = _p3(sc["fraction"])
    ve, k2 = _p3(sc["velocity"])
    us, k3 = _p3(sc["solid_velocity"])
    jp = None
    if sc.get("jump") is not None:
        jp, k4 = _p3(sc["jump"])
    g2m_flat = np.ascontiguousarray(g2m.ravel())
    m2g_c = np.ascontiguousarray(m2g) if n else np.zeros(1, np.int32)
    nnz = lib().orc_pressure_matrix(nx, ny, nz, _d(phi), fr, ve, us,
                                    C.cast(jp, C.c_void_p) if jp is not None else None,
                                    float(sc.get("vol_error", 0.0)), _i(g2m_flat), _i(m2g_c), n,
                                    _i(rowptr), _i(col), _d(val), _d(rhs))
    if nnz < 0:
        raise ValueError("open face on the grid boundary next to a liquid cell")
    A = sp.csr_matrix((val[:nnz], col[:nnz], rowptr), shape=(n, n))
    return dict(grid2mat=g2m, mat2grid=m2g, A=A, rhs=rhs[:n].copy(), n=n)


def pressure_rhs(sc, mat2grid):
    phi = sc["phi"]
    nx, ny, nz = phi.shape
    n = mat2grid.size
    rhs = np.zeros(max(n, 1))
    fr, k1 = _p3(sc["fraction"])
    ve, k2 = _p3(sc["velocity"])
    us, k3 = _p3(sc["solid_velocity"])
    lib().orc_pressure_rhs(nx, ny, nz, fr, ve, us, float(sc.get("vol_error", 0.0)),
                           _i(np.ascontiguousarray(mat2grid, dtype=np.int32)), n, _d(rhs))
    return rhs[:n]


def pressure_apply(sc, grid2mat, p):
    """Pressure::ApplyProjection on copies of sc['velocity']; returns new velocities."""
    phi = np.ascontiguousarray(sc["phi"], dtype=np.float64)
    nx, ny, nz = phi.shape
    vel = [np.array(v, dtype=np.float64, copy=True) for v in sc["velocity"]]
    vp = (_dp * 3)(*[_d(v) for v in vel])
    fr, k1 = _p3(sc["fraction"])
    jp = None
    if sc.get("jump") is not None:
        jp, k4 = _p3(sc["jump"])
    p = np.ascontiguousarray(p, dtype=np.float64)
    if p.size == 0:
        p = np.zeros(1)
    rc = lib().orc_pressure_apply(nx, ny, nz, _d(phi), fr,
                                  C.cast(jp, C.c_void_p) if jp is not None else None,
                                  _i(np.ascontiguousarray(grid2mat.ravel(), dtype=np.int32)),
                                  _d(p), vp)
    if rc != 0:
        raise ValueError("open face on the grid boundary")
    return vel


# ---------------------------------------------------------------- magnetic
def magnetic_coeff(sc):
    geo = sc["geo"]
    phi = np.ascontiguousarray(sc["phi"], dtype=np.float64)
    nx, ny, nz = phi.shape
    mu = [np.empty(geo.face_shape(a)) for a in range(3)]
    mp = (_dp * 3)(*[_d(m) for m in mu])
    org = np.array(geo.origin, dtype=np.float64)
    lib().orc_magnetic_coeff(nx, ny, nz, geo.spacing, _d(org), _d(phi), float(sc["chi"]),
                             float(sc["band"]), mp)
    return mu


def magnetic_system(sc, mu=None, bc=0):
    """MagneticField::BuildMatrix + AssembleRHS. Returns dict with mu, A, rhs.
    bc = 1: the library's opt-in Dirichlet box instead of the reference's
    boundary (sfp_magnetic_bc)."""
    geo = sc["geo"]
    nx, ny, nz = geo.n
    if mu is None:
        mu = magnetic_coeff(sc)
    n = nx * ny * nz
    rowptr = np.zeros(n + 1, np.int32)
    col = np.zeros(7 * n, np.int32)
    val = np.zeros(7 * n, np.float64)
    mp, k1 = _p3(mu)
    nnz = lib().orc_magnetic_matrix(nx, ny, nz, mp, int(bc), _i(rowptr), _i(col), _d(val))
    A = sp.csr_matrix((val[:nnz], col[:nnz], rowptr), shape=(n, n))
    rhs = np.zeros(n)
    hp, k2 = _p3(sc["hext"])
    lib().orc_magnetic_rhs(nx, ny, nz, geo.spacing, mp, hp, int(bc), _d(rhs))
    return dict(mu=mu, A=A, rhs=rhs, n=n)


def magnetic_rhs(sc, mu, bc=0):
    """MagneticField::AssembleRHS alone (N^3 doubles), for grids whose CSR
    matrix would not fit in memory."""
    geo = sc["geo"]
    nx, ny, nz = geo.n
    rhs = np.zeros(geo.ncells)
    mp, k1 = _p3(mu)
    hp, k2 = _p3(sc["hext"])
    lib().orc_magnetic_rhs(nx, ny, nz, geo.spacing, mp, hp, int(bc), _d(rhs))
    return rhs


def magnetic_rel_residual(sc, mu, rhs, x, bc=0):
    """||rhs - A x||_2 / ||rhs||_2 of the magnetic system, row by row without
    the CSR (any grid size)."""
    geo = sc["geo"]
    nx, ny, nz = geo.n
    mp, k1 = _p3(mu)
    bb = C.c_double(0)
    rr = lib().orc_magnetic_residual(nx, ny, nz, mp, int(bc), _d(np.ascontiguousarray(rhs)),
                                     _d(np.ascontiguousarray(np.asarray(x, np.float64).ravel())),
                                     C.byref(bb))
    return float(np.sqrt(rr / bb.value)) if bb.value > 0 else float(np.sqrt(rr))


def magnetic_total_field(sc, potential):
    geo = sc["geo"]
    nx, ny, nz = geo.n
    H = [np.empty(geo.face_shape(a)) for a in range(3)]
    hp_out = (_dp * 3)(*[_d(h) for h in H])
    hp, k = _p3(sc["hext"])
    lib().orc_magnetic_total_field(nx, ny, nz, geo.spacing,
                                   _d(np.ascontiguousarray(potential.ravel(), dtype=np.float64)),
                                   hp, hp_out)
    return H


def magnetic_vertex_pressures(sc, htotal, positions, normals):
    geo = sc["geo"]
    nx, ny, nz = geo.n
    positions = np.ascontiguousarray(positions, dtype=np.float64)
    normals = np.ascontiguousarray(normals, dtype=np.float64)
    nv = positions.shape[0]
    out = np.empty(max(nv, 1))
    hp, k = _p3(htotal)
    org = np.array(geo.origin, dtype=np.float64)
    lib().orc_magnetic_vertex_pressures(nx, ny, nz, geo.spacing, _d(org), hp, float(sc["chi"]),
                                        nv, _d(positions), _d(normals), _d(out))
    return out[:nv]


# ---------------------------------------------------------------- after the projection
def _u3(arrs):
    arrs = [np.ascontiguousarray(a, dtype=np.uint8) for a in arrs]
    return (C.POINTER(C.c_ubyte) * 3)(*[a.ctypes.data_as(C.POINTER(C.c_ubyte)) for a in arrs]), arrs


def velocity_valid(sc):
    """Extrapolation mask of Simulation::ProjectVelocity, per face grid (uint8)."""
    geo = sc["geo"]
    nx, ny, nz = geo.n
    out = [np.zeros(geo.face_shape(a), np.uint8) for a in range(3)]
    fp, k1 = _p3(sc["fraction"])
    up, k2 = _u3(out)
    lib().orc_velocity_valid(nx, ny, nz, _d(np.ascontiguousarray(sc["phi"], dtype=np.float64)), fp, up)
    return k2


def extrapolate_faces(geo, data, valid, clear=0.0, steps=6):
    """Extrapolation::Solve on the three face grids; returns new arrays."""
    nx, ny, nz = geo.n
    dp, arrs = _p3([np.array(a, dtype=np.float64, copy=True) for a in data])
    vp, k = _u3(valid)
    lib().orc_extrapolate_faces(nx, ny, nz, dp, vp, float(clear), int(steps))
    return arrs


def collider_enforce(geo, vel, frac, normal, colvel):
    """Collider::Enforce; returns new velocity arrays."""
    nx, ny, nz = geo.n
    vp, arrs = _p3([np.array(a, dtype=np.float64, copy=True) for a in vel])
    fp, k1 = _p3(frac)
    npp, k2 = _p3(normal)
    cp, k3 = _p3(colvel)
    org = np.array(geo.origin, dtype=np.float64)
    lib().orc_collider_enforce(nx, ny, nz, geo.spacing, _d(org), vp, fp, npp, cp)
    return arrs


# ---------------------------------------------------------------- solvers
def _csr_parts(A):
    A = A.tocsr()
    return (np.ascontiguousarray(A.indptr, dtype=np.int32),
            np.ascontiguousarray(A.indices, dtype=np.int32),
            np.ascontiguousarray(A.data, dtype=np.float64))


def spmv(A, x):
    ptr, col, val = _csr_parts(A)
    y = np.empty(A.shape[0])
    lib().orc_spmv(A.shape[0], _i(ptr), _i(col), _d(val), _d(np.ascontiguousarray(x, np.float64)),
                   _d(y))
    return y


def rel_residual(A, b, x):
    ptr, col, val = _csr_parts(A)
    return lib().orc_rel_residual(A.shape[0], _i(ptr), _i(col), _d(val),
                                  _d(np.ascontiguousarray(b, np.float64)),
                                  _d(np.ascontiguousarray(x, np.float64)))


PRECOND = {"none": 0, "jacobi": 1, "amg": 2}


def solve(A, b, method="cg", precond="amg", tol=1e-12, maxit=1000, x0=None):
    """Returns (x, iterations, rel_residual, setup_s, solve_s)."""
    ptr, col, val = _csr_parts(A)
    n = A.shape[0]
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    b = np.ascontiguousarray(b, dtype=np.float64)
    rel = C.c_double(0)
    ts = C.c_double(0)
    tv = C.c_double(0)
    f = lib().orc_cg if method == "cg" else lib().orc_bicgstab
    if n == 0:
        return x, 0, 0.0, 0.0, 0.0
    it = f(n, _i(ptr), _i(col), _d(val), _d(b), _d(x), tol, maxit, PRECOND[precond],
           C.byref(rel), C.byref(ts), C.byref(tv))
    return x, it, rel.value, ts.value, tv.value


# ---------------------------------------------------------------- U-Net
def unet_param_count():
    return int(lib().orc_unet_param_count())


def _opt(a):
    return None if a is None else C.cast(_d(np.ascontiguousarray(a, dtype=np.float64)), C.c_void_p)


def unet_forward(params, inp, target=None, mask=None):
    """inp: [B, X, Y, Z, 4]; returns (output [B, X, Y, Z], loss)."""
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    B, X, Y, Z, _ = inp.shape
    out = np.zeros((B, X, Y, Z))
    t = None if target is None else np.ascontiguousarray(target, dtype=np.float64)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float64)
    loss = lib().orc_unet_forward(B, X, Y, Z, _d(np.ascontiguousarray(params, np.float64)), _d(inp),
                                  _opt(t), _opt(m), _d(out))
    return out, loss


def unet_gradients(params, inp, target, mask=None):
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    B, X, Y, Z, _ = inp.shape
    g = np.zeros(unet_param_count())
    t = np.ascontiguousarray(target, dtype=np.float64)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float64)
    loss = lib().orc_unet_gradients(B, X, Y, Z, _d(np.ascontiguousarray(params, np.float64)), _d(inp),
                                    C.cast(_d(t), C.c_void_p), _opt(m), _d(g))
    return loss, g
