"""Shared helpers of the parity tests: mapping the oracle's compact CSR systems
onto the full-grid 7-point representation the GPU library stores."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def full_grid_operator(A, rhs, mat2grid, shape):
    """(diag, [cpl_x, cpl_y, cpl_z], rhs) on the full grid from a CSR system
    whose row r is cell mat2grid[r]. Non-unknown cells get identity rows;
    cpl_a[c] = -A[c, c - e_a]."""
    n = int(np.prod(shape))
    nx, ny, nz = shape
    strides = (ny * nz, nz, 1)
    diag = np.ones(n)
    cpl = [np.zeros(n) for _ in range(3)]
    b = np.zeros(n)
    A = A.tocoo()
    rows = mat2grid[A.row]
    cols = mat2grid[A.col]
    on = rows == cols
    diag[rows[on]] = A.data[on]
    b[mat2grid] = rhs
    diff = rows - cols
    for a in range(3):
        m = diff == strides[a]
        cpl[a][rows[m]] = -A.data[m]
    return (diag.reshape(shape), [c.reshape(shape) for c in cpl], b.reshape(shape))


def csr_from_full_grid(diag, cpl, active=None):
    """Sparse matrix of a 7-point operator given on the full grid."""
    shape = diag.shape
    n = diag.size
    nx, ny, nz = shape
    strides = (ny * nz, nz, 1)
    idx = np.arange(n).reshape(shape)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [diag.ravel()]
    for a in range(3):
        sl = [slice(None)] * 3
        sl[a] = slice(1, None)
        c = idx[tuple(sl)].ravel()
        v = cpl[a][tuple(sl)].ravel()
        nz_ = v != 0
        c, v = c[nz_], v[nz_]
        rows += [c, c - strides[a]]
        cols += [c - strides[a], c]
        vals += [-v, -v]
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))


def aggregation_prolongation(shape, active):
    """Piecewise-constant prolongation from the 2x2x2 aggregates onto the
    active fine cells."""
    nx, ny, nz = shape
    cs = ((nx + 1) // 2, (ny + 1) // 2, (nz + 1) // 2)
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    parent = ((I // 2) * cs[1] + J // 2) * cs[2] + K // 2
    fine = np.arange(nx * ny * nz)
    m = active.ravel()
    P = sp.csr_matrix((np.ones(m.sum()), (fine[m], parent.ravel()[m])),
                      shape=(nx * ny * nz, int(np.prod(cs))))
    return P, cs


def ulp_diff(a, b):
    """Largest distance in units of the last place between two float64 arrays."""
    a = np.ascontiguousarray(a, dtype=np.float64).view(np.int64)
    b = np.ascontiguousarray(b, dtype=np.float64).view(np.int64)
    a = np.where(a < 0, np.int64(-0x8000000000000000) - a, a)
    b = np.where(b < 0, np.int64(-0x8000000000000000) - b, b)
    return int(np.max(np.abs(a - b))) if a.size else 0
