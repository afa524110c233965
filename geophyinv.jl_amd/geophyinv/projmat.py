"""Source-spray and receiver-interpolation weights (src/proj_mat.jl:69-247).

For every point the reference builds a sparse column of bilinear (2-D) or trilinear (3-D)
weights on the staggered grid of the field it injects into or samples
(src/fdtd/ageom.jl:19-54).  Here each column is returned as ``nnz`` (index, weight) pairs
in ascending index order, which is the order of the CSC column ``sparsevec`` builds.
"""
from __future__ import annotations

import numpy as np

from .grid import Grid

COMPAT_SWAP3D = 1


def neighbours(g: Grid, v: float):
    """``get_neighbour_indices`` (src/proj_mat.jl:69-79) with ``searchsortedfirst``, 0-based."""
    vals = g.values()
    idx = int(np.searchsorted(vals, v, side="left"))  # first index with vals[idx] >= v
    n = g.n
    if idx == 0:
        return 0, 1
    if idx >= n - 1:
        return n - 2, n - 1
    return idx - 1, idx


def _fraction(g: Grid, lo: int, hi: int, v: float) -> float:
    if v < g.first:
        return 0.0
    if v > g.last:
        return 1.0
    x1, x2 = g.values()[[lo, hi]]
    return (v - x1) / (x2 - x1)


def point_weights(mgrid_f: list, pos, compat: int = 0):
    """Weights of one point ``pos`` = (z, [y,] x) on a field grid ``mgrid_f``."""
    nd = len(mgrid_f)
    if nd == 2:
        gz, gx = mgrid_f
        pz, px = float(pos[0]), float(pos[1])
        z1, z2 = neighbours(gz, pz)
        x1, x2 = neighbours(gx, px)
        dz = _fraction(gz, z1, z2, pz)
        dx = _fraction(gx, x1, x2, px)
        nz = gz.n
        # sparsevec order by linear index: (z1,x1), (z2,x1), (z1,x2), (z2,x2)
        idx = [z1 + nz * x1, z2 + nz * x1, z1 + nz * x2, z2 + nz * x2]
        w = [(1 - dz) * (1 - dx), dz * (1 - dx), (1 - dz) * dx, dz * dx]
        return np.array(idx, np.int64), np.array(w, np.float64).astype(np.float32)
    gz, gy, gx = mgrid_f
    pz, py, px = (float(p) for p in pos)
    if compat & COMPAT_SWAP3D:
        # the reference calls bilinear_interp(mz,my,mx, Pz,Py,Px) against the signature
        # (x,y,z, zi,yi,xi), so z is searched with the x coordinate and vice versa
        pz, px = px, pz
    z = neighbours(gz, pz)
    y = neighbours(gy, py)
    x = neighbours(gx, px)
    fz = _fraction(gz, z[0], z[1], pz)
    fy = _fraction(gy, y[0], y[1], py)
    fx = _fraction(gx, x[0], x[1], px)
    nz, ny = gz.n, gy.n
    idx, w = [], []
    for bx in (0, 1):
        for by in (0, 1):
            for bz in (0, 1):
                idx.append(z[bz] + nz * (y[by] + ny * x[bx]))
                wz = fz if bz else 1 - fz
                wy = fy if by else 1 - fy
                wx = fx if bx else 1 - fx
                w.append((wz * wy) * wx)
    return np.array(idx, np.int64), np.array(w, np.float64).astype(np.float32)


def _axis(g: Grid, v: np.ndarray):
    """Vectorised ``neighbours`` and ``_fraction`` for many coordinates on one axis."""
    vals = g.values()
    n = g.n
    idx = np.searchsorted(vals, v, side="left")
    lo = np.where(idx == 0, 0, np.where(idx >= n - 1, n - 2, idx - 1))
    hi = lo + 1
    x1, x2 = vals[lo], vals[hi]
    f = np.where(v < g.first, 0.0, np.where(v > g.last, 1.0, (v - x1) / (x2 - x1)))
    return lo, hi, f


def get_proj_matrix(points, mgrid_f: list, compat: int = 0):
    """All columns for a list of points: ``idx`` (npts, nnz) int64 and ``w`` (npts, nnz) f32.

    The same arithmetic as ``point_weights`` (FP64 products rounded once to FP32), for all
    points at once."""
    pts = np.asarray(points, np.float64).reshape(len(points), len(mgrid_f))
    if len(mgrid_f) == 2:
        gz, gx = mgrid_f
        z1, z2, dz = _axis(gz, pts[:, 0])
        x1, x2, dx = _axis(gx, pts[:, 1])
        nz = gz.n
        idx = np.stack([z1 + nz * x1, z2 + nz * x1, z1 + nz * x2, z2 + nz * x2], axis=1)
        w = np.stack([(1 - dz) * (1 - dx), dz * (1 - dx), (1 - dz) * dx, dz * dx], axis=1)
        return idx.astype(np.int64), w.astype(np.float32)
    gz, gy, gx = mgrid_f
    pz, py, px = pts[:, 0], pts[:, 1], pts[:, 2]
    if compat & COMPAT_SWAP3D:
        pz, px = px, pz
    z, y, x = _axis(gz, pz), _axis(gy, py), _axis(gx, px)
    nz, ny = gz.n, gy.n
    idx, w = [], []
    for bx in (0, 1):
        for by in (0, 1):
            for bz in (0, 1):
                idx.append(z[bz] + nz * (y[by] + ny * x[bx]))
                wz = z[2] if bz else 1 - z[2]
                wy = y[2] if by else 1 - y[2]
                wx = x[2] if bx else 1 - x[2]
                w.append((wz * wy) * wx)
    return np.stack(idx, axis=1).astype(np.int64), np.stack(w, axis=1).astype(np.float32)
