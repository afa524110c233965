"""Komatitsch–Martin CPML coefficient vectors (src/fdtd/cpml.jl:8-155).

Each derivative field gets (a, b, kI) vectors of length 2*npml on its own staggered 1-D
grid along its derivative axis: the first npml entries for the min face, the last npml
for the max face.  The profile is evaluated in float64 and stored as float32, as in the
reference's ``copyto!`` into ``Data.Array``.
"""
from __future__ import annotations

import math

import numpy as np

from .fields import derivative_fields, deriv_axis, get_mgrid
from .grid import Grid


def pml_profile(exgrid_f: Grid, mgrid: Grid, fmin: bool, fmax: bool, dt32, velavg32,
                freqpeak32, npml: int):
    """``update_pml!`` for one derivative field (src/fdtd/cpml.jl:8-103)."""
    nx = exgrid_f.n
    x = exgrid_f.values()
    ds = mgrid.step
    xl = mgrid.first - 2.0 * ds
    xr = mgrid.last + 2.0 * ds
    NPOWER = 2.0
    K_MAX = 1.0
    alpha_max = math.pi * float(freqpeak32)
    thick = (npml - 1) * ds
    R = 0.001
    d0 = -(NPOWER + 1.0) * float(velavg32) * math.log(R) / (2.0 * thick)
    k = np.ones(nx)
    d = np.zeros(nx)
    alpha = np.zeros(nx)
    if fmin:
        ab = xl - x
        sel = ab >= 0.0
        an = ab[sel] / thick
        d[sel] = d0 * an ** NPOWER
        k[sel] = 1.0 + (K_MAX - 1.0) * an ** NPOWER
        alpha[sel] = alpha_max * (1.0 - an) + 0.01 * alpha_max
    if fmax:
        ab = x - xr
        sel = ab >= 0.0
        an = ab[sel] / thick
        d[sel] = d0 * an ** NPOWER
        k[sel] = 1.0 + (K_MAX - 1.0) * an ** NPOWER
        alpha[sel] = alpha_max * (1.0 - an) + 0.01 * alpha_max
    alpha[alpha < 0.0] = 0.0
    b = np.exp(-(d / k + alpha) * float(dt32))
    a = np.zeros(nx)
    sel = np.abs(d) > 1e-6
    a[sel] = d[sel] * (b[sel] - 1.0) / (k[sel] * (d[sel] + k[sel] * alpha[sel]))
    sl = np.r_[0:npml, nx - npml:nx]
    return (a[sl].astype(np.float32), b[sl].astype(np.float32), (1.0 / k[sl]).astype(np.float32))


def pml_vectors(physics: int, exgrid: list, mgrid: list, pml_faces, order: int, npml: int,
                dt32, velavg32, freqpeak32) -> dict:
    """CPML (a, b, kI) for every derivative field of the physics (src/fdtd/cpml.jl:125-155)."""
    ndim = len(exgrid)
    names = ["z", "x"] if ndim == 2 else ["z", "y", "x"]
    out = {}
    for df in derivative_fields(physics, ndim):
        ax = deriv_axis(df)
        i = names.index("zyx"[ax])
        gf = get_mgrid(df, physics, exgrid, order)[i]
        dim = names[i]
        out[df] = pml_profile(gf, mgrid[i], f"{dim}min" in pml_faces, f"{dim}max" in pml_faces,
                              dt32, velavg32, freqpeak32, npml)
    return out
