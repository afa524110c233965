"""Source wavelets and frequency scanning.

``ricker`` restates src/srcwav/wavelets.jl:14-72 and ``findfreq`` restates
src/Utils/freq.jl:28-57 (the ``:peak``, ``:min`` and ``:max`` attributes used to set the
CPML alpha profile and the stability check).
"""
from __future__ import annotations

import math

import numpy as np

from .grid import Grid


def ricker(fqdom: float, tgrid: Grid, tpeak: float | None = None, maxamp: float = 1.0) -> np.ndarray:
    """Ricker wavelet ``(1 - 2 pf tau^2) exp(-pf tau^2) * maxamp`` with pf = (pi f)^2, float64."""
    if tpeak is None:
        tpeak = tgrid.first + 1.5 / fqdom
    if tpeak < tgrid.first + 1.5 / fqdom or tpeak > tgrid.last - 1.5 / fqdom:
        raise ValueError("cannot output Ricker for given tgrid and tpeak")
    if math.isclose(fqdom, 0.0, abs_tol=1e-300):
        raise ValueError("dominant frequency cannot be zero")
    pf = (math.pi * math.pi) * (fqdom ** 2.0)
    t = tgrid.values()
    tsq = (t - tpeak) * (t - tpeak)
    return (1.0 - 2.0 * pf * tsq) * np.exp(-1.0e0 * pf * tsq) * maxamp


def findfreq(x: np.ndarray, tgrid: Grid, attrib: str = "peak", threshold: float = -50.0) -> float:
    """Peak / min / max frequency of the stacked power spectrum of ``x`` (nt x ns)."""
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    nt = x.shape[0]
    cx = np.fft.rfft(x, axis=0)
    # FFTW.rfftfreq(n, fs) = (0:n÷2) * (fs/n)
    fgrid = np.arange(nt // 2 + 1) * ((1.0 / tgrid.step) / nt)
    ax = np.abs(cx) ** 2
    mx = ax.max()
    if mx == 0.0:
        return 0.0
    ax = 10.0 * np.log10(ax / mx)
    flat = ax.ravel(order="F")   # column-major linear order of Julia arrays
    if attrib == "max":
        ii = np.nonzero(flat >= threshold)[0][-1]
    elif attrib == "min":
        ii = np.nonzero(flat >= threshold)[0][0]
    elif attrib == "peak":
        ii = int(np.argmax(flat))
    else:
        raise ValueError(attrib)
    return float(fgrid[ii % ax.shape[0]])
