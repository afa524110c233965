"""Source-wavelet and record containers, acquisition geometry and the misfit.

``Srcs``/``Recs`` restate src/database/database.jl:18-55 (``n``, ``grid``, ``fields`` and
``d[field]`` = nt x n matrices, FP32); ``SrcWav``/``Records`` are lists of them, one per
supersource (src/srcwav/srcwav.jl:1, src/records/records.jl:2).  ``AGeomss``/``AGeom``
restate src/ageom/ageom.jl:4-67.  ``lossvalue``/``gradient`` restate the L2 misfit and its
adjoint source (src/database/database.jl:354-380), accumulated in FP64.
"""
from __future__ import annotations

import numpy as np

from .grid import Grid


class _Data:
    def __init__(self, n: int, grid: Grid, fields=("vz",), d=None):
        self.n = int(n)
        self.grid = grid
        self.fields = list(fields)
        if d is None:
            d = {f: np.zeros((len(grid), self.n), np.float32, order="F") for f in self.fields}
        else:
            d = {f: np.asfortranarray(np.asarray(d[f], np.float32)) for f in self.fields}
        for f, a in d.items():
            if a.shape != (len(grid), self.n):
                raise ValueError("inconsistent data size")
        self.d = d

    def __getitem__(self, key):
        if key in ("n", "ns", "nr"):
            return self.n
        return self.d[key]

    def copy(self):
        return type(self)(self.n, self.grid, self.fields, {f: a.copy() for f, a in self.d.items()})

    def reversed(self):
        """The same data with time reversed, as views (no copy); read-only use."""
        r = object.__new__(type(self))
        r.__dict__.update(self.__dict__)
        r.d = {f: a[::-1] for f, a in self.d.items()}
        return r

    def fill(self, value: float = 0.0):
        for a in self.d.values():
            a.fill(value)

    def update(self, fields, w):
        """``update!(d, fields, w)``: set the wavelet of every point (vector ``w``) or per
        point (nt x n matrix ``w``) for the given fields (src/database/database.jl:566-609)."""
        w = np.asarray(w)
        for f in fields:
            if f not in self.d:
                raise KeyError(f)
            if w.ndim == 1:
                if w.shape[0] != len(self.grid):
                    raise ValueError("wavelet length differs from the time grid")
                self.d[f][:, :] = w.astype(np.float32)[:, None]
            else:
                if w.shape != (len(self.grid), self.n):
                    raise ValueError("wavelet matrix must be nt x n")
                self.d[f][:, :] = w.astype(np.float32)
        return self

    def vec(self) -> np.ndarray:
        return np.concatenate([self.d[f].ravel(order="F") for f in self.fields])


class Srcs(_Data):
    pass


class Recs(_Data):
    pass


def SrcWav(tgrid: Grid, ageom, fields) -> list:
    """``Srcs(tgrid, ageom, fields)`` for every supersource (src/srcwav/srcwav.jl:16-22)."""
    return [Srcs(a.ns, tgrid, fields) for a in ageom]


def Records(tgrid: Grid, ageom, fields) -> list:
    return [Recs(a.nr, tgrid, fields) for a in ageom]


class AGeomss:
    """Source and receiver coordinates of one supersource (src/ageom/ageom.jl:4-16).

    ``s`` and ``r`` map axis names ("z", ["y",] "x") to coordinate vectors (metres)."""

    def __init__(self, s: dict, r: dict, ns: int | None = None, nr: int | None = None):
        s = {k: np.asarray(v, np.float64).ravel() for k, v in s.items()}
        r = {k: np.asarray(v, np.float64).ravel() for k, v in r.items()}
        ns = len(next(iter(s.values()))) if ns is None else ns
        nr = len(next(iter(r.values()))) if nr is None else nr
        if set(s) != set(r) or any(len(v) != ns for v in s.values()) or any(len(v) != nr for v in r.values()):
            raise ValueError("AGeomss construct")
        self.s, self.r, self.ns, self.nr = s, r, int(ns), int(nr)

    def dims(self) -> list:
        return ["z", "x"] if len(self.s) == 2 else ["z", "y", "x"]

    def src_points(self) -> np.ndarray:
        return np.stack([self.s[k] for k in self.dims()], axis=1)

    def rec_points(self) -> np.ndarray:
        return np.stack([self.r[k] for k in self.dims()], axis=1)

    def inside(self, grid: list) -> bool:
        """``ageom ∈ mgrid`` (src/ageom/ageom.jl:291-322)."""
        for k, g in zip(self.dims(), grid):
            for v in (self.s[k], self.r[k]):
                if np.any(v < g.first) or np.any(v > g.last):
                    return False
        return True


class L2DistLoss:
    """``LossFunctions.L2DistLoss`` (LossFunctions 0.11.2): value (output - target)^2 and
    derivative 2 (output - target)."""

    def value(self, output, target):
        return (output - target) ** 2

    def deriv(self, output, target):
        return 2.0 * (output - target)


def _loss(loss):
    if loss is None or isinstance(loss, L2DistLoss) or loss == "L2DistLoss":
        return L2DistLoss()
    raise NotImplementedError("only L2DistLoss is available")


def lossvalue(loss, dobs, dcal) -> float:
    """``lossvalue(loss, dobs, dcal)`` (src/database/database.jl:357-368): the sum over
    supersources, fields, samples and receivers of ``loss(dobs, dcal)``, accumulated in FP64."""
    L = _loss(loss)
    if isinstance(dobs, list):
        return float(sum(lossvalue(L, a, b) for a, b in zip(dobs, dcal)))
    total = 0.0
    for f in dobs.fields:
        r = np.subtract(dobs.d[f], dcal.d[f], dtype=np.float64).ravel(order="K")
        total += float(np.dot(r, r))   # sum of L2DistLoss values, FP64
    return total


def gradient(buffer, loss, dobs, dcal):
    """``gradient!(buffer, loss, dobs, dcal)`` (src/database/database.jl:371-380): the adjoint
    source ``deriv(loss, dobs, dcal)`` of every sample into ``buffer`` (FP32).

    LossFunctions 0.11 orders the arguments as (output, target), so the call gives
    2 (dobs - dcal).  With that sign the reference's time-reversal gradient agrees with
    finite differences of the misfit."""
    L = _loss(loss)
    if isinstance(dobs, list):
        for g, a, b in zip(buffer, dobs, dcal):
            gradient(g, L, a, b)
        return buffer
    for f in dobs.fields:
        # 2 (dobs - dcal): the FP32 difference rounds once, as the FP64 one cast to FP32 does
        np.subtract(dobs.d[f], dcal.d[f], out=buffer.d[f], dtype=np.float32)
        buffer.d[f] *= np.float32(2.0)
    return buffer

