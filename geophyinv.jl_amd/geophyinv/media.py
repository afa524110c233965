"""Seismic media and the FP32 coefficient preparation of the FDTD engine.

Restates src/media/media.jl (``Medium`` types, derived parameters, replicate padding) and
src/fdtd/medium.jl:143-221 (``update_dmod!``: the dt-scaled coefficient arrays the time
loop reads).  Arrays are numpy arrays indexed like the reference, ``m[iz, ix]`` or
``m[iz, iy, ix]``, held in Fortran (column-major) order so the memory layout is the
reference's z-fastest layout that the C ABI expects.
"""
from __future__ import annotations

import numpy as np

from .grid import Grid

F32 = np.float32


def _f(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=F32))


class MediumParameter:
    """A medium parameter with its reference value and bounds (src/media/media.jl:18-37)."""

    def __init__(self, m, ref=None, bounds=None):
        m = _f(m)
        if np.any(m < 0):
            raise ValueError("negative medium parameters")
        self.m = m
        # Statistics.mean: the float64 mean rounded to the array's element type
        self.ref = F32(np.mean(m, dtype=np.float64)) if ref is None else F32(ref)
        if bounds is None:
            frac = F32(0.1)
            lo = F32(m.min()) - frac * self.ref
            hi = F32(m.max()) + frac * self.ref
            bounds = [max(F32(0.0), lo), hi]
        self.bounds = [F32(bounds[0]), F32(bounds[1])]

    def copy(self) -> "MediumParameter":
        out = MediumParameter.__new__(MediumParameter)   # validated already: copy as is
        out.m = self.m.copy(order="K")
        out.ref = self.ref
        out.bounds = list(self.bounds)
        return out


class Medium:
    """Base of ``AcousticMedium`` / ``ElasticMedium`` (src/media/media.jl:39-66)."""

    params: tuple = ()

    def __init__(self, grid, **params):
        self.grid = [g if isinstance(g, Grid) else Grid(*g) for g in grid]
        shape = tuple(len(g) for g in self.grid)
        for name in self.params:
            p = params[name]
            p = p if isinstance(p, MediumParameter) else MediumParameter(p)
            if p.m.shape != shape:
                raise ValueError(f"{name} has shape {p.m.shape}, grid is {shape}")
            setattr(self, name, p)

    @property
    def ndim(self) -> int:
        return len(self.grid)

    def __getitem__(self, name: str):
        """``medium[:vp]`` returns the parameter; derived names return arrays."""
        if name in self.params:
            return getattr(self, name)
        return derived(self, name)

    def copy(self) -> "Medium":
        return type(self)(list(self.grid), **{k: getattr(self, k).copy() for k in self.params})


class AcousticMedium(Medium):
    params = ("vp", "rho")
    physics = 0

    def __init__(self, grid, vp, rho):
        super().__init__(grid, vp=vp, rho=rho)


class ElasticMedium(Medium):
    params = ("vp", "vs", "rho")
    physics = 1

    def __init__(self, grid, vp, vs, rho):
        super().__init__(grid, vp=vp, vs=vs, rho=rho)


def derived(medium: Medium, name: str) -> np.ndarray:
    """Derived parameters in FP32 (src/media/media.jl:111-130)."""
    vp = medium.vp.m
    rho = medium.rho.m
    if isinstance(medium, ElasticMedium):
        vs = medium.vs.m
        if name == "invlambda":
            return _f(F32(1) / (((vp * vp) - F32(2) * (vs * vs)) * rho))
        if name == "invmu":
            return _f(F32(1) / ((vs * vs) * rho))
        if name == "lambda":
            return _f(((vp * vp) - F32(2) * (vs * vs)) * rho)
        if name == "mu":
            return _f((vs * vs) * rho)
        if name == "M":
            return _f((vp * vp) * rho)
    else:
        if name == "invK":
            return _f(F32(1) / ((vp * vp) * rho))
        if name == "K":
            return _f((vp * vp) * rho)
    if name == "invrho":
        return _f(F32(1) / rho)
    raise KeyError(name)


def _pad_widths(ndim: int, npml: int, faces) -> list:
    axes = ["z", "x"] if ndim == 2 else ["z", "y", "x"]
    return [((npml if f"{a}min" in faces else 0), (npml if f"{a}max" in faces else 0)) for a in axes]


def padmgrid(grid: list, npml: int, faces) -> list:
    """Extended grid (src/media/media.jl:278-296)."""
    out = []
    for g, (lo, hi) in zip(grid, _pad_widths(len(grid), npml, faces)):
        n = g.n + lo + hi
        start, stop = g.first - lo * g.step, g.last + hi * g.step
        out.append(Grid.range(start, stop, length=n))
    return out


def padarray(medium: Medium, npml: int, faces) -> Medium:
    """Replicate-pad the medium by ``npml`` cells on the PML faces (src/media/media.jl:260-304).

    The padded parameters get fresh ``ref``/``bounds`` from the padded arrays, as
    ``padarray!`` rebuilds them with ``eval(mname)(number.(m11))``."""
    widths = _pad_widths(medium.ndim, npml, faces)
    grid = padmgrid(medium.grid, npml, faces)
    params = {k: _f(np.pad(getattr(medium, k).m, widths, mode="edge")) for k in medium.params}
    return type(medium)(grid, **params)


# ------------------------------------------------------------------------------------------
# dt-scaled coefficient arrays (update_dmod!, src/fdtd/medium.jl:143-221)
# ------------------------------------------------------------------------------------------

def _avg2(a, b):
    # (A[..] + A[..]) * 0.5 with the literal taken as FP32 inside a ParallelStencil kernel
    return (a + b) * F32(0.5)


# independent parameters held per experiment (MediumParameters, src/fdtd/medium.jl:81-95)
MPARAMS = {"acoustic": ["invK", "rho"], "elastic": ["invlambda", "invmu", "rho"]}


def independent(exmedium: Medium) -> dict:
    """``pac.mod``: the independent FP32 parameters of the padded medium."""
    kind = "elastic" if isinstance(exmedium, ElasticMedium) else "acoustic"
    return {k: (exmedium.rho.m if k == "rho" else derived(exmedium, k)) for k in MPARAMS[kind]}


def coefficients(exmedium: Medium, dt32: np.float32, order: int) -> dict:
    """Return the coefficient arrays keyed by the reference's names (``mod`` and ``dmod``)."""
    return dependent(independent(exmedium), dt32, order, exmedium.ndim)


def dtk_of(invK: np.ndarray, dt32: np.float32) -> np.ndarray:
    """``dtK = dt * inv(invK)`` (src/fdtd/medium.jl:147-148)."""
    return _f((F32(1) / invK) * dt32)


def dependent(mod: dict, dt32: np.float32, order: int, nd: int) -> dict:
    """``update_dmod!`` (src/fdtd/medium.jl:143-186): ``mod`` plus the dt-scaled arrays."""
    h = order - 1
    rho = mod["rho"]
    out = dict(mod)
    if nd == 2:
        nz, nx = rho.shape
        out["dtinvavxirho"] = _f(dt32 / _avg2(rho[h:nz - h, 0:nx - h], rho[h:nz - h, 1:nx - h + 1]))
        out["dtinvavzirho"] = _f(dt32 / _avg2(rho[0:nz - h, h:nx - h], rho[1:nz - h + 1, h:nx - h]))
    else:
        nz, ny, nx = rho.shape
        out["dtinvavxirho"] = _f(dt32 / _avg2(rho[h:nz - h, h:ny - h, 0:nx - h],
                                             rho[h:nz - h, h:ny - h, 1:nx - h + 1]))
        out["dtinvavyirho"] = _f(dt32 / _avg2(rho[h:nz - h, 0:ny - h, h:nx - h],
                                             rho[h:nz - h, 1:ny - h + 1, h:nx - h]))
        out["dtinvavzirho"] = _f(dt32 / _avg2(rho[0:nz - h, h:ny - h, h:nx - h],
                                             rho[1:nz - h + 1, h:ny - h, h:nx - h]))
    if "invK" in mod:
        out["dtK"] = dtk_of(mod["invK"], dt32)
        return out
    il = mod["invlambda"]
    im = mod["invmu"]
    out["dtlambda"] = _f((F32(1) / il) * dt32)
    # inv(invλ) + 2.0*inv(invμ) is a host broadcast: the Float64 literal promotes
    m64 = (F32(1) / il).astype(np.float64) + 2.0 * (F32(1) / im).astype(np.float64)
    out["dtM"] = _f(m64.astype(F32) * dt32)
    q = F32(0.25)
    if nd == 2:
        nz, nx = im.shape
        av = (((im[0:nz - h, 0:nx - h] + im[1:nz - h + 1, 0:nx - h]) + im[0:nz - h, 1:nx - h + 1])
              + im[1:nz - h + 1, 1:nx - h + 1]) * q
        out["dtavmu"] = _f(dt32 / av)
    else:
        nz, ny, nx = im.shape
        # @av_xzi on the dvxdz grid (H, I, H)
        a = im[0:nz - h, h:ny - h, 0:nx - h]
        b = im[1:nz - h + 1, h:ny - h, 0:nx - h]
        c = im[0:nz - h, h:ny - h, 1:nx - h + 1]
        d = im[1:nz - h + 1, h:ny - h, 1:nx - h + 1]
        out["dtavxzimu"] = _f(dt32 / ((((a + b) + c) + d) * q))
        # @av_xyi on the dvxdy grid (I, H, H)
        a = im[h:nz - h, 0:ny - h, 0:nx - h]
        b = im[h:nz - h, 1:ny - h + 1, 0:nx - h]
        c = im[h:nz - h, 0:ny - h, 1:nx - h + 1]
        d = im[h:nz - h, 1:ny - h + 1, 1:nx - h + 1]
        out["dtavxyimu"] = _f(dt32 / ((((a + b) + c) + d) * q))
        # @av_yzi on the dvydz grid (H, H, I)
        a = im[0:nz - h, 0:ny - h, h:nx - h]
        b = im[1:nz - h + 1, 0:ny - h, h:nx - h]
        c = im[0:nz - h, 1:ny - h + 1, h:nx - h]
        d = im[1:nz - h + 1, 1:ny - h + 1, h:nx - h]
        out["dtavyzimu"] = _f(dt32 / ((((a + b) + c) + d) * q))
    return out
