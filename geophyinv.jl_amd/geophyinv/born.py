"""Born modelling and its linear map (src/fdtd/born.jl, src/fdtd/func_grad.jl:51-120).

``FdtdAcoustic[Born]()`` is ``FdtdAcoustic{Born}()``: two wavefields, the sources drive the
background one and the receivers sample the scattered one (propagate.jl:53-60).
``update(pa, medium, medium_pert)`` stores the perturbation ``dmod = mod(medium_pert) -
mod(medium)`` (medium.jl:106-116) on the unpadded region; ``update(pa)`` then models the
Born data of every supersource into ``pa["data"]``.

The scattered wavefield's secondary sources are the linearisation of the velocity and
pressure updates in their coefficients (the reference's own calls do not match the kernels
they call): d(dtK) = -dt dinvK / invK^2 on the stress grid and d(dtinvav*irho) =
-dt @av(drho) / @av(rho)^2 on the velocity grids, zero where a CPML memory variable acts.

``LinearMap(pa)`` is the linearised forward operator F (model perturbation -> scattered
velocity records) and its transpose, the migration operator, computed with the gradient
machinery: adjoint sources = the data, time reversed (adjoint_map!, func_grad.jl:78-101),
and F^T d = -(imaging of the adjoint run) (the gradient of the misfit is -F^T of its
adjoint sources).  The transpose uses the checkpointed exact-wavefield adjoint when the
experiment has ``checkpoint > 0`` and the reference's time reversal otherwise (2-D).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .fields import FIELD_ID
from .media import independent, padarray

F32 = np.float32


class Born:
    """Marker of single-scattering modelling: ``FdtdAcoustic[Born]()``."""


def _avg2(a, b):
    return (a + b) * 0.5


def perturbation(pa, medium_pert) -> dict:
    """``dmod = mod(padarray(medium_pert)) - mod`` on the unpadded region, zero outside,
    FP64 (medium.jl:106-116)."""
    if [g.n for g in medium_pert.grid] != [g.n for g in pa.medium.grid]:
        raise ValueError("the perturbed medium must live on the experiment's grid")
    mp = independent(padarray(medium_pert, pa.npml, pa.pml_faces))
    out = {}
    for k in ("invK", "rho"):
        d = np.zeros(pa.ext_shape, np.float64, order="F")
        sl = pa.view_inner(d)
        d[sl] = mp[k][sl].astype(np.float64) - pa.mod[k][sl].astype(np.float64)
        out[k] = d
    return out


def dcoefs(pa, dmod: dict) -> dict:
    """Linearised coefficient perturbations (FP32) of ``dmod`` (padded grid, FP64):
    d(dtK) and d(dtinvav{x,y,z}irho), zero on the CPML strips of their own axes."""
    h = pa.order - 1
    dt = float(pa.dt32)
    invK = pa.mod["invK"].astype(np.float64)
    rho = pa.mod["rho"].astype(np.float64)
    dK = dmod.get("invK", np.zeros_like(invK))
    dR = dmod.get("rho", np.zeros_like(rho))
    out = {"dtK": -dt * dK / (invK * invK)}
    if pa.ndim == 2:
        nz, nx = rho.shape
        sx = (slice(h, nz - h), slice(0, nx - h)), (slice(h, nz - h), slice(1, nx - h + 1))
        sz = (slice(0, nz - h), slice(h, nx - h)), (slice(1, nz - h + 1), slice(h, nx - h))
        grids = {"dtinvavxirho": (sx, 1), "dtinvavzirho": (sz, 0)}
    else:
        nz, ny, nx = rho.shape
        sx = ((slice(h, nz - h), slice(h, ny - h), slice(0, nx - h)),
              (slice(h, nz - h), slice(h, ny - h), slice(1, nx - h + 1)))
        sy = ((slice(h, nz - h), slice(0, ny - h), slice(h, nx - h)),
              (slice(h, nz - h), slice(1, ny - h + 1), slice(h, nx - h)))
        sz = ((slice(0, nz - h), slice(h, ny - h), slice(h, nx - h)),
              (slice(1, nz - h + 1), slice(h, ny - h), slice(h, nx - h)))
        grids = {"dtinvavxirho": (sx, 2), "dtinvavyirho": (sy, 1), "dtinvavzirho": (sz, 0)}
    npml = pa.npml
    names = ["z", "x"] if pa.ndim == 2 else ["z", "y", "x"]
    for name, ((a, b), ax) in grids.items():
        av = _avg2(rho[a], rho[b])
        d = -dt * _avg2(dR[a], dR[b]) / (av * av)
        # the CPML memory of this derivative acts on its first/last npml points along ax
        n = d.shape[ax]
        idx = [slice(None)] * d.ndim
        if f"{names[ax]}min" in pa.pml_faces:
            idx[ax] = slice(0, min(npml, n))
            d[tuple(idx)] = 0.0
        if f"{names[ax]}max" in pa.pml_faces:
            idx[ax] = slice(max(n - npml, 0), n)
            d[tuple(idx)] = 0.0
        out[name] = d
    return {k: np.asfortranarray(v, dtype=F32) for k, v in out.items()}


def run_born(pa, iss: int) -> dict:
    """Born records of supersource ``iss`` (gpi_born): dict field -> nt x nr."""
    from .fdtd import COEF_ID

    if pa.dmod is None:
        raise RuntimeError("set the perturbation with update(pa, medium, medium_pert) first")
    pa._ensure_model()
    ctx = pa._ctx
    nt = len(pa.tgrid)
    keep = []
    src_arr, wav_arr, nsrc = pa._points(pa.spray[iss], pa.wavelets[iss], keep)
    recs, outs, names = [], [], []
    for f, (idx, w) in pa.interp[iss].items():
        idx = np.ascontiguousarray(idx)
        w = np.ascontiguousarray(w)
        out = np.zeros((nt, idx.shape[0]), F32, order="F")
        keep += [idx, w, out]
        p = _lib.gpi_points()
        p.field, p.npts, p.nnz = FIELD_ID[f], idx.shape[0], idx.shape[1]
        p.idx, p.w = _lib.iptr(idx), _lib.fptr(w)
        recs.append(p)
        outs.append(out)
        names.append(f)
    shot = _lib.gpi_shot()
    shot.nt = nt
    shot.nsrc = nsrc
    shot.nrec = len(recs)
    shot.src = src_arr
    shot.wavelets = wav_arr
    shot.rec = (_lib.gpi_points * max(1, len(recs)))(*recs)
    shot.records = (C.POINTER(C.c_float) * max(1, len(outs)))(*[_lib.fptr(o) for o in outs])
    dc = pa.born_dcoefs
    arrs = []
    for name, a in dc.items():
        x = _lib.gpi_array()
        x.id = COEF_ID[name]
        x.count = a.size
        x.data = _lib.fptr(a)
        arrs.append(x)
    arr = (_lib.gpi_array * len(arrs))(*arrs)
    ctx.check(ctx.lib.gpi_born(ctx.h, pa._model, C.byref(shot), arr, len(arrs)))
    pa.timing = ctx.timing()
    return dict(zip(names, outs))


class LinearMap:
    """``LinearMap(pa)`` for ``FdtdAcoustic[Born]`` (func_grad.jl:106-120): ``F @ x`` models
    the scattered velocity records of every supersource for the model perturbation
    ``x = [dinvK; drho]`` (unpadded region, column-major, ``mparams`` order) and
    ``F.T @ d`` migrates records ``d`` (same layout as ``F @ x``) back to the model."""

    def __init__(self, pa, mparams=("invK", "rho")):
        if not getattr(pa, "born", False):
            raise ValueError("LinearMap needs an FdtdAcoustic[Born] experiment")
        self.pa = pa
        self.mparams = list(mparams)
        self.vfields = [f for f in pa.rfields if f in ("vx", "vy", "vz")]
        if not self.vfields:
            raise ValueError("the linear map records velocity fields (the adjoint sources of the reference)")
        nt = len(pa.tgrid)
        self.inner = pa.view_inner(np.zeros(pa.ext_shape))
        self.ncell = int(np.prod([s.stop - s.start for s in self.inner]))
        self.nd = sum(nt * a.nr * len(self.vfields) for a in pa.ageom)
        self.shape = (self.nd, self.ncell * len(self.mparams))
        self.dtype = np.dtype(np.float64)

    # --- forward_map!
    def matvec(self, x):
        pa = self.pa
        x = np.asarray(x, np.float64).ravel()
        dmod = {}
        for k, chunk in zip(self.mparams, np.split(x, len(self.mparams))):
            d = np.zeros(pa.ext_shape, np.float64, order="F")
            d[self.inner] = chunk.reshape(d[self.inner].shape, order="F")
            dmod[k] = d
        pa.set_perturbation(dmod)
        out = []
        for iss in range(len(pa.ageom)):
            recs = run_born(pa, iss)
            out += [recs[f].astype(np.float64).ravel(order="F") for f in self.vfields]
        return np.concatenate(out)

    # --- adjoint_map!
    def rmatvec(self, y):
        pa = self.pa
        y = np.asarray(y, np.float64).ravel()
        nt = len(pa.tgrid)
        grads = {k: np.zeros(pa.ext_shape, np.float64, order="F") for k in self.mparams}
        off = 0
        for iss, a in enumerate(pa.ageom):
            adj = {}
            for f in self.vfields:
                n = nt * a.nr
                adj[f] = y[off:off + n].reshape(nt, a.nr, order="F")[::-1].astype(F32)
                off += n
            g = pa.migrate(iss, adj)
            for k in self.mparams:
                grads[k] -= g[k]
        return np.concatenate([grads[k][self.inner].ravel(order="F") for k in self.mparams])

    def __matmul__(self, x):
        return self.matvec(x)

    @property
    def T(self):
        outer = self

        class _T:
            shape = (outer.shape[1], outer.shape[0])

            def __matmul__(self, y):
                return outer.rmatvec(y)

        return _T()

    def aslinearoperator(self):
        """The same map as a ``scipy.sparse.linalg.LinearOperator`` (for lsqr/lsmr)."""
        from scipy.sparse.linalg import LinearOperator
        return LinearOperator(self.shape, matvec=self.matvec, rmatvec=self.rmatvec, dtype=np.float64)
