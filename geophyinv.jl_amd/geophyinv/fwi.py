"""Full-waveform inversion on a coarse inversion grid (src/fwi/fwi.jl:1-75,
src/fwi/func_grad.jl:1-68).

``SeisInvExpt(paf, dobs, migrid, mparams)`` wraps an npw = 2 ``SeisForwExpt``: the model
vector lives on the inversion grid ``migrid``; the modelling-grid log-parameters are its
separable linear interpolation ``mfull = P' m Q`` (``apply_proj_matrix!``,
src/proj_mat.jl:208-226, with the 1-D matrices of ``get_proj_matrix``, :208-226 and
:176-200), and the gradient on the inversion grid is the transposed map of the
modelling-grid gradient.  ``update(paw)`` minimises the misfit with L-BFGS.
"""
from __future__ import annotations

import numpy as np

from . import func_grad
from .grid import Grid

F32 = np.float32


def _neighbours(vals: np.ndarray, v: float):
    """``get_neighbour_indices`` (src/proj_mat.jl:69-79), 0-based."""
    n = len(vals)
    idx = int(np.searchsorted(vals, v, side="left"))
    if idx == 0:
        return 0, 1
    if idx >= n - 1:
        return n - 2, n - 1
    return idx - 1, idx


def interp_matrix(points, grid) -> np.ndarray:
    """``get_proj_matrix([points], [grid])`` in 1-D: a dense (len(grid), len(points)) FP32
    matrix whose column j holds the linear-interpolation weights of point j on ``grid``
    (``bilinear_interp(x, xi)``, src/proj_mat.jl:180-200: clamped outside the grid)."""
    x = grid.values() if isinstance(grid, Grid) else np.asarray(grid, np.float64)
    p = points.values() if isinstance(points, Grid) else np.asarray(points, np.float64)
    M = np.zeros((len(x), len(p)), F32)
    for j, xi in enumerate(p):
        i1, i2 = _neighbours(x, float(xi))
        if xi < x.min():
            dx = 0.0
        elif xi > x.max():
            dx = 1.0
        else:
            dx = (float(xi) - x[i1]) / (x[i2] - x[i1])
        M[i1, j] = F32(1.0 - dx)
        M[i2, j] = F32(dx)
    return M


def apply_proj(m: np.ndarray, mats) -> np.ndarray:
    """``m1[i, j(, o)] = P[k, i] m[k, l(, n)] Q[l, j](R[n, o])`` (src/proj_mat.jl:208-226), in
    FP64 and stored FP32."""
    out = np.asarray(m, np.float64)
    for ax, P in enumerate(mats):
        out = np.moveaxis(np.tensordot(np.asarray(P, np.float64), out, axes=([0], [ax])), 0, ax)
    return out.astype(F32)


class SeisInvExpt:
    """``SeisInvExpt(paf, dobs, migrid, mparams; loss)`` (src/fwi/fwi.jl:10-35).

    ``migrid`` is a list of ``Grid`` (one per axis) or a list of point counts, in which case
    each axis spans the modelling grid shrunk by 4 cells at both ends (fwi.jl:28-35)."""

    def __init__(self, paf, dobs, migrid=None, mparams=None, loss=None):
        if paf.npw != 2:
            raise ValueError("the modelling experiment needs npw = 2")
        mmgrid = paf.medium.grid
        if migrid is None:
            migrid = list(mmgrid)
        elif all(isinstance(n, (int, np.integer)) for n in migrid):
            migrid = [Grid.range(m.first + 4 * m.step, m.last - 4 * m.step, length=int(n))
                      for m, n in zip(mmgrid, migrid)]
        if len(migrid) != len(mmgrid):
            raise ValueError("migrid must have one grid per axis")
        self.paf = paf
        self.dobs = dobs
        self.dobs0 = [d.copy() for d in dobs]
        self.migrid = list(migrid)
        self.mparams = list(mparams or paf.mparams)
        self.loss = loss
        # inversion -> modelling interpolation: P (n_mi, n_mm), applied transposed
        self.P = [interp_matrix(mm, mi) for mm, mi in zip(mmgrid, self.migrid)]
        # modelling -> inversion sampling used by get_modelvector
        self.Pm = [interp_matrix(mi, mm) for mm, mi in zip(mmgrid, self.migrid)]
        self.mfull = paf.model_vector(self.mparams)
        self.gmfull = np.zeros_like(self.mfull)

    @property
    def nm(self) -> int:
        return len(self.mparams) * int(np.prod([g.n for g in self.migrid]))

    def _shapes(self):
        return tuple(g.n for g in self.paf.medium.grid), tuple(g.n for g in self.migrid)

    def _to_full(self, m):
        smm, smi = self._shapes()
        chunks = np.split(np.asarray(m, np.float64), len(self.mparams))
        full = [apply_proj(c.reshape(smi, order="F"), self.P).ravel(order="F") for c in chunks]
        self.mfull[...] = np.concatenate(full)

    def get_modelvector(self) -> np.ndarray:
        """``get_modelvector(pa)`` (src/fwi/func_grad.jl:1-21)."""
        smm, smi = self._shapes()
        self.mfull[...] = self.paf.model_vector(self.mparams)
        chunks = np.split(self.mfull, len(self.mparams))
        out = [apply_proj(c.reshape(smm, order="F"), self.Pm).ravel(order="F") for c in chunks]
        return np.concatenate(out).astype(np.float64)

    def lossvalue(self, m) -> float:
        """``lossvalue(m, pa)`` (src/fwi/func_grad.jl:23-39)."""
        self._to_full(m)
        return func_grad.lossvalue(self.mfull, self.loss, self.dobs, self.paf, self.mparams)

    def gradient(self, g, m) -> float:
        """``gradient!(g, m, pa)`` (src/fwi/func_grad.jl:41-68); returns the misfit of the
        modelled data (the modelling experiment is switched to compat = 0 semantics for the
        returned value only)."""
        smm, smi = self._shapes()
        self._to_full(m)
        compat = self.paf.compat
        self.paf.compat = 0
        try:
            J = func_grad.gradient(self.gmfull, self.mfull, self.loss, self.dobs, self.paf, self.mparams)
        finally:
            self.paf.compat = compat
        PT = [np.ascontiguousarray(P.T) for P in self.P]
        chunks = np.split(self.gmfull, len(self.mparams))
        out = [apply_proj(c.reshape(smm, order="F"), PT).ravel(order="F") for c in chunks]
        g[...] = np.concatenate(out)
        return J

    def update_source(self, maxiter: int = 100) -> dict:
        """``update!(paw, pac)`` (src/fwi/fwi.jl:37-64): model the data at the current model,
        then per supersource estimate the filter w that best maps the original observed data
        onto the modelled data (LSMR on the convolution operator with g = dobs0, all fields'
        traces sharing one w) and replace the observed data by dobs0 * w.  Returns the
        misfits before and after."""
        from .conv import estimate_filter

        m = self.get_modelvector()
        loss_initial = self.lossvalue(m)
        for iss, (d0, d, dcal) in enumerate(zip(self.dobs0, self.dobs, self.paf.data)):
            fields = list(dcal.fields)
            g = np.concatenate([np.asarray(d0.d[f], np.float64) for f in fields], axis=1)
            dc = np.concatenate([np.asarray(dcal.d[f], np.float64) for f in fields], axis=1)
            _, conv = estimate_filter(g, dc, maxiter=maxiter)
            off = 0
            for f in fields:
                n = d.d[f].shape[1]
                d.d[f][:, :] = conv[:, off:off + n].astype(d.d[f].dtype)
                off += n
        loss_final = self.lossvalue(m)
        return {"loss_initial": loss_initial, "loss_final": loss_final}

    def update(self, maxiter: int = 20, callback=None):
        """``update!(paw)`` (src/fwi/fwi.jl:66-75): L-BFGS on the misfit from the current
        model.  The reference drives Optim.jl's LBFGS; this uses SciPy's L-BFGS-B without
        bounds.  Returns the SciPy result; the final model stays in the experiment."""
        from scipy.optimize import minimize

        m0 = self.get_modelvector()
        g = np.zeros_like(m0)

        def fg(m):
            J = self.gradient(g, m)
            return J, g.astype(np.float64)   # a new FP64 array per call

        res = minimize(fg, m0, jac=True, method="L-BFGS-B", callback=callback,
                       options={"maxiter": maxiter, "gtol": 1e-25, "ftol": 0.0})
        self.lossvalue(res.x)   # leave the experiment at the final model
        return res
