"""Full-waveform inversion on a coarse inversion grid (src/fwi/fwi.jl:1-75,
src/fwi/func_grad.jl:1-68).

``SeisInvExpt(paf, dobs, migrid, mparams)`` wraps an npw = 2 ``SeisForwExpt``: the model
vector lives on the inversion grid ``migrid``; the modelling-grid log-parameters are its
separable linear interpolation ``mfull = P' m Q`` (``apply_proj_matrix!``,
src/proj_mat.jl:208-226, with the 1-D matrices of ``get_proj_matrix``, :208-226 and
:176-200), and the gradient on the inversion grid is the transposed map of the
modelling-grid gradient.  Both maps run on the device (``gpi_interp_apply``).
``update(paw)`` minimises the misfit with L-BFGS whose two-loop recursion runs on the
device with the pair history in HBM (``gpi_lbfgs_*``); the host keeps the line search.
"""
from __future__ import annotations

import ctypes as C
import time
from types import SimpleNamespace

import numpy as np

from . import _lib, func_grad
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


def apply_proj(ctx, m: np.ndarray, mats) -> np.ndarray:
    """``m1[i, j(, o)] = P[k, i] m[k, l(, n)] Q[l, j](R[n, o])`` (src/proj_mat.jl:208-226) on
    the device (``gpi_interp_apply``): FP64 sums, stored FP32 as the reference's arrays are."""
    m = np.asfortranarray(m, dtype=np.float64)
    nd = len(mats)
    if m.ndim != nd or any(m.shape[a] != P.shape[0] for a, P in enumerate(mats)):
        raise ValueError("the model and the projection matrices disagree in shape")
    nin = (C.c_int64 * 3)(*(list(m.shape) + [1] * (3 - nd)))
    nout = (C.c_int64 * 3)(*([P.shape[1] for P in mats] + [1] * (3 - nd)))
    ws = [np.asfortranarray(P, dtype=F32) for P in mats]
    wp = (C.POINTER(C.c_float) * 3)(*([_lib.fptr(w) for w in ws] + [C.POINTER(C.c_float)()] * (3 - nd)))
    out = np.zeros(tuple(P.shape[1] for P in mats), np.float64, order="F")
    ctx.check(ctx.lib.gpi_interp_apply(ctx.h, nd, nin, nout, wp, _lib.dptr(m), _lib.dptr(out)))
    return out.astype(F32)


class DeviceLBFGS:
    """The L-BFGS two-loop recursion on the device (``gpi_lbfgs_*``): ``direction(m, g)``
    folds the last step's pair into the history held in HBM and returns -H g."""

    def __init__(self, ctx, n: int, memory: int = 10):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(ctx.lib.gpi_lbfgs_create(ctx.h, n, memory, C.byref(h)))
        self.h = h
        self.n = n
        self.pairs = 0

    def direction(self, m: np.ndarray, g: np.ndarray) -> np.ndarray:
        m = np.ascontiguousarray(m, dtype=np.float64)
        g = np.ascontiguousarray(g, dtype=np.float64)
        if m.size != self.n or g.size != self.n:
            raise ValueError("model and gradient must have the optimizer's size")
        d = np.empty(self.n, np.float64)
        k = C.c_int32()
        self.ctx.check(self.ctx.lib.gpi_lbfgs_direction(self.h, _lib.dptr(m), _lib.dptr(g), _lib.dptr(d),
                                                         C.byref(k)))
        self.pairs = k.value
        return d

    def reset(self):
        self.ctx.check(self.ctx.lib.gpi_lbfgs_reset(self.h))
        self.pairs = 0

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.gpi_lbfgs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def wolfe_search(fg, x, f0, g0, d, alpha0=1.0, c1=1e-4, c2=0.9, maxfev=20):
    """Step along d satisfying the strong Wolfe conditions (Nocedal & Wright, Algorithms 3.5
    and 3.6, cubic interpolation in the zoom): returns (alpha, f, g, evaluations); alpha is
    None when no such step was found within maxfev evaluations of fg (one gradient each)."""
    dg0 = float(np.dot(g0, d))
    if not dg0 < 0:
        raise ValueError("d is not a descent direction")
    nfev = 0

    def phi(a):
        nonlocal nfev
        nfev += 1
        f, g = fg(x + a * d)
        return f, g, float(np.dot(g, d))

    def cubic(a0, f0_, d0, a1, f1, d1, lo, hi):
        d1_ = d0 + d1 - 3.0 * (f0_ - f1) / (a0 - a1)
        r = d1_ * d1_ - d0 * d1
        if r >= 0.0 and a1 != a0:
            d2 = np.sign(a1 - a0) * np.sqrt(r)
            a = a1 - (a1 - a0) * (d1 + d2 - d1_) / (d1 - d0 + 2.0 * d2)
            if min(lo, hi) + 0.1 * abs(hi - lo) <= a <= max(lo, hi) - 0.1 * abs(hi - lo):
                return a
        return 0.5 * (lo + hi)

    def zoom(alo, flo, dlo, ahi, fhi, dhi):
        while nfev < maxfev:
            a = cubic(alo, flo, dlo, ahi, fhi, dhi, alo, ahi)
            f, g, dg = phi(a)
            if f > f0 + c1 * a * dg0 or f >= flo:
                ahi, fhi, dhi = a, f, dg
            else:
                if abs(dg) <= -c2 * dg0:
                    return a, f, g
                if dg * (ahi - alo) >= 0:
                    ahi, fhi, dhi = alo, flo, dlo
                alo, flo, dlo = a, f, dg
        return None, None, None

    a_prev, f_prev, d_prev = 0.0, f0, dg0
    a = alpha0
    while nfev < maxfev:
        f, g, dg = phi(a)
        if f > f0 + c1 * a * dg0 or (nfev > 1 and f >= f_prev):
            out = zoom(a_prev, f_prev, d_prev, a, f, dg)
            return out + (nfev,)
        if abs(dg) <= -c2 * dg0:
            return a, f, g, nfev
        if dg >= 0:
            out = zoom(a, f, dg, a_prev, f_prev, d_prev)
            return out + (nfev,)
        a_prev, f_prev, d_prev = a, f, dg
        a = 2.0 * a
    return None, None, None, nfev


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
    def ctx(self):
        """The modelling experiment's context (the library's default one until it has one)."""
        if self.paf._ctx is None:
            self.paf._ctx = _lib.default_context()
        return self.paf._ctx

    @property
    def nm(self) -> int:
        return len(self.mparams) * int(np.prod([g.n for g in self.migrid]))

    def _shapes(self):
        return tuple(g.n for g in self.paf.medium.grid), tuple(g.n for g in self.migrid)

    def _to_full(self, m):
        smm, smi = self._shapes()
        chunks = np.split(np.asarray(m, np.float64), len(self.mparams))
        full = [apply_proj(self.ctx, c.reshape(smi, order="F"), self.P).ravel(order="F") for c in chunks]
        self.mfull[...] = np.concatenate(full)

    def get_modelvector(self) -> np.ndarray:
        """``get_modelvector(pa)`` (src/fwi/func_grad.jl:1-21)."""
        smm, smi = self._shapes()
        self.mfull[...] = self.paf.model_vector(self.mparams)
        chunks = np.split(self.mfull, len(self.mparams))
        out = [apply_proj(self.ctx, c.reshape(smm, order="F"), self.Pm).ravel(order="F") for c in chunks]
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
        out = [apply_proj(self.ctx, c.reshape(smm, order="F"), PT).ravel(order="F") for c in chunks]
        g[...] = np.concatenate(out)
        return J

    def update_source(self, maxiter: int = 100) -> dict:
        """``update!(paw, pac)`` (src/fwi/fwi.jl:37-64): model the data at the current model,
        then per supersource estimate the filter w that best maps the original observed data
        onto the modelled data (LSMR on the convolution operator with g = dobs0, all fields'
        traces sharing one w; on the GPU, ``conv.estimate_filter_device``) and replace the
        observed data by dobs0 * w.  Returns the misfits before and after."""
        from .conv import estimate_filter_device

        m = self.get_modelvector()
        loss_initial = self.lossvalue(m)
        for iss, (d0, d, dcal) in enumerate(zip(self.dobs0, self.dobs, self.paf.data)):
            fields = list(dcal.fields)
            g = np.concatenate([np.asarray(d0.d[f], np.float64) for f in fields], axis=1)
            dc = np.concatenate([np.asarray(dcal.d[f], np.float64) for f in fields], axis=1)
            _, conv, _, _ = estimate_filter_device(self.ctx, g, dc, maxiter=maxiter)
            off = 0
            for f in fields:
                n = d.d[f].shape[1]
                d.d[f][:, :] = conv[:, off:off + n].astype(d.d[f].dtype)
                off += n
        loss_final = self.lossvalue(m)
        return {"loss_initial": loss_initial, "loss_final": loss_final}

    def update(self, maxiter: int = 20, callback=None, optimizer: str = "lbfgs", memory: int = 10):
        """``update!(paw)`` (src/fwi/fwi.jl:66-75): L-BFGS on the misfit from the current
        model (Optim.jl's LBFGS in the reference, memory 10).  ``optimizer="lbfgs"``: the
        two-loop recursion on the device (history in HBM), a strong-Wolfe line search on the
        host, alpha = 1 once the history holds a pair (1/||g|| before); ``"lbfgsb"``: SciPy's
        L-BFGS-B without bounds, whose history and compact-form products stay on the host.
        Returns a result with ``x``, ``fun``, ``nit``, ``nfev`` and ``host_s`` (seconds of
        optimizer work per iteration outside the misfit/gradient evaluations); the final
        model stays in the experiment."""
        m0 = self.get_modelvector()
        g = np.zeros_like(m0)
        t_eval = [0.0]
        nfev = [0]

        def fg(m):
            t0 = time.perf_counter()
            J = self.gradient(g, m)
            t_eval[0] += time.perf_counter() - t0
            nfev[0] += 1
            return J, g.astype(np.float64)   # a new FP64 array per call

        t_start = time.perf_counter()
        if optimizer == "lbfgsb":
            from scipy.optimize import minimize
            res = minimize(fg, m0, jac=True, method="L-BFGS-B", callback=callback,
                           options={"maxiter": maxiter, "gtol": 1e-25, "ftol": 0.0})
            x, fun, nit = res.x, float(res.fun), int(res.nit)
        elif optimizer == "lbfgs":
            opt = DeviceLBFGS(self.ctx, m0.size, memory)
            x = m0.copy()
            fun, gx = fg(x)
            nit = 0
            try:
                for nit in range(1, maxiter + 1):
                    d = opt.direction(x, gx)
                    if not np.dot(gx, d) < 0:   # not a descent direction: drop the history
                        opt.reset()
                        d = opt.direction(x, gx)
                    a0 = 1.0 if opt.pairs else min(1.0, 1.0 / max(float(np.linalg.norm(gx)), 1e-300))
                    a, f, gnew, _ = wolfe_search(fg, x, fun, gx, d, a0)
                    if a is None:
                        nit -= 1
                        break
                    x = x + a * d
                    fun, gx = f, gnew
                    if callback is not None:
                        callback(x)
            finally:
                opt.close()
        else:
            raise ValueError(f"unknown optimizer {optimizer!r}")
        total = time.perf_counter() - t_start
        self.lossvalue(x)   # leave the experiment at the final model
        return SimpleNamespace(x=x, fun=fun, nit=nit, nfev=nfev[0],
                               host_s=(total - t_eval[0]) / max(nit, 1))
