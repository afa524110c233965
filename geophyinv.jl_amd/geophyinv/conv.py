"""FFT convolution with lags and its linear operators (src/Conv.jl), for source-signature
estimation (``update!(paw, pac)``, src/fwi/fwi.jl:37-64).

``Pconv(dsize, ssize, gsize)`` holds d = g * s (convolution along the first axis).  Each of
d, g, s is a signal of ``nplags + nnlags + 1`` samples: first the negative lags in
decreasing order, then lag 0 at index ``nnlags``, then the positive lags (Conv.jl:154-168).
Defaults: d and g have only non-negative lags, s has as many negative as positive lags
(Conv.jl:83-101); the FFT length is the next 2-3-5-7-smooth number at or above
``len(s) + len(g) - 1`` (DSP.nextfastfft), so the circular convolution is the linear one.
A dimension beyond the first is a trace axis; when one of the three signals has a single
trace it is shared by all traces (and the products into it are summed over traces),
Conv.jl:296-420.  ``Pconv`` works on the host in FP64 with NumPy's real FFTs;
``apply_device`` and ``estimate_filter_device`` run the shared-filter operator and its
LSMR solve on the GPU (``gpi_conv_apply``, ``gpi_conv_lsmr``: batched hipFFT, FP64).
"""
from __future__ import annotations

import numpy as np


def nextfastfft(n: int) -> int:
    """Smallest m >= n whose prime factors are 2, 3, 5 and 7 (DSP.nextfastfft)."""
    m = max(int(n), 1)
    while True:
        k = m
        for p in (2, 3, 5, 7):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def _default_lags(n: int, symmetric: bool):
    if symmetric:
        p = (n - 1) // 2
        return p, n - 1 - p
    return n - 1, 0


def pad(x: np.ndarray, nplags: int, nnlags: int, npow2: int) -> np.ndarray:
    """``pad!`` (Conv.jl:170-197): lag 0 at index 0, positive lags after it, negative lags
    wrapped to the end."""
    if x.shape[0] != nplags + nnlags + 1:
        raise ValueError("pad: signal length differs from its lags")
    out = np.zeros((npow2,) + x.shape[1:], np.float64)
    out[0:nplags + 1] = x[nnlags:nnlags + nplags + 1]
    if nnlags:
        out[npow2 - nnlags:] = x[0:nnlags]
    return out


def truncate(xpow2: np.ndarray, nplags: int, nnlags: int) -> np.ndarray:
    """``truncate!`` (Conv.jl:199-224), the inverse of ``pad``."""
    n = nplags + nnlags + 1
    npow2 = xpow2.shape[0]
    out = np.zeros((n,) + xpow2.shape[1:], np.float64)
    out[nnlags:nnlags + nplags + 1] = xpow2[0:nplags + 1]
    if nnlags:
        out[0:nnlags] = xpow2[npow2 - nnlags:]
    return out


class Pconv:
    """``Pconv(T; dsize, ssize, gsize, d, s, g, slags, dlags, glags, np2)`` (Conv.jl:59-150)."""

    def __init__(self, dsize, ssize, gsize, d=None, s=None, g=None, slags=None, dlags=None, glags=None, np2=None):
        self.d = np.zeros(dsize) if d is None else np.array(d, np.float64)
        self.s = np.zeros(ssize) if s is None else np.array(s, np.float64)
        self.g = np.zeros(gsize) if g is None else np.array(g, np.float64)
        if self.d.shape != tuple(dsize) or self.s.shape != tuple(ssize) or self.g.shape != tuple(gsize):
            raise ValueError("Pconv: arrays and sizes differ")
        self.slags = tuple(slags) if slags is not None else _default_lags(ssize[0], True)
        self.dlags = tuple(dlags) if dlags is not None else _default_lags(dsize[0], False)
        self.glags = tuple(glags) if glags is not None else _default_lags(gsize[0], False)
        for name, lags, n in (("glags", self.glags, gsize[0]), ("dlags", self.dlags, dsize[0]),
                              ("slags", self.slags, ssize[0])):
            if sum(lags) + 1 != n:
                raise ValueError(name)
        self.np2 = int(np2) if np2 is not None else nextfastfft(ssize[0] + gsize[0] - 1)
        nd, ng, ns = len(dsize), len(gsize), len(ssize)
        if not (nd == ng == ns or (nd, ng, ns) in ((1, 2, 2), (2, 1, 2), (2, 2, 1))):
            raise ValueError("Pconv: unsupported combination of dimensions")
        self.kind = (nd, ng, ns)

    def _f(self, x, lags):
        return np.fft.rfft(pad(x, lags[0], lags[1], self.np2), axis=0)

    def _i(self, X, lags):
        return truncate(np.fft.irfft(X, self.np2, axis=0), lags[0], lags[1])

    @staticmethod
    def _b(X, ndim_target):
        """Broadcast a single-trace spectrum against trace spectra."""
        return X if X.ndim == ndim_target else X[:, None]

    def mod_D(self):
        """d = g * s (``mod!(pa, D())``)."""
        G, S = self._f(self.g, self.glags), self._f(self.s, self.slags)
        if self.kind == (1, 2, 2):   # dsum: the traces of g * s are summed into d
            D = np.sum(G * S, axis=1)
        else:
            D = self._b(G, 2 if self.d.ndim == 2 else 1) * self._b(S, 2 if self.d.ndim == 2 else 1)
        self.d = self._i(D, self.dlags)
        return self.d

    def mod_G(self):
        """g = d x s (correlation, ``mod!(pa, G())``)."""
        D, S = self._f(self.d, self.dlags), np.conj(self._f(self.s, self.slags))
        if self.kind == (2, 1, 2):   # gsum
            Gf = np.sum(D * S, axis=1)
        else:
            Gf = self._b(D, 2 if self.g.ndim == 2 else 1) * self._b(S, 2 if self.g.ndim == 2 else 1)
        self.g = self._i(Gf, self.glags)
        return self.g

    def mod_S(self):
        """s = d x g (correlation, ``mod!(pa, S())``)."""
        D, Gc = self._f(self.d, self.dlags), np.conj(self._f(self.g, self.glags))
        if self.kind == (2, 2, 1):   # ssum
            Sf = np.sum(D * Gc, axis=1)
        else:
            Sf = self._b(D, 2 if self.s.ndim == 2 else 1) * self._b(Gc, 2 if self.s.ndim == 2 else 1)
        self.s = self._i(Sf, self.slags)
        return self.s


def operator(pa: Pconv, unknown: str):
    """``Conv.operator(pa, G())`` (unknown ``"s"``: s -> d with g fixed) or ``Conv.operator(pa,
    S())`` (unknown ``"g"``: g -> d with s fixed), as a SciPy LinearOperator on flattened
    (column-major) vectors (Conv.jl:423-480)."""
    from scipy.sparse.linalg import LinearOperator

    if unknown not in ("s", "g"):
        raise ValueError("unknown must be 's' or 'g'")
    xshape = pa.s.shape if unknown == "s" else pa.g.shape
    dshape = pa.d.shape

    def fw(x):
        setattr(pa, unknown, np.asarray(x, np.float64).reshape(xshape, order="F"))
        return pa.mod_D().ravel(order="F")

    def bk(y):
        pa.d = np.asarray(y, np.float64).reshape(dshape, order="F")
        out = pa.mod_S() if unknown == "s" else pa.mod_G()
        return out.ravel(order="F")

    n = int(np.prod(xshape))
    m = int(np.prod(dshape))
    return LinearOperator((m, n), matvec=fw, rmatvec=bk, dtype=np.float64)


# The stopping tolerances of IterativeSolvers.lsmr! (version 0.9.4, pinned by the
# reference's Manifest.toml and not vendored; its documented keyword defaults, those of
# Fong and Saunders' LSMR): atol = btol = 1e-6, conlim = 1e8.  update!(paw, pac) passes
# only maxiter = 100 (src/fwi/fwi.jl:52).
LSMR_TOL = 1e-6
LSMR_CONLIM = 1e8


def estimate_filter(g: np.ndarray, d: np.ndarray, maxiter: int = 100, tol: float | None = None,
                    conlim: float = LSMR_CONLIM):
    """The filter s minimising ||g * s - d|| by LSMR from zero (IterativeSolvers.lsmr!,
    maxiter 100, its default tolerances), and g * s, on the host
    (SciPy's LSMR on ``operator``).  ``g``, ``d``: nt x ntraces; s has nt samples with
    symmetric lags."""
    from scipy.sparse.linalg import lsmr

    nt = g.shape[0]
    pa = Pconv(dsize=d.shape, ssize=(nt,), gsize=g.shape, g=g)
    A = operator(pa, "s")
    t = LSMR_TOL if tol is None else tol
    w = lsmr(A, np.asarray(d, np.float64).ravel(order="F"), atol=t, btol=t, conlim=conlim, maxiter=maxiter)[0]
    pa.s = w
    return w, pa.mod_D()


def _traces(a: np.ndarray) -> np.ndarray:
    a = np.asfortranarray(a, dtype=np.float64)
    return a.reshape(a.shape[0], -1, order="F")


def apply_device(ctx, g: np.ndarray, x: np.ndarray, mode: str) -> np.ndarray:
    """The shared-filter operator on the GPU: ``mode="D"`` d = g * x (x: nt filter samples,
    symmetric lags), ``mode="S"`` the adjoint s = sum over traces of x correlated with g
    (x: nt x ntraces); as ``Pconv.mod_D`` / ``mod_S`` with dsize = gsize = (nt, ntraces)."""
    from . import _lib

    g = _traces(g)
    nt, ntr = g.shape
    if mode == "D":
        xin = np.ascontiguousarray(x, dtype=np.float64).ravel()
        if xin.size != nt:
            raise ValueError("apply_device: the filter needs nt samples")
        out = np.zeros((nt, ntr), np.float64, order="F")
    elif mode == "S":
        xin = _traces(x)
        if xin.shape != (nt, ntr):
            raise ValueError("apply_device: data and g differ in shape")
        out = np.zeros(nt, np.float64)
    else:
        raise ValueError("mode must be 'D' or 'S'")
    ctx.check(ctx.lib.gpi_conv_apply(ctx.h, nt, ntr, _lib.dptr(g), 0 if mode == "D" else 1, _lib.dptr(xin),
                                     _lib.dptr(out)))
    return out


def estimate_filter_device(ctx, g: np.ndarray, d: np.ndarray, maxiter: int = 100, tol: float | None = None,
                           conlim: float = LSMR_CONLIM):
    """``estimate_filter`` on the GPU (``gpi_conv_lsmr``): the operator by batched hipFFT
    and LSMR with the long vectors in HBM, in FP64.  Returns (s, g * s, iterations, istop)."""
    import ctypes as C

    from . import _lib

    g = _traces(g)
    dd = _traces(d)
    if dd.shape != g.shape:
        raise ValueError("estimate_filter_device: g and d differ in shape")
    nt, ntr = g.shape
    t = LSMR_TOL if tol is None else float(tol)
    s = np.zeros(nt, np.float64)
    gs = np.zeros((nt, ntr), np.float64, order="F")
    it, stop = C.c_int32(0), C.c_int32(0)
    ctx.check(ctx.lib.gpi_conv_lsmr(ctx.h, nt, ntr, _lib.dptr(g), _lib.dptr(dd), int(maxiter), t, t, float(conlim),
                                    _lib.dptr(s), _lib.dptr(gs), C.byref(it), C.byref(stop)))
    return s, gs, int(it.value), int(stop.value)
