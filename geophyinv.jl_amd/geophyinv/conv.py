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
Conv.jl:296-420.  Work is done on the host in FP64 with NumPy's real FFTs.
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


def estimate_filter(g: np.ndarray, d: np.ndarray, maxiter: int = 100, tol: float | None = None):
    """The filter s minimising ||g * s - d|| by LSMR from zero (IterativeSolvers.lsmr!,
    maxiter 100; its default tolerances sqrt(eps) of the Float32 number type), and g * s.
    ``g``, ``d``: nt x ntraces; s has nt samples with symmetric lags."""
    from scipy.sparse.linalg import lsmr

    nt = g.shape[0]
    pa = Pconv(dsize=d.shape, ssize=(nt,), gsize=g.shape, g=g)
    A = operator(pa, "s")
    t = float(np.sqrt(np.finfo(np.float32).eps)) if tol is None else tol
    w = lsmr(A, np.asarray(d, np.float64).ravel(order="F"), atol=t, btol=t, maxiter=maxiter)[0]
    pa.s = w
    return w, pa.mod_D()
