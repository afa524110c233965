"""Convolution with lags and source-signature estimation (geophyinv/conv.py, src/Conv.jl),
CPU only: padding round trip, the FFT convolution against direct linear convolution with
the lag conventions, adjoint tests of both operators (Conv.jl's adj_test), the notebook's
recovery of a 3-sample filter by LSMR, and recovery of a smooth filter from many traces."""
import numpy as np
import pytest

from geophyinv.conv import Pconv, estimate_filter, nextfastfft, operator, pad, truncate


def test_nextfastfft():
    assert [nextfastfft(n) for n in (1, 11, 13, 17, 97, 121, 1999)] == [1, 12, 14, 18, 98, 125, 2000]


@pytest.mark.parametrize("lags", [(4, 0), (0, 4), (2, 3)])
def test_pad_truncate_round_trip(lags):
    rng = np.random.Generator(np.random.PCG64(1))
    x = rng.standard_normal((sum(lags) + 1, 3))
    xp = pad(x, lags[0], lags[1], 16)
    assert np.array_equal(truncate(xp, lags[0], lags[1]), x)
    assert xp[0, 0] == x[lags[1], 0]   # lag 0 first


def direct(g, s, slags, nt):
    q = slags[1]
    out = np.zeros((nt, g.shape[1]))
    for r in range(g.shape[1]):
        full = np.convolve(g[:, r], s)
        out[:, r] = full[q:q + nt]
    return out


@pytest.mark.parametrize("ns", [3, 8, 41])
def test_fft_convolution_equals_direct(ns):
    rng = np.random.Generator(np.random.PCG64(ns))
    nt, nr = 60, 4
    g = rng.standard_normal((nt, nr))
    s = rng.standard_normal(ns)
    pa = Pconv(dsize=(nt, nr), ssize=(ns,), gsize=(nt, nr), g=g, s=s)
    d = pa.mod_D()
    assert np.allclose(d, direct(g, s, pa.slags, nt), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("unknown", ["s", "g"])
def test_operator_adjoint(unknown):
    rng = np.random.Generator(np.random.PCG64(7))
    nt, nr, ns = 50, 3, 9
    pa = Pconv(dsize=(nt, nr), ssize=(ns,), gsize=(nt, nr), g=rng.standard_normal((nt, nr)),
               s=rng.standard_normal(ns))
    A = operator(pa, unknown)
    x = rng.standard_normal(A.shape[1])
    y = rng.standard_normal(A.shape[0])
    a = float(np.dot(y, A.matvec(x)))
    b = float(np.dot(x, A.rmatvec(y)))
    c = float(np.dot(x, A.rmatvec(A.matvec(x))))
    assert abs(a - b) <= 1e-10 * abs(a) and c > 0


def test_notebook_filter_recovery():
    """Conv.jl's notebook: Pconv(dsize=(10,2), gsize=(10,2), ssize=(3,)); dtrue = A*strue;
    lsmr! recovers strue."""
    from scipy.sparse.linalg import lsmr
    rng = np.random.Generator(np.random.PCG64(3))
    pa = Pconv(dsize=(10, 2), ssize=(3,), gsize=(10, 2), g=rng.standard_normal((10, 2)))
    A = operator(pa, "s")
    strue = rng.standard_normal(3)
    dtrue = A.matvec(strue)
    w = lsmr(A, dtrue, atol=1e-12, btol=1e-12)[0]
    assert np.allclose(w, strue, rtol=1e-8, atol=1e-10)


def test_estimate_filter_recovers_a_smooth_filter():
    rng = np.random.Generator(np.random.PCG64(5))
    nt, nr = 120, 16
    g = rng.standard_normal((nt, nr))
    lag = np.arange(nt) - (nt - 1 - (nt - 1) // 2)      # symmetric lags of the filter
    strue = np.exp(-((lag - 3.0) / 2.5) ** 2) * 0.8      # a delayed, scaled pulse
    pa = Pconv(dsize=(nt, nr), ssize=(nt,), gsize=(nt, nr), g=g, s=strue)
    d = pa.mod_D()
    w, dfit = estimate_filter(g, d, maxiter=200, tol=1e-10)
    assert np.linalg.norm(w - strue) / np.linalg.norm(strue) < 1e-6
    assert np.linalg.norm(dfit - d) / np.linalg.norm(d) < 1e-8
