"""Source-signature estimation on the GPU (gpi_conv_apply, gpi_conv_lsmr: batched hipFFT,
FP64) against the host restatement of src/Conv.jl (conv.Pconv with NumPy FFTs) and SciPy's
LSMR on conv.operator with the same tolerances (src/fwi/fwi.jl:37-64)."""
import numpy as np
import pytest
from scipy.sparse.linalg import lsmr

from geophyinv import conv

pytestmark = pytest.mark.gpu


def pconv(g, s=None):
    nt, ntr = g.shape
    return conv.Pconv(dsize=(nt, ntr), ssize=(nt,), gsize=(nt, ntr), g=g, s=s)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("nt,ntr", [(1, 1), (2, 3), (57, 4), (300, 6), (4000, 2403)])
def test_operator_equals_the_host(gpu_ctx, nt, ntr):
    """d = g * s and its adjoint against Pconv.mod_D / mod_S, and the dot-product test;
    (4000, 2403): the 2001^2 elastic workload's 4000 samples and 3 x 801 traces."""
    rng = np.random.Generator(np.random.PCG64(nt + ntr))
    g = rng.standard_normal((nt, ntr))
    s = rng.standard_normal(nt)
    d = rng.standard_normal((nt, ntr))
    pa = pconv(g, s)
    ref_d = pa.mod_D().copy()
    pa.d = d
    ref_s = pa.mod_S().copy()
    dev_d = conv.apply_device(gpu_ctx, g, s, "D")
    dev_s = conv.apply_device(gpu_ctx, g, d, "S")
    assert rel(dev_d, ref_d) <= 1e-12
    assert rel(dev_s, ref_s) <= 1e-12
    assert abs(np.vdot(dev_d, d) - np.vdot(s, dev_s)) <= 1e-12 * abs(np.vdot(dev_d, d))


def planted(nt, ntr, rng, band_limited):
    sl = (nt - 1) - (nt - 1) // 2
    w = np.zeros(nt)
    w[sl + 7], w[sl + 8] = 1.7, -0.4
    if band_limited:
        t = np.arange(31)
        pulse = np.exp(-0.5 * ((t - 15) / 4.0) ** 2)
        g = np.stack([np.convolve(rng.standard_normal(nt), pulse)[:nt] for _ in range(ntr)], axis=1)
    else:
        g = rng.standard_normal((nt, ntr))
    d = pconv(g, w).mod_D() + 1e-3 * rng.standard_normal((nt, ntr))
    return g, d, w


def solve_both(ctx, g, d, maxiter):
    sd, gs, it, stop = conv.estimate_filter_device(ctx, g, d, maxiter=maxiter)
    ref = lsmr(conv.operator(pconv(g), "s"), d.ravel(order="F"), atol=conv.LSMR_TOL, btol=conv.LSMR_TOL,
               conlim=conv.LSMR_CONLIM, maxiter=maxiter)
    assert rel(gs, pconv(g, sd).mod_D()) <= 1e-12
    return sd, gs, it, stop, ref


@pytest.mark.parametrize("maxiter", [3, 100])
def test_lsmr_equals_scipy_when_well_conditioned(gpu_ctx, maxiter):
    g, d, w = planted(240, 5, np.random.Generator(np.random.PCG64(81)), band_limited=False)
    sd, gs, it, stop, ref = solve_both(gpu_ctx, g, d, maxiter)
    assert (stop, it) == (ref[1], ref[2])
    assert rel(sd, ref[0]) <= 1e-9
    if maxiter == 100:
        assert stop in (1, 2) and rel(sd, w) < 0.01


def device_operator(ctx, g):
    from scipy.sparse.linalg import LinearOperator
    nt, ntr = g.shape
    return LinearOperator((nt * ntr, nt), matvec=lambda x: conv.apply_device(ctx, g, x, "D").ravel(order="F"),
                          rmatvec=lambda y: conv.apply_device(ctx, g, y.reshape(nt, ntr, order="F"), "S"),
                          dtype=np.float64)


def test_lsmr_ill_conditioned(gpu_ctx):
    """Band-limited traces make an ill-conditioned operator, on which LSMR's iterates depend
    on the operator's rounding once orthogonality is lost.  So: the first iterations equal
    SciPy's on the host operator; 20 iterations equal SciPy's LSMR driving the device
    operator (the recurrences are the same, step for step); the full solve fits the data
    about as well as SciPy's on the host operator."""
    g, d, w = planted(240, 5, np.random.Generator(np.random.PCG64(82)), band_limited=True)
    sd, gs, it, stop, ref = solve_both(gpu_ctx, g, d, 6)
    assert (stop, it) == (7, 6) == (ref[1], ref[2])
    assert rel(sd, ref[0]) <= 1e-12
    sd, gs, it, stop = conv.estimate_filter_device(gpu_ctx, g, d, maxiter=20)
    ref = lsmr(device_operator(gpu_ctx, g), d.ravel(order="F"), atol=conv.LSMR_TOL, btol=conv.LSMR_TOL,
               conlim=conv.LSMR_CONLIM, maxiter=20)
    assert (stop, it) == (ref[1], ref[2])
    assert rel(sd, ref[0]) <= 1e-9
    sd, gs, it, stop, ref = solve_both(gpu_ctx, g, d, 100)
    assert np.linalg.norm(gs - d) <= 1.3 * ref[3]


def test_lsmr_at_the_drop_in_size(gpu_ctx):
    """4000 samples, 2403 traces: the planted filter is recovered and the normal equations
    hold at the stopping point (SciPy on the host takes minutes at this size)."""
    g, d, w = planted(4000, 2403, np.random.Generator(np.random.PCG64(83)), band_limited=False)
    sd, gs, it, stop = conv.estimate_filter_device(gpu_ctx, g, d, maxiter=100)
    assert stop in (1, 2) and it < 100
    assert rel(sd, w) < 1e-3
    pa = pconv(g)
    pa.d = gs - d
    normal = pa.mod_S()
    pa.d = d
    assert np.linalg.norm(normal) <= 1e-3 * np.linalg.norm(pa.mod_S())
