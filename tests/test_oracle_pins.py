"""Pins of the CPU oracle by mathematics that does not restate its code.

- stencils: exact on polynomials of degree < order, for both staggered variants;
- 2-D acoustic records against the analytic Green's function (Hankel H0^(2));
- 2-D elastic explosive source against the acoustic solution: on a staggered grid in a
  homogeneous medium the explosive P-wave velocity equals minus the acoustic velocity
  (discrete Helmholtz decomposition), so the two oracles must agree to FP32 rounding
  until the wavefront reaches the CPML;
- CPML coefficient invariants; linearity in the wavelet; energy decay in the CPML.
"""
import numpy as np
import pytest
from scipy.special import hankel2

import geophyinv as G
import oracle as O
from geophyinv.grid import Grid
from helpers import rel_l2, wavelet


def homogeneous(physics, n=121, ds=5.0, vp=2000.0, vs=1150.0, rho=2000.0):
    grid = [Grid(0.0, ds, n), Grid(0.0, ds, n)]
    a = np.full((n, n), vp, np.float32)
    r = np.full((n, n), rho, np.float32)
    if physics == 0:
        return G.AcousticMedium(grid, a, r)
    return G.ElasticMedium(grid, a, np.full((n, n), vs, np.float32), r)


def run(medium, *, sfield, rfields, spos, rpos, nt, dt, w, order=2, npml=20, faces=None):
    faces = ["zmin", "zmax", "xmin", "xmax"] if faces is None else faces
    orc = O.Oracle(physics=0 if isinstance(medium, G.AcousticMedium) else 1, order=order,
                   grid=medium.grid, vp=medium.vp.m, rho=medium.rho.m,
                   vs=medium.vs.m if isinstance(medium, G.ElasticMedium) else None,
                   npml=npml, pml_faces=faces, rigid_faces=faces, nt=nt, dt=dt,
                   src_pos=[spos], src_fields=[sfield], wavelets=[w[:, None]], rec_pos=rpos,
                   rec_fields=rfields)
    orc.run()
    return orc, {f: orc.records(q) for q, f in enumerate(rfields)}


# ------------------------------------------------------------------------------ stencils
@pytest.mark.parametrize("order", [2, 4, 6, 8])
@pytest.mark.parametrize("deriv,src,axis,kind", [
    ("dpdx", "p", 1, "i"), ("dpdz", "p", 0, "i"), ("dvxdx", "vx", 1, "a"), ("dvzdz", "vz", 0, "a")])
def test_stencil_polynomial_exactness(order, deriv, src, axis, kind):
    """D applied to x^k (k < order) equals k x^(k-1) at the staggered midpoint."""
    med = homogeneous(0, n=30, ds=1.0)
    nt = 8
    orc, _ = run(med, sfield="p", rfields=["p"], spos=(10.0, 10.0), rpos=[(5.0, 5.0)], nt=nt, dt=1e-4,
                 w=np.zeros(nt, np.float32), order=order, npml=4)
    gz, gx = med.grid
    ext = orc.n3
    h = order - 1
    grids = G.get_mgrid(src, 0, [Grid(-4.0, 1.0, ext[0]), Grid(-4.0, 1.0, ext[2])], order)
    dgr = G.get_mgrid(deriv, 0, [Grid(-4.0, 1.0, ext[0]), Grid(-4.0, 1.0, ext[2])], order)
    for k in range(order):
        coord = grids[axis].values()
        f = np.broadcast_to((coord / 8.0) ** k if axis == 1 else ((coord / 8.0) ** k)[:, None],
                            (len(grids[0]), len(grids[1])))
        orc.set_field(src, f)
        d = orc.derivative(deriv).astype(np.float64)
        dc = dgr[axis].values()
        exact = (k * (dc / 8.0) ** (k - 1) / 8.0) if k > 0 else np.zeros_like(dc)
        exact = np.broadcast_to(exact if axis == 1 else exact[:, None], d.shape)
        scale = max(1.0, float(np.abs(exact).max()))
        assert np.abs(d - exact).max() <= 2e-4 * scale * order, (order, k)
    # a wrong shift of the orthogonal axis ('i' variant) would sample a different row/column
    assert h >= 1


# ------------------------------------------------------------------------------ Green's function
def green_errors(order, ds, offs=(150.0, 250.0)):
    """Relative L2 misfit of oracle p records of a point pressure source in a homogeneous
    medium against the analytic 2-D solution P(w) = rho w F(w) H0^(2)(w r / c) / 4 with
    F = w(t - dt/2) dx dz (the source enters the stress update half a step late)."""
    c, rho = 2000.0, 2000.0
    n = int(round(800.0 / ds)) + 1
    med = homogeneous(0, n=n, ds=ds, vp=c, rho=rho)
    dt = 0.35 * 5.0 / c * (ds / 5.0)
    nt = int(0.32 / dt)
    w = wavelet(nt, dt, 15.0, 0.08).astype(np.float32)
    L = (n - 1) * ds
    recs = [(L / 2, L / 2 + o) for o in offs]
    _, r = run(med, sfield="p", rfields=["p"], spos=(L / 2, L / 2), rpos=recs, nt=nt, dt=dt, w=w,
               order=order)
    nfft = 8 * nt
    om = 2 * np.pi * np.fft.rfftfreq(nfft, dt)
    F = np.fft.rfft(w.astype(np.float64), nfft) * ds * ds * np.exp(-1j * om * dt / 2)
    errs = []
    for k, o in enumerate(offs):
        P = np.zeros_like(F)
        P[1:] = rho * om[1:] * F[1:] * hankel2(0, om[1:] * o / c) / 4.0
        p = np.fft.irfft(P, nfft)[:nt]
        m = int(0.3 / dt)   # before any CPML/boundary return reaches the receiver
        errs.append(rel_l2(r["p"][:m, k], p[:m]))
        # polarity: a sign mistake gives a negative correlation
        assert np.dot(r["p"][:m, k], p[:m]) > 0
    return errs


@pytest.mark.parametrize("order,tol", [(2, 8e-2), (4, 1e-2)])
def test_acoustic_green_function(order, tol):
    coarse = green_errors(order, 5.0)
    assert max(coarse) < tol, coarse
    # grid refinement: the misfit falls at least like the scheme's spatial order allows
    fine = green_errors(order, 2.5)
    assert all(f < c / 3.0 for f, c in zip(fine, coarse)), (coarse, fine)


# ------------------------------------------------------------------------------ elastic vs acoustic
@pytest.mark.parametrize("order", [2, 4])
def test_elastic_explosion_equals_minus_acoustic(order):
    ds, n = 5.0, 121
    c, rho = 2000.0, 2000.0
    ac = homogeneous(0, n=n, ds=ds, vp=c, rho=rho)
    el = homogeneous(1, n=n, ds=ds, vp=c, vs=1150.0, rho=rho)
    dt = 0.3 * ds / c
    nt = 240
    w = wavelet(nt, dt, 25.0, 0.04).astype(np.float32)
    L = (n - 1) * ds
    src = (L / 2 + 1.3, L / 2 - 2.1)
    recs = [(L / 2 - 60.0, L / 2 + 2.0), (L / 2 + 3.0, L / 2 + 75.0), (L / 2 + 40.0, L / 2 - 50.0)]
    _, ra = run(ac, sfield="p", rfields=["vx", "vz"], spos=src, rpos=recs, nt=nt, dt=dt, w=w, order=order)
    _, re = run(el, sfield="tauxx", rfields=["vx", "vz"], spos=src, rpos=recs, nt=nt, dt=dt, w=w, order=order)
    # window: the P front needs (L/2 - npml-free margin) / c to reach the CPML strips
    m = int(((L / 2) - 2 * ds) / c / dt)
    for f in ("vx", "vz"):
        e = rel_l2(re[f][:m], -ra[f][:m])
        assert e < 2e-4, (order, f, e)
        assert np.abs(ra[f][:m]).max() > 0


# ------------------------------------------------------------------------------ CPML
def test_cpml_coefficient_invariants():
    med = homogeneous(0, n=61, ds=5.0)
    nt = 100
    orc, _ = run(med, sfield="p", rfields=["p"], spos=(150.0, 150.0), rpos=[(0.0, 0.0)], nt=nt, dt=1e-3,
                 w=wavelet(nt, 1e-3, 20.0).astype(np.float32), npml=12)
    npml = 12
    for df in ("dpdx", "dpdz", "dvxdx", "dvzdz"):
        a, b, k = orc.pml(df)
        assert np.all(k == 1.0)                       # K_MAX = 1
        assert np.all((b > 0) & (b <= 1.0)) and np.all(a <= 0.0)
        # a/(b-1) = d/(d+alpha): zero where d = 0, below 1, and growing towards the outer
        # edge (d grows as x^2 while alpha falls linearly); index 0 is the outer min edge
        damp = b < 1.0
        r = np.zeros_like(a, dtype=np.float64)
        r[damp] = a[damp].astype(np.float64) / (b[damp].astype(np.float64) - 1.0)
        assert np.all((r >= 0) & (r < 1))
        assert np.all(np.diff(r[:npml]) <= 1e-6) and np.all(np.diff(r[npml:]) >= -1e-6)
        # no damping term next to the medium (the profile starts 2 cells outside it)
        assert a[npml - 1] == 0.0 and a[npml] == 0.0
        # outermost cells are strongly damped
        assert r[0] > 0.5 and r[-1] > 0.5
        # symmetric grid: min and max profiles of the node-centred fields mirror each other
        if df in ("dvxdx", "dvzdz"):
            np.testing.assert_allclose(b[:npml], b[npml:][::-1], rtol=1e-6)
            np.testing.assert_allclose(a[:npml], a[npml:][::-1], rtol=1e-5, atol=1e-7)


def test_linearity_in_the_wavelet():
    med = homogeneous(0, n=61, ds=5.0)
    nt = 150
    w = wavelet(nt, 1e-3, 20.0).astype(np.float32)
    kw = dict(sfield="p", rfields=["p", "vz"], spos=(100.0, 140.0), rpos=[(20.0, 50.0), (200.0, 250.0)],
              nt=nt, dt=1e-3, npml=10)
    _, r1 = run(med, w=w, **kw)
    _, r2 = run(med, w=2 * w, **kw)
    for f in r1:
        assert np.array_equal(2 * r1[f], r2[f])    # scaling by 2 is exact in binary FP


def test_energy_decays_in_the_cpml():
    med = homogeneous(0, n=81, ds=5.0)
    nt = 900
    dt = 1e-3
    w = wavelet(nt, dt, 20.0, 0.05).astype(np.float32)
    ener = []
    for steps in (250, 500, 900):
        orc = O.Oracle(physics=0, order=2, grid=med.grid, vp=med.vp.m, rho=med.rho.m, npml=15,
                       pml_faces=["zmin", "zmax", "xmin", "xmax"], rigid_faces=[], nt=nt, dt=dt,
                       src_pos=[(200.0, 200.0)], src_fields=["p"], wavelets=[w[:, None]],
                       rec_pos=[(0.0, 0.0)], rec_fields=["p"])
        orc.run(steps)
        ener.append(float(np.sum(orc.field("p").astype(np.float64) ** 2)))
    assert ener[1] < 1e-2 * ener[0] and ener[2] < 0.5 * ener[1]


def test_oracle_rejects_bad_order():
    with pytest.raises(RuntimeError):
        O.Oracle(physics=0, order=3, grid=[Grid(0, 1, 10), Grid(0, 1, 10)], vp=np.ones((10, 10)),
                 rho=np.ones((10, 10)), npml=2, pml_faces=[], rigid_faces=[], nt=4, dt=1e-3,
                 src_pos=[(1.0, 1.0)], src_fields=["p"], wavelets=[np.zeros((4, 1))],
                 rec_pos=[(1.0, 1.0)], rec_fields=["p"])
