"""Host-side preparation (numpy) against the oracle's own preparation (C) on the same raw
inputs: coefficient arrays, CPML vectors, spray/interpolation weights, dt/dsI, freqpeak and
velavg.  Two independent restatements of the reference's host code must agree bit for bit."""
import numpy as np
import pytest

import geophyinv as G
import oracle as O
from geophyinv.grid import Grid
from helpers import experiment, layered_acoustic, layered_elastic, wavelet


def check_prep(pa):
    orc = O.from_experiment(pa)
    sc = orc.scalars()
    assert sc["dt"] == pa.dt32
    assert sc["dzI"] == pa.dsI[0] and sc["dxI"] == pa.dsI[-1]
    assert sc["freqpeak"] == pa.freqpeak
    assert sc["velavg"] == pa.velavg
    for name, a in pa.coefs.items():
        b = orc.coef(name)
        if name == "rho":
            continue
        assert b is not None, name
        assert a.shape == b.shape, name
        assert np.array_equal(a, b), (name, float(np.abs(a - b).max()))
    for df, (a, b, k) in pa.pml.items():
        oa, ob, ok = orc.pml(df)
        assert np.array_equal(a, oa) and np.array_equal(b, ob) and np.array_equal(k, ok), df
    for q, f in enumerate(pa.srcwav[0].fields):
        i, w = orc.weights(1, q)
        hi, hw = pa.spray[0][f]
        assert np.array_equal(i, hi) and np.array_equal(w, hw), f
    for q, f in enumerate(pa.rfields):
        i, w = orc.weights(0, q)
        hi, hw = pa.interp[0][f]
        assert np.array_equal(i, hi) and np.array_equal(w, hw), f
    return orc


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_acoustic_prep(order):
    check_prep(experiment(layered_acoustic(), order=order, sfields=("p", "vx", "vz"), ns=3, nr=7))


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_elastic_prep(order):
    check_prep(experiment(layered_elastic(), order=order, sfields=("tauxx", "vz"), rfields=("vx", "vz"),
                          ns=2, nr=7))


@pytest.mark.parametrize("faces", [["zmax", "xmin"], ["xmax"], []])
def test_prep_face_subsets(faces):
    check_prep(experiment(layered_acoustic(), pml_faces=faces, ns=1, nr=5))


def _experiment3d(physics, order, compat):
    rng = np.random.Generator(np.random.PCG64(5))
    n = (14, 12, 16)
    grid = [Grid(0.0, 10.0, n[0]), Grid(0.0, 10.0, n[1]), Grid(0.0, 10.0, n[2])]
    vp = (2000 + 300 * rng.random(n)).astype(np.float32)
    rho = (1800 + 300 * rng.random(n)).astype(np.float32)
    if physics == 0:
        med = G.AcousticMedium(grid, vp, rho)
        attrib, sf, rf = G.FdtdAcoustic(), ["p", "vy"], ["p", "vx", "vy", "vz"]
    else:
        med = G.ElasticMedium(grid, vp, (vp / 1.8).astype(np.float32), rho)
        attrib, sf, rf = G.FdtdElastic(), ["tauxx", "vy"], ["vx", "vy", "vz"]
    nt = 80
    dt = 1e-3
    tg = Grid(0.0, dt, nt)
    ag = G.AGeomss({"z": [33.0, 41.5], "y": [52.0, 60.2], "x": [71.0, 64.4]},
                   {"z": [0.0, 3.3, 7.1], "y": [12.0, 77.7, 100.0], "x": [150.0, 0.0, 99.9]})
    s = G.Srcs(2, tg, sf)
    s.update(sf, np.stack([wavelet(nt, dt, 30.0), wavelet(nt, dt, 28.0)], axis=1))
    return G.SeisForwExpt(attrib, medium=med, tgrid=tg, ageom=[ag], srcwav=[s], rfields=rf,
                          order=order, npml=4, compat=compat)


@pytest.mark.parametrize("physics", [0, 1])
@pytest.mark.parametrize("order", [2, 4])
@pytest.mark.parametrize("compat", [0, 1])
def test_prep_3d(physics, order, compat):
    check_prep(_experiment3d(physics, order, compat))


def test_vectorised_weights_equal_pointwise():
    """get_proj_matrix (all points at once) equals point_weights (one point at a time), inside
    and outside the grid, on nodes, 2-D and 3-D, both 3-D coordinate conventions."""
    from geophyinv.grid import Grid
    from geophyinv.projmat import get_proj_matrix, point_weights
    rng = np.random.default_rng(0)
    for nd in (2, 3):
        g = ([Grid(-3.0, 2.5, 40), Grid(1.0, 5.0, 33)] if nd == 2
             else [Grid(-3.0, 2.5, 20), Grid(0.0, 4.0, 17), Grid(1.0, 5.0, 23)])
        lo = np.array([x.first for x in g]) - 10
        hi = np.array([x.last for x in g]) + 10
        pts = rng.uniform(lo, hi, (300, nd))
        pts[:20] = np.array([x.values()[3] for x in g])
        for compat in (0, 1):
            I, W = get_proj_matrix(pts, g, compat)
            for k, p in enumerate(pts):
                i, w = point_weights(g, p, compat)
                assert np.array_equal(I[k], i) and np.array_equal(W[k], w), (nd, compat, k)


def test_illum_flag_checks():
    """illum_flag: the stack has the unpadded grid's shape and starts at zero; elastic media
    (no pressure field) and Born modelling reject it."""
    med = layered_acoustic(nz=30, nx=34)
    base = experiment(med, nt=120)
    pa = G.SeisForwExpt(base.attrib_mod, medium=med, tgrid=base.tgrid, ageom=base.ageom, srcwav=base.srcwav,
                        rfields=base.rfields, npml=base.npml, illum_flag=True)
    assert pa["illum_stack"].shape == (30, 34) and pa["illum_stack"].dtype == np.float64
    assert not pa["illum_stack"].any()
    emed = layered_elastic(nz=30, nx=34)
    eb = experiment(emed, nt=120, sfields=("tauxx",), rfields=("vz",))
    with pytest.raises(ValueError):
        G.SeisForwExpt(eb.attrib_mod, medium=emed, tgrid=eb.tgrid, ageom=eb.ageom, srcwav=eb.srcwav,
                       rfields=eb.rfields, npml=eb.npml, illum_flag=True)
    with pytest.raises(ValueError):
        G.SeisForwExpt(G.FdtdAcoustic[G.Born](), medium=med, tgrid=base.tgrid, ageom=base.ageom,
                       srcwav=base.srcwav, rfields=base.rfields, npml=base.npml, illum_flag=True)
