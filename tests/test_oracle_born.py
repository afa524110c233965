"""Pins of the oracle's Born modelling (tests/helpers.py: oracle_born), CPU only.

The Born records of a perturbation are the derivative of the nonlinear records along it:
they equal the central finite difference of the oracle's forward records along smooth
interior perturbations of invK (at fixed rho) and of rho (at fixed invK), to the
difference's truncation and FP32 noise.  A planted sign error in either secondary source
fails the comparison.
"""
import numpy as np
import pytest

import geophyinv as G
import oracle as O
from helpers import layered_acoustic, oracle_born, rel_l2, wavelet
from geophyinv.grid import Grid


def born_experiment(medium, nt=260, rfields=("p", "vx", "vz"), order=2, npml=10, seed=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    gz, gx = medium.grid
    Lz, Lx = gz.last, gx.last
    dt = 0.4 * gz.step / (1.6 * float(medium.vp.m.max()))
    tgrid = Grid(0.0, dt, nt)
    ag = G.AGeomss({"z": [0.2 * Lz + rng.uniform(-1, 1) * gz.step], "x": [0.5 * Lx]},
                   {"z": np.full(9, 0.8 * Lz), "x": np.linspace(0.1 * Lx, 0.9 * Lx, 9)})
    src = G.Srcs(1, tgrid, ["p"])
    src.update(["p"], wavelet(nt, dt, 25.0, 0.03) * 1e3)
    return G.SeisForwExpt(G.FdtdAcoustic[G.Born](), medium=medium, tgrid=tgrid, ageom=[ag], srcwav=[src],
                          rfields=list(rfields), order=order, npml=npml)


def forward_records(pa, medium):
    o = O.from_experiment(pa)
    o2 = O.Oracle(physics=0, order=pa.order, grid=medium.grid, vp=medium.vp.m, rho=medium.rho.m, npml=pa.npml,
                  pml_faces=pa.pml_faces, rigid_faces=pa.rigid_faces, nt=len(pa.tgrid), dt=pa.tgrid.step,
                  src_pos=pa.ageom[0].src_points(), src_fields=pa.srcwav[0].fields,
                  wavelets=[pa.srcwav[0].d[f] for f in pa.srcwav[0].fields], rec_pos=pa.ageom[0].rec_points(),
                  rec_fields=pa.rfields, freqpeak=float(pa.pml_freqpeak))
    del o
    o2.run()
    return {f: o2.records(q).astype(np.float64) for q, f in enumerate(pa.rfields)}


def blob(shape, cz, cx, rad):
    zz, xx = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    return np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2.0 * rad ** 2))


@pytest.mark.parametrize("param", ["invK", "rho"])
def test_born_records_are_the_derivative_of_the_records(param):
    med = layered_acoustic(nz=40, nx=52, seed=7)
    pa = born_experiment(med)
    n = med.vp.m.shape
    b = blob(n, 22, 26, 3.5)
    vp0 = med.vp.m.astype(np.float64)
    rho0 = med.rho.m.astype(np.float64)
    invK0 = 1.0 / (vp0 ** 2 * rho0)
    K0 = 1.0 / invK0
    eps = 1e-2

    def medium_at(t):
        m = med.copy()
        if param == "invK":   # invK (1 + t b) at fixed rho
            m.vp.m[:] = (vp0 * (1 + t * b) ** -0.5).astype(np.float32)
        else:                 # rho (1 + t b) at fixed invK
            r = rho0 * (1 + t * b)
            m.rho.m[:] = r.astype(np.float32)
            m.vp.m[:] = np.sqrt(K0 / r).astype(np.float32)
        return m

    dp, dm = forward_records(pa, medium_at(eps)), forward_records(pa, medium_at(-eps))
    dmod = {k: np.zeros(pa.ext_shape, np.float64, order="F") for k in ("invK", "rho")}
    inner = pa.view_inner(dmod["invK"])
    base = pa.mod[param][inner].astype(np.float64)
    dmod[param][inner] = base * b
    pa.set_perturbation(dmod)
    born = oracle_born(pa)[0]
    for f in pa.rfields:
        fd = (dp[f] - dm[f]) / (2 * eps)
        e = rel_l2(born[f], fd)
        assert np.abs(fd).max() > 0
        print(f"{param} {f}: Born vs central difference rel L2 {e:.2e}")
        assert e < 2e-3, (param, f, e)
    # teeth: a flipped secondary source is far off
    pa.born_dcoefs = {k: (-v if (param == "invK") == (k == "dtK") else v) for k, v in pa.born_dcoefs.items()}
    flipped = oracle_born(pa)[0]
    assert rel_l2(flipped["vz"], (dp["vz"] - dm["vz"]) / (2 * eps)) > 0.5
