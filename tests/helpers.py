"""Shared builders for the tests: small experiments and comparison metrics."""
from __future__ import annotations

import numpy as np

import geophyinv as G
from geophyinv.grid import Grid


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def wavelet(nt: int, dt: float, f: float = 20.0, delay: float | None = None) -> np.ndarray:
    """Ricker-shaped pulse that needs no minimum record length (any nt)."""
    t = np.arange(nt) * dt
    t0 = 1.2 / f if delay is None else delay
    a = (np.pi * f * (t - t0)) ** 2
    return (1 - 2 * a) * np.exp(-a)


def layered_acoustic(nz=48, nx=56, ds=5.0, seed=7):
    rng = np.random.Generator(np.random.PCG64(seed))
    grid = [Grid(0.0, ds, nz), Grid(0.0, ds, nx)]
    vp = 1800.0 + 600.0 * (np.arange(nz)[:, None] / nz) + 80.0 * rng.standard_normal((nz, nx))
    rho = 1900.0 + 300.0 * rng.random((nz, nx))
    return G.AcousticMedium(grid, vp.astype(np.float32), rho.astype(np.float32))


def layered_elastic(nz=48, nx=56, ds=5.0, seed=11):
    rng = np.random.Generator(np.random.PCG64(seed))
    grid = [Grid(0.0, ds, nz), Grid(0.0, ds, nx)]
    vp = 2000.0 + 900.0 * (np.arange(nz)[:, None] / nz) + 60.0 * rng.standard_normal((nz, nx))
    vs = vp / (1.7 + 0.2 * rng.random((nz, nx)))
    rho = 310.0 * vp ** 0.25
    return G.ElasticMedium(grid, vp.astype(np.float32), vs.astype(np.float32), rho.astype(np.float32))


def experiment(medium, *, nt=240, sfields=("p",), rfields=("p", "vx", "vz"), ns=1, nr=9, order=2,
               npml=10, pml_faces=None, stressfree_faces=(), seed=3, ctx=None, src_depth=None,
               rec_depth=0.0, jitter=True):
    """A single-supersource experiment with random (seeded) off-node positions."""
    rng = np.random.Generator(np.random.PCG64(seed))
    gz, gx = medium.grid
    Lz, Lx = gz.last, gx.last
    dt = 0.4 * gz.step / (1.6 * float(medium.vp.m.max()))
    tgrid = Grid(0.0, dt, nt)
    sz = np.full(ns, src_depth if src_depth is not None else 0.3 * Lz)
    sz = sz + (rng.uniform(-1, 1, ns) * gz.step if jitter else 0.0)
    sx = np.linspace(0.35 * Lx, 0.65 * Lx, ns) + (rng.uniform(-1, 1, ns) * gx.step if jitter else 0)
    rz = np.full(nr, rec_depth) + (rng.uniform(0, 1, nr) * gz.step if jitter else 0)
    rx = np.linspace(0.0, Lx, nr)
    ag = G.AGeomss({"z": sz, "x": sx}, {"z": rz, "x": rx})
    src = G.Srcs(ns, tgrid, list(sfields))
    w = np.stack([wavelet(nt, dt, 25.0, 0.03 + 0.004 * k) for k in range(ns)], axis=1)
    src.update(list(sfields), w)
    attrib = G.FdtdAcoustic() if isinstance(medium, G.AcousticMedium) else G.FdtdElastic()
    kw = {}
    if pml_faces is not None:
        kw["pml_faces"] = pml_faces
    return G.SeisForwExpt(attrib, medium=medium, tgrid=tgrid, ageom=[ag], srcwav=[src],
                          rfields=list(rfields), order=order, npml=npml,
                          stressfree_faces=list(stressfree_faces), context=ctx, **kw)


def medium3d(physics, n=(22, 18, 24), ds=10.0, seed=5):
    rng = np.random.Generator(np.random.PCG64(seed))
    grid = [Grid(0.0, ds, n[0]), Grid(0.0, ds, n[1]), Grid(0.0, ds, n[2])]
    vp = (2000.0 + 500.0 * (np.arange(n[0])[:, None, None] / n[0]) + 60.0 * rng.standard_normal(n)).astype(np.float32)
    rho = (1800.0 + 300.0 * rng.random(n)).astype(np.float32)
    if physics == 0:
        return G.AcousticMedium(grid, vp, rho)
    vs = (vp / (1.7 + 0.2 * rng.random(n))).astype(np.float32)
    return G.ElasticMedium(grid, vp, vs, rho)


def experiment3d(medium, *, nt=120, sfields=("p",), rfields=("p", "vx", "vy", "vz"), ns=2, nr=6,
                 order=2, npml=5, pml_faces=None, stressfree_faces=(), compat=1, seed=9, ctx=None):
    rng = np.random.Generator(np.random.PCG64(seed))
    gz, gy, gx = medium.grid
    dt = 0.3 * gz.step / (1.7 * float(medium.vp.m.max()))
    tgrid = Grid(0.0, dt, nt)
    s = {"z": gz.last * rng.uniform(0.3, 0.6, ns), "y": gy.last * rng.uniform(0.3, 0.7, ns),
         "x": gx.last * rng.uniform(0.3, 0.7, ns)}
    r = {"z": gz.step * rng.uniform(0, 1.5, nr), "y": gy.last * rng.uniform(0, 1, nr),
         "x": gx.last * rng.uniform(0, 1, nr)}
    ag = G.AGeomss(s, r)
    src = G.Srcs(ns, tgrid, list(sfields))
    w = np.stack([wavelet(nt, dt, 18.0, 0.05 + 0.005 * k) for k in range(ns)], axis=1)
    src.update(list(sfields), w)
    attrib = G.FdtdAcoustic() if isinstance(medium, G.AcousticMedium) else G.FdtdElastic()
    kw = {}
    if pml_faces is not None:
        kw["pml_faces"] = pml_faces
    return G.SeisForwExpt(attrib, medium=medium, tgrid=tgrid, ageom=[ag], srcwav=[src],
                          rfields=list(rfields), order=order, npml=npml, compat=compat,
                          stressfree_faces=list(stressfree_faces), context=ctx, **kw)
