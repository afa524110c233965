"""Synthetic workloads of the FDTD engine, generated from fixed seeds (numpy PCG64).

``acoustic_anomaly_2d`` is the 201x201 acoustic single-shot case and ``elastic_layers_2d``
the 2001x2001 P-SV drop-in case at 2.5 m; both build (medium, ageom, srcwav, tgrid, kw)
ready for ``SeisForwExpt``.  ``scale`` shrinks a workload for quick checks.
"""
from __future__ import annotations

import math

import numpy as np

from .database import AGeomss, Srcs
from .grid import Grid
from .media import AcousticMedium, ElasticMedium
from .wavelets import ricker


def _box_smooth(a: np.ndarray, k: int) -> np.ndarray:
    """k x k moving average with edge replication."""
    r = k // 2
    p = np.pad(a, r, mode="edge")
    c = np.cumsum(np.cumsum(p, axis=0), axis=1)
    c = np.pad(c, ((1, 0), (1, 0)))
    s = c[k:, k:] - c[:-k, k:] - c[k:, :-k] + c[:-k, :-k]
    return s / (k * k)


def acoustic_anomaly_2d(n: int = 201, ds: float = 5.0, nt: int = 1000, npml: int = 20,
                        freq: float = 15.0, nrec: int = 101):
    """2-D acoustic, vp = 2000 m/s with a +10% Gaussian anomaly (radius 50 m) at the centre,
    rho = 2000, Ricker on p five cells below the surface at mid-x, receivers along the
    surface recording p, vx, vz, dt at CFL 0.5."""
    L = (n - 1) * ds
    grid = [Grid.range(0.0, L, length=n), Grid.range(0.0, L, length=n)]
    z = grid[0].values()[:, None]
    x = grid[1].values()[None, :]
    cz, cx = L / 2, L / 2
    r2 = (z - cz) ** 2 + (x - cx) ** 2
    vp = (2000.0 * (1.0 + 0.10 * np.exp(-r2 / (2.0 * 50.0 ** 2)))).astype(np.float32)
    rho = np.full((n, n), 2000.0, np.float32)
    medium = AcousticMedium(grid, vp, rho)
    vmax = 2200.0
    dt = 0.5 * ds / vmax
    tgrid = Grid(0.0, dt, nt)
    sx = L / 2
    ageom = [AGeomss({"z": [5 * ds], "x": [sx]},
                     {"z": np.zeros(nrec), "x": np.linspace(0.0, L, nrec)})]
    src = Srcs(1, tgrid, ["p"])
    src.update(["p"], ricker(freq, tgrid))
    kw = dict(rfields=["p", "vx", "vz"], npml=npml)
    return medium, ageom, [src], tgrid, kw


def elastic_layers_2d(n: int = 2001, ds: float = 2.5, nt: int = 4000, npml: int = 40,
                      freq: float = 20.0, nrec: int = 801, nlayers: int = 8, seed: int = 2002):
    """2-D P-SV: equal-thickness flat layers vp 1500->4500 m/s, vs = vp/1.8, Gardner density,
    a +-3% white-noise perturbation smoothed over 5 cells, explosive (tauxx) Ricker source,
    surface receivers recording vx and vz, dt at CFL 0.5 of sqrt(vp^2 + vs^2)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = (n - 1) * ds
    grid = [Grid.range(0.0, L, length=n), Grid.range(0.0, L, length=n)]
    layer = np.minimum((np.arange(n) * nlayers) // n, nlayers - 1)
    vp1 = 1500.0 + layer * (3000.0 / (nlayers - 1))
    vp = np.repeat(vp1[:, None], n, axis=1)
    eps = _box_smooth(rng.uniform(-0.03, 0.03, size=(n, n)), 5)
    vp = vp * (1.0 + eps)
    vs = vp / 1.8
    rho = 310.0 * vp ** 0.25
    medium = ElasticMedium(grid, vp.astype(np.float32), vs.astype(np.float32), rho.astype(np.float32))
    dt = 0.5 * ds / math.sqrt(4500.0 ** 2 + 2500.0 ** 2)
    tgrid = Grid(0.0, dt, nt)
    ageom = [AGeomss({"z": [10 * ds], "x": [L / 2]},
                     {"z": np.zeros(nrec), "x": np.linspace(0.0, L, nrec)})]
    src = Srcs(1, tgrid, ["tauxx"])
    # a run shorter than the wavelet keeps its first nt samples
    nw = max(nt, int(math.ceil(3.2 / freq / dt)))
    src.update(["tauxx"], ricker(freq, Grid(0.0, dt, nw))[:nt])
    kw = dict(rfields=["vx", "vz"], npml=npml)
    return medium, ageom, [src], tgrid, kw


def serving_stream(nshots: int = 512, seed: int = 3003, rate: float = 200.0, ds: float = 5.0,
                   side=(256, 2048), steps=(1000, 6000), freq=(10.0, 30.0), receivers=(64, 512),
                   npml: int = 20):
    """A stream of 2-D acoustic shots arriving as a Poisson process.

    Each shot: square grid with side ~ U{side}, dx = ds; nt ~ U{steps}; Ricker with peak
    ~ U(freq) on p five cells below the surface at a random x; nr ~ U{receivers} evenly
    spaced along the surface recording p and vz; a layered medium (3-10 layers, vp
    increasing within [1500, 4500] m/s, Gardner density) with one Gaussian anomaly of
    +-10 % and radius ~ U(50, 250) m; CFL 0.5; npml CPML cells on every face.

    Returns a list of dicts with ``arrival`` (seconds from the first arrival) and the
    inputs of ``SeisForwExpt`` (``medium``, ``ageom``, ``srcwav``, ``tgrid``, ``kw``)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    gaps = rng.exponential(1.0 / rate, nshots)
    gaps[0] = 0.0
    arrivals = np.cumsum(gaps)
    out = []
    for k in range(nshots):
        s = int(rng.integers(side[0], side[1] + 1))
        nt = int(rng.integers(steps[0], steps[1] + 1))
        f = float(rng.uniform(*freq))
        nr = int(rng.integers(receivers[0], receivers[1] + 1))
        nl = int(rng.integers(3, 11))
        L = (s - 1) * ds
        grid = [Grid.range(0.0, L, length=s), Grid.range(0.0, L, length=s)]
        tops = np.sort(rng.uniform(0.0, L, nl - 1))
        vl = np.sort(rng.uniform(1500.0, 4500.0, nl))
        z = grid[0].values()
        vcol = vl[np.searchsorted(tops, z, side="right")]
        amp = float(rng.choice([-0.10, 0.10]))
        rad = float(rng.uniform(50.0, 250.0))
        cz, cx = rng.uniform(0.2 * L, 0.8 * L, 2)
        zz = z[:, None]
        xx = grid[1].values()[None, :]
        vp = np.asfortranarray(vcol[:, None] * (1.0 + amp * np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2)
                                                                   / (2.0 * rad ** 2))), dtype=np.float32)
        rho = np.asfortranarray(310.0 * vp.astype(np.float64) ** 0.25, dtype=np.float32)
        medium = AcousticMedium(grid, vp, rho)
        dt = 0.5 * ds / float(vp.max())
        tgrid = Grid(0.0, dt, nt)
        ageom = [AGeomss({"z": [5 * ds], "x": [float(rng.uniform(0.1 * L, 0.9 * L))]},
                         {"z": np.zeros(nr), "x": np.linspace(0.0, L, nr)})]
        src = Srcs(1, tgrid, ["p"])
        src.update(["p"], ricker(f, tgrid))
        out.append(dict(arrival=float(arrivals[k]), medium=medium, ageom=ageom, srcwav=[src], tgrid=tgrid,
                        kw=dict(rfields=["p", "vz"], npml=npml), side=s, nt=nt))
    return out


def acoustic_gradient_3d(n: int = 400, ds: float = 10.0, nt: int = 3000, npml: int = 20, freq: float = 12.0,
                         nrec_side: int = 201, seed: int = 4004):
    """3-D acoustic: vp = 1500 + 2500 z/L m/s with three spherical anomalies of +-8 % (centres
    in the middle half of the cube, radii 150-400 m), rho = 2000; 12 Hz Ricker on p five cells
    below the surface at the centre; nrec_side^2 surface receivers at 20 m spacing recording
    p; dt at CFL 0.5.  The keywords carry compat=0: each coordinate is searched on its own
    axis, so the source and the receiver patch are where this describes them (the
    reference's 3-D spray swaps z and x, src/proj_mat.jl:240)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = (n - 1) * ds
    grid = [Grid.range(0.0, L, length=n) for _ in range(3)]
    z = grid[0].values()[:, None, None]
    y = grid[1].values()[None, :, None]
    x = grid[2].values()[None, None, :]
    vp = 1500.0 + 2500.0 * z / L + 0.0 * y + 0.0 * x
    for _ in range(3):
        c = rng.uniform(0.25 * L, 0.75 * L, 3)
        r = rng.uniform(150.0, 400.0)
        a = float(rng.choice([-0.08, 0.08]))
        inside = (z - c[0]) ** 2 + (y - c[1]) ** 2 + (x - c[2]) ** 2 <= r * r
        vp = np.where(inside, vp * (1.0 + a), vp)
    vp = np.asfortranarray(vp, dtype=np.float32)
    rho = np.full((n, n, n), 2000.0, np.float32, order="F")
    medium = AcousticMedium(grid, vp, rho)
    dt = 0.5 * ds / float(vp.max())
    tgrid = Grid(0.0, dt, nt)
    span = min(L, (nrec_side - 1) * 20.0)
    ry, rx = np.meshgrid(np.linspace(L / 2 - span / 2, L / 2 + span / 2, nrec_side),
                         np.linspace(L / 2 - span / 2, L / 2 + span / 2, nrec_side), indexing="ij")
    ageom = [AGeomss({"z": [5 * ds], "y": [L / 2], "x": [L / 2]},
                     {"z": np.zeros(ry.size), "y": ry.ravel(), "x": rx.ravel()})]
    src = Srcs(1, tgrid, ["p"])
    src.update(["p"], _long_ricker(freq, tgrid))
    return medium, ageom, [src], tgrid, dict(rfields=["p"], npml=npml, compat=0)


def acoustic_fwi_3d(n: int = 400, ds: float = 10.0, nt: int = 3000, npml: int = 20, freq: float = 12.0,
                    nrec_side: int = 201, sigma_cells: float = 5.0, seed: int = 4004):
    """The 3-D acoustic FWI gradient problem: the true model of ``acoustic_gradient_3d``,
    a starting model that is the true velocity smoothed by a Gaussian of ``sigma_cells``
    cells (edges replicated), the same source and surface receiver patch recording vz
    (the velocity records drive the adjoint sources).

    Returns (true medium, starting medium, ageom, srcwav, tgrid, kw)."""
    from scipy.ndimage import gaussian_filter

    true, ageom, srcwav, tgrid, kw = acoustic_gradient_3d(n=n, ds=ds, nt=nt, npml=npml, freq=freq,
                                                          nrec_side=nrec_side, seed=seed)
    vp0 = gaussian_filter(true.vp.m.astype(np.float64), sigma_cells, mode="nearest")
    start = AcousticMedium(true.grid, np.asfortranarray(vp0, dtype=np.float32), true.rho.m.copy())
    kw = dict(kw, rfields=["vz"])
    return true, start, ageom, srcwav, tgrid, kw


def elastic_fwi_3d(n: int = 256, ds: float = 5.0, nt: int = 2000, npml: int = 20, freq: float = 25.0,
                   nrec_side: int = 64, seed: int = 5005):
    """A 3-D elastic FWI gradient problem: three layers (vp 1800 / 2600 / 3500 m/s, vs =
    vp/1.9 / vp/1.8 / vp/1.75, Gardner density 310 vp^0.25), the deeper interface dipping
    along x at a seeded angle in [10, 30] degrees; an explosive (tauxx) 25 Hz Ricker five
    cells below the surface at the centre; an nrec_side x nrec_side receiver patch on the
    surface spanning the model, recording vx, vy and vz; dt at CFL 0.5 of sqrt(vp^2+vs^2).
    The starting model is the same medium with the dipping interface one cell shallower.
    The keywords carry compat=0 (coordinates on their own axes, no z<->x swap of
    src/proj_mat.jl:240), so the source and the patch sit at the surface.

    Returns (true medium, starting medium, ageom, srcwav, tgrid, kw)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = (n - 1) * ds
    grid = [Grid.range(0.0, L, length=n) for _ in range(3)]
    theta = np.deg2rad(rng.uniform(10.0, 30.0))
    z1 = rng.uniform(0.25, 0.35) * L
    z2c = rng.uniform(0.55, 0.65) * L
    z = grid[0].values()[:, None, None]
    x = grid[2].values()[None, None, :]

    def medium(shift):
        z2 = z2c + np.tan(theta) * (x - 0.5 * L) - shift * ds
        layer = (z >= z1).astype(np.int32) + (z >= z2).astype(np.int32)   # 0, 1, 2
        layer = np.broadcast_to(layer, (n, n, n))
        vp = np.choose(layer, [1800.0, 2600.0, 3500.0])
        vs = vp / np.choose(layer, [1.9, 1.8, 1.75])
        rho = 310.0 * vp ** 0.25
        f = lambda a: np.asfortranarray(a, dtype=np.float32)
        return ElasticMedium(grid, f(vp), f(vs), f(rho))

    true, start = medium(0.0), medium(1.0)
    dt = 0.5 * ds / float(np.sqrt(3500.0 ** 2 + (3500.0 / 1.75) ** 2))
    tgrid = Grid(0.0, dt, nt)
    rp = np.linspace(0.0, L, nrec_side)
    ry, rx = np.meshgrid(rp, rp, indexing="ij")
    ageom = [AGeomss({"z": [5 * ds], "y": [0.5 * L], "x": [0.5 * L]},
                     {"z": np.zeros(rx.size), "y": ry.ravel(), "x": rx.ravel()})]
    src = Srcs(1, tgrid, ["tauxx"])
    src.update(["tauxx"], _long_ricker(freq, tgrid) * 1e6)
    kw = dict(rfields=["vx", "vy", "vz"], npml=npml, compat=0)
    return true, start, ageom, [src], tgrid, kw


def _long_ricker(freq: float, tgrid: Grid) -> np.ndarray:
    """Ricker on tgrid, computed on a grid long enough for its delay and then cut."""
    need = int(math.ceil(3.0 / freq / tgrid.step)) + 2
    if tgrid.n >= need:
        return ricker(freq, tgrid)
    return ricker(freq, Grid(tgrid.first, tgrid.step, need))[:tgrid.n]
