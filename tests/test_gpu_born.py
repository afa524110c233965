"""GPU parity of Born modelling (gpi_born) against the oracle's Born composition, the
linearity of the Born map, and the adjoint (dot-product) test of LinearMap(pa), whose
transpose is the imaging of the gradient path (checkpointed, or the reference's time
reversal).

Parity bar: relative L2 1e-5 (the kernels keep the oracle's FP32 evaluation order, so the
records agree bit for bit).  Dot-product test: the invK imaging is the adjoint of the
linearised map up to the PML, so <F x, y> = <x, F^T y> to 1e-4 for a perturbation inside
the medium and away from the sources (a pressure source's own term is scaled by dtK, which
the imaging sees and the secondary sources, like the reference's, leave out); the reference's density imaging averages velocity products one cell off
(gradient.jl:31-56) and is not the exact adjoint, so rho is held to 0.1.
"""
import numpy as np
import pytest

import geophyinv as G
from helpers import layered_acoustic, medium3d, oracle_born, rel_l2, wavelet
from geophyinv.grid import Grid

pytestmark = pytest.mark.gpu
TOL = 1e-5


def born2d(ctx, order=2, nss=2, rfields=("p", "vx", "vz"), nt=260, checkpoint=0, seed=3, sfield="p"):
    med = layered_acoustic(nz=40, nx=52, seed=7)
    rng = np.random.Generator(np.random.PCG64(seed))
    gz, gx = med.grid
    dt = 0.4 * gz.step / (1.6 * float(med.vp.m.max()))
    tgrid = Grid(0.0, dt, nt)
    ageom, srcwav = [], []
    for k in range(nss):
        ageom.append(G.AGeomss({"z": [0.2 * gz.last + rng.uniform(-1, 1) * gz.step],
                                "x": [gx.last * (0.35 + 0.3 * k / max(nss - 1, 1))]},
                               {"z": np.full(9, 0.8 * gz.last), "x": np.linspace(0.1, 0.9, 9) * gx.last}))
        s = G.Srcs(1, tgrid, [sfield])
        s.update([sfield], wavelet(nt, dt, 25.0, 0.03 + 0.003 * k) * 1e3)
        srcwav.append(s)
    return G.SeisForwExpt(G.FdtdAcoustic[G.Born](), medium=med, tgrid=tgrid, ageom=ageom, srcwav=srcwav,
                          rfields=list(rfields), order=order, npml=10 + order, context=ctx, checkpoint=checkpoint)


def born3d(ctx, order=2):
    med = medium3d(0, n=(24, 20, 22), seed=5)
    gz, gy, gx = med.grid
    nt = 150
    dt = 0.3 * gz.step / (1.7 * float(med.vp.m.max()))
    tgrid = Grid(0.0, dt, nt)
    ag = G.AGeomss({"z": [0.25 * gz.last], "y": [0.5 * gy.last], "x": [0.5 * gx.last]},
                   {"z": np.full(8, 0.75 * gz.last), "y": np.linspace(0.2, 0.8, 8) * gy.last,
                    "x": np.linspace(0.3, 0.7, 8) * gx.last})
    s = G.Srcs(1, tgrid, ["p"])
    s.update(["p"], wavelet(nt, dt, 18.0, 0.05) * 1e3)
    return G.SeisForwExpt(G.FdtdAcoustic[G.Born](), medium=med, tgrid=tgrid, ageom=[ag], srcwav=[s],
                          rfields=["p", "vx", "vy", "vz"], order=order, npml=5 + order, context=ctx)


def blob_dmod(pa, centre, rad, scale=(1e-2, 5e-3)):
    shape = pa.medium.vp.m.shape
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    b = np.exp(-sum((g - c) ** 2 for g, c in zip(grids, centre)) / (2.0 * rad ** 2))
    dmod = {}
    for (k, sc) in zip(("invK", "rho"), scale):
        d = np.zeros(pa.ext_shape, np.float64, order="F")
        inner = pa.view_inner(d)
        d[inner] = sc * b * pa.mod[k][inner]
        dmod[k] = d
    return dmod


CASES = [dict(ndim=2, order=2, label="2d o2, 2 shots"), dict(ndim=2, order=4, label="2d o4"),
         dict(ndim=2, order=8, label="2d o8"), dict(ndim=3, order=2, label="3d o2"),
         dict(ndim=3, order=4, label="3d o4")]


@pytest.mark.parametrize("case", CASES, ids=[c["label"] for c in CASES])
def test_born_parity(gpu_ctx, case):
    print("born parity", case["label"], flush=True)
    if case["ndim"] == 2:
        pa = born2d(gpu_ctx, order=case["order"], nss=2 if case["order"] == 2 else 1)
        dmod = blob_dmod(pa, (20, 26), 4.0)
    else:
        pa = born3d(gpu_ctx, order=case["order"])
        dmod = blob_dmod(pa, (12, 10, 11), 3.0)
    pa.set_perturbation(dmod)
    pa.update()
    ref = oracle_born(pa)
    worst = 0.0
    for iss, r in enumerate(ref):
        for f in pa.rfields:
            assert np.abs(r[f]).max() > 0, f
            e = rel_l2(pa.data[iss].d[f], r[f])
            worst = max(worst, e)
            assert e <= TOL, (iss, f, e)
    print(f"  worst rel L2 {worst:.3e}", flush=True)


def test_born_linearity_and_update_with_perturbed_medium(gpu_ctx):
    pa = born2d(gpu_ctx, nss=1)
    dmod = blob_dmod(pa, (18, 30), 3.0)
    pa.set_perturbation(dmod)
    pa.update()
    d1 = {f: pa.data[0].d[f].copy() for f in pa.rfields}
    pa.set_perturbation({k: 2.0 * v for k, v in dmod.items()})
    pa.update()
    for f in pa.rfields:   # doubling is exact in FP32 except where intermediates underflow
        assert rel_l2(pa.data[0].d[f], 2 * d1[f]) < 1e-7, f
    # update!(pa, medium, medium_pert): dmod = mod(pert) - mod
    pert = pa.medium.copy()
    pert.vp.m[15:22, 24:32] *= np.float32(1.02)
    G.update(pa, pa.medium, pert)
    assert np.abs(pa.dmod["invK"]).max() > 0 and not np.any(pa.dmod["invK"][:pa.npml])
    pa.update()
    assert np.abs(pa.data[0].d["vz"]).max() > 0


@pytest.mark.parametrize("checkpoint", [50, 0])
def test_linear_map_dot_product(gpu_ctx, checkpoint):
    # the reference's time reversal (checkpoint 0) needs velocity sources (type -1 sink)
    pa = born2d(gpu_ctx, nss=2, rfields=("vx", "vz"), checkpoint=checkpoint, sfield="p" if checkpoint else "vz")
    F = G.LinearMap(pa)
    rng = np.random.Generator(np.random.PCG64(11))
    shape = pa.medium.vp.m.shape
    zz, xx = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    # away from the sources: the secondary sources do not rescale a pressure source term
    b = np.exp(-((zz - 28) ** 2 + (xx - 26) ** 2) / (2.0 * 3.0 ** 2)).ravel(order="F")
    n = F.shape[1] // 2
    # band-limited data: white noise filtered in time by the source wavelet
    nt = len(pa.tgrid)
    w = pa.srcwav[0].d[pa.srcwav[0].fields[0]][:, 0].astype(np.float64)
    noise = rng.standard_normal((nt, F.shape[0] // nt))
    spec = np.fft.rfft(noise, 2 * nt, axis=0) * np.fft.rfft(w, 2 * nt)[:, None]
    y = np.fft.irfft(spec, 2 * nt, axis=0)[:nt].ravel(order="F")
    Fty = F.T @ y
    for part, tol in ((0, 1e-4), (1, 0.1)):
        x = np.zeros(F.shape[1])
        x[part * n:(part + 1) * n] = b * (1e-12 if part == 0 else 10.0)
        lhs = float(np.dot(F @ x, y))
        rhs = float(np.dot(x, Fty))
        print(f"  checkpoint {checkpoint} {'invK' if part == 0 else 'rho'}: <Fx,y> {lhs:.6e} <x,F^T y> {rhs:.6e} "
              f"rel {abs(lhs - rhs) / abs(lhs):.2e}", flush=True)
        assert abs(lhs - rhs) <= tol * abs(lhs)


def test_linear_map_additivity_and_normal_operator(gpu_ctx):
    """test/fwi/born_map.jl: F (x1 + x2) = F x1 + F x2 (to FP32 rounding here) and
    <x, F^T F x> > 0."""
    pa = born2d(gpu_ctx, nss=1, rfields=("vz",), checkpoint=40)
    F = G.LinearMap(pa)
    rng = np.random.Generator(np.random.PCG64(5))
    n = F.shape[1] // 2
    scale = np.concatenate([np.full(n, 1e-12), np.full(n, 1.0)])   # invK ~ 1e-10, rho ~ 1e3
    x1 = rng.standard_normal(F.shape[1]) * scale
    x2 = rng.standard_normal(F.shape[1]) * scale
    d12 = F @ (x1 + x2)
    d1, d2 = F @ x1, F @ x2
    e = np.linalg.norm(d12 - (d1 + d2)) / np.linalg.norm(d12)
    print(f"  additivity rel {e:.2e}", flush=True)
    assert e < 1e-5
    c = float(np.dot(x1, F.T @ (F @ x1)))
    assert c > 0.0
