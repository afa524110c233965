"""GPU: the elastic checkpointed gradient (gpi_adjoint_checkpoint_elastic) against finite
differences of the FP64 misfit, 2-D P-SV and 3-D.

The reference has no elastic gradient.  The gradient here is the discrete adjoint of the
staggered update away from the CPML, so directional derivatives along smooth bumps of
log(invlambda), log(invmu) and log(rho) placed between the source and the receivers (clear
of the CPML and the free surface) agree with fourth-order central differences of the misfit
to 1e-3 relative (the bar written here), also for a bump on the explosive source, whose
dtM-scaled injection makes the source term part of the gradient.  The chain rule to
(vp, vs, rho) is checked the same way through the medium's own parameterization.
"""
import numpy as np
import pytest

import geophyinv as G
from geophyinv import func_grad
from helpers import layered_elastic, medium3d, wavelet

pytestmark = pytest.mark.gpu
TOL = 1e-3


def elastic_experiment(med, ctx, *, nt, every, mode="forward_save", seed=4, nr=11):
    rng = np.random.Generator(np.random.PCG64(seed))
    grid = med.grid
    step = grid[0].step
    vmax = float(np.sqrt(med.vp.m.astype(np.float64) ** 2 + med.vs.m.astype(np.float64) ** 2).max())
    dt = 0.35 * step / vmax
    tg = G.Grid(0.0, dt, nt)
    if med.ndim == 2:
        gz, gx = grid
        src = {"z": [0.25 * gz.last], "x": [0.5 * gx.last + rng.uniform(-1, 1) * gx.step]}
        rec = {"z": np.full(nr, 0.75 * gz.last) + rng.uniform(0, 1, nr) * gz.step,
               "x": np.linspace(0.15 * gx.last, 0.85 * gx.last, nr)}
        rf = ["vx", "vz"]
    else:
        gz, gy, gx = grid
        src = {"z": [0.25 * gz.last], "y": [0.5 * gy.last], "x": [0.5 * gx.last]}
        rec = {"z": np.full(nr, 0.75 * gz.last), "y": gy.last * rng.uniform(0.2, 0.8, nr),
               "x": gx.last * rng.uniform(0.2, 0.8, nr)}
        rf = ["vx", "vy", "vz"]
    ag = G.AGeomss(src, rec)
    s = G.Srcs(1, tg, ["tauxx"])
    s.update(["tauxx"], wavelet(nt, dt, 30.0 if med.ndim == 2 else 18.0, 0.03) * 1e6)
    return G.SeisForwExpt(G.FdtdElastic(mode, 2), medium=med, tgrid=tg, ageom=[ag], srcwav=[s], rfields=rf,
                          order=2, npml=8 if med.ndim == 2 else 5, context=ctx, checkpoint=every)


def observed(true, pa, ctx):
    po = elastic_experiment(true, ctx, nt=len(pa.tgrid), every=pa.checkpoint, mode="forward")
    G.update(po)
    out = [d.copy() for d in po.data]
    po.free()
    return out


def bump(shape, centre, radius):
    idx = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    r2 = sum((a - c) ** 2 for a, c in zip(idx, centre))
    return np.exp(-r2 / (2.0 * radius ** 2))


def fd4(J, m, d, e):
    return (8.0 * (J(m + e * d) - J(m - e * d)) - (J(m + 2 * e * d) - J(m - 2 * e * d))) / (12.0 * e)


CASES = [dict(ndim=2, every=64, nt=360, at_source=False, label="2-D P-SV, every 64"),
         dict(ndim=2, every=64, nt=360, at_source=True, label="2-D P-SV, bump on the source"),
         dict(ndim=3, every=50, nt=170, at_source=False, label="3-D, every 50"),
         dict(ndim=3, every=50, nt=170, at_source=True, label="3-D, bump on the source")]


@pytest.mark.parametrize("case", CASES, ids=[c["label"] for c in CASES])
def test_elastic_gradient_matches_finite_differences(gpu_ctx, case):
    print("elastic gradient", case["label"], flush=True)
    if case["ndim"] == 2:
        true = layered_elastic(nz=44, nx=52, seed=11)
    else:
        true = medium3d(1, n=(24, 20, 22), seed=5)
    start = true.copy()
    inner = true.vp.m.shape
    ctr = [0.5 * (n - 1) for n in inner]
    b = bump(inner, ctr, 3.0).astype(np.float32)
    start.vp.m[...] *= (1 - 0.04 * b).astype(np.float32)
    start.vs.m[...] *= (1 - 0.03 * b).astype(np.float32)
    start.rho.m[...] *= (1 + 0.03 * b).astype(np.float32)
    pa = elastic_experiment(start, gpu_ctx, nt=case["nt"], every=case["every"])
    dobs = observed(true, pa, gpu_ctx)
    loss = G.L2DistLoss()
    m = pa.model_vector()
    g = np.zeros_like(m)
    J0 = G.gradient(g, m, loss, dobs, pa)
    assert J0 > 0 and np.all(np.isfinite(g)) and np.abs(g).max() > 0
    J = lambda x: G.lossvalue(x, loss, dobs, pa)
    n = m.size // 3
    if case["at_source"]:
        # on the explosive source: its dtM scaling makes the injected stress depend on the
        # model there (the source term of the gradient)
        src = pa.ageom[0].src_points()[0]
        centre = [(x - ax.first) / ax.step for x, ax in zip(src, pa.medium.grid)]
    else:   # off the centre, between source and receivers
        centre = [0.55 * (inner[0] - 1)] + [0.45 * (k - 1) for k in inner[1:]]
    dshape = bump(inner, centre, 2.5).ravel(order="F")
    for k, name in enumerate(("invlambda", "invmu", "rho")):
        d = np.zeros_like(m)
        d[k * n:(k + 1) * n] = dshape
        fd = fd4(J, m, d, 5e-3)
        ad = float(np.dot(g.astype(np.float64), d))
        rel = abs(ad / fd - 1)
        print(f"  {name}: fd {fd:.6e} adjoint {ad:.6e} rel {rel:.2e}", flush=True)
        assert rel < TOL, (name, ad, fd)
    pa.free()


def test_vp_vs_rho_chain_rule(gpu_ctx):
    """g w.r.t. (vp, vs, rho) from the (invlambda, invmu, rho) gradient against finite
    differences of the misfit in the (vp, vs, rho) parameterization, to the 1e-3 bar.  The
    step of each direction moves the misfit by about 1e-3 of itself (within 1e-3 .. 5e-2 of
    the parameter's mean at the bump's peak): the rho direction at fixed velocities is about
    150 times less sensitive than vp, and a fixed step leaves its misfit change near the
    FP32 rounding noise of the wavefields."""
    true = layered_elastic(nz=40, nx=48, seed=12)
    start = true.copy()
    b = bump(true.vp.m.shape, [19.5, 23.5], 3.0).astype(np.float32)
    start.vp.m[...] *= (1 - 0.04 * b).astype(np.float32)
    pa = elastic_experiment(start, gpu_ctx, nt=320, every=80)
    dobs = observed(true, pa, gpu_ctx)
    loss = G.L2DistLoss()
    m = pa.model_vector()
    g = np.zeros_like(m)
    J0 = G.gradient(g, m, loss, dobs, pa)
    n = m.size // 3
    sl = pa.view_inner(pa.mod["rho"])
    gm = {k: g[i * n:(i + 1) * n].astype(np.float64) / pa.mod[k][sl].ravel(order="F").astype(np.float64)
          for i, k in enumerate(("invlambda", "invmu", "rho"))}   # d/dparam from d/dlog(param)
    gv = func_grad.chainrule_elastic(gm["invlambda"], gm["invmu"], gm["rho"], start.vp.m.ravel(order="F"),
                                     start.vs.m.ravel(order="F"), start.rho.m.ravel(order="F"))

    def J_of(vp, vs, rho):
        med = G.ElasticMedium(start.grid, vp, vs, rho)
        pa.update_medium(med)
        return G.lossvalue(pa.model_vector(), loss, dobs, pa)

    d = bump(true.vp.m.shape, [22.0, 21.0], 2.5)
    base = {k: getattr(start, k).m.astype(np.float64) for k in ("vp", "vs", "rho")}
    for k, name in enumerate(("vp", "vs", "rho")):
        scale = base[name].mean()
        def J(t):
            p = {q: base[q].copy() for q in base}
            p[name] = p[name] + t * scale * d
            return J_of(*(p[q].astype(np.float32) for q in ("vp", "vs", "rho")))
        ad = float(np.dot(gv[k], (scale * d).ravel(order="F")))
        e = float(np.clip(1e-3 * J0 / max(abs(ad), 1e-30), 1e-3, 5e-2))
        fd = (8.0 * (J(e) - J(-e)) - (J(2 * e) - J(-2 * e))) / (12.0 * e)
        rel = abs(ad / fd - 1)
        print(f"  {name}: step {e:.1e} fd {fd:.6e} adjoint {ad:.6e} rel {rel:.2e}", flush=True)
        assert rel < TOL, (name, ad, fd)
    pa.free()
