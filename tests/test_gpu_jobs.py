"""GPU: the job queue (gpi_jobs_*): forward and gradient jobs of 2-D and 3-D experiments
run concurrently on worker streams give the plain calls' results.

- forward jobs: records equal ``run_shot``'s bit for bit;
- gradient jobs: the per-supersource gradients equal those of ``gradient!``'s :adjoint run
  bit for bit (same kernels; the adjoint sources 2 (dobs - dcal) are formed in FP32 the
  same way), the misfit to FP64 rounding (the host sums in another order);
- a budget that holds one job at a time still completes every job, one after another.
"""
import numpy as np
import pytest

import geophyinv as G
from geophyinv import _lib
from geophyinv.serve import JobServer
from helpers import fwi_experiment, layered_acoustic, medium3d, wavelet

pytestmark = pytest.mark.gpu


def elastic3d(ctx, seed, nt=140, every=40, mode="forward_save"):
    rng = np.random.Generator(np.random.PCG64(seed))
    med = medium3d(1, n=(20, 18, 22), seed=seed)
    gz, gy, gx = med.grid
    vmax = float(np.sqrt(med.vp.m.astype(np.float64) ** 2 + med.vs.m.astype(np.float64) ** 2).max())
    dt = 0.35 * gz.step / vmax
    tg = G.Grid(0.0, dt, nt)
    ag = G.AGeomss({"z": [0.3 * gz.last], "y": [gy.last * rng.uniform(0.4, 0.6)], "x": [gx.last * rng.uniform(0.4, 0.6)]},
                   {"z": np.full(8, 0.7 * gz.last), "y": gy.last * rng.uniform(0.2, 0.8, 8),
                    "x": gx.last * rng.uniform(0.2, 0.8, 8)})
    s = G.Srcs(1, tg, ["tauxx"])
    s.update(["tauxx"], wavelet(nt, dt, 20.0, 0.04) * 1e6)
    return G.SeisForwExpt(G.FdtdElastic(mode, 2), medium=med, tgrid=tg, ageom=[ag], srcwav=[s],
                          rfields=["vx", "vy", "vz"], order=2, npml=5, context=ctx, checkpoint=every)


def observed(pa, ctx, scale=1.02):
    """Observed records: the experiment's medium with vp scaled, run forward."""
    med = pa.medium.copy()
    med.vp.m[...] *= np.float32(scale)
    po = G.SeisForwExpt(G.FdtdElastic() if pa.physics else G.FdtdAcoustic(), medium=med, tgrid=pa.tgrid,
                        ageom=pa.ageom, srcwav=pa.srcwav, rfields=pa.rfields, order=pa.order, npml=pa.npml,
                        context=ctx)
    G.update(po)
    out = [d.copy() for d in po.data]
    po.free()
    return out


def plain_gradient(pa, dobs):
    loss = G.L2DistLoss()
    m = pa.model_vector()
    G.gradient(np.zeros_like(m), m, loss, dobs, pa)
    grads = {k: v.copy() for k, v in pa.gradients.items()}
    J = G.lossvalue(m, loss, dobs, pa)
    return grads, J


def test_jobs_equal_plain_calls(gpu_ctx):
    print("job queue: 3-D elastic forward + gradient, 2-D acoustic gradient", flush=True)
    fwd = [elastic3d(gpu_ctx, 30 + k, mode="forward") for k in range(2)]
    plain_recs = [pa.run_shot(0)[0] for pa in fwd]
    pe = elastic3d(gpu_ctx, 40)
    dobs_e = observed(pe, gpu_ctx)
    ge, Je = plain_gradient(pe, dobs_e)
    med = layered_acoustic(nz=40, nx=50, seed=9)
    pac = fwi_experiment(med, nss=1, nt=260, ctx=gpu_ctx, checkpoint=50, grad_rho=True)
    dobs_a = observed(pac, gpu_ctx, scale=0.97)
    ga, Ja = plain_gradient(pac, dobs_a)
    with JobServer(workers=3) as js:
        tf = [js.submit(pa, 0) for pa in fwd]
        te = js.submit(pe, 0, dobs=dobs_e[0])
        ta = js.submit(pac, 0, dobs=dobs_a[0])
        out_f = [js.wait(t, timeout=300.0) for t in tf]
        oe = js.wait(te, timeout=300.0)
        oa = js.wait(ta, timeout=300.0)
        st = js.stats()
    for a, b in zip(plain_recs, out_f):
        for f in a:
            assert np.array_equal(a[f], b["records"][f]), f
    for out, ref, J in ((oe, ge, Je), (oa, ga, Ja)):
        for k, g in out["gradients"].items():
            assert np.abs(ref[k]).max() > 0, k
            assert np.array_equal(g, ref[k]), (k, np.abs(g - ref[k]).max())
        assert abs(out["misfit"] / J - 1) < 1e-12
    print(f"  {st['completed']} jobs, p50 {st['p50_ms']:.1f} ms, forward p50 {st['forward_p50_ms']:.1f} ms, "
          f"gradient p50 {st['gradient_p50_ms']:.1f} ms", flush=True)
    assert st["completed"] == 4 and st["forward_done"] == 2 and st["gradient_done"] == 2 and st["failed"] == 0
    for pa in fwd + [pe, pac]:
        pa.free()


def test_jobs_at_the_elastic_serving_scale(gpu_ctx):
    """The 3-D elastic serving workload's problem at its smallest size (128^3 + CPML =
    168^3, dipping interface, explosive source, surface patch recording vx, vy, vz): a
    forward job and a gradient job in one queue; the forward records against the oracle
    (relative L2 1e-5), the gradient job's directional derivatives along a bump between
    the source and the dipping interface against fourth-order central differences of the
    FP64 misfit (1e-3)."""
    import oracle as O
    from geophyinv import gallery
    print("job queue at 168^3: elastic forward vs oracle, gradient vs finite differences", flush=True)
    nt = 300
    true, _, ageom, srcwav, tgrid, kw = gallery.elastic_fwi_3d(n=128, nt=nt, nrec_side=32)
    # the starting model differs from the true one within reach of 300 steps (the dipping
    # interface's reflection returns later): a smooth anomaly 20 cells down
    inner = true.vp.m.shape
    idx = np.meshgrid(*[np.arange(s) for s in inner], indexing="ij")
    blob = np.exp(-sum((a - c) ** 2 for a, c in zip(idx, (20.0, 66.0, 60.0))) / (2.0 * 5.0 ** 2))
    start = true.copy()
    start.vp.m[...] *= (1 - 0.05 * blob).astype(np.float32)
    start.vs.m[...] *= (1 - 0.04 * blob).astype(np.float32)
    start.rho.m[...] *= (1 + 0.03 * blob).astype(np.float32)
    fwd = G.SeisForwExpt(G.FdtdElastic(), medium=true, tgrid=tgrid, ageom=ageom, srcwav=srcwav, context=gpu_ctx,
                         **kw)
    pg = G.SeisForwExpt(G.FdtdElastic("forward_save", 2), medium=start, tgrid=tgrid, ageom=ageom, srcwav=srcwav,
                        context=gpu_ctx, checkpoint=50, **kw)
    po = G.SeisForwExpt(G.FdtdElastic(), medium=true, tgrid=tgrid, ageom=ageom, srcwav=srcwav, context=gpu_ctx,
                        **kw)
    G.update(po)
    dobs = [d.copy() for d in po.data]
    po.free()
    # gradient! first: it sets the coefficients from the model vector (a log/exp round
    # trip), which the gradient job then runs on as well
    loss = G.L2DistLoss()
    m = pg.model_vector()
    g = np.zeros_like(m)
    G.gradient(g, m, loss, dobs, pg)
    plain = {k: v.copy() for k, v in pg.gradients.items()}
    with JobServer(workers=2) as js:
        tf = js.submit(fwd, 0)
        tg = js.submit(pg, 0, dobs=dobs[0])
        of = js.wait(tf, timeout=600.0)
        og = js.wait(tg, timeout=600.0)
    orc = O.from_experiment(fwd)
    orc.run()
    for q, f in enumerate(fwd.rfields):
        r = of["records"][f]
        e = float(np.linalg.norm((r - orc.records(q)).astype(np.float64)) /
                  np.linalg.norm(orc.records(q).astype(np.float64)))
        print(f"  forward {f}: rel L2 vs oracle {e:.2e}", flush=True)
        assert np.abs(r).max() > 0 and e <= 1e-5, (f, e)
    for k, v in og["gradients"].items():
        assert np.abs(plain[k]).max() > 0, k
        assert np.array_equal(v, plain[k]), (k, np.abs(v - plain[k]).max())
    J = lambda x: G.lossvalue(x, loss, dobs, pg)
    J0 = J(m)
    n = m.size // 3
    # a bump (radius 3 cells, cut off at 9) at the depth of the anomaly, off to one side:
    # on the paths of its reflections.  Cells outside the cut-off keep their model values
    # bit for bit (a Gaussian tail of 1e-20 would still flip the last bit of some FP32
    # coefficients, whose effect on the misfit is noise, not derivative)
    r2 = sum((a - c) ** 2 for a, c in zip(idx, (20.0, 45.0, 50.0)))
    d0 = np.where(r2 <= 81.0, np.exp(-r2 / (2.0 * 3.0 ** 2)), 0.0).ravel(order="F")
    for k, name in enumerate(("invlambda", "invmu", "rho")):
        d = np.zeros_like(m)
        d[k * n:(k + 1) * n] = d0
        ad = float(np.dot(g.astype(np.float64), d))
        # a step that moves the misfit by about 1e-3 of itself: above the noise of FP32
        # wavefields in the misfit (about 1e-6 of it here), inside the linear range
        e = min(max(1e-3 * J0 / abs(ad), 1e-3), 0.1)
        fd = (8.0 * (J(m + e * d) - J(m - e * d)) - (J(m + 2 * e * d) - J(m - 2 * e * d))) / (12.0 * e)
        rel = abs(ad / fd - 1)
        print(f"  gradient {name}: J {J0:.4e} eps {e:.1e} fd {fd:.6e} adjoint {ad:.6e} rel {rel:.2e}", flush=True)
        assert rel < 1e-3, (name, ad, fd)
    fwd.free()
    pg.free()


def test_jobs_one_at_a_time_under_a_small_budget(gpu_ctx):
    fwd = [elastic3d(gpu_ctx, 50 + k, mode="forward") for k in range(3)]
    plain = [pa.run_shot(0)[0] for pa in fwd]
    with JobServer(workers=3, hbm_budget_bytes=120 << 20) as js:   # about one of these jobs
        ts = [js.submit(pa, 0) for pa in fwd]
        outs = [js.wait(t, timeout=300.0) for t in ts]
        st = js.stats()
    for a, b in zip(plain, outs):
        for f in a:
            assert np.array_equal(a[f], b["records"][f])
    assert st["completed"] == 3
    for pa in fwd:
        pa.free()


def test_jobs_reject(gpu_ctx):
    pa = elastic3d(gpu_ctx, 60, mode="forward", every=0)
    with JobServer(workers=1, hbm_budget_bytes=1 << 20) as js:
        with pytest.raises(_lib.GpiError):
            js.submit(pa, 0)                          # larger than the whole budget
    with JobServer(workers=1) as js:
        with pytest.raises(ValueError):
            js.submit(pa, 0, dobs=pa.data[0])         # a gradient job needs checkpoint > 0
    pa.free()
