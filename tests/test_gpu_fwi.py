"""GPU: the inversion-grid gradient of SeisInvExpt (src/fwi/func_grad.jl) against central
finite differences of the misfit, a short L-BFGS run that lowers the misfit, and the two
device pieces of the outer loop against their host restatements: the projection P / P'
(gpi_interp_apply) and the L-BFGS two-loop direction with the history in HBM (gpi_lbfgs_*).

The bulk-modulus part of the reference gradient is the discrete adjoint of the pressure
update away from the boundary strips, so the directional derivative along an invK
direction that vanishes next to the PML agrees with finite differences to 1e-3 (the bar
written here).  Within about three cells of the strips the reference images a wavefield
that is re-forced only at the next step, so directions reaching there are not used."""
import numpy as np
import pytest

import geophyinv as G
from geophyinv.fwi import DeviceLBFGS, apply_proj, interp_matrix
from geophyinv.grid import Grid
from helpers import apply_proj_host, fwi_experiment, lbfgs_two_loop, layered_acoustic, oracle_records

pytestmark = pytest.mark.gpu


def setup(ctx, nss=2):
    true = layered_acoustic(nz=44, nx=60, seed=9)
    med0 = true.copy()
    med0.vp.m[16:28, 22:40] *= np.float32(0.95)
    pa = fwi_experiment(med0, nss=nss, nt=320, ctx=ctx)
    dobs = oracle_records(true, pa)
    return G.SeisInvExpt(pa, dobs, [10, 12])


def test_inversion_grid_gradient_matches_finite_differences(gpu_ctx):
    paw = setup(gpu_ctx)
    m = paw.get_modelvector()
    g = np.zeros_like(m)
    J0 = paw.gradient(g, m)
    assert J0 > 0 and np.all(np.isfinite(g))
    rng = np.random.Generator(np.random.PCG64(11))
    n = m.size // 2
    shape = tuple(gr.n for gr in paw.migrid)
    ring = np.zeros(shape)
    ring[1:-1, 1:-1] = 1.0                      # zero on the outermost inversion nodes
    for trial in range(2):
        d = np.zeros_like(m)
        d[:n] = (rng.standard_normal(shape) * ring).ravel(order="F")   # invK part only
        eps = 1e-3
        Jp = paw.lossvalue(m + eps * d)
        Jm = paw.lossvalue(m - eps * d)
        fd = (Jp - Jm) / (2 * eps)
        ad = float(np.dot(g, d))
        print(f"trial {trial}: fd {fd:.6e} adjoint {ad:.6e} rel {abs(ad / fd - 1):.2e}", flush=True)
        assert abs(ad / fd - 1) < 1e-3


def test_lbfgs_lowers_the_misfit(gpu_ctx):
    paw = setup(gpu_ctx, nss=1)
    m0 = paw.get_modelvector()
    J0 = paw.lossvalue(m0)
    res = paw.update(maxiter=3)
    print(f"misfit {J0:.6e} -> {res.fun:.6e} after {res.nit} iterations, {res.nfev} evaluations, "
          f"{res.host_s:.3f} s of optimizer work per iteration", flush=True)
    assert res.nit == 3 and res.fun < 0.9 * J0
    assert res.fun == pytest.approx(paw.lossvalue(res.x))


def test_projection_on_the_device_equals_the_host(gpu_ctx):
    """P (inversion -> modelling grid) and P' (its transpose) in 2-D and 3-D: the device map
    equals the host restatement (FP64 sums, FP32 results) to one FP32 rounding."""
    rng = np.random.Generator(np.random.PCG64(21))
    for mm, mi in (([Grid(0.0, 5.0, 44), Grid(0.0, 5.0, 61)], [Grid.range(20.0, 195.0, length=9),
                                                                 Grid.range(20.0, 280.0, length=13)]),
                   ([Grid(0.0, 10.0, 30), Grid(0.0, 10.0, 26), Grid(0.0, 10.0, 35)],
                    [Grid.range(40.0, 250.0, length=7), Grid.range(40.0, 210.0, length=6),
                     Grid.range(40.0, 300.0, length=8)])):
        P = [interp_matrix(a, b) for a, b in zip(mm, mi)]            # (n_mi, n_mm)
        PT = [np.ascontiguousarray(p.T) for p in P]
        m = rng.standard_normal(tuple(g.n for g in mi))
        y = rng.standard_normal(tuple(g.n for g in mm))
        for x, mats in ((m, P), (y, PT)):
            dev = apply_proj(gpu_ctx, x, mats)
            host = apply_proj_host(x, mats)
            assert dev.shape == host.shape and dev.dtype == np.float32
            err = np.abs(dev.astype(np.float64) - host) / np.maximum(np.abs(host), 1e-30)
            print(f"projection {x.shape} -> {dev.shape}: max rel diff {err.max():.1e}", flush=True)
            assert err.max() <= 1.2e-7


def test_lbfgs_direction_on_the_device_equals_the_host(gpu_ctx):
    """A sequence of 14 points of a quadratic with a diagonal Hessian (n = 300000, memory 5,
    so the history wraps), with one step whose pair has negative curvature (skipped by
    both): every device direction equals the host two-loop on the same pairs to 1e-12."""
    rng = np.random.Generator(np.random.PCG64(5))
    n, mem = 300000, 5
    h = rng.uniform(0.5, 50.0, n)
    b = rng.standard_normal(n)
    grad = lambda x: h * x - b
    opt = DeviceLBFGS(gpu_ctx, n, mem)
    pairs = []
    x = rng.standard_normal(n)
    xp = gp = None
    for k in range(14):
        g = grad(x)
        if k == 6:
            g = gp - 0.5 * (x - xp)              # this step's pair: s'y < 0
        if xp is not None:
            s, y = x - xp, g - gp
            if s @ y > 2.220446049250313e-16 * (y @ y):
                pairs = (pairs + [(s, y)])[-mem:]
        d_dev = opt.direction(x, g)
        d_host = lbfgs_two_loop(pairs, g)
        rel = np.linalg.norm(d_dev - d_host) / np.linalg.norm(d_host)
        print(f"step {k}: {opt.pairs} pairs, rel diff {rel:.1e}", flush=True)
        assert opt.pairs == len(pairs) and rel <= 1e-12
        xp, gp = x, g
        x = x + (0.7 if k % 3 else 1.0) * d_dev
    opt.reset()
    assert np.array_equal(opt.direction(x, grad(x)), -grad(x))
    opt.close()


def test_seisinvexpt_model_vector_round_trip(gpu_ctx):
    med = layered_acoustic(nz=40, nx=52, seed=2)
    pa = fwi_experiment(med, nss=1, nt=50, ctx=gpu_ctx)
    paw = G.SeisInvExpt(pa, [G.Recs(a.nr, pa.tgrid, pa.rfields) for a in pa.ageom], [8, 10])
    assert [g.n for g in paw.migrid] == [8, 10]
    assert paw.migrid[0].first == pytest.approx(pa.medium.grid[0].first + 4 * pa.medium.grid[0].step)
    m = paw.get_modelvector()
    assert m.shape == (paw.nm,)
    # log-parameters of a layered model are O(0.1) around the reference value
    assert np.all(np.isfinite(m)) and np.abs(m).max() < 2.0
    paw._to_full(m)
    assert paw.mfull.shape == pa.model_vector().shape


def test_source_estimation_matches_the_modelled_data(gpu_ctx):
    """update!(paw, pac): observed data made with a delayed, scaled source wavelet in the
    same medium; the estimated filter maps them onto the modelled data, so the misfit drops
    by orders of magnitude, and the filtered data equal dobs0 convolved with the filter."""
    from geophyinv.conv import Pconv
    med = layered_acoustic(nz=44, nx=60, seed=9)
    pa = fwi_experiment(med, nss=2, nt=320, ctx=gpu_ctx)
    # observed data: the same experiment with every source wavelet delayed 4 samples, x0.7
    obs_src = [s.copy() for s in pa.srcwav]
    for s in obs_src:
        for f in s.fields:
            w = s.d[f].copy()
            s.d[f][:] = 0.0
            s.d[f][4:] = 0.7 * w[:-4]
    pa_obs = fwi_experiment(med, nss=2, nt=320, ctx=gpu_ctx)
    pa_obs.update_srcwav([obs_src, pa_obs.srcwav2], [1, 0])
    dobs = oracle_records(med, pa_obs)
    paw = G.SeisInvExpt(pa, dobs)   # inversion grid = modelling grid: the model is the true one
    out = paw.update_source()
    print(f"misfit {out['loss_initial']:.6e} -> {out['loss_final']:.6e}", flush=True)
    assert out["loss_final"] < 1e-3 * out["loss_initial"]
    for d0, d in zip(paw.dobs0, paw.dobs):
        assert not np.array_equal(d0.d["vz"], d.d["vz"])
