"""GPU: the inversion-grid gradient of SeisInvExpt (src/fwi/func_grad.jl) against central
finite differences of the misfit, and a short L-BFGS run that lowers the misfit.

The bulk-modulus part of the reference gradient is the discrete adjoint of the pressure
update away from the boundary strips, so the directional derivative along an invK
direction that vanishes next to the PML agrees with finite differences to 1e-3 (the bar
written here).  Within about three cells of the strips the reference images a wavefield
that is re-forced only at the next step, so directions reaching there are not used."""
import numpy as np
import pytest

import geophyinv as G
from helpers import fwi_experiment, layered_acoustic, oracle_records

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
    print(f"misfit {J0:.6e} -> {res.fun:.6e} after {res.nit} iterations", flush=True)
    assert res.fun < 0.9 * J0


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
