"""Pins of the oracle's time reversal and reference gradient (CPU only).

- time reversal: re-propagating a :forward_save run backwards from its final snapshots,
  with the boundary strips forced and the type -1 source sink, reproduces the forward
  records (p with the sign of the stored -p snapshot, v unchanged) to FP32 rounding;
- the invK gradient equals the central finite difference of the misfit along smooth
  perturbations (the imaging condition of compute_gmodKI! is the discrete adjoint of the
  pressure update);
- the rho gradient agrees with the central finite difference (at fixed bulk modulus)
  within 10 percent: the reference's density imaging (compute_gmodrho!) averages
  velocity products on neighbouring cells, which is not the exact discrete adjoint.
"""
import numpy as np
import pytest

import geophyinv as G
from helpers import fwi_experiment, layered_acoustic, oracle_gradient, oracle_pair, oracle_records as records


@pytest.mark.parametrize("sfield", ["vz", "vx"])
def test_time_reversal_reconstructs_records(sfield):
    med = layered_acoustic(nz=40, nx=56, seed=7)
    pa = fwi_experiment(med, nss=1, nt=300, sfields=(sfield,), rfields=("p", "vz"))
    s1, s2, get_source = oracle_pair(pa, 0)
    s1.run_save()
    fwd = {f: s1.records(q).astype(np.float64) for q, f in enumerate(pa.rfields)}
    s1.set_wavelets(0, get_source(pa.srcwav[0].d[sfield], sfield, -1))
    s1.adjoint(s2)   # the first wavefield records while it runs backwards
    bwd = {f: s1.records(q).astype(np.float64) for q, f in enumerate(pa.rfields)}
    nt = len(pa.tgrid)
    k = np.arange(1, nt - 1)
    peak_p = np.abs(fwd["p"]).max()
    peak_v = np.abs(fwd["vz"]).max()
    assert peak_p > 0 and peak_v > 0
    # p recorded at the start of backward step k is -p of forward step nt - k
    e_p = np.abs(bwd["p"][k] + fwd["p"][nt - k]).max() / peak_p
    # vz recorded after the backward velocity update of step k is vz of forward step nt-1-k
    e_v = np.abs(bwd["vz"][k] - fwd["vz"][nt - 1 - k]).max() / peak_v
    assert e_p < 5e-6 and e_v < 5e-6, (e_p, e_v)
    # a wrong alignment is far off: the check has teeth
    assert np.abs(bwd["p"][k] + fwd["p"][nt - k - 1]).max() / peak_p > 1e-2


def blob(shape, cz, cx, rad):
    zz, xx = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    return np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2.0 * rad ** 2))


def test_gradient_matches_directional_finite_differences():
    true = layered_acoustic(nz=40, nx=56, seed=7)
    med0 = true.copy()
    med0.vp.m[18:24, 25:35] *= np.float32(0.97)
    pa = fwi_experiment(med0, nss=2, nt=300)
    dobs = records(true, pa)
    J, gk, gr, _ = oracle_gradient(pa, dobs)
    assert J > 0
    npml = pa.npml
    inner = (slice(npml, npml + 40), slice(npml, npml + 56))
    vp0 = med0.vp.m.astype(np.float64)
    rho0 = med0.rho.m.astype(np.float64)
    invK = 1.0 / (vp0 ** 2 * rho0)
    eps = 1e-3

    def misfit(m):
        return G.lossvalue(None, dobs, records(m, pa))

    for (cz, cx, rad) in [(20, 28, 4.0), (12, 20, 3.0)]:
        b = blob((40, 56), cz, cx, rad)
        # invK <- invK (1 + eps b) at fixed rho
        Js = []
        for s in (1, -1):
            m = med0.copy()
            m.vp.m[:] = (vp0 * (1 + s * eps * b) ** -0.5).astype(np.float32)
            Js.append(misfit(m))
        fd = (Js[0] - Js[1]) / (2 * eps)
        ad = float(np.sum(gk[inner] * invK * b))
        assert abs(ad / fd - 1) < 1e-3, ("invK", cz, cx, ad, fd)
        # rho <- rho (1 + eps b) at fixed bulk modulus
        Js = []
        for s in (1, -1):
            m = med0.copy()
            f = 1 + s * eps * b
            m.rho.m[:] = (rho0 * f).astype(np.float32)
            m.vp.m[:] = (vp0 / np.sqrt(f)).astype(np.float32)
            Js.append(misfit(m))
        fd = (Js[0] - Js[1]) / (2 * eps)
        ad = float(np.sum(gr[inner] * rho0 * b))
        assert abs(ad / fd - 1) < 1e-1, ("rho", cz, cx, ad, fd)


def test_gradient_sign_is_pinned():
    """A plausible mistake, the adjoint source with the opposite sign, flips the gradient."""
    true = layered_acoustic(nz=40, nx=56, seed=7)
    med0 = true.copy()
    med0.vp.m[18:24, 25:35] *= np.float32(0.97)
    pa = fwi_experiment(med0, nss=1, nt=300)
    dobs = records(true, pa)
    _, gk, _, dcal = oracle_gradient(pa, dobs)
    # dobs' = 2 dcal - dobs mirrors the residual: the gradient must change sign
    mirror = [G.Recs(a.nr, pa.tgrid, pa.rfields,
                     {f: (2.0 * dcal[i][f].astype(np.float64) - dobs[i].d[f]).astype(np.float32)
                      for f in pa.rfields}) for i, a in enumerate(pa.ageom)]
    _, gk2, _, _ = oracle_gradient(pa, mirror)
    sl = (slice(pa.npml, -pa.npml), slice(pa.npml, -pa.npml))
    corr = np.sum(gk[sl].astype(np.float64) * gk2[sl]) / (np.linalg.norm(gk[sl]) * np.linalg.norm(gk2[sl]))
    assert corr < -0.99
