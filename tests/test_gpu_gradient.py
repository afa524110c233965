"""GPU parity of the time-reversal path: :forward_save, :adjoint and the model-vector
misfit/gradient (src/fdtd/func_grad.jl:1-49) against the oracle's reference flow.

The gradients keep the oracle's FP32 evaluation order, so they agree bit for bit; the bar
written here is the relative-L2 tolerance of 1e-5.
"""
import numpy as np
import pytest

import geophyinv as G
from geophyinv import _lib
from helpers import (fwi_experiment, layered_acoustic, layered_elastic, medium3d, experiment3d, oracle_gradient,
                     oracle_records, rel_l2)

pytestmark = pytest.mark.gpu
TOL = 1e-5


def test_forward_save_records_equal_forward(gpu_ctx):
    med = layered_acoustic(nz=48, nx=60, seed=3)
    pa = fwi_experiment(med, nss=2, nt=280, rfields=("p", "vz"), ctx=gpu_ctx)
    pa.update()   # :forward_save
    saved = [{f: d.d[f].copy() for f in pa.rfields} for d in pa.data]
    pa.attrib_mod.mode = "forward"
    pa.update()
    for iss, d in enumerate(pa.data):
        for f in pa.rfields:
            assert np.array_equal(saved[iss][f], d.d[f]), (iss, f)
            assert np.abs(d.d[f]).max() > 0


CASES = [
    dict(order=2, npml=10, nss=2, sfields=("vz",), rfields=("vz",), pml_faces=None, label="o2 vz/vz"),
    dict(order=4, npml=12, nss=2, sfields=("vx",), rfields=("p", "vx", "vz"), pml_faces=None, label="o4 vx/p,vx,vz"),
    dict(order=2, npml=10, nss=1, sfields=("vz", "vx"), rfields=("vz",), pml_faces=["zmax", "xmin", "xmax"],
         label="o2 open top"),
    dict(order=8, npml=14, nss=1, sfields=("vz",), rfields=("vx",), pml_faces=None, label="o8 vz/vx"),
]


@pytest.mark.parametrize("case", CASES, ids=[c["label"] for c in CASES])
def test_gradient_parity(gpu_ctx, case):
    print("gradient parity", case["label"], flush=True)
    true = layered_acoustic(nz=44, nx=60, seed=9)
    med0 = true.copy()
    med0.vp.m[18:26, 24:36] *= np.float32(0.96)
    kw = {k: case[k] for k in ("order", "npml", "nss", "sfields", "rfields", "pml_faces")}
    pa = fwi_experiment(med0, nt=320, ctx=gpu_ctx, **kw)
    dobs = oracle_records(true, pa)
    m = pa.model_vector()
    g = np.zeros_like(m)
    J_gpu = G.gradient(g, m, G.L2DistLoss(), dobs, pa)
    J, gk, gr, dcal = oracle_gradient(pa, dobs)
    e_k = rel_l2(pa.gradients["invK"], gk)
    e_r = rel_l2(pa.gradients["rho"], gr)
    print(f"  rel L2 invK {e_k:.3e} rho {e_r:.3e}", flush=True)
    assert np.abs(gk).max() > 0 and np.abs(gr).max() > 0
    assert e_k <= TOL and e_r <= TOL
    # chain rule on the unpadded region (func_grad.jl:33-45)
    x = np.split(m, 2)
    exp = []
    for xi, name, full in zip(x, ["invK", "rho"], [gk, gr]):
        inner = full[pa.view_inner(full)]
        exp.append((inner.astype(np.float64) * np.exp(xi.reshape(inner.shape, order="F"))
                    * np.float64(pa.ref_mod[name])).astype(np.float32).ravel(order="F"))
    assert rel_l2(g, np.concatenate(exp)) <= TOL
    # compat = 1: the reference returns the misfit of the zeroed data after :adjoint
    assert J_gpu == pytest.approx(G.lossvalue(None, dobs, [G.Recs(a.nr, pa.tgrid, pa.rfields) for a in pa.ageom]))
    # the forward_save misfit of the same model vector
    J2 = G.lossvalue(m, None, dobs, pa)
    assert J2 == pytest.approx(J, rel=1e-12)
    for iss in range(len(pa.ageom)):
        for f in pa.rfields:
            assert rel_l2(pa.data[iss].d[f], dcal[iss][f]) <= TOL


def test_gradient_compat0_returns_misfit(gpu_ctx):
    true = layered_acoustic(nz=40, nx=52, seed=4)
    med0 = true.copy()
    med0.rho.m[15:22, 20:30] *= np.float32(1.05)
    pa = fwi_experiment(med0, nss=1, nt=260, ctx=gpu_ctx)
    pa.compat = 0
    dobs = oracle_records(true, pa)
    m = pa.model_vector()
    g = np.zeros_like(m)
    J = G.gradient(g, m, None, dobs, pa)
    J_o, _, _, _ = oracle_gradient(pa, dobs)
    assert J == pytest.approx(J_o, rel=1e-12) and J > 0
    # a second gradient call reuses the experiment (stores are replaced, not leaked)
    g2 = np.zeros_like(m)
    G.gradient(g2, m, None, dobs, pa)
    assert np.array_equal(g, g2)


def test_adjoint_needs_forward_save(gpu_ctx):
    med = layered_acoustic(nz=40, nx=52, seed=4)
    pa = fwi_experiment(med, nss=1, nt=120, mode="adjoint", ctx=gpu_ctx)
    with pytest.raises(RuntimeError):
        pa.update()


def test_no_gradient_for_elastic(gpu_ctx):
    med = layered_elastic(nz=40, nx=52)
    gz, gx = med.grid
    dt = 0.4 * gz.step / (1.6 * float(med.vp.m.max()))
    tg = G.Grid(0.0, dt, 100)
    ag = G.AGeomss({"z": [60.0], "x": [120.0]}, {"z": np.full(5, 150.0), "x": np.linspace(20, 240, 5)})
    s = G.Srcs(1, tg, ["vz"])
    pa = G.SeisForwExpt(G.FdtdElastic("forward_save", 2), medium=med, tgrid=tg, ageom=[ag], srcwav=[s],
                        rfields=["vz"], npml=10, context=gpu_ctx)
    pa.update()   # boundary saving of 2-D elastic exists
    pa.attrib_mod.mode = "adjoint"
    with pytest.raises(NotImplementedError):
        pa.update()


def test_forward_save_rejects_3d_elastic(gpu_ctx):
    pa = experiment3d(medium3d(1), nt=40, sfields=("vz",), rfields=("vx", "vy", "vz"), ctx=gpu_ctx)
    pa.attrib_mod = G.FdtdElastic("forward_save", 2)
    pa.npw = 2
    pa.srcwav2 = []
    with pytest.raises(_lib.GpiError):
        pa.run_shot(0, save=True)


def test_forward_save_3d_acoustic_and_2d_elastic(gpu_ctx):
    """The boundary strips of every axis (3-D) and of the three stored stresses (elastic)."""
    pa = experiment3d(medium3d(0), nt=60, ctx=gpu_ctx)
    rec_s, _ = pa.run_shot(0, save=True)
    rec_f, _ = pa.run_shot(0)
    for f in pa.rfields:
        assert np.array_equal(rec_s[f], rec_f[f]), f
    med = layered_elastic(nz=40, nx=52)
    gz, gx = med.grid
    dt = 0.4 * gz.step / (1.6 * float(med.vp.m.max()))
    tg = G.Grid(0.0, dt, 160)
    ag = G.AGeomss({"z": [60.0], "x": [120.0]}, {"z": np.full(5, 150.0), "x": np.linspace(20, 240, 5)})
    s = G.Srcs(1, tg, ["vz"])
    s.update(["vz"], np.sin(np.arange(160) * dt * 2 * np.pi * 20.0) * np.exp(-((np.arange(160) - 40) / 12.0) ** 2))
    pa = G.SeisForwExpt(G.FdtdElastic("forward_save", 2), medium=med, tgrid=tg, ageom=[ag], srcwav=[s],
                        rfields=["vx", "vz"], npml=10, context=gpu_ctx)
    rec_s, _ = pa.run_shot(0, save=True)
    rec_f, _ = pa.run_shot(0)
    for f in pa.rfields:
        assert np.array_equal(rec_s[f], rec_f[f]), f
        assert np.abs(rec_f[f]).max() > 0
