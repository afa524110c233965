"""GPU parity of the checkpointed gradient (gpi_forward_checkpoint + gpi_adjoint_checkpoint)
against the oracle's exact-wavefield gradient (tests/helpers.py: oracle_gradient_exact), 2-D
and 3-D acoustic.

Both form each step's imaging term in FP32 with the same expressions and accumulate in
FP64 in the same order, so they agree bit for bit; the bar written here is the relative-L2
tolerance of 1e-5.  Checkpoint intervals cover one step, an interval that does not divide
nt, and one longer than the run (a single segment).
"""
import numpy as np
import pytest

import geophyinv as G
from helpers import (fwi_experiment, fwi_experiment3d, layered_acoustic, medium3d, oracle_gradient_exact,
                     oracle_records, rel_l2)

pytestmark = pytest.mark.gpu
TOL = 1e-5

CASES = [
    dict(ndim=2, order=2, every=64, nss=2, rho=True, rfields=("vz",), label="2d o2 every 64, 2 shots"),
    dict(ndim=2, order=4, every=1, nss=1, rho=True, rfields=("p", "vx", "vz"), label="2d o4 every 1"),
    dict(ndim=2, order=2, every=1000, nss=1, rho=True, rfields=("vx",), label="2d o2 one segment"),
    dict(ndim=2, order=8, every=37, nss=1, rho=False, rfields=("vz",), label="2d o8 every 37 invK only"),
    dict(ndim=3, order=2, every=50, nss=1, rho=True, rfields=("vz",), label="3d o2 every 50"),
    dict(ndim=3, order=4, every=33, nss=1, rho=True, rfields=("vx", "vz"), label="3d o4 every 33"),
    dict(ndim=3, order=2, every=40, nss=2, rho=False, rfields=("vz",), label="3d o2 invK only, 2 shots"),
]


def build(case, ctx):
    kw = dict(checkpoint=case["every"], grad_rho=case["rho"], ctx=ctx, order=case["order"], nss=case["nss"],
              rfields=case["rfields"])
    if case["ndim"] == 2:
        true = layered_acoustic(nz=44, nx=60, seed=9)
        med0 = true.copy()
        med0.vp.m[18:26, 24:36] *= np.float32(0.96)
        pa = fwi_experiment(med0, nt=300, npml=10 + case["order"], **kw)
    else:
        true = medium3d(0, n=(26, 22, 24), seed=5)
        med0 = true.copy()
        med0.vp.m[10:16, 8:14, 9:15] *= np.float32(0.95)
        pa = fwi_experiment3d(med0, nt=160, npml=5 + case["order"], **kw)
    return true, pa


@pytest.mark.parametrize("case", CASES, ids=[c["label"] for c in CASES])
def test_checkpoint_gradient_parity(gpu_ctx, case):
    print("checkpoint gradient", case["label"], flush=True)
    true, pa = build(case, gpu_ctx)
    dobs = oracle_records(true, pa)
    m = pa.model_vector()
    g = np.zeros_like(m)
    G.gradient(g, m, G.L2DistLoss(), dobs, pa)
    J, gk, gr, dcal = oracle_gradient_exact(pa, dobs, grad_rho=case["rho"])
    assert pa.gradients["invK"].dtype == np.float64
    e_k = rel_l2(pa.gradients["invK"], gk)
    assert np.abs(gk).max() > 0
    msg = f"  rel L2 invK {e_k:.3e}"
    if case["rho"]:
        e_r = rel_l2(pa.gradients["rho"], gr)
        msg += f" rho {e_r:.3e}"
        assert np.abs(gr).max() > 0 and e_r <= TOL
    else:
        assert not np.any(pa.gradients["rho"])
    print(msg, flush=True)
    assert e_k <= TOL
    # the :forward_save misfit of the same model vector (records of the checkpointed run)
    assert G.lossvalue(m, None, dobs, pa) == pytest.approx(J, rel=1e-12)
    for iss in range(len(pa.ageom)):
        for f in pa.rfields:
            assert rel_l2(pa.data[iss].d[f], dcal[iss][f]) <= TOL
    # the same experiment again: stores are replaced, results repeat exactly
    g2 = np.zeros_like(m)
    G.gradient(g2, m, G.L2DistLoss(), dobs, pa)
    assert np.array_equal(g, g2)


def test_checkpoint_matches_reference_flow_in_the_interior(gpu_ctx):
    """2-D: the checkpointed gradient equals the reference flow's (time reversal with
    boundary strips) away from the strips, where the reversal reconstructs the forward
    wavefield up to FP32 rounding."""
    true = layered_acoustic(nz=44, nx=60, seed=9)
    med0 = true.copy()
    med0.vp.m[18:26, 24:36] *= np.float32(0.96)
    pa_c = fwi_experiment(med0, nt=300, ctx=gpu_ctx, checkpoint=50)
    pa_r = fwi_experiment(med0, nt=300, ctx=gpu_ctx)
    dobs = oracle_records(true, pa_c)
    for pa in (pa_c, pa_r):
        m = pa.model_vector()
        G.gradient(np.zeros_like(m), m, None, dobs, pa)
    n = pa_c.npml + 4
    inner = (slice(n, -n), slice(n, -n))
    for k in ("invK", "rho"):
        e = rel_l2(pa_r.gradients[k][inner], pa_c.gradients[k][inner])
        print(f"  {k}: reference flow vs checkpointed, interior rel L2 {e:.3e}", flush=True)
        assert e < 1e-5


def test_elastic_checkpoint_needs_velocity_records(gpu_ctx):
    """Elastic gradients (tests/test_gpu_elastic_gradient.py) take velocity records only: the
    adjoint of a stress or p record is not a velocity source."""
    from helpers import layered_elastic
    kw = dict(medium=layered_elastic(), tgrid=G.Grid(0.0, 1e-4, 10),
              ageom=[G.AGeomss({"z": [50.0], "x": [100.0]}, {"z": [0.0], "x": [100.0]})],
              srcwav=[G.Srcs(1, G.Grid(0.0, 1e-4, 10), ["vz"])], checkpoint=10, context=gpu_ctx)
    with pytest.raises(ValueError):
        G.SeisForwExpt(G.FdtdElastic("forward_save", 2), rfields=["p", "vz"], **kw)
    G.SeisForwExpt(G.FdtdElastic("forward_save", 2), rfields=["vx", "vz"], **kw).free()
