"""GPU: the reverse rule of the misfit (geophyinv.chainrules) against ``gradient!``.

``rrule(m, loss, dobs, pa)`` runs the :forward_save mode once and its pullback runs only the
adjoint half, so its gradient must equal, bit for bit, what ``gradient!`` returns for the
same model vector (src/fdtd/func_grad.jl:11-49), scaled by the cotangent; the torch
``Misfit`` function must hand the same array to autograd.  A pullback taken after the
experiment ran again, or taken twice, raises.
"""
import numpy as np
import pytest

import geophyinv as G
from geophyinv import chainrules
from helpers import fwi_experiment, layered_acoustic, oracle_records

pytestmark = pytest.mark.gpu


def setup(ctx, **kw):
    true = layered_acoustic(nz=44, nx=60, seed=9)
    med0 = true.copy()
    med0.vp.m[16:28, 22:40] *= np.float32(0.95)
    pa = fwi_experiment(med0, nss=2, nt=300, ctx=ctx, **kw)
    return pa, oracle_records(true, pa)


@pytest.mark.parametrize("kw", [{}, {"checkpoint": 40, "grad_rho": True}], ids=["boundary-saved", "checkpointed"])
def test_pullback_equals_gradient(gpu_ctx, kw):
    print("rrule", kw, flush=True)
    pa, dobs = setup(gpu_ctx, **kw)
    loss = G.L2DistLoss()
    m = pa.model_vector()
    g = np.zeros_like(m)
    G.gradient(g, m, loss, dobs, pa)
    J_ref = G.lossvalue(m, loss, dobs, pa)

    J, pullback = G.rrule(m, loss, dobs, pa)
    assert J == J_ref
    dm = pullback(2.5)
    assert dm.dtype == np.float64 and dm.shape == m.shape
    assert np.abs(g).max() > 0
    assert np.array_equal(dm, 2.5 * g.astype(np.float64))
    with pytest.raises(RuntimeError, match="already"):
        pullback(1.0)


def test_pullback_rejects_a_changed_experiment(gpu_ctx):
    pa, dobs = setup(gpu_ctx)
    m = pa.model_vector()
    _, pullback = G.rrule(m, G.L2DistLoss(), dobs, pa)
    pa.update()
    with pytest.raises(RuntimeError, match="changed"):
        pullback(1.0)
    _, pullback = G.rrule(m, G.L2DistLoss(), dobs, pa)
    pa.update_model(m + 0.01)
    with pytest.raises(RuntimeError, match="changed"):
        pullback(1.0)


def test_torch_misfit_backward(gpu_ctx):
    import torch
    pa, dobs = setup(gpu_ctx)
    loss = G.L2DistLoss()
    m = pa.model_vector()
    g = np.zeros_like(m)
    G.gradient(g, m, loss, dobs, pa)
    J_ref = G.lossvalue(m, loss, dobs, pa)

    mt = torch.tensor(m, dtype=torch.float64, requires_grad=True)
    J = chainrules.Misfit.apply(mt, loss, dobs, pa, None)
    assert float(J) == J_ref
    (3.0 * J).backward()
    assert np.array_equal(mt.grad.numpy(), 3.0 * g.astype(np.float64))
