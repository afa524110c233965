"""Misfit and gradient of a model vector (src/fdtd/func_grad.jl:1-49).

``lossvalue(m, loss, dobs, pa, mparams)`` puts the log-parameter vector ``m`` into the
experiment (``mod = exp(m) * ref`` on the unpadded region), models the data in the
:forward_save mode and returns the misfit.  ``gradient(g, m, loss, dobs, pa, mparams)``
also back-propagates the adjoint sources in the :adjoint mode and writes
``g = g_mod * exp(m) * ref`` (the chain rule of func_grad.jl:33-45).
"""
from __future__ import annotations

import numpy as np

from . import database

F32 = np.float32


def _forward_save(pa):
    mode_save = pa.attrib_mod.mode
    pa.attrib_mod.mode = "forward_save"
    pa.update_srcwav([pa.srcwav, pa.srcwav2], [1, 0])
    pa.update()
    return mode_save


def lossvalue(m, loss, dobs, pa, mparams=None) -> float:
    """``lossvalue(m, loss, dobs, pa, mparams)`` (func_grad.jl:1-9)."""
    if pa.npw != 2:
        raise ValueError("the experiment needs npw = 2")
    pa.update_model(m, mparams)
    mode_save = _forward_save(pa)
    pa.attrib_mod.mode = mode_save
    return database.lossvalue(loss, dobs, pa.data)


def gradient(g, m, loss, dobs, pa, mparams=None) -> float:
    """``gradient!(g, m, loss, dobs, pa, mparams)`` (func_grad.jl:11-49); fills ``g`` in place
    and returns the misfit.

    With ``pa.compat == 1`` the return value is the reference's: its :adjoint ``update!``
    zeroes ``pa.c.data`` (``initialize!``, types.jl:169-176) before ``lossvalue`` reads it,
    so it returns the misfit of zero data.  With ``compat == 0`` it returns the misfit of
    the modelled data."""
    if pa.npw != 2:
        raise ValueError("the experiment needs npw = 2")
    mparams = list(mparams or pa.mparams)
    pa.update_model(m, mparams)
    mode_save = _forward_save(pa)
    J = database.lossvalue(loss, dobs, pa.data)
    backprop(g, m, loss, dobs, pa, mparams)
    pa.attrib_mod.mode = mode_save
    return database.lossvalue(loss, dobs, pa.data) if pa.compat == 1 else J


def backprop(g, m, loss, dobs, pa, mparams=None) -> None:
    """The adjoint half of ``gradient!`` (func_grad.jl:21-45): from the records and stored
    boundary values of the :forward_save run just made at ``m``, back-propagate the adjoint
    sources and write the chain-ruled gradient into ``g``.  Leaves ``pa`` in the :adjoint
    mode with its data zeroed, as the reference's ``update!`` does."""
    mparams = list(mparams or pa.mparams)
    # adjoint sources, time reversed: gradient!(buffer, loss, dobs, dcal) of the time-reversed
    # records, the same FP32 values as reversing its output, in one pass
    database.gradient(pa.srcwav2, loss, [d.reversed() for d in dobs], [d.reversed() for d in pa.data])
    # the reference's first wavefield runs backwards with the type -1 sink (velocity sources
    # only); the checkpointed adjoint recomputes it from the stored run's own sources, so
    # any source field works there and the first wavelets are left as they are
    pa.update_srcwav([pa.srcwav, pa.srcwav2], [1 if pa.checkpoint else -1, 1])
    pa.attrib_mod.mode = "adjoint"
    pa.update()
    x = np.split(np.asarray(m, np.float64), len(mparams))
    flat = g.reshape(-1)
    if not np.shares_memory(flat, g):
        flat = np.empty(g.size, g.dtype)
    gs = np.split(flat, len(mparams))
    for xi, gi, name in zip(x, gs, mparams):
        gm = pa.gradients[name][pa.view_inner(pa.gradients[name])]
        # gm1 * exp(x1) * r in FP64, stored back into the FP32 gradient
        t = np.exp(xi.reshape(gm.shape, order="F"))
        t *= gm
        t *= np.float64(pa.ref_mod[name])
        gi.reshape(gm.shape, order="F")[...] = t if gi.dtype == F32 else t.astype(F32)
    if not np.shares_memory(flat, g):
        g[...] = flat.reshape(g.shape)


def chainrule_elastic(g_invlambda, g_invmu, g_rho, vp, vs, rho):
    """Gradients w.r.t. (vp, vs, rho) from gradients w.r.t. the independent elastic
    parameters (invlambda, invmu, rho) of the same cells, FP64.

    The elastic extension of old/chainrule.jl:43-45 (``grad_of_vp``/``grad_of_rho`` for
    invK = 1/(vp^2 rho)): with invlambda = 1/((vp^2 - 2 vs^2) rho) and invmu = 1/(vs^2 rho)
    (src/media/media.jl:111-130),
      d invlambda/d vp = -2 vp rho invlambda^2,  d invlambda/d vs = 4 vs rho invlambda^2,
      d invlambda/d rho = -invlambda/rho,        d invmu/d vs = -2 vs rho invmu^2,
      d invmu/d rho = -invmu/rho.
    Returns (g_vp, g_vs, g_rho)."""
    vp, vs, rho = (np.asarray(a, np.float64) for a in (vp, vs, rho))
    gl, gm, gr = (np.asarray(a, np.float64) for a in (g_invlambda, g_invmu, g_rho))
    il = 1.0 / ((vp * vp - 2.0 * vs * vs) * rho)
    im = 1.0 / (vs * vs * rho)
    g_vp = gl * (-2.0 * vp * rho * il * il)
    g_vs = gl * (4.0 * vs * rho * il * il) + gm * (-2.0 * vs * rho * im * im)
    g_rho = gr + gl * (-il / rho) + gm * (-im / rho)
    return g_vp, g_vs, g_rho
