"""Reverse rule of the misfit of a model vector.

The Python counterpart of the ChainRulesCore rule a Julia caller would register for the
reference's ``lossvalue(m, loss, dobs, pa, mparams)`` (src/fdtd/func_grad.jl:1-9): the
primal runs the :forward_save mode (records plus the stored boundary values or
checkpoints) and returns the misfit; the pullback runs only the adjoint half of
``gradient!`` (func_grad.jl:21-45) on what that run stored, so the forward modelling is not
repeated.  ``Misfit`` wraps the same pair as a ``torch.autograd.Function`` for callers that
differentiate with torch (the tensors stay on the host: the experiment owns the device
state).

The pullback reads the experiment's records and stores, so it must run before the
experiment is run again or its model, sources or geometry change (``pa.generation``); it
raises otherwise, and each pullback runs once (it leaves the experiment in the :adjoint
mode with its records zeroed, as ``gradient!`` does).
"""
from __future__ import annotations

import numpy as np

from . import database, func_grad

F32 = np.float32


def rrule(m, loss, dobs, pa, mparams=None):
    """``rrule(lossvalue, m, loss, dobs, pa, mparams)``: returns ``(J, pullback)``, where
    ``pullback(dJ)`` returns ``dJ * dJ/dm`` (FP64 array of the shape of ``m``)."""
    mparams = list(mparams or pa.mparams)
    m = np.array(m, dtype=np.float64)
    J = func_grad.lossvalue(m, loss, dobs, pa, mparams)
    gen = pa.generation
    mode = pa.attrib_mod.mode
    state = {"used": False}

    def pullback(dJ=1.0):
        if state["used"]:
            raise RuntimeError("this pullback has run already; evaluate the rule again")
        if pa.generation != gen:
            raise RuntimeError("the experiment changed after the primal run (another run or a new "
                               "model, sources or geometry); its stored wavefield no longer matches m")
        state["used"] = True
        g = np.zeros(m.size, F32)
        func_grad.backprop(g, m, loss, dobs, pa, mparams)
        pa.attrib_mod.mode = mode
        return np.float64(dJ) * g.astype(np.float64)

    return J, pullback


def _torch_function():
    import torch

    class Misfit(torch.autograd.Function):
        """``Misfit.apply(m, loss, dobs, pa, mparams)``: the misfit as a 0-d FP64 tensor whose
        backward is the pullback of :func:`rrule`."""

        @staticmethod
        def forward(ctx, m, loss, dobs, pa, mparams=None):
            J, pb = rrule(m.detach().cpu().numpy(), loss, dobs, pa, mparams)
            ctx.pullback = pb
            ctx.m_dtype, ctx.m_device = m.dtype, m.device
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, dJ):
            g = ctx.pullback(float(dJ))
            ctx.pullback = None
            return (torch.from_numpy(g).to(dtype=ctx.m_dtype, device=ctx.m_device), None, None, None, None)

    return Misfit


_MISFIT = None


def __getattr__(name):
    # torch is imported on first use of ``Misfit`` only
    global _MISFIT
    if name == "Misfit":
        if _MISFIT is None:
            _MISFIT = _torch_function()
        return _MISFIT
    raise AttributeError(name)
