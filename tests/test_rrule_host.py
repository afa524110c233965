"""Host logic of the misfit's reverse rule (geophyinv.chainrules) with the modelling stubbed:
the pullback scales the gradient by the cotangent, runs once, and refuses an experiment that
changed after the primal run.  The modelling itself is checked on the GPU
(tests/test_gpu_rrule.py)."""
import types

import numpy as np
import pytest

from geophyinv import chainrules


class FakeExperiment:
    def __init__(self):
        self.generation = 0
        self.mparams = ["invK", "rho"]
        self.attrib_mod = types.SimpleNamespace(mode="forward")


@pytest.fixture
def stubbed(monkeypatch):
    calls = []

    def lossvalue(m, loss, dobs, pa, mparams):
        pa.generation += 1
        calls.append(("loss", list(mparams)))
        return float(np.sum(m * m))

    def backprop(g, m, loss, dobs, pa, mparams):
        pa.generation += 1
        pa.attrib_mod.mode = "adjoint"
        calls.append(("backprop", list(mparams)))
        g[...] = (2.0 * m).astype(np.float32)

    monkeypatch.setattr(chainrules.func_grad, "lossvalue", lossvalue)
    monkeypatch.setattr(chainrules.func_grad, "backprop", backprop)
    return calls


def test_pullback_scales_and_runs_once(stubbed):
    pa = FakeExperiment()
    m = np.array([0.5, -1.0, 2.0, 0.25])
    J, pb = chainrules.rrule(m, None, None, pa)
    assert J == pytest.approx(5.3125)
    g = pb(3.0)
    assert np.array_equal(g, 3.0 * (2.0 * m))
    assert pa.attrib_mod.mode == "forward"     # the caller's mode comes back
    assert [c[0] for c in stubbed] == ["loss", "backprop"]
    assert stubbed[0][1] == ["invK", "rho"]
    with pytest.raises(RuntimeError, match="already"):
        pb(1.0)


def test_pullback_refuses_changed_experiment(stubbed):
    pa = FakeExperiment()
    m = np.ones(4)
    _, pb = chainrules.rrule(m, None, None, pa)
    pa.generation += 1                          # any run or model/source/geometry change
    with pytest.raises(RuntimeError, match="changed"):
        pb(1.0)
    assert [c[0] for c in stubbed] == ["loss"]


def test_primal_input_is_copied(stubbed):
    pa = FakeExperiment()
    m = np.ones(4)
    _, pb = chainrules.rrule(m, None, None, pa)
    m[:] = 7.0                                  # the caller reuses its buffer
    assert np.array_equal(pb(1.0), np.full(4, 2.0))


def test_torch_function_hands_the_pullback_to_autograd(stubbed):
    torch = pytest.importorskip("torch")
    pa = FakeExperiment()
    mt = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64, requires_grad=True)
    J = chainrules.Misfit.apply(mt, None, None, pa, None)
    assert J.dtype == torch.float64 and float(J) == pytest.approx(5.25)
    (0.5 * J).backward()
    assert torch.equal(mt.grad, 0.5 * 2.0 * mt.detach())
