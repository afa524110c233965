"""GPU: source illumination (illum_flag): the per-supersource sum over the time steps of
p^2 after each step, FP32 square and FP64 sum (compute_illum!, src/fdtd/fdtd.jl:570-581),
stacked over the supersources on the unpadded grid (stack_illums!, fdtd.jl:556-566).

The oracle is stepped one time step at a time and its pressure squared and summed in the
same order, so the per-supersource illumination is equal bit for bit; the stack is the sum
of the inner views.  The records are unchanged by the flag.
"""
import numpy as np
import pytest

import geophyinv as G
import oracle as O
from helpers import fwi_experiment, fwi_experiment3d, layered_acoustic, medium3d

pytestmark = pytest.mark.gpu


def oracle_illum(pa, iss):
    orc = O.from_experiment(pa, iss)
    acc = None
    for it in range(len(pa.tgrid)):
        orc.step(it)
        p = orc.field("p")
        sq = (p * p).astype(np.float64)
        acc = sq if acc is None else acc + sq
    return acc


@pytest.mark.parametrize("ndim", [2, 3])
def test_illumination_matches_oracle(gpu_ctx, ndim):
    print(f"illumination {ndim}-D", flush=True)
    if ndim == 2:
        med = layered_acoustic(nz=44, nx=52, seed=21)
        pa = fwi_experiment(med, nss=2, nt=220, mode="forward", ctx=gpu_ctx, illum_flag=True)
        plain = fwi_experiment(med, nss=2, nt=220, mode="forward", ctx=gpu_ctx)
    else:
        med = medium3d(0, n=(22, 18, 20), seed=22)
        pa = fwi_experiment3d(med, nss=1, nt=120, ctx=gpu_ctx, illum_flag=True)   # :forward_save
        plain = fwi_experiment3d(med, nss=1, nt=120, ctx=gpu_ctx)
    pa.update()
    plain.update()
    stack = np.zeros_like(pa["illum_stack"])
    for iss in range(len(pa.ageom)):
        ref = oracle_illum(pa, iss)
        got = pa.illums[iss]
        assert ref.max() > 0
        assert np.array_equal(got, ref), (iss, np.abs(got - ref).max())
        stack += ref[pa.view_inner(ref)]
        for f in pa.rfields:
            assert np.array_equal(pa.data[iss].d[f], plain.data[iss].d[f])
    assert pa["illum_stack"].shape == tuple(len(g) for g in med.grid)
    assert np.array_equal(pa["illum_stack"], stack)
    pa.update()   # the stack starts from zero on every update
    assert np.array_equal(pa["illum_stack"], stack)
    pa.free()
    plain.free()
