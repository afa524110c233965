"""GPU: receiver sampling by wavefront shuffles (`gather_block`, pass_extras.h) at receiver
counts around its workgroup size (2-D: 4 taps per receiver, 64 receivers per gather
workgroup; 3-D: 8 taps, 32 per workgroup), with several receiver sets per pass (p in the
velocity pass, the velocities in the stress pass), acoustic and elastic: records against
the oracle's interpolation (src/fdtd/receiver.jl:3-14) on the same inputs.
"""
import numpy as np
import pytest

import oracle as O
from helpers import experiment, experiment3d, layered_acoustic, layered_elastic, medium3d, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


def check(pa):
    recs, _ = pa.run_shot(0)
    orc = O.from_experiment(pa)
    orc.run()
    for q, f in enumerate(pa.rfields):
        ref = orc.records(q)
        assert recs[f].shape == ref.shape, f
        assert np.abs(ref).max() > 0, f
        assert rel_l2(recs[f], ref) <= TOL, (f, rel_l2(recs[f], ref))
    pa.free()


@pytest.mark.parametrize("nr", [63, 64, 65, 129])
@pytest.mark.parametrize("physics", ["acoustic", "elastic"])
def test_gather_2d_receiver_counts(gpu_ctx, physics, nr):
    print(f"gather: 2-D {physics}, {nr} receivers", flush=True)
    if physics == "acoustic":
        med = layered_acoustic(nz=30, nx=44, seed=12)
        pa = experiment(med, nt=90, nr=nr, rfields=("p", "vx", "vz"), ctx=gpu_ctx, seed=50 + nr, rec_depth=20.0)
    else:
        med = layered_elastic(nz=30, nx=44, seed=13)
        pa = experiment(med, nt=90, nr=nr, sfields=("tauzz",), rfields=("vx", "vz"), ctx=gpu_ctx, seed=60 + nr,
                        rec_depth=20.0)
    check(pa)


@pytest.mark.parametrize("nr", [31, 32, 33, 97])
@pytest.mark.parametrize("physics", [0, 1])
def test_gather_3d_receiver_counts(gpu_ctx, physics, nr):
    print(f"gather: 3-D physics {physics}, {nr} receivers", flush=True)
    med = medium3d(physics, seed=14 + physics)
    if physics == 0:
        pa = experiment3d(med, nt=70, nr=nr, rfields=("p", "vx", "vy", "vz"), ctx=gpu_ctx, seed=70 + nr)
    else:
        pa = experiment3d(med, nt=70, nr=nr, sfields=("tauzz",), rfields=("vx", "vy", "vz"), ctx=gpu_ctx,
                          seed=80 + nr)
    check(pa)
