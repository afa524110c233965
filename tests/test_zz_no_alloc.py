"""gpi_forward_static, the forward entry point that allocates no device memory (its per-shot
buffers come from the library's static scratch array), against gpi_forward on the same
shots: records and final wavefields bit-identical, on consecutive calls whose second input
is smaller than the first (the array keeps the previous call's contents), and a shot whose
buffers exceed the array is refused before any step runs."""
import numpy as np
import pytest

from geophyinv import _lib
from helpers import experiment, experiment3d, layered_acoustic, layered_elastic, medium3d

pytestmark = pytest.mark.gpu

STATIC_BYTES = 64 << 20   # GPI_STATIC_SCRATCH_BYTES


def _pair(pa, finals):
    print(f"gpi_forward vs gpi_forward_static: {pa.physics} {pa.ndim}-D {pa.ext_shape} nt={len(pa.tgrid)}",
          flush=True)
    ref, ref_fin = pa.run_shot(0, final_fields=finals)
    got, got_fin = pa.run_shot(0, final_fields=finals, static=True)
    assert set(ref) == set(got)
    for f in ref:
        assert np.all(np.isfinite(ref[f])) and np.abs(ref[f]).max() > 0, f
        assert np.array_equal(ref[f], got[f]), f
    for f in finals:
        assert np.array_equal(ref_fin[f], got_fin[f]), f


def test_static_equals_forward_on_consecutive_calls(gpu_ctx):
    big = experiment(layered_acoustic(nz=96, nx=120), nt=420, sfields=("p", "vx"), ns=2, nr=41, npml=12,
                     ctx=gpu_ctx)
    small = experiment(layered_elastic(nz=40, nx=44), nt=160, sfields=("tauxx",), rfields=("vx", "vz"), ns=1,
                       nr=7, npml=8, ctx=gpu_ctx, stressfree_faces=["zmin"], pml_faces=["zmax", "xmin", "xmax"])
    _pair(big, ("p", "vz"))
    _pair(small, ("vx",))        # second call, smaller input: nothing left over from the first
    _pair(big, ())                # and back to the larger one


def test_static_equals_forward_3d(gpu_ctx):
    pa = experiment3d(medium3d(0, n=(22, 18, 24)), nt=90, order=4, ctx=gpu_ctx)
    _pair(pa, ("p",))
    pa = experiment3d(medium3d(1, n=(16, 14, 18)), nt=60, sfields=("tauxx",), rfields=("vx", "vy", "vz"),
                      ctx=gpu_ctx)
    _pair(pa, ())


def test_static_refuses_shots_beyond_the_array(gpu_ctx):
    # 18000 x 1000 FP32 records on the device alone exceed 64 MiB
    nt, nr = 18000, 1000
    assert nt * nr * 4 > STATIC_BYTES
    pa = experiment(layered_acoustic(nz=40, nx=44), nt=nt, rfields=("p",), nr=nr, npml=8, ctx=gpu_ctx)
    with pytest.raises(_lib.GpiError) as e:
        pa.run_shot(0, static=True)
    assert e.value.code == -2 and "static scratch" in str(e.value)
    # the context stays usable, and a fitting shot after the refusal is still exact
    _pair(experiment(layered_acoustic(nz=40, nx=44), nt=200, nr=9, npml=8, ctx=gpu_ctx), ("p",))
