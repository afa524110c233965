"""Models built from exactly the arrays the Julia binding passes (geophyinv.jl_amd/julia/
GPI.jl, `coef_arrays`): every name of MediumParameters (src/fdtd/medium.jl:81-95), i.e.
pac.dmod plus pac.mod with ids 0-13, for 2-D and 3-D, acoustic and elastic; records of
gpi_forward against the oracle.  The array set of the earlier documentation (dmod plus rho
only) must be refused for elastic models, since the elastic passes read invlambda and
invmu."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from geophyinv import _lib
from geophyinv.fields import FIELD_ID
from helpers import experiment, experiment3d, layered_acoustic, layered_elastic, medium3d, rel_l2

pytestmark = pytest.mark.gpu

# (independent, dependent) names of MediumParameters, src/fdtd/medium.jl:81-95
MEDIUM_PARAMETERS = {
    (0, 2): (["invK", "rho"], ["dtinvavxirho", "dtinvavzirho", "dtK"]),
    (0, 3): (["invK", "rho"], ["dtinvavxirho", "dtinvavyirho", "dtinvavzirho", "dtK"]),
    (1, 2): (["invlambda", "invmu", "rho"], ["dtinvavxirho", "dtinvavzirho", "dtavmu", "dtM", "dtlambda"]),
    (1, 3): (["invlambda", "invmu", "rho"], ["dtinvavxirho", "dtinvavyirho", "dtinvavzirho", "dtavxzimu",
                                               "dtavxyimu", "dtavyzimu", "dtM", "dtlambda"]),
}


def binding_model(pa, names):
    """gpi_model_create with exactly `names` (the binding's coef_arrays) and pac.pml."""
    ctx = pa._ctx
    keep, arrs = pa.coef_arrays(names=set(names))
    assert len(arrs) == len(names), (len(arrs), names)
    pmls = []
    for df, (a, b, k) in pa.pml.items():
        p = _lib.gpi_pml()
        p.field, p.len = FIELD_ID[df], a.size
        p.a, p.b, p.kI = _lib.fptr(a), _lib.fptr(b), _lib.fptr(k)
        pmls.append(p)
    parr = (_lib.gpi_pml * len(pmls))(*pmls)
    h = C.c_uint64()
    g = pa.grid_struct()
    ctx.check(ctx.lib.gpi_model_create(ctx.h, C.byref(g), arrs, len(arrs), parr, len(pmls), C.byref(h)))
    return h.value


def build(physics, ndim, ctx):
    if ndim == 2:
        med = layered_acoustic(nz=44, nx=50) if physics == 0 else layered_elastic(nz=44, nx=50)
        return experiment(med, nt=220, sfields=("p",) if physics == 0 else ("tauxx", "vz"),
                          rfields=("p", "vx", "vz") if physics == 0 else ("vx", "vz"), ns=2, nr=9, npml=8, ctx=ctx)
    med = medium3d(physics, n=(22, 20, 24))
    return experiment3d(med, nt=110, sfields=("p",) if physics == 0 else ("tauxx",),
                        rfields=("p", "vx", "vy", "vz") if physics == 0 else ("vx", "vy", "vz"), npml=5, ctx=ctx)


@pytest.mark.parametrize("physics,ndim", sorted(MEDIUM_PARAMETERS))
def test_binding_array_set_forward_matches_oracle(gpu_ctx, physics, ndim):
    pa = build(physics, ndim, gpu_ctx)
    indep, dep = MEDIUM_PARAMETERS[(physics, ndim)]
    print(f"binding array set: physics {physics} ndim {ndim}", flush=True)
    h = binding_model(pa, indep + dep)
    try:
        pa._model = h
        recs, _ = pa.run_shot(0)
    finally:
        pa._model = None
        gpu_ctx.check(gpu_ctx.lib.gpi_model_free(gpu_ctx.h, h))
    orc = O.from_experiment(pa)
    orc.run()
    for q, f in enumerate(pa.rfields):
        assert np.abs(recs[f]).max() > 0
        assert rel_l2(recs[f], orc.records(q)) <= 1e-5, f


@pytest.mark.parametrize("ndim", [2, 3])
def test_elastic_model_without_invlambda_invmu_is_refused(gpu_ctx, ndim):
    pa = build(1, ndim, gpu_ctx)
    indep, dep = MEDIUM_PARAMETERS[(1, ndim)]
    with pytest.raises(_lib.GpiError) as e:
        binding_model(pa, dep + ["rho"])
    assert e.value.code == -1
