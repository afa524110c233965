"""GPU: a shot whose dt violates the stability condition (src/fdtd/stability.jl:8-60, which
only warns) grows without bound; its records hold Inf/NaN and every path returns
GPI_ENONFINITE instead of status 0: the plain call, the server and the job queue.  The
kernels keep running on non-finite values, so nothing hangs.  A stable shot in the same
server still completes with its plain records."""
import warnings

import numpy as np
import pytest

import geophyinv as G
from geophyinv import _lib
from geophyinv.serve import JobServer, Server
from helpers import layered_acoustic, wavelet

pytestmark = pytest.mark.gpu
ENONFINITE = -4


def shot(ctx, cfl, nt=400, seed=1):
    """2-D acoustic shot with dt = cfl * ds / vmax (stable below 1/sqrt(2) at order 2)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    med = layered_acoustic(nz=50, nx=60, seed=seed)
    gz, gx = med.grid
    dt = cfl * gz.step / float(med.vp.m.max())
    tg = G.Grid(0.0, dt, nt)
    ag = G.AGeomss({"z": [gz.last * 0.4], "x": [gx.last * 0.5]},
                   {"z": np.zeros(7), "x": np.linspace(0, gx.last, 7)})
    s = G.Srcs(1, tg, ["p"])
    s.update(["p"], wavelet(nt, dt, 25.0, 0.03) * (1.0 + rng.uniform()))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return G.SeisForwExpt(G.FdtdAcoustic(), medium=med, tgrid=tg, ageom=[ag], srcwav=[s],
                              rfields=["p", "vz"], npml=8, context=ctx)


def test_plain_path_returns_enonfinite(gpu_ctx):
    print("plain: unstable shot (cfl 1.4)", flush=True)
    bad = shot(gpu_ctx, 1.4)
    with pytest.raises(_lib.GpiError) as e:
        bad.run_shot(0)
    assert e.value.code == ENONFINITE
    assert "non-finite" in str(e.value)
    bad.free()
    good = shot(gpu_ctx, 0.4)
    recs, _ = good.run_shot(0)
    assert all(np.isfinite(r).all() for r in recs.values())
    good.free()


def test_server_returns_enonfinite_and_serves_the_rest(gpu_ctx):
    print("server: unstable + stable shots", flush=True)
    bad = shot(gpu_ctx, 1.4, seed=2)
    good = shot(gpu_ctx, 0.4, seed=3)
    ref, _ = good.run_shot(0)
    with Server(0, 2) as srv:
        tb = srv.submit(bad)
        tg = srv.submit(good)
        with pytest.raises(_lib.GpiError) as e:
            srv.wait(tb, timeout=120.0)
        assert e.value.code == ENONFINITE
        out = srv.wait(tg, timeout=120.0)
    for f in good.rfields:
        assert np.array_equal(out[f], ref[f]), f


def test_job_queue_returns_enonfinite(gpu_ctx):
    print("job queue: unstable shot", flush=True)
    bad = shot(gpu_ctx, 1.4, seed=4)
    with JobServer(workers=1) as q:
        t = q.submit(bad)
        with pytest.raises(_lib.GpiError) as e:
            q.wait(t, timeout=120.0)
    assert e.value.code == ENONFINITE
