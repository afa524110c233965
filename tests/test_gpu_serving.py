"""GPU: device coefficient preparation and the serving scheduler (continuous batching).

- coefficients derived on the device from the unpadded medium equal the host's (the
  reference's update_dmod! arithmetic) bit for bit, and so do the records of a shot run
  on them;
- every shot run through the server, whatever else is resident, whatever workgroup shape
  the batched kernels use and however the HBM budget staggers admission, gives the same
  records as the same shot run alone through the plain path (the bar written here is the
  relative-L2 tolerance of 1e-5; the observed difference is 0).
"""
import numpy as np
import pytest

import geophyinv as G
from geophyinv import _lib
from geophyinv.serve import Server
from helpers import layered_acoustic, layered_elastic, rel_l2, wavelet

pytestmark = pytest.mark.gpu
TOL = 1e-5


def shot(ctx, seed, physics=0, order=2, coefficients="host", faces=None):
    """A seeded shot of random size, duration, source field and receiver count."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nz, nx = int(rng.integers(40, 130)), int(rng.integers(40, 150))
    med = layered_acoustic(nz=nz, nx=nx, seed=seed) if physics == 0 else layered_elastic(nz=nz, nx=nx, seed=seed)
    gz, gx = med.grid
    nt = int(rng.integers(60, 260))
    dt = 0.4 * gz.step / (1.6 * float(med.vp.m.max()))
    tg = G.Grid(0.0, dt, nt)
    nr = int(rng.integers(3, 40))
    sf = ["p", "vz"][seed % 2] if physics == 0 else ["tauxx", "vx"][seed % 2]
    ag = G.AGeomss({"z": [gz.last * rng.uniform(0.2, 0.6)], "x": [gx.last * rng.uniform(0.2, 0.8)]},
                   {"z": rng.uniform(0, 0.3, nr) * gz.last, "x": np.linspace(0, gx.last, nr)})
    s = G.Srcs(1, tg, [sf])
    s.update([sf], wavelet(nt, dt, 25.0, 0.02 + 0.01 * rng.uniform()) * 1e3)
    rf = ["p", "vx", "vz"] if physics == 0 else ["vx", "vz"]
    attrib = G.FdtdAcoustic() if physics == 0 else G.FdtdElastic()
    kw = {} if faces is None else {"pml_faces": faces}
    return G.SeisForwExpt(attrib, medium=med, tgrid=tg, ageom=[ag], srcwav=[s], rfields=rf, order=order,
                          npml=int(rng.integers(6, 14)), context=ctx, coefficients=coefficients, **kw)


@pytest.mark.parametrize("physics,order", [(0, 2), (0, 4), (1, 2), (1, 8)])
def test_device_coefficients_equal_host(gpu_ctx, physics, order):
    pa = shot(gpu_ctx, 3 + order, physics, order, coefficients="device",
              faces=["zmax", "xmin", "xmax"] if order == 4 else None)
    host = G.SeisForwExpt.__new__(G.SeisForwExpt)
    host.__dict__.update(pa.__dict__)
    host._mod = host._coefs = host._ref_mod = None
    for name, a in host.coefs.items():
        if name in ("invK", "invlambda", "invmu") and physics == (0 if name != "invK" else 1):
            continue
        d = pa.device_coef(name)
        assert d.shape == a.shape, name
        assert np.array_equal(d, a), (name, np.abs(d - a).max())
    rec_d, _ = pa.run_shot(0)
    pb = shot(gpu_ctx, 3 + order, physics, order, coefficients="host",
              faces=["zmax", "xmin", "xmax"] if order == 4 else None)
    rec_h, _ = pb.run_shot(0)
    for f in pa.rfields:
        assert np.array_equal(rec_d[f], rec_h[f]), f
    pa.free()
    pb.free()


def run_server(ctx, shots, **kw):
    plain = []
    for pa in shots:
        r, _ = pa.run_shot(0)
        plain.append(r)
        pa.free()
    with Server(shots[0].physics, shots[0].order, **kw) as srv:
        tickets = [srv.submit(pa, 0) for pa in shots]
        got = [srv.wait(t, timeout=120.0) for t in tickets]
        st = srv.stats()
    worst = 0.0
    for k, (a, b) in enumerate(zip(plain, got)):
        for f in a:
            e = rel_l2(b[f], a[f])
            worst = max(worst, e)
            assert e <= TOL, (k, f, e)
            assert np.abs(a[f]).max() > 0 or f == "p"
    return st, worst


@pytest.mark.parametrize("tile", [0, 1, 2, 3])
def test_server_alone_equals_mix_fixed_tiles(gpu_ctx, tile):
    shots = [shot(gpu_ctx, 100 + k) for k in range(6)]
    st, worst = run_server(gpu_ctx, shots, autotune=False, tile=tile, chunk_steps=16)
    print(f"tile {tile}: {st['completed']} shots, mean batch {st['mean_batch']:.2f}, worst rel L2 {worst:.1e}, "
          f"{st['graph_chunks']} graph chunks / {st['graph_captures']} captures", flush=True)
    assert st["completed"] == 6 and st["failed"] == 0
    assert st["mean_batch"] > 1.5    # shots did overlap on the device
    # chunks whose batch did not change replayed captured graphs, some graphs more than once
    assert st["graph_captures"] >= 1 and st["graph_chunks"] > st["graph_captures"]
    assert st["host_admit_ms"] > 0 and st["host_launch_ms"] > 0 and st["host_retire_ms"] > 0


def test_server_direct_launch_equals_plain(gpu_ctx):
    shots = [shot(gpu_ctx, 100 + k) for k in range(6)]
    st, _ = run_server(gpu_ctx, shots, autotune=False, tile=0, chunk_steps=16, graphs=False)
    assert st["completed"] == 6 and st["graph_chunks"] == 0 and st["graph_captures"] == 0


@pytest.mark.skip(reason="persistent cooperative chunk launches hung on the MI355X (the first case, 2-D "
                         "acoustic order 2, made no progress for 180 s); no cause in this code was shown, so "
                         "the mode was taken out of the library and this case stays off the GPU")
@pytest.mark.parametrize("physics,order", [(0, 2), (0, 4), (1, 2)])
def test_server_persistent_launch_equals_plain(gpu_ctx, physics, order):
    """Every chunk as one cooperative launch (grid barriers between the passes): each served
    shot equals the plain path, with staggered admission (at most 3 resident) and the
    autotuner switching workgroup shapes between chunks."""
    shots = [shot(gpu_ctx, 500 + k, physics=physics, order=order) for k in range(5)]
    st, _ = run_server(gpu_ctx, shots, chunk_steps=16, max_shots=3, persistent=True)
    assert st["completed"] == 5 and st["failed"] == 0
    assert st["graph_chunks"] == 0 and st["graph_captures"] == 0
    assert st["launches"] < 2 * st["steps"]   # one launch per chunk


def test_server_autotune_orders_and_staggered_admission(gpu_ctx):
    shots = [shot(gpu_ctx, 200 + k, order=4) for k in range(8)]
    st, worst = run_server(gpu_ctx, shots, autotune=True, chunk_steps=8, max_shots=3)
    print(f"autotune: tile {st['tile_ty']}x{st['tile_rx']}, mean batch {st['mean_batch']:.2f}, p50 "
          f"{st['p50_ms']:.1f} ms p99 {st['p99_ms']:.1f} ms", flush=True)
    assert st["completed"] == 8 and st["mean_batch"] <= 3.0 + 1e-9
    # a budget that holds about two of these shots still serves all of them
    shots = [shot(gpu_ctx, 300 + k) for k in range(5)]
    st, _ = run_server(gpu_ctx, shots, hbm_budget_bytes=5 << 20, chunk_steps=8)
    assert st["completed"] == 5


def test_server_elastic(gpu_ctx):
    shots = [shot(gpu_ctx, 400 + k, physics=1, order=4) for k in range(5)]
    st, _ = run_server(gpu_ctx, shots, chunk_steps=12)
    assert st["completed"] == 5


def test_served_stream_shots_against_the_oracle(gpu_ctx):
    """Shots of the serving-stream generator (layered medium + anomaly, 20-cell CPML, p
    source five cells down, p and vz surface receivers) with sides of 256-512 cells and
    1000-1500 steps, served together; each shot's records against the oracle directly."""
    import oracle as O
    from geophyinv import gallery
    stream = gallery.serving_stream(3, seed=3003, side=(256, 512), steps=(1000, 1500))
    pas = [G.SeisForwExpt(G.FdtdAcoustic(), medium=s["medium"], tgrid=s["tgrid"], ageom=s["ageom"],
                          srcwav=s["srcwav"], context=gpu_ctx, **s["kw"]) for s in stream]
    print("served stream shots vs oracle:", [(s["side"], s["nt"]) for s in stream], flush=True)
    with Server(0, 2, chunk_steps=32) as srv:
        tickets = [srv.submit(pa, 0) for pa in pas]
        got = [srv.wait(t, timeout=120.0) for t in tickets]
        st = srv.stats()
    assert st["completed"] == 3 and st["mean_batch"] > 1.5
    for k, pa in enumerate(pas):
        orc = O.from_experiment(pa)
        orc.run()
        for q, f in enumerate(pa.rfields):
            e = rel_l2(got[k][f], orc.records(q))
            print(f"  shot {k} ({stream[k]['side']}^2, {stream[k]['nt']} steps) {f}: rel L2 {e:.1e}", flush=True)
            assert np.abs(got[k][f]).max() > 0 and e <= TOL, (k, f, e)
        pa.free()


def test_server_rejects(gpu_ctx):
    with Server(0, 2) as srv:
        pa = shot(gpu_ctx, 7, order=4)
        with pytest.raises(_lib.GpiError):
            srv.submit(pa)            # order differs from the server's
    with Server(0, 2, hbm_budget_bytes=1 << 20) as srv:
        pa = shot(gpu_ctx, 8)
        with pytest.raises(_lib.GpiError):
            srv.submit(pa)            # larger than the whole budget
