"""GPU parity: libgpi (HIP, gfx950) against the CPU oracle on the same seeded inputs.

The bar is the relative-L2 tolerance of 1e-5 on records and final wavefields; the kernels
keep the oracle's FP32 evaluation order, so the observed difference is usually exactly 0.
"""
import numpy as np
import pytest

import geophyinv as G
import oracle as O
from geophyinv import gallery
from helpers import experiment, layered_acoustic, layered_elastic, rel_l2, wavelet

pytestmark = pytest.mark.gpu
TOL = 1e-5


def compare(pa, finals, label, nsteps=-1):
    orc = O.from_experiment(pa)
    orc.run(nsteps)
    recs, fin = pa.run_shot(0, final_fields=finals)
    worst = 0.0
    for q, f in enumerate(pa.rfields):
        e = rel_l2(recs[f], orc.records(q))
        worst = max(worst, e)
        assert np.all(np.isfinite(recs[f])), (label, f)
        assert e <= TOL, f"{label}: records {f} rel L2 {e:.3e}"
    for f in finals:
        e = rel_l2(fin[f], orc.field(f))
        worst = max(worst, e)
        assert e <= TOL, f"{label}: final {f} rel L2 {e:.3e}"
    print(f"{label}: worst rel L2 {worst:.3e}", flush=True)
    return recs, fin, orc


def test_acoustic_201_anomaly_npml20(gpu_ctx):
    med, ageom, srcwav, tgrid, kw = gallery.acoustic_anomaly_2d()
    pa = G.SeisForwExpt(G.FdtdAcoustic(), medium=med, tgrid=tgrid, ageom=ageom, srcwav=srcwav,
                        context=gpu_ctx, **kw)
    recs, _, _ = compare(pa, ["p", "vx", "vz"], "acoustic 201^2 npml 20")
    assert np.abs(recs["p"]).max() > 0


def test_acoustic_201_anomaly_reference_npml(gpu_ctx):
    med, ageom, srcwav, tgrid, kw = gallery.acoustic_anomaly_2d()
    kw["npml"] = 41   # 40 + order - 1, the reference's compile-time PML width
    pa = G.SeisForwExpt(G.FdtdAcoustic(), medium=med, tgrid=tgrid, ageom=ageom, srcwav=srcwav,
                        context=gpu_ctx, **kw)
    compare(pa, ["p", "vx", "vz"], "acoustic 201^2 npml 41")


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_acoustic_orders_mixed_sources(gpu_ctx, order):
    med = layered_acoustic(nz=52, nx=60)
    pa = experiment(med, nt=260, sfields=("p", "vx", "vz"), ns=3, nr=11, order=order, npml=9,
                    ctx=gpu_ctx, seed=order)
    compare(pa, ["p", "vx", "vz"], f"acoustic order {order}")


@pytest.mark.parametrize("faces", [["zmax", "xmin", "xmax"], ["zmin", "zmax"], []])
def test_acoustic_pml_face_subsets(gpu_ctx, faces):
    med = layered_acoustic(nz=40, nx=44)
    pa = experiment(med, nt=200, sfields=("p",), ns=2, nr=7, npml=8, pml_faces=faces, ctx=gpu_ctx)
    compare(pa, ["p", "vx", "vz"], f"acoustic faces {faces}")


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_elastic_orders(gpu_ctx, order):
    med = layered_elastic(nz=54, nx=58)
    pa = experiment(med, nt=260, sfields=("tauxx",), rfields=("vx", "vz"), ns=2, nr=9, order=order,
                    npml=9, ctx=gpu_ctx, seed=10 + order)
    compare(pa, ["tauxx", "tauzz", "tauxz", "vx", "vz"], f"elastic order {order}")


@pytest.mark.parametrize("order", [2, 4])
def test_elastic_free_surface_velocity_sources(gpu_ctx, order):
    med = layered_elastic(nz=50, nx=52)
    pa = experiment(med, nt=240, sfields=("vz", "vx", "tauzz"), rfields=("vx", "vz"), ns=3,
                    nr=8, order=order, npml=8, pml_faces=["zmax", "xmin", "xmax"],
                    stressfree_faces=["zmin"], ctx=gpu_ctx, seed=20 + order, src_depth=12.0)
    compare(pa, ["tauxx", "tauzz", "tauxz", "vx", "vz"], f"elastic free surface o{order}")


def test_shared_source_cells(gpu_ctx):
    """Several sources on the same cells: contributions must add in source order."""
    med = layered_acoustic(nz=40, nx=40)
    pa = experiment(med, nt=160, sfields=("p", "vx"), ns=4, nr=5, npml=8, ctx=gpu_ctx, jitter=False)
    a = pa.ageom[0]
    a.s["z"][:] = a.s["z"][0] + np.array([0.0, 1.0, 2.5, 0.0])
    a.s["x"][:] = a.s["x"][0] + np.array([0.0, 1.5, 0.5, 0.0])
    pa.update_ageom([a])
    compare(pa, ["p", "vx", "vz"], "shared source cells")


def test_elastic_full_size_grid_short(gpu_ctx):
    """The 2001^2 P-SV workload at its full grid (2081^2 with CPML), 40 steps."""
    med, ageom, srcwav, tgrid, kw = gallery.elastic_layers_2d()
    nt = 40
    tg = G.Grid(0.0, tgrid.step, nt)
    s = G.Srcs(1, tg, ["tauxx"])
    s.update(["tauxx"], wavelet(nt, tgrid.step, 300.0, 0.002))
    pa = G.SeisForwExpt(G.FdtdElastic(), medium=med, tgrid=tg, ageom=ageom, srcwav=[s],
                        context=gpu_ctx, **kw)
    compare(pa, ["tauxx", "vz"], "elastic 2081^2 x 40 steps")


def test_elastic_drop_in_workload_full_length(gpu_ctx):
    """The 2001^2 P-SV drop-in workload as the bench runs it: 2081^2 with CPML, all 4000
    steps, its own 20 Hz explosive Ricker, 801 surface receivers (vx, vz), against the
    oracle on the host (about 17 G cell-steps, some 25 s on 16 host threads)."""
    med, ageom, srcwav, tgrid, kw = gallery.elastic_layers_2d()
    assert len(tgrid) == 4000
    pa = G.SeisForwExpt(G.FdtdElastic(), medium=med, tgrid=tgrid, ageom=ageom, srcwav=srcwav,
                        context=gpu_ctx, **kw)
    recs, _, _ = compare(pa, ["vx", "vz"], "elastic 2081^2 x 4000 steps")
    assert np.abs(recs["vz"]).max() > 0


def test_medium_update_resident(gpu_ctx):
    """update!(pa, medium) re-uploads coefficients in place; results follow the new medium."""
    med = layered_acoustic(nz=36, nx=40)
    pa = experiment(med, nt=150, ns=1, nr=5, npml=8, ctx=gpu_ctx)
    r1, _ = pa.run_shot(0)
    med2 = med.copy()
    # stay inside the original [min, max] so the CPML profile (kept from construction, as
    # in the reference) is the one a fresh experiment would build
    med2.vp.m[10:20, 10:30] = med.vp.m.max()
    pa.update_medium(med2)
    r2, _ = pa.run_shot(0)
    assert rel_l2(r1["p"], r2["p"]) > 1e-3
    compare(pa, ["p"], "after medium update")


# ------------------------------------------------------------------------------ 3-D
from helpers import experiment3d, medium3d  # noqa: E402


@pytest.mark.parametrize("order", [2, 4, 8])
@pytest.mark.parametrize("compat", [1, 0])
def test_acoustic_3d(gpu_ctx, order, compat):
    n = (22, 18, 24) if order < 8 else (32, 30, 34)
    pa = experiment3d(medium3d(0, n=n), order=order, sfields=("p", "vy"), compat=compat,
                      ctx=gpu_ctx, seed=order + compat)
    compare(pa, ["p", "vx", "vy", "vz"], f"acoustic 3-D order {order} compat {compat}")


@pytest.mark.parametrize("order", [2, 4, 8])
def test_elastic_3d(gpu_ctx, order):
    n = (22, 18, 24) if order < 8 else (32, 30, 34)
    pa = experiment3d(medium3d(1, n=n), order=order, sfields=("tauxx", "vx"), rfields=("vx", "vy", "vz"),
                      ctx=gpu_ctx, seed=30 + order)
    compare(pa, ["tauxx", "tauyy", "tauzz", "tauxy", "tauxz", "tauyz", "vx", "vy", "vz"],
            f"elastic 3-D order {order}")


def test_elastic_3d_free_surface(gpu_ctx):
    pa = experiment3d(medium3d(1), order=2, sfields=("tauzz", "vz"), rfields=("vx", "vz"),
                      pml_faces=["zmax", "ymin", "ymax", "xmin", "xmax"], stressfree_faces=["zmin"],
                      ctx=gpu_ctx, seed=41)
    compare(pa, ["tauzz", "tauxz", "tauyz", "vz"], "elastic 3-D free surface")


def test_acoustic_3d_large_grid_short(gpu_ctx):
    """A 160^3 medium (170^3 with CPML) for 12 steps."""
    pa = experiment3d(medium3d(0, n=(160, 160, 160), ds=10.0), nt=12, npml=5, ns=1, nr=40,
                      ctx=gpu_ctx, seed=3)
    compare(pa, ["p", "vz"], "acoustic 3-D 170^3")


@pytest.mark.parametrize("physics,field", [(0, "p"), (0, "vz"), (1, "tauxz")])
def test_snapshots(gpu_ctx, physics, field):
    """pa[:snaps, i]: the field at the steps nearest to tsnaps equals the oracle's state
    after that many steps (snaps_save! at the end of the step, propagate.jl:237-244)."""
    med = layered_acoustic(nz=50, nx=58) if physics == 0 else layered_elastic(nz=50, nx=58)
    base = experiment(med, nt=200, sfields=("p",) if physics == 0 else ("tauxx",),
                      rfields=("vz",), ctx=gpu_ctx, seed=4)
    tv = base.tgrid.values()
    tsnaps = [tv[37] + 0.2 * base.tgrid.step, tv[120], tv[-1] + 1.0]
    pa = G.SeisForwExpt(base.attrib_mod, medium=med, tgrid=base.tgrid, ageom=base.ageom, srcwav=base.srcwav,
                        rfields=base.rfields, npml=base.npml, context=gpu_ctx, tsnaps=tsnaps, snaps_field=field)
    assert pa.itsnaps == [37, 120, 199]
    pa.update()
    snaps = pa["snaps", 0]
    for k, it in enumerate(pa.itsnaps):
        orc = O.from_experiment(pa)
        orc.run(it + 1)
        ref = orc.field(field)
        assert np.abs(ref).max() > 0
        assert rel_l2(snaps[k], ref) <= TOL, (field, it)
