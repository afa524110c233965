"""Benchmark: the 2-D elastic P-SV drop-in workload on one MI355X.

Workload: 2001 x 2001 grid at 2.5 m, 8 flat layers (vp 1500 -> 4500 m/s, vs = vp/1.8,
Gardner density, +-3% smoothed perturbation), 40-cell CPML (2081 x 2081 extended grid),
20 Hz explosive Ricker, 801 surface receivers recording vx and vz, 4000 time steps, FP32.
One bench "step" is one whole shot (4000 time steps) with the medium resident in HBM.

Metric: grid-point updates per second (extended-grid cells x time steps / time).  The
roofline of the dominant kernel (the stress pass) is measured live with HIP events on the
library's stream; bytes per launch are the algorithmic bytes of that pass.  The CPU baseline
is the oracle (a C restatement of the reference's unfused pass structure, OpenMP) timed on
a bounded number of time steps of the same workload on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "geophyinv.jl_amd"))

HBM_PEAK_GBS = 8000.0


def strip_cells(n_cells, n_axis, npml, faces_on_axis):
    """Cells of the CPML strips along one axis (2*npml of every n_axis lines)."""
    return n_cells * (npml * faces_on_axis) / n_axis


def algorithmic_bytes(nz, nx, npml, physics):
    """Compulsory FP32 bytes of one velocity pass and one stress pass (each array read once,
    each output written once; derivatives never stored; CPML memory read+written in strips)."""
    N = nz * nx
    sx = strip_cells(N, nx, npml, 2)
    sz = strip_cells(N, nz, npml, 2)
    if physics == 1:
        # velocity: r tauxx,tauzz,tauxz,vx,vz,2 coefs; w vx,vz = 9 floats; 2 psi per strip cell
        vel = 36 * N + 16 * (sx + sz)
        # stress: r vx,vz,tauxx,tauzz,tauxz,dtM,dtlambda,dtavmu; w 3 taus = 11 floats
        stress = 44 * N + 16 * (sx + sz)
    else:
        vel = 28 * N + 8 * (sx + sz)
        stress = 20 * N + 8 * (sx + sz)
    return vel, stress


def cpu_baseline(pa, seconds):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    orc = O.from_experiment(pa)
    nz, nx = pa.ext_shape
    t0 = time.perf_counter()
    orc.run(8)
    probe = time.perf_counter() - t0
    steps = max(8, int(seconds / max(probe / 8, 1e-6)))
    steps = min(steps, len(pa.tgrid))
    orc2 = O.from_experiment(pa)
    t0 = time.perf_counter()
    orc2.run(steps)
    dt = time.perf_counter() - t0
    return {"value": nz * nx * steps / dt / 1e9, "unit": "Gpt-updates/s", "cores": O.threads(),
            "kind": "port",
            "sample": f"oracle (unfused C restatement, OpenMP) on the full {nz}x{nx} grid, "
                      f"first {steps} of {len(pa.tgrid)} time steps, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nt", type=int, default=4000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the single-GPU path only "
              "(shots would shard one per GPU; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import geophyinv as G
    from geophyinv import _lib, gallery

    med, ageom, srcwav, tgrid, kw = gallery.elastic_layers_2d(nt=args.nt)
    ctx = _lib.Context(0)
    ctx.set_pass_sampling(64)
    pa = G.SeisForwExpt(G.FdtdElastic(), medium=med, tgrid=tgrid, ageom=ageom, srcwav=srcwav,
                        context=ctx, **kw)
    nz, nx = pa.ext_shape
    nt = len(tgrid)
    for _ in range(args.warmup):
        pa.run_shot(0)
    dev_ms, vel_ms, str_ms = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        recs, _ = pa.run_shot(0)
        t = pa.timing
        dev_ms.append(t["total_ms"])
        vel_ms.append(t["velocity_avg_ms"])
        str_ms.append(t["stress_avg_ms"])
    wall = time.perf_counter() - t0
    finite = all(np.all(np.isfinite(r)) for r in recs.values())
    gpt = nz * nx * nt * args.steps
    value = gpt / wall / 1e9
    vel_b, str_b = algorithmic_bytes(nz, nx, pa.npml, 1)
    # Each time step is two launches (velocity pass, stress pass).  Event brackets around a
    # sampled pass add the marker latency, so the measured per-step device time is split
    # between the two passes in the ratio of their bracket means: the per-launch figure is
    # the pass's share of the real loop (kernel plus its dispatch gap).
    s_br = float(np.mean(str_ms)) * 1e-3
    v_br = float(np.mean(vel_ms)) * 1e-3
    step_s = float(np.mean(dev_ms)) * 1e-3 / nt
    s_avg = step_s * s_br / (s_br + v_br)
    v_avg = step_s * v_br / (s_br + v_br)
    achieved = str_b / s_avg / 1e9
    step_bytes = vel_b + str_b
    line = {
        "metric": "Gpt-updates/s (2-D elastic FDTD, FP32)",
        "value": round(value, 3),
        "unit": "Gpt-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "2-D elastic P-SV FdtdElastic, 2001x2001 at 2.5 m, 40-cell CPML "
                               "(2081x2081), 8 layers + 3% smoothed noise, explosive 20 Hz Ricker, "
                               "801 receivers (vx, vz), one shot per step",
                   "time_steps": nt, "grid": [nz, nx], "order": pa.order},
        "device_ms_per_step": round(float(np.mean(dev_ms)), 3),
        "device_gpt_per_s": round(nz * nx * nt / (float(np.mean(dev_ms)) * 1e-3) / 1e9, 3),
        "step_algorithmic_GBps": round(step_bytes * nt / (float(np.mean(dev_ms)) * 1e-3) / 1e9, 1),
        "roofline": {"bound": "hbm", "kernel": "str2d_elastic<2> (stress pass)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": int(str_b), "avg_launch_ms": round(s_avg * 1e3, 5),
                     "bracket_ms": round(s_br * 1e3, 5),
                     "velocity_pass": {"bytes_per_launch": int(vel_b), "avg_launch_ms": round(v_avg * 1e3, 5),
                                       "bracket_ms": round(v_br * 1e3, 5),
                                       "achieved": round(vel_b / v_avg / 1e9, 1)}},
        "records_finite": bool(finite),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(pa, args.cpu_seconds)
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
