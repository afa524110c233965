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

The default line also carries two serving measurements: ``serving`` (128 2-D acoustic
shots of at most 512 cells per side through the batched server and through plain calls)
and ``serving_stream`` (the 512-shot serving stream at its own sizes, 256-2048 cells per
side, through the job queue and through plain calls); both compare the records of the two
paths.  Other workloads: ``--workload acoustic3d | elastic3d | gradient3d |
gradient3d_elastic | serve | serve_jobs | serve3d | fwi_optimizer | source_estimation``.
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


def algorithmic_bytes_3d(n, npml):
    """3-D acoustic: velocity pass reads p, vx, vy, vz and three density coefficients and writes
    vx, vy, vz (40 B per cell); stress pass reads vx, vy, vz, p, dtK and writes p (24 B per
    cell); one CPML memory variable read and written per strip cell per axis (8 B)."""
    N = n[0] * n[1] * n[2]
    strips = sum(strip_cells(N, n[a], npml, 2) for a in range(3))
    return 40 * N + 8 * strips, 24 * N + 8 * strips


def algorithmic_bytes_3d_elastic(n, npml):
    """3-D elastic: velocity pass reads the six stresses, vx, vy, vz and rho and writes the
    velocities (52 B per cell); stress pass reads vx, vy, vz, the six stresses, invlambda and
    invmu and writes the stresses (68 B per cell); three CPML memory variables read and
    written per strip cell per axis in each pass (24 B)."""
    N = n[0] * n[1] * n[2]
    strips = sum(strip_cells(N, n[a], npml, 2) for a in range(3))
    return 52 * N + 24 * strips, 68 * N + 24 * strips


def algorithmic_bytes(nz, nx, npml, physics):
    """Compulsory FP32 bytes of one velocity pass and one stress pass (each array read once,
    each output written once; derivatives never stored; CPML memory read+written in strips;
    the dt-scaled coefficients are formed in the passes from rho, invlambda and invmu)."""
    N = nz * nx
    sx = strip_cells(N, nx, npml, 2)
    sz = strip_cells(N, nz, npml, 2)
    if physics == 1:
        # velocity: r tauxx,tauzz,tauxz,vx,vz,rho; w vx,vz = 8 floats; 2 psi per strip cell
        vel = 32 * N + 16 * (sx + sz)
        # stress: r vx,vz,tauxx,tauzz,tauxz,invlambda,invmu; w 3 taus = 10 floats
        stress = 40 * N + 16 * (sx + sz)
    else:
        vel = 24 * N + 8 * (sx + sz)
        stress = 20 * N + 8 * (sx + sz)
    return vel, stress


def reference_bytes(shape, npml, physics):
    """Algorithmic bytes of one velocity pass and one stress pass by the per-cell figures of
    the two-pass staggered leapfrog as the reference stores its coefficients (every array
    read once per pass that uses it, outputs written once, one CPML memory variable read and
    written per derivative along its axis in the strips): 2-D acoustic 28 / 20 B, 2-D elastic
    36 / 44 B (dtinvav{x,z}irho; dtM, dtlambda, dtavmu), 3-D acoustic 40 / 24 B, 3-D elastic
    60 / 80 B (three density coefficients; dtM, dtlambda and the three averaged moduli), plus
    8 B (acoustic, 2-D elastic: 16 B; 3-D elastic: 24 B) per strip cell per axis in each pass.
    The passes here read rho, invlambda and invmu instead and form the dt-scaled coefficients
    in registers, so their own compulsory bytes (the functions above) are smaller for the
    elastic passes."""
    N = 1
    for n_axis in shape:
        N *= n_axis
    strips = sum(strip_cells(N, n_axis, npml, 2) for n_axis in shape)
    if len(shape) == 2:
        (vc, sc, e) = (36, 44, 16) if physics == 1 else (28, 20, 8)
    else:
        (vc, sc, e) = (60, 80, 24) if physics == 1 else (40, 24, 8)
    return vc * N + e * strips, sc * N + e * strips


def measured_traffic(summary, kernel):
    """HBM bytes per launch of a kernel from a committed counter summary (FETCH_SIZE +
    WRITE_SIZE runs, calibrated with a copy of known size), or None."""
    import re
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch %s (\d+)$" % re.escape(kernel), open(path).read(), re.M)
    return int(m.group(1)) if m else None


def cpu_baseline(pa, seconds):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    import numpy as np

    orc = O.from_experiment(pa)
    cells = int(np.prod(pa.ext_shape))
    t0 = time.perf_counter()
    orc.run(8)
    probe = time.perf_counter() - t0
    steps = max(8, int(seconds / max(probe / 8, 1e-6)))
    steps = min(steps, len(pa.tgrid))
    orc2 = O.from_experiment(pa)
    t0 = time.perf_counter()
    orc2.run(steps)
    dt = time.perf_counter() - t0
    return {"value": cells * steps / dt / 1e9, "unit": "Gpt-updates/s", "cores": O.threads(),
            "kind": "port",
            "sample": f"oracle (unfused C restatement, OpenMP) on the full {'x'.join(map(str, pa.ext_shape))} grid, "
                      f"first {steps} of {len(pa.tgrid)} time steps, {dt:.1f} s"}


def _digest(recs):
    """Checksums of one shot's records (compared between the plain and served runs)."""
    import xxhash
    return tuple(xxhash.xxh3_64_intdigest(memoryview(recs[f].reshape(-1, order="A")).cast("B")) for f in sorted(recs))


def _experiment(G, rq, **kw):
    return G.SeisForwExpt(G.FdtdAcoustic(), medium=rq["medium"], tgrid=rq["tgrid"], ageom=rq["ageom"],
                          srcwav=rq["srcwav"], **kw, **rq["kw"])


def run_plain(reqs, ctx, G, prebuilt=False):
    """The drop-in path on the stream: one shot per call, in arrival order, each prepared on
    the host as the reference does (coefficients included) and run by gpi_forward.  With
    ``prebuilt`` the experiments (geometry weights, wavelets, CPML vectors) are made before
    the stream starts; the coefficients are still derived and uploaded per shot."""
    import numpy as np

    if isinstance(prebuilt, list):   # experiments made by the caller
        pas, prebuilt = list(prebuilt), True
    else:
        pas = [_experiment(G, rq, context=ctx, coefficients="host") for rq in reqs] if prebuilt else None
    t0 = time.perf_counter()
    lat, out = [], []
    for k, rq in enumerate(reqs):
        t_arr = t0 + rq["arrival"]
        now = time.perf_counter()
        if now < t_arr:
            time.sleep(t_arr - now)
        pa = pas[k] if prebuilt else _experiment(G, rq, context=ctx, coefficients="host")
        recs, _ = pa.run_shot(0)
        pa.free()
        lat.append(time.perf_counter() - t_arr)
        out.append(recs)
        if len(reqs) >= 256 and (k + 1) % 64 == 0:
            print(f"plain path: {k + 1}/{len(reqs)} shots", flush=True)
    span = time.perf_counter() - t0
    return span, np.array(lat), [_digest(r) for r in out]


def run_served(reqs, G, opts, prebuilt=False):
    """The same stream through the serving scheduler: each shot is prepared on the host
    (weights, CPML vectors; coefficients on the device) and submitted at its arrival; with
    ``prebuilt`` the host preparation happens before the stream starts, as in run_plain."""
    import numpy as np
    from geophyinv.serve import Server

    pas = [_experiment(G, rq, coefficients="device") for rq in reqs] if prebuilt else None
    with Server(0, 2, **opts) as srv:
        t0 = time.perf_counter()
        tickets = []
        for k, rq in enumerate(reqs):
            t_arr = t0 + rq["arrival"]
            now = time.perf_counter()
            if now < t_arr:
                time.sleep(t_arr - now)
            pa = pas[k] if prebuilt else _experiment(G, rq, coefficients="device")
            if prebuilt:
                pas[k] = None
            lag = time.perf_counter() - t_arr
            tickets.append((srv.submit(pa, 0), lag))
            del pa
        lat, out = [], []
        for t, lag in tickets:
            out.append(srv.wait(t))
            lat.append(lag + srv.latency_ms(t) * 1e-3)
        span = time.perf_counter() - t0
        st = srv.stats()
    return span, np.array(lat), [_digest(r) for r in out], st


def run_jobs_stream(reqs, G, workers, prebuilt=False):
    """The same stream through the job queue (gpi_jobs_*): each shot, prepared on the host as
    for the plain path (coefficients included), is submitted at its arrival; worker threads,
    each with its own stream, run gpi_forward on the shots in arrival order as HBM allows."""
    import numpy as np
    from geophyinv.serve import JobServer

    if isinstance(prebuilt, list):   # experiments made by the caller
        pas, prebuilt = list(prebuilt), True
    else:
        pas = [_experiment(G, rq, coefficients="host") for rq in reqs] if prebuilt else None
    with JobServer(workers=workers) as js:
        t0 = time.perf_counter()
        tickets = []
        for k, rq in enumerate(reqs):
            t_arr = t0 + rq["arrival"]
            now = time.perf_counter()
            if now < t_arr:
                time.sleep(t_arr - now)
            pa = pas[k] if prebuilt else _experiment(G, rq, coefficients="host")
            if prebuilt:
                pas[k] = None
            tickets.append((js.submit(pa, 0), t_arr))
            del pa
        lat, out = [], []
        for k, (t, t_arr) in enumerate(tickets):
            out.append(js.wait(t)["records"])
            lat.append(time.perf_counter() - t_arr)
            if len(tickets) >= 256 and (k + 1) % 64 == 0:
                print(f"job queue: {k + 1}/{len(tickets)} shots", flush=True)
        span = time.perf_counter() - t0
        st = js.stats()
    return span, np.array(lat), [_digest(r) for r in out], st


def run_fwi_optimizer(args, G, _lib):
    """Optimizer work per FWI iteration at the 3-D gradient's model size (400^3 = 64 M FP64
    log-parameters), outside the misfit/gradient evaluations: L-BFGS (memory 10) with the
    two-loop recursion on the device and a strong-Wolfe line search on the host, against
    SciPy's L-BFGS-B (history and compact-form products on the host).  The objective is a
    separable quadratic evaluated with numpy (its time is excluded from both figures)."""
    import numpy as np
    from scipy.optimize import minimize

    from geophyinv.fwi import DeviceLBFGS, wolfe_search

    n = args.opt_n
    iters = 6
    rng = np.random.Generator(np.random.PCG64(1))
    h = rng.uniform(0.5, 50.0, n)
    b = rng.standard_normal(n)
    t_eval = [0.0]
    nfev = [0]

    def fg(x):
        t0 = time.perf_counter()
        hx = h * x
        f = 0.5 * float(np.dot(hx, x)) - float(np.dot(b, x))
        g = hx - b
        t_eval[0] += time.perf_counter() - t0
        nfev[0] += 1
        return f, g

    out = {}
    fmin = -0.5 * float(np.dot(b, b / h))     # the minimum; progress = (f - fmin) / (f0 - fmin)
    ctx = _lib.Context(0)
    x = np.ones(n)
    t_eval[0], nfev[0] = 0.0, 0
    t0 = time.perf_counter()
    f0, g = fg(x)
    opt = DeviceLBFGS(ctx, n, 10)
    f = f0
    for _ in range(iters):
        d = opt.direction(x, g)
        a0 = 1.0 if opt.pairs else min(1.0, 1.0 / float(np.linalg.norm(g)))
        a, f, g, _ = wolfe_search(fg, x, f, g, d, a0)
        x = x + a * d
    opt.close()
    total = time.perf_counter() - t0
    out["device_lbfgs"] = {"host_s_per_iteration": round((total - t_eval[0]) / iters, 3), "iterations": iters,
                           "evaluations": nfev[0], "gap_ratio": (f - fmin) / (f0 - fmin)}
    ctx.close()
    x = np.ones(n)
    t_eval[0], nfev[0] = 0.0, 0
    t0 = time.perf_counter()
    res = minimize(fg, x, jac=True, method="L-BFGS-B", options={"maxiter": iters, "gtol": 1e-25, "ftol": 0.0})
    total = time.perf_counter() - t0
    out["scipy_lbfgsb"] = {"host_s_per_iteration": round((total - t_eval[0]) / max(res.nit, 1), 3),
                           "iterations": int(res.nit), "evaluations": nfev[0],
                           "gap_ratio": (float(res.fun) - fmin) / (f0 - fmin)}
    return {"metric": "optimizer host seconds per FWI iteration (64 M parameters)",
            "value": out["device_lbfgs"]["host_s_per_iteration"], "unit": "s", "n_gpus": 1, "steps": iters,
            "warmup": 0, "ms_per_step": out["device_lbfgs"]["host_s_per_iteration"] * 1e3,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"L-BFGS memory 10 on a separable quadratic of {n} FP64 parameters, {iters} "
                                   "iterations; evaluation time excluded", "parameters": n},
            **out}


def run_source_estimation(args, G, _lib):
    """Source-signature estimation (update!(paw, pac)) for one supersource of the 2001^2
    elastic workload's shape: 4000 samples x 2403 traces (801 receivers x vx, vz, p),
    synthetic band-limited records convolved with a planted filter plus noise; LSMR with at
    most 100 iterations.  Device: gpi_conv_lsmr (batched hipFFT, FP64).  Host: SciPy's LSMR
    on conv.operator (NumPy FFTs), timed for a bounded number of iterations and reported
    per iteration beside the device's."""
    import numpy as np
    from scipy.sparse.linalg import lsmr

    from geophyinv import conv

    nt, ntr = 4000, 2403
    rng = np.random.Generator(np.random.PCG64(9))
    t = np.arange(61)
    pulse = np.exp(-0.5 * ((t - 30) / 6.0) ** 2) * np.cos(0.3 * (t - 30))
    g = np.asfortranarray(np.stack([np.convolve(rng.standard_normal(nt), pulse)[:nt] for _ in range(ntr)], axis=1))
    sl = (nt - 1) - (nt - 1) // 2
    w = np.zeros(nt)
    w[sl + 12], w[sl + 13], w[sl + 20] = 1.3, -0.5, 0.2
    pa = conv.Pconv(dsize=(nt, ntr), ssize=(nt,), gsize=(nt, ntr), g=g, s=w)
    d = np.asfortranarray(pa.mod_D() + 1e-3 * rng.standard_normal((nt, ntr)))
    ctx = _lib.Context(0)
    conv.estimate_filter_device(ctx, g, d, maxiter=2)        # FFT plans and kernels built once
    times, res = [], None
    for _ in range(max(args.steps, 1)):
        t0 = time.perf_counter()
        res = conv.estimate_filter_device(ctx, g, d, maxiter=100)
        times.append(time.perf_counter() - t0)
    ctx.close()
    sd, gs, it, stop = res
    dev_s = min(times)
    out = {"device": {"seconds": round(dev_s, 4), "iterations": it, "istop": stop,
                      "ms_per_iteration": round(dev_s / max(it, 1) * 1e3, 3),
                      "relative_residual": float(np.linalg.norm(gs - d) / np.linalg.norm(d))}}
    if not args.no_cpu:
        host_iters = 3
        pa2 = conv.Pconv(dsize=(nt, ntr), ssize=(nt,), gsize=(nt, ntr), g=g)
        t0 = time.perf_counter()
        ref = lsmr(conv.operator(pa2, "s"), d.ravel(order="F"), atol=conv.LSMR_TOL, btol=conv.LSMR_TOL,
                   conlim=conv.LSMR_CONLIM, maxiter=host_iters)
        host_s = time.perf_counter() - t0
        out["host_scipy"] = {"iterations_timed": int(ref[2]), "seconds": round(host_s, 3),
                             "ms_per_iteration": round(host_s / max(int(ref[2]), 1) * 1e3, 1),
                             "threads": os.cpu_count()}
    return {"metric": "source-signature estimation seconds per supersource", "value": round(dev_s, 4), "unit": "s",
            "n_gpus": 1, "steps": len(times), "warmup": 1, "ms_per_step": round(dev_s * 1e3, 2),
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "LSMR (at most 100 iterations) for one filter shared by 2403 traces of 4000 "
                                   "samples (the 2001^2 elastic workload's records), FFT length 8000",
                       "samples": nt, "traces": ntr}, **out}


def run_serve_jobs(args, G, _lib):
    """The serving-stream workload at its own sizes (512 2-D acoustic shots, sides 256-2048,
    1000-6000 steps, Poisson arrivals at 200/s) through the job queue against the plain
    path.  Latencies are arrival to records in the caller's arrays; the plain path's
    shots wait for the previous ones (one call after another)."""
    from geophyinv import gallery
    import numpy as np

    nshots = args.serve_shots or 512
    side = args.serve_side or [256, 2048]
    steps = args.serve_steps or [1000, 6000]
    t = time.perf_counter()
    reqs = gallery.serving_stream(nshots=nshots, side=tuple(side), steps=tuple(steps), rate=args.serve_rate)
    gen_s = time.perf_counter() - t
    cells = float(sum((r["side"] + 2 * r["kw"]["npml"]) ** 2 * r["nt"] for r in reqs))
    print(f"serve_jobs: {nshots} shots, {cells / 1e12:.2f} T cell-steps", flush=True)
    ctx = _lib.Context(0)
    # with host preparation before the stream, both paths take the same prepared experiments
    prep = False if args.serve_host_prep else [_experiment(G, rq, context=ctx, coefficients="host") for rq in reqs]
    p_span, p_lat, p_sums = run_plain(reqs, ctx, G, prep)
    print(f"serve_jobs: plain path {p_span:.1f} s", flush=True)
    j_span, j_lat, j_sums, st = run_jobs_stream(reqs, G, args.workers, prep)
    del prep
    ctx.close()
    print(f"serve_jobs: job queue {j_span:.1f} s", flush=True)
    same = sum(1 for a, b in zip(p_sums, j_sums) if a == b)
    pct = lambda a, q: round(float(np.percentile(a, q)) * 1e3, 1)
    return {
        "metric": "shots/s (2-D acoustic serving stream through the job queue)",
        "value": round(nshots / j_span, 3), "unit": "shots/s", "n_gpus": 1, "steps": nshots, "warmup": 0,
        "ms_per_step": round(j_span / nshots * 1e3, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{nshots} 2-D acoustic shots (order 2), Poisson arrivals at {args.serve_rate:g}/s, "
                               f"side U{{{side[0]}..{side[1]}}} at 5 m + 20-cell CPML, nt U{{{steps[0]}..{steps[1]}}}, "
                               "10-30 Hz Ricker on p, 64-512 surface receivers (p, vz), layered medium + Gaussian "
                               "anomaly; seed 3003" + ("; host preparation of each shot at its arrival"
                                                       if args.serve_host_prep else
                                                       "; host preparation of each shot before the stream starts"),
                   "workers": args.workers},
        "arrivals_span_s": round(reqs[-1]["arrival"], 3),
        "gpt_updates": cells,
        "plain": {"shots_per_s": round(nshots / p_span, 3), "span_s": round(p_span, 2),
                  "p50_ms": pct(p_lat, 50), "p99_ms": pct(p_lat, 99), "gpt_per_s": round(cells / p_span / 1e9, 2)},
        "jobs": {"shots_per_s": round(nshots / j_span, 3), "span_s": round(j_span, 2),
                 "p50_ms": pct(j_lat, 50), "p99_ms": pct(j_lat, 99), "gpt_per_s": round(cells / j_span / 1e9, 2),
                 "queue_p50_ms": round(st["p50_ms"], 1), "queue_p99_ms": round(st["p99_ms"], 1)},
        "speedup_shots_per_s": round(p_span / j_span, 3),
        "records_identical": f"{same}/{nshots}",
        "generation_s": round(gen_s, 1),
    }


def serving_section(nshots, opts, G, _lib, side=(256, 2048), steps=(1000, 6000), compare_direct=False,
                    rate=200.0, prebuilt=False):
    from geophyinv import gallery
    import numpy as np

    if side[1] > 512:
        raise SystemExit("shots above 512 cells per side are kept off the batched server (a GPU fault on such "
                         "shots has no cause shown); --workload serve_jobs runs the full-size stream")
    t = time.perf_counter()
    reqs = gallery.serving_stream(nshots=nshots, side=tuple(side), steps=tuple(steps), rate=rate)
    gen_s = time.perf_counter() - t
    cells = float(sum((r["side"] + 2 * r["kw"]["npml"]) ** 2 * r["nt"] for r in reqs))
    ctx = _lib.Context(0)
    p_span, p_lat, p_sums = run_plain(reqs, ctx, G, prebuilt)
    ctx.close()
    s_span, s_lat, s_sums, st = run_served(reqs, G, opts, prebuilt)
    same = sum(1 for a, b in zip(p_sums, s_sums) if a == b)
    pct = lambda a, q: round(float(np.percentile(a, q)) * 1e3, 1)
    direct = None
    if compare_direct:   # the same stream with every pass launched directly (no graphs)
        d_span, d_lat, d_sums, d_st = run_served(reqs, G, dict(opts, graphs=False), prebuilt)
        direct = {"shots_per_s": round(nshots / d_span, 3), "span_s": round(d_span, 2),
                  "p50_ms": pct(d_lat, 50), "p99_ms": pct(d_lat, 99),
                  "device_busy_s": round(d_st["busy_ms"] * 1e-3, 2),
                  "records_identical": f"{sum(1 for a, b in zip(p_sums, d_sums) if a == b)}/{nshots}"}
    return {
        "workload": f"{nshots} 2-D acoustic shots (order 2), Poisson arrivals at {rate:g}/s, side "
                    f"U{{{side[0]}..{side[1]}}} at 5 m + 20-cell CPML, nt U{{{steps[0]}..{steps[1]}}}, 10-30 Hz "
                    "Ricker on p, 64-512 surface receivers (p, vz), layered medium + Gaussian anomaly; seed 3003"
                    + ("; host preparation of each shot before the stream starts" if prebuilt else ""),
        "arrivals_span_s": round(reqs[-1]["arrival"], 3),
        "shots": nshots,
        "gpt_updates": cells,
        "plain": {"shots_per_s": round(nshots / p_span, 3), "span_s": round(p_span, 2),
                  "p50_ms": pct(p_lat, 50), "p99_ms": pct(p_lat, 99),
                  "gpt_per_s": round(cells / p_span / 1e9, 2)},
        "served": {"shots_per_s": round(nshots / s_span, 3), "span_s": round(s_span, 2),
                   "p50_ms": pct(s_lat, 50), "p99_ms": pct(s_lat, 99),
                   "gpt_per_s": round(cells / s_span / 1e9, 2), "device_busy_s": round(st["busy_ms"] * 1e-3, 2),
                   "mean_batch": round(st["mean_batch"], 2), "tile": [st["tile_ty"], st["tile_rx"]],
                   "graph_chunks": st["graph_chunks"], "graph_captures": st["graph_captures"],
                   "batched_steps": st["steps"], "options": opts,
                   "host_ms": {k[5:-3]: round(st[k], 1) for k in st if k.startswith("host_")}},
        "served_direct_launch": direct,
        "speedup_shots_per_s": round(p_span / s_span, 3),
        "records_identical": f"{same}/{nshots}",
        "generation_s": round(gen_s, 1),
    }


def run_gradient3d(args, G, _lib, gallery):
    """The 3-D acoustic FWI gradient: 400^3 at 10 m (+20-cell CPML = 440^3), vp 1500->4000
    m/s with three +-8% spheres, observed vz on a 201 x 201 surface patch from the true
    model, the gradient w.r.t. log(invK) at the starting model (true vp smoothed by a
    Gaussian of 5 cells), checkpoints every --checkpoint steps, FP64 gradient accumulation.
    One bench step is one whole gradient: the forward run keeping checkpoints, the host
    residual and adjoint sources, and the adjoint run that recomputes each segment."""
    import numpy as np

    nt = args.nt or 3000
    t0 = time.perf_counter()
    true, start, ageom, srcwav, tgrid, kw = gallery.acoustic_fwi_3d(n=args.grid_n, nt=nt,
                                                                    nrec_side=min(201, args.grid_n // 2 + 1))
    ctx = _lib.Context(0)
    pt = G.SeisForwExpt(G.FdtdAcoustic(), medium=true, tgrid=tgrid, ageom=ageom, srcwav=srcwav, context=ctx, **kw)
    pt.update()
    dobs = pt.data
    pt.free()
    pa = G.SeisForwExpt(G.FdtdAcoustic("forward_save", 2), medium=start, tgrid=tgrid, ageom=ageom, srcwav=srcwav,
                        context=ctx, checkpoint=args.checkpoint, grad_rho=False, **kw)
    setup_s = time.perf_counter() - t0
    mp = ["invK"]
    loss = G.L2DistLoss()
    m = pa.model_vector(mp)
    g = np.zeros_like(m)
    for _ in range(args.warmup):
        G.gradient(g, m, loss, dobs, pa, mp)
    wall, adj_ms = [], []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        J = G.gradient(g, m, loss, dobs, pa, mp)
        wall.append(time.perf_counter() - t1)
        adj_ms.append(pa.timing["total_ms"])
    # device time of the forward run with checkpoints (same model vector)
    G.lossvalue(m, loss, dobs, pa, mp)
    fwd_ms = pa.timing["total_ms"]
    J = G.lossvalue(m, loss, dobs, pa, mp)
    shape = pa.ext_shape
    cells = int(np.prod(shape))
    K = (nt + args.checkpoint - 1) // args.checkpoint
    slab = 4 * (-(-(shape[0] + 1) // 32) * 32) * (shape[1] + 1) * (shape[2] + 1)
    mean_wall = float(np.mean(wall))
    line = {
        "metric": "Gpt-updates/s (3-D acoustic FWI gradient, checkpointed, FP32 wavefields, FP64 reductions)",
        "value": round(3 * cells * nt / mean_wall / 1e9, 3),
        "unit": "Gpt-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(mean_wall * 1e3, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "3-D acoustic FWI gradient w.r.t. log(invK): 400^3 at 10 m, 20-cell CPML (440^3), "
                               "true vp 1500->4000 m/s gradient + three +-8% spheres, start = true smoothed by a "
                               "5-cell Gaussian, 12 Hz Ricker on p, 201x201 surface receivers (vz), "
                               f"checkpoints every {args.checkpoint} steps",
                   "time_steps": nt, "grid": list(shape), "order": pa.order, "compat": pa.compat},
        "propagations_per_gradient": 3,
        "gradient_s": round(mean_wall, 2),
        "forward_device_ms": round(fwd_ms, 1),
        "adjoint_device_ms": round(float(np.mean(adj_ms)), 1),
        "misfit": J,
        "checkpoint_wavefields_gb": round((K - 1) * 4 * slab / 1e9, 1),
        "segment_gb": round((args.checkpoint + 1) * slab / 1e9, 1),
        "setup_s": round(setup_s, 1),
    }
    if args.fd_cells:
        line["fd_check"] = fd_check_3d(args, pa, m, g, J, dobs, loss, mp, start, G)
    ctx.close()
    return line


def run_gradient3d_elastic(args, G, _lib, gallery):
    """The 3-D elastic FWI gradient: a (n+40)^3 three-layer medium at 5 m with a dipping
    interface, an explosive 25 Hz source, vx/vy/vz on a 64 x 64 surface patch from the true
    model, the gradient w.r.t. log(invlambda, invmu, rho) at the start model (interface one
    cell shallower), checkpoints every --checkpoint steps, FP64 accumulation.  One bench
    step is one whole gradient (forward run with checkpoints, residuals and adjoint sources
    on the host, adjoint run with recomputation)."""
    import numpy as np

    nt = args.nt or 2000
    n = args.grid_n if args.grid_n != 400 else 256
    t0 = time.perf_counter()
    true, start, ageom, srcwav, tgrid, kw = gallery.elastic_fwi_3d(n=n, nt=nt)
    ctx = _lib.Context(0)
    pt = G.SeisForwExpt(G.FdtdElastic(), medium=true, tgrid=tgrid, ageom=ageom, srcwav=srcwav, context=ctx, **kw)
    pt.update()
    dobs = pt.data
    pt.free()
    pa = G.SeisForwExpt(G.FdtdElastic("forward_save", 2), medium=start, tgrid=tgrid, ageom=ageom, srcwav=srcwav,
                        context=ctx, checkpoint=args.checkpoint, **kw)
    setup_s = time.perf_counter() - t0
    mp = ["invlambda", "invmu", "rho"]
    loss = G.L2DistLoss()
    m = pa.model_vector(mp)
    g = np.zeros_like(m)
    for _ in range(args.warmup):
        G.gradient(g, m, loss, dobs, pa, mp)
    wall, adj_ms = [], []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        G.gradient(g, m, loss, dobs, pa, mp)
        wall.append(time.perf_counter() - t1)
        adj_ms.append(pa.timing["total_ms"])
    J = G.lossvalue(m, loss, dobs, pa, mp)
    fwd_ms = pa.timing["total_ms"]
    shape = pa.ext_shape
    cells = int(np.prod(shape))
    mean_wall = float(np.mean(wall))
    K = (nt + args.checkpoint - 1) // args.checkpoint
    slab = 4 * (-(-(shape[0] + 1) // 32) * 32) * (shape[1] + 1) * (shape[2] + 1)
    line = {
        "metric": "Gpt-updates/s (3-D elastic FWI gradient, checkpointed, FP32 wavefields, FP64 reductions)",
        "value": round(3 * cells * nt / mean_wall / 1e9, 3),
        "unit": "Gpt-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(mean_wall * 1e3, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"3-D elastic FWI gradient w.r.t. log(invlambda, invmu, rho): {n}^3 at 5 m, "
                               f"20-cell CPML ({shape[0]}^3), three layers vp 1800/2600/3500 m/s with a seeded "
                               "dipping interface, start = interface one cell shallower, 25 Hz explosive source, "
                               f"64x64 surface receivers (vx, vy, vz), checkpoints every {args.checkpoint} steps",
                   "time_steps": nt, "grid": list(shape), "order": pa.order, "compat": pa.compat},
        "propagations_per_gradient": 3,
        "gradient_s": round(mean_wall, 2),
        "forward_device_ms": round(fwd_ms, 1),
        "adjoint_device_ms": round(float(np.mean(adj_ms)), 1),
        "misfit": J,
        "gradient_norms": {k: float(np.linalg.norm(x)) for k, x in zip(mp, np.split(g.astype(np.float64), 3))},
        "checkpoint_gb": round((K - 1) * state_floats_3d_elastic(shape, pa.npml) * 4 / 1e9, 1),
        "segment_gb": round((args.checkpoint + 1) * 9 * slab / 1e9, 1),
        "setup_s": round(setup_s, 1),
    }
    if args.fd_cells:
        line["fd_check"] = fd_check_3d(args, pa, m, g, J, dobs, loss, mp, start, G, blocks=(0, 1, 2))
    ctx.close()
    return line


def run_serve3d(args, G, _lib, gallery):
    """3-D elastic shots, a fraction of them gradient requests, through the job queue
    (worker streams, HBM admission) against the plain path (one call after another).
    Every shot is a gallery.elastic_fwi_3d problem of seeded size and length: forward
    shots model the true medium, gradient shots the start medium against the true one's
    records (made before the timed region).  Both paths start from prepared experiments;
    coefficient derivation and upload are inside the timed span of both."""
    import numpy as np
    from geophyinv.serve import JobServer

    rng = np.random.Generator(np.random.PCG64(5005))
    ns = args.serve_shots or 16
    side = args.serve_side or [96, 160]
    steps = args.serve_steps or [600, 1200]
    ctx = _lib.Context(0)
    shots = []
    t0 = time.perf_counter()
    for k in range(ns):
        n = int(rng.integers(side[0], side[1] + 1))
        nt = int(rng.integers(steps[0], steps[1] + 1))
        grad = bool(rng.random() < args.grad_fraction)
        true, start, ageom, srcwav, tgrid, kw = gallery.elastic_fwi_3d(n=n, nt=nt, seed=5005 + k)
        if grad:
            pt = G.SeisForwExpt(G.FdtdElastic(), medium=true, tgrid=tgrid, ageom=ageom, srcwav=srcwav, context=ctx,
                                **kw)
            pt.update()
            dobs = [d.copy() for d in pt.data]
            pt.free()
            mk = lambda start=start, tgrid=tgrid, ageom=ageom, srcwav=srcwav, kw=kw: G.SeisForwExpt(
                G.FdtdElastic("forward_save", 2), medium=start, tgrid=tgrid, ageom=ageom, srcwav=srcwav,
                context=ctx, checkpoint=args.checkpoint, **kw)
        else:
            dobs = None
            mk = lambda true=true, tgrid=tgrid, ageom=ageom, srcwav=srcwav, kw=kw: G.SeisForwExpt(
                G.FdtdElastic(), medium=true, tgrid=tgrid, ageom=ageom, srcwav=srcwav, context=ctx, **kw)
        shots.append(dict(n=n + 2 * kw["npml"], nt=nt, grad=grad, dobs=dobs, mk=mk))
        print(f"serve3d: shot {k + 1}/{ns} made ({n}^3, {nt} steps{', gradient' if grad else ''})", flush=True)
    gen_s = time.perf_counter() - t0
    loss = G.L2DistLoss()
    # plain path
    pas = [sh["mk"]() for sh in shots]
    lat_p, out_p = [], []
    t0 = time.perf_counter()
    for sh, pa in zip(shots, pas):
        if sh["grad"]:
            m = pa.model_vector()
            G.gradient(np.zeros_like(m), m, loss, sh["dobs"], pa)
            out_p.append({k: v.copy() for k, v in pa.gradients.items()})
        else:
            out_p.append(pa.run_shot(0)[0])
        pa.free()
        lat_p.append(time.perf_counter() - t0)
        print(f"serve3d plain path: {len(lat_p)}/{ns} shots at {lat_p[-1]:.1f} s", flush=True)
    span_p = time.perf_counter() - t0
    # job queue: every shot submitted at t = 0 (from the same model state as the plain run)
    pas = [sh["mk"]() for sh in shots]
    for sh, pa in zip(shots, pas):
        if sh["grad"]:
            pa.update_model(pa.model_vector())   # the model round trip the plain gradient made
    with JobServer(workers=args.workers) as js:
        t0 = time.perf_counter()
        tickets = [js.submit(pa, 0, dobs=sh["dobs"][0] if sh["grad"] else None) for sh, pa in zip(shots, pas)]
        out_s = []
        for t in tickets:
            out_s.append(js.wait(t))
            print(f"serve3d job queue: {len(out_s)}/{ns} shots at {time.perf_counter() - t0:.1f} s", flush=True)
        span_s = time.perf_counter() - t0
        st = js.stats()
    same = 0
    for sh, a, b in zip(shots, out_p, out_s):
        if sh["grad"]:
            same += all(np.array_equal(a[k], b["gradients"][k]) for k in b["gradients"])
        else:
            same += all(np.array_equal(a[f], b["records"][f]) for f in a)
    cells = float(sum(sh["n"] ** 3 * sh["nt"] * (3 if sh["grad"] else 1) for sh in shots))
    pct = lambda a, q: round(float(np.percentile(np.asarray(a) * 1e3, q)), 1)
    ctx.close()
    return {
        "metric": "shots/s (3-D elastic job queue, forward and gradient requests)",
        "value": round(ns / span_s, 3), "unit": "shots/s", "n_gpus": 1, "steps": ns, "warmup": 0,
        "ms_per_step": round(span_s / ns * 1e3, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{ns} 3-D elastic shots of gallery.elastic_fwi_3d, side U{{{side[0]}..{side[1]}}} "
                               f"+ 40 CPML cells, nt U{{{steps[0]}..{steps[1]}}}, {args.grad_fraction:.0%} "
                               f"gradient requests (checkpoints every {args.checkpoint}), all queued at t = 0",
                   "workers": args.workers, "compat": 0},
        "gradient_shots": int(sum(sh["grad"] for sh in shots)),
        "plain": {"shots_per_s": round(ns / span_p, 3), "span_s": round(span_p, 2), "p50_ms": pct(lat_p, 50),
                  "p99_ms": pct(lat_p, 99)},
        "queued": {"shots_per_s": round(ns / span_s, 3), "span_s": round(span_s, 2), "p50_ms": round(st["p50_ms"], 1),
                   "p99_ms": round(st["p99_ms"], 1), "forward_p50_ms": round(st["forward_p50_ms"], 1),
                   "gradient_p50_ms": round(st["gradient_p50_ms"], 1)},
        "speedup_shots_per_s": round(span_p / span_s, 3),
        "results_identical": f"{same}/{ns}",
        "gpt_updates_per_s": round(cells / span_s / 1e9, 2),
        "generation_s": round(gen_s, 1),
    }


def state_floats_3d_elastic(shape, npml):
    """Floats of one 3-D elastic state: nine wavefield slabs and the CPML memory of the
    twelve derivative fields (two strips of npml per derivative along its own axis)."""
    nz, ny, nx = shape
    slab = (-(-(nz + 1) // 32) * 32) * (ny + 1) * (nx + 1)
    return 9 * slab + 12 * 2 * npml * max(nz * ny, ny * nx, nz * nx)


def fd_check_3d(args, pa, m, g, J, dobs, loss, mp, start, G, blocks=(0,)):
    """Finite differences of the misfit along Gaussian bumps (radius 3 cells) of log(invK)
    centred on seeded interior cells, against the gradient's directional derivative.
    Fourth-order central differences from plain forward runs at m +- eps*b, m +- 2 eps*b;
    eps is chosen per bump so that the misfit moves by about fd_target of itself (the
    misfit of FP32 wavefields carries a relative noise of ~1e-7, so weakly sensitive bumps
    need a larger step), within [1e-3, 5e-2].

    Bumps are chosen above that noise floor: of 4 x fd_cells seeded candidate centres, the
    fd_cells whose directional derivative |<g, b>| is largest are checked, and each
    bump's predicted relative misfit change at its step (|<g, b>| eps / J) is reported."""
    import numpy as np

    from geophyinv import database

    rng = np.random.Generator(np.random.PCG64(4004))
    n = start.vp.m.shape
    lo = min(60, n[0] // 3)
    w = 9
    nb = b_len = int(np.prod(n))
    zz, yy, xx = np.meshgrid(*[np.arange(-w, w + 1)] * 3, indexing="ij")
    local = np.exp(-(zz ** 2 + yy ** 2 + xx ** 2) / (2.0 * 3.0 ** 2))

    def bump_vector(c, blk):
        b = np.zeros(n, np.float64, order="F")
        b[tuple(slice(int(ci) - w, int(ci) + w + 1) for ci in c)] = local
        bv = np.zeros(m.size)
        bv[blk * nb:blk * nb + b_len] = b.ravel(order="F")
        return bv

    # depth from 20 cells (clear of the source five cells down and its bump window) to 150,
    # where the surface acquisition's waves carry the misfit; y and x clear of the CPML
    zlo, zhi = min(20, n[0] // 4), max(min(150, n[0] - lo), min(20, n[0] // 4) + 1)
    cands = np.stack([rng.integers(zlo, zhi, 4 * args.fd_cells)] +
                     [rng.integers(lo, s - lo, 4 * args.fd_cells) for s in n[1:]], axis=1)
    sens = []
    for c in cands:
        for k in blocks:
            sens.append((abs(float(np.dot(g, bump_vector(c, k)))), tuple(int(x) for x in c), k))
    sens.sort(key=lambda t: -t[0])
    chosen = sens[:args.fd_cells * len(blocks)]
    mode = pa.attrib_mod.mode
    pa.attrib_mod.mode = "forward"
    out = []
    for _, c, blk in chosen:
        bv = bump_vector(c, blk)
        ad = float(np.dot(g, bv))
        eps = min(max(args.fd_target * J / abs(ad), 1e-3), 5e-2) if ad else 1e-2
        Js = {}
        for k in (1, -1, 2, -2):
            pa.update_model(m + k * eps * bv, mp)
            pa.update()
            Js[k] = database.lossvalue(loss, dobs, pa.data)
        fd = (8.0 * (Js[1] - Js[-1]) - (Js[2] - Js[-2])) / (12.0 * eps)
        out.append({"cell": [int(x) for x in c], "param": mp[blk], "eps": eps, "fd": fd, "adjoint": ad,
                    "misfit_change": abs(ad) * eps / J, "rel_err": abs(ad / fd - 1) if fd else None})
        print(json.dumps({"fd": out[-1]}), flush=True)
    pa.update_model(m, mp)
    pa.attrib_mod.mode = mode
    errs = [o["rel_err"] for o in out if o["rel_err"] is not None]
    return {"fd_target": args.fd_target, "stencil": "4th-order central", "bump_radius_cells": 3,
            "selection": f"the {len(chosen)} of {len(sens)} seeded candidate bumps with the largest |<g, b>|",
            "cells": out, "max_rel_err": max(errs) if errs else None,
            "median_rel_err": float(np.median(errs)) if errs else None,
            "within_1e-3": f"{sum(e <= 1e-3 for e in errs)}/{len(errs)}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nt", type=int, default=0, help="time steps (default: the workload's own)")
    ap.add_argument("--workload", choices=["elastic2d", "acoustic3d", "elastic3d", "gradient3d", "gradient3d_elastic",
                                           "serve", "serve3d", "serve_jobs", "fwi_optimizer", "source_estimation"],
                    default="elastic2d")
    ap.add_argument("--opt-n", type=int, default=64_000_000, help="fwi_optimizer: parameters")
    ap.add_argument("--grad-fraction", type=float, default=0.25, help="serve3d: share of gradient requests")
    ap.add_argument("--workers", type=int, default=2, help="serve3d, serve_jobs: jobs in flight")
    ap.add_argument("--checkpoint", type=int, default=100, help="gradient3d: checkpoint interval")
    ap.add_argument("--fd-cells", type=int, default=0, help="gradient3d: finite-difference checks")
    ap.add_argument("--fd-target", type=float, default=1e-3, help="gradient3d: relative misfit change per FD step")
    ap.add_argument("--grid-n", type=int, default=400, help="gradient3d: cells per side of the medium")
    ap.add_argument("--grid-2d", type=int, default=2001,
                    help="elastic2d: cells per side of the medium (2001; 4001 puts each pass's working set in HBM)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    # Serving stream (the default 2-D elastic run appends it; --no-serve leaves it out).  The
    # defaults are 128 shots of at most 512 cells per side, all queued at once, prepared on
    # the host before the stream: a 48-shot sample of the full 256..2048-cell stream faulted
    # the GPU once and the cause has not been shown, so sizes beyond 512 need the flags.
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-stream", action="store_true",
                    help="default line: leave out the 512-shot serving stream through the job queue")
    ap.add_argument("--serve-shots", type=int, default=None, help="serving: shots (128; serve3d 16)")
    ap.add_argument("--max-shots", type=int, default=64)
    ap.add_argument("--chunk-steps", type=int, default=32)
    ap.add_argument("--serve-side", type=int, nargs=2, default=None, help="serving: grid side range (128 512)")
    ap.add_argument("--serve-steps", type=int, nargs=2, default=None, help="serving: time step range (500 1500)")
    ap.add_argument("--serve-direct", action="store_true", help="serving: also run without graph replay")
    ap.add_argument("--serve-rate", type=float, default=None,
                    help="serving: Poisson arrival rate (1/s; serve 1e6, serve_jobs 200)")
    ap.add_argument("--serve-host-prep", action="store_true",
                    help="serving: prepare each shot on the host at its arrival, inside the timed span")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the single-GPU path only "
              "(shots would shard one per GPU, one process per GPU)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import geophyinv as G
    from geophyinv import _lib, gallery

    if args.workload == "serve":
        opts = dict(max_shots=args.max_shots, chunk_steps=args.chunk_steps, autotune=True)
        sec = serving_section(args.serve_shots or 128, opts, G, _lib, args.serve_side or [128, 512],
                              args.serve_steps or [500, 1500], args.serve_direct, args.serve_rate or 1e6,
                              not args.serve_host_prep)
        print(json.dumps({
            "metric": "shots/s (2-D acoustic serving stream, continuous batching)",
            "value": sec["served"]["shots_per_s"], "unit": "shots/s", "n_gpus": 1, "steps": sec["shots"],
            "warmup": 0, "ms_per_step": round(1e3 / max(sec["served"]["shots_per_s"], 1e-9), 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": sec["workload"], "max_shots": args.max_shots, "chunk_steps": args.chunk_steps},
            "serving": sec}), flush=True)
        return
    if args.workload == "serve3d":
        print(json.dumps(run_serve3d(args, G, _lib, gallery)), flush=True)
        return
    if args.workload == "fwi_optimizer":
        print(json.dumps(run_fwi_optimizer(args, G, _lib)), flush=True)
        return
    if args.workload == "source_estimation":
        print(json.dumps(run_source_estimation(args, G, _lib)), flush=True)
        return
    if args.workload == "serve_jobs":
        if args.serve_rate is None:
            args.serve_rate = 200.0
        print(json.dumps(run_serve_jobs(args, G, _lib)), flush=True)
        return
    if args.workload == "gradient3d_elastic":
        print(json.dumps(run_gradient3d_elastic(args, G, _lib, gallery)), flush=True)
        return
    if args.workload == "gradient3d":
        print(json.dumps(run_gradient3d(args, G, _lib, gallery)), flush=True)
        return
    three = args.workload in ("acoustic3d", "elastic3d")
    el3 = args.workload == "elastic3d"
    if el3:
        med, _, ageom, srcwav, tgrid, kw = gallery.elastic_fwi_3d(n=args.grid_n if args.grid_n != 400 else 320,
                                                                  nt=args.nt or 1000)
        attrib = G.FdtdElastic()
    elif three:
        med, ageom, srcwav, tgrid, kw = gallery.acoustic_gradient_3d(nt=args.nt or 3000)
        attrib = G.FdtdAcoustic()
    else:
        med, ageom, srcwav, tgrid, kw = gallery.elastic_layers_2d(n=args.grid_2d, nt=args.nt or 4000)
        attrib = G.FdtdElastic()
    ctx = _lib.Context(0)
    ctx.set_pass_sampling(64)
    pa = G.SeisForwExpt(attrib, medium=med, tgrid=tgrid, ageom=ageom, srcwav=srcwav, context=ctx, **kw)
    shape = pa.ext_shape
    cells = int(np.prod(shape))
    nt = len(tgrid)
    # the profiled workloads: their measured traffic per launch is in profiles/
    profiled = pa.order == 2 and shape in ((2081, 2081), (440, 440, 440), (360, 360, 360))
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
    gpt = cells * nt * args.steps
    value = gpt / wall / 1e9
    # roofline bytes: the per-cell figures of the reference's coefficient layout; the passes'
    # own compulsory bytes beside them
    vel_b, str_b = reference_bytes(shape, pa.npml, 1 if (el3 or not three) else 0)
    if el3:
        vel_k, str_k = algorithmic_bytes_3d_elastic(shape, pa.npml)
    elif three:
        vel_k, str_k = algorithmic_bytes_3d(shape, pa.npml)
    else:
        vel_k, str_k = algorithmic_bytes(shape[0], shape[1], pa.npml, 1)
    # Each time step is two launches (velocity pass, stress pass).  Event brackets around a
    # sampled pass add the marker latency, so the measured per-step device time is split
    # between the two passes in the ratio of their bracket means: the per-launch figure is
    # the pass's share of the real loop (kernel plus its dispatch gap).
    s_br = float(np.mean(str_ms)) * 1e-3
    v_br = float(np.mean(vel_ms)) * 1e-3
    step_s = float(np.mean(dev_ms)) * 1e-3 / nt
    s_avg = step_s * s_br / (s_br + v_br)
    v_avg = step_s * v_br / (s_br + v_br)
    # the dominant kernel: the elastic stress passes; the 3-D acoustic velocity pass
    vdom = three and not el3
    dom_b, dom_t, dom_k = (vel_b, v_avg, vel_k) if vdom else (str_b, s_avg, str_k)
    other_b, other_t, other_br, other_k = (str_b, s_avg, s_br, str_k) if vdom else (vel_b, v_avg, v_br, vel_k)
    achieved = dom_b / dom_t / 1e9
    step_bytes = vel_b + str_b
    if el3:
        metric = "Gpt-updates/s (3-D elastic FDTD, FP32)"
        workload = (f"3-D elastic FdtdElastic, {shape[0] - 40}^3 at 5 m, 20-cell CPML ({shape[0]}^3), three layers "
                    "vp 1800/2600/3500 m/s with a seeded dipping interface, explosive 25 Hz Ricker, 64x64 surface "
                    "receivers (vx, vy, vz), one shot per step")
        kname, other_name = "str3d_elastic_staged<2, 4, true> (stress pass)", "velocity_pass"
        trf = measured_traffic("fdtd3d_elastic_360_traffic.md", "str3d_elastic_staged")
    elif three:
        metric = "Gpt-updates/s (3-D acoustic FDTD, FP32)"
        workload = ("3-D acoustic FdtdAcoustic, 400^3 at 10 m, 20-cell CPML (440^3), vp 1500->4000 m/s "
                    "gradient + three +-8% spheres, 12 Hz Ricker on p, 201x201 surface receivers (p), "
                    "one shot per step")
        kname, other_name = "vel3d_acoustic<2> (velocity pass)", "stress_pass"
        trf = measured_traffic("fdtd3d_acoustic_440_traffic.md", "vel3d_acoustic")
    else:
        metric = "Gpt-updates/s (2-D elastic FDTD, FP32)"
        workload = (f"2-D elastic P-SV FdtdElastic, {args.grid_2d}x{args.grid_2d} at 2.5 m, 40-cell CPML "
                    f"({shape[0]}x{shape[1]}), 8 layers + 3% smoothed noise, explosive 20 Hz Ricker, 801 receivers "
                    "(vx, vz), one shot per step")
        kname, other_name = "str2d_elastic<2> (stress pass)", "velocity_pass"
        trf = measured_traffic("fdtd2d_elastic_2081x2081_traffic.md", "str2d_elastic")
    line = {
        "metric": metric,
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
        "config": {"workload": workload, "time_steps": nt, "grid": list(shape), "order": pa.order,
                   "compat": pa.compat},
        "device_ms_per_step": round(float(np.mean(dev_ms)), 3),
        "device_gpt_per_s": round(cells * nt / (float(np.mean(dev_ms)) * 1e-3) / 1e9, 3),
        "step_algorithmic_GBps": round(step_bytes * nt / (float(np.mean(dev_ms)) * 1e-3) / 1e9, 1),
        "roofline": {"bound": "hbm", "kernel": kname,
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": trf if profiled else None,
                     "bytes_per_launch": int(dom_b), "avg_launch_ms": round(dom_t * 1e3, 5),
                     "bracket_ms": round((v_br if vdom else s_br) * 1e3, 5),
                     "kernel_bytes_per_launch": int(dom_k), "kernel_achieved": round(dom_k / dom_t / 1e9, 1),
                     other_name: {"bytes_per_launch": int(other_b), "avg_launch_ms": round(other_t * 1e3, 5),
                                  "bracket_ms": round(other_br * 1e3, 5),
                                  "achieved": round(other_b / other_t / 1e9, 1),
                                  "kernel_bytes_per_launch": int(other_k),
                                  "kernel_achieved": round(other_k / other_t / 1e9, 1)}},
        "records_finite": bool(finite),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(pa, args.cpu_seconds)
    ctx.close()
    if not args.no_serve and not three:
        opts = dict(max_shots=args.max_shots, chunk_steps=args.chunk_steps, autotune=True)
        line["serving"] = serving_section(args.serve_shots or 128, opts, G, _lib, args.serve_side or [128, 512],
                                          args.serve_steps or [500, 1500], args.serve_direct,
                                          args.serve_rate or 1e6, not args.serve_host_prep)
    if not args.no_serve and not three and not args.no_stream:
        # the serving-stream workload at its own sizes (512 shots of 256-2048 cells, 1000-6000
        # steps, 200/s), through the job queue against the plain path
        js = argparse.Namespace(**vars(args))
        js.serve_shots, js.serve_side, js.serve_steps, js.serve_rate, js.workers = 512, [256, 2048], [1000, 6000], \
            200.0, 4
        r = run_serve_jobs(js, G, _lib)
        line["serving_stream"] = {"workload": r["config"]["workload"], "workers": 4, "path": "job queue",
                                  "arrivals_span_s": r["arrivals_span_s"], "gpt_updates": r["gpt_updates"],
                                  "plain": r["plain"], "jobs": r["jobs"],
                                  "speedup_shots_per_s": r["speedup_shots_per_s"],
                                  "records_identical": r["records_identical"]}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
