"""A sample of the 2-D serving stream through a CPU build of libgpi (this directory's
Makefile): shots arrive as the seeded Poisson process of ``gallery.serving_stream``, go
through the library's server (``gpi_server_*``: admission stream, compute stream, batched
kernels, graph replay, pool) and, with ``--plain``, through ``gpi_forward`` one after
another; every served record must equal the plain one bit for bit.  The CPU build makes a
freed block inaccessible at the point the library may hand it to another shot, and puts
every allocation against a guard page, so an access out of bounds or after release ends
the run with status 86 and names the kernel, workgroup and allocation.

Records are compared through SHA-256 digests of each field's bytes: ``--digests`` writes
the digests of this run, ``--compare`` checks the served records against the plain digests
of an earlier run (the plain path of the GPU build gives the same bits as this build, so
a GPU run of ``--plain-only`` with the product library can supply them).

``--pkg`` runs another copy of the host package (with ``--lib`` built from its sources),
so an earlier version of the server can be put through the same sample.  Arrival times
are multiplied by ``--time-scale`` (the CPU build is some 600 times slower than the GPU,
so such a scale keeps arrivals interleaved with the batch as they are on the GPU).

Test infrastructure: run by hand or from tests/test_cpu_emu.py.
"""
from __future__ import annotations

import argparse
import hashlib
import importlib.util
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def digests(recs: dict) -> dict:
    return {f: hashlib.sha256(a.tobytes(order="F")).hexdigest() for f, a in sorted(recs.items())}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", default=os.path.join(HERE, "_build", "libgpi_emu.so"))
    ap.add_argument("--pkg", default=os.path.join(ROOT, "geophyinv.jl_amd"))
    ap.add_argument("--shots", type=int, default=48)
    ap.add_argument("--seed", type=int, default=3003)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--side", type=int, nargs=2, default=(256, 2048))
    ap.add_argument("--steps", type=int, nargs=2, default=(1000, 6000))
    ap.add_argument("--time-scale", type=float, default=600.0)
    ap.add_argument("--max-shots", type=int, default=64)
    ap.add_argument("--chunk-steps", type=int, default=32)
    ap.add_argument("--plain", action="store_true", help="also run gpi_forward and compare records")
    ap.add_argument("--plain-only", action="store_true", help="run gpi_forward only (no server)")
    ap.add_argument("--digests", default=None, help="write the record digests of this run here (JSON)")
    ap.add_argument("--compare", default=None, help="plain digests to check the served records against (JSON)")
    args = ap.parse_args()

    sys.path.insert(0, args.pkg)
    from geophyinv import _lib
    _lib.load(args.lib)
    import numpy as np

    import geophyinv as G

    # the stream generator of this checkout, bound to the package under test
    spec = importlib.util.spec_from_file_location("geophyinv._sample_stream",
                                                  os.path.join(ROOT, "geophyinv.jl_amd", "geophyinv", "gallery.py"))
    gal = importlib.util.module_from_spec(spec)
    gal.__package__ = "geophyinv"
    spec.loader.exec_module(gal)
    stream = gal.serving_stream(args.shots, seed=args.seed, rate=args.rate, side=tuple(args.side),
                                steps=tuple(args.steps))
    cells = sum((s["side"] + 40) ** 2 * s["nt"] for s in stream)
    print(f"sample: {len(stream)} shots, sides {min(s['side'] for s in stream)}-{max(s['side'] for s in stream)}, "
          f"steps {min(s['nt'] for s in stream)}-{max(s['nt'] for s in stream)}, {cells / 1e9:.1f} G cell-steps",
          flush=True)
    ctx = _lib.Context(0)
    pas = [G.SeisForwExpt(G.FdtdAcoustic(), medium=s["medium"], tgrid=s["tgrid"], ageom=s["ageom"],
                          srcwav=s["srcwav"], context=ctx, **s["kw"]) for s in stream]
    out = {"shots": len(stream), "seed": args.seed}
    rc = 0
    if not args.plain_only:
        from geophyinv.serve import Server
        srv = Server(0, 2, max_shots=args.max_shots, chunk_steps=args.chunk_steps)
        t0 = time.time()
        tickets, pending = [], set()
        for k, s in enumerate(stream):
            while time.time() - t0 < s["arrival"] * args.time_scale:
                time.sleep(0.01)
            tickets.append(srv.submit(pas[k], 0))
            pending.add(k)
        done = 0
        while pending:
            for k in sorted(pending):
                st = srv.poll(tickets[k])
                if st != 1:
                    pending.discard(k)
                    done += 1
                    print(f"served {done}/{len(stream)}: shot {k} side {stream[k]['side']} nt {stream[k]['nt']} "
                          f"status {st} at {time.time() - t0:.1f} s", flush=True)
                    if st != 0:
                        return 1
            time.sleep(0.05)
        out["served"] = [digests(srv.wait(t)) for t in tickets]
        srv.close()
        print(f"served all {len(stream)} shots in {time.time() - t0:.1f} s", flush=True)
    if args.plain or args.plain_only:
        out["plain"] = []
        for k, pa in enumerate(pas):
            recs, _ = pa.run_shot(0)
            out["plain"].append(digests(recs))
            print(f"plain {k + 1}/{len(pas)} done", flush=True)
    if args.digests:
        with open(args.digests, "w") as f:
            json.dump(out, f, indent=1)
    ref = None
    if args.compare:
        with open(args.compare) as f:
            ref = json.load(f)["plain"]
    elif "plain" in out:
        ref = out["plain"]
    if ref is not None and "served" in out:
        equal = sum(a == b for a, b in zip(out["served"], ref))
        print(f"records equal: {equal}/{len(stream)}", flush=True)
        rc = 0 if equal == len(stream) == len(ref) else 1
    return rc


if __name__ == "__main__":
    sys.exit(main())
