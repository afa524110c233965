"""The CPU build of libgpi (tests/cpu_emu): the HIP sources compiled for the host against a
runtime with one worker thread per stream, events, stream waits, graph capture, and
guarded memory that becomes inaccessible when it is freed (hipFree, hipFreeAsync at its
stream position, or a pool block at release).

- the runtime catches an out-of-bounds store and a read after a stream-ordered free, and
  lets the same free pass when it is ordered after the read (so a run without a fault
  means something);
- shots through the CPU build equal the oracle bit for bit (the kernels keep the oracle's
  FP32 operation order on the host too), gpi_forward_static included;
- a small serving stream (arrivals interleaved with the batch, graph replay, pool blocks
  released while other shots run) gives records equal to the plain path's.
Each case runs in a child process, since a caught fault ends the process (status 86)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "cpu_emu")
LIB = os.path.join(EMU, "_build", "libgpi_emu.so")


@pytest.fixture(scope="module")
def built():
    jobs = str(min(8, os.cpu_count() or 1))
    subprocess.run(["make", "-s", "-j", jobs, "-C", EMU, "all", "selftest"], check=True)
    return LIB


def run(args, timeout=600):
    return subprocess.run(args, capture_output=True, text=True, timeout=timeout, cwd=ROOT)


@pytest.mark.parametrize("case,faults", [("oob", True), ("after-free", True), ("ordered", False)])
def test_runtime_catches_bad_accesses(built, case, faults):
    r = run([os.path.join(EMU, "_build", "selftest"), case], timeout=60)
    if faults:
        assert r.returncode == 86, (r.returncode, r.stderr[-2000:])
        assert ("past its end" if case == "oob" else "ALREADY FREED") in r.stderr
    else:
        assert r.returncode == 0 and "no fault" in r.stdout, r.stderr[-2000:]


SHOTS = r"""
import sys
sys.path[:0] = [{pkg!r}, {orc!r}, {tests!r}]
from geophyinv import _lib
_lib.load({lib!r})
import numpy as np
import oracle as O
from helpers import experiment, experiment3d, layered_acoustic, layered_elastic, medium3d, rel_l2
ctx = _lib.Context(0)
cases = [
    experiment(layered_acoustic(nz=40, nx=44), nt=160, sfields=("p", "vx"), ns=2, nr=7, npml=8, ctx=ctx),
    experiment(layered_elastic(nz=40, nx=44), nt=160, sfields=("tauxx",), rfields=("vx", "vz"), ns=1, nr=7,
               npml=8, ctx=ctx, stressfree_faces=["zmin"], pml_faces=["zmax", "xmin", "xmax"]),
    experiment(layered_acoustic(nz=36, nx=40), nt=120, order=4, ns=1, nr=5, npml=6, ctx=ctx),
    experiment3d(medium3d(1, n=(14, 12, 16)), nt=50, sfields=("tauxx",), rfields=("vx", "vy", "vz"), ctx=ctx),
]
for pa in cases:
    recs, _ = pa.run_shot(0)
    stat, _ = pa.run_shot(0, static=True)
    orc = O.from_experiment(pa)
    orc.run()
    for q, f in enumerate(pa.rfields):
        assert np.abs(recs[f]).max() > 0, f
        assert rel_l2(recs[f], orc.records(q)) == 0.0, (f, rel_l2(recs[f], orc.records(q)))
        assert np.array_equal(recs[f], stat[f]), f
print("shots equal the oracle")
"""


def test_shots_equal_the_oracle(built):
    code = SHOTS.format(pkg=os.path.join(ROOT, "geophyinv.jl_amd"), orc=os.path.join(ROOT, "oracle"),
                        tests=os.path.join(ROOT, "tests"), lib=built)
    r = run([sys.executable, "-c", code])
    assert r.returncode == 0 and "shots equal the oracle" in r.stdout, (r.returncode, r.stderr[-3000:])


ELASTIC = r"""
import sys
sys.path[:0] = [{pkg!r}, {orc!r}, {tests!r}]
from geophyinv import _lib
_lib.load({lib!r})
import numpy as np
import oracle as O
from helpers import experiment, layered_elastic, rel_l2
ctx = _lib.Context(0)
FIN = ["tauxx", "tauzz", "tauxz", "vx", "vz"]
cases = []
for order in (2, 4, 6, 8):
    # nz + npml padding not a multiple of 4 (partial four-row groups at both ends of a column)
    cases.append(experiment(layered_elastic(nz=43 + order, nx=30), nt=150, sfields=("tauxx",), rfields=("vx", "vz"),
                            ns=2, nr=7, order=order, npml=7, ctx=ctx, seed=40 + order))
    cases.append(experiment(layered_elastic(nz=41 + order, nx=28), nt=150, sfields=("vz", "vx", "tauzz"),
                            rfields=("vx", "vz"), ns=3, nr=6, order=order, npml=6, pml_faces=["zmax", "xmin", "xmax"],
                            stressfree_faces=["zmin"], ctx=ctx, seed=50 + order, src_depth=6.0))
# a column longer than 256 rows, sources near both ends
cases.append(experiment(layered_elastic(nz=291, nx=18), nt=90, sfields=("vz", "tauxx"), rfields=("vx", "vz"), ns=2,
                        nr=5, order=2, npml=5, ctx=ctx, seed=61))
for k, pa in enumerate(cases):
    orc = O.from_experiment(pa)
    orc.run()
    recs, fin = pa.run_shot(0, final_fields=FIN)
    for q, f in enumerate(pa.rfields):
        assert rel_l2(recs[f], orc.records(q)) == 0.0, (k, f, rel_l2(recs[f], orc.records(q)))
    for f in FIN:
        assert np.abs(fin[f]).max() > 0, (k, f)
        assert np.array_equal(fin[f], orc.field(f)), (k, f, rel_l2(fin[f], orc.field(f)))
print("elastic cases equal the oracle", len(cases))
"""


def test_elastic_orders_equal_the_oracle(built):
    """The 2-D elastic kernels against the oracle through the CPU build: orders 2-8, free
    surface with velocity and stress sources beside it, rigid faces, padded columns whose
    length is not a multiple of 4, a column longer than one 256-row workgroup.  Records and
    every final wavefield (ghost cells included) equal bit for bit."""
    code = ELASTIC.format(pkg=os.path.join(ROOT, "geophyinv.jl_amd"), orc=os.path.join(ROOT, "oracle"),
                       tests=os.path.join(ROOT, "tests"), lib=built)
    r = run([sys.executable, "-c", code], timeout=900)
    assert r.returncode == 0 and "elastic cases equal the oracle" in r.stdout, (r.returncode, r.stderr[-3000:])


JOBS = r"""
import gc, sys
sys.path[:0] = [{pkg!r}, {tests!r}]
from geophyinv import _lib
_lib.load({lib!r})
import numpy as np
from geophyinv.serve import JobServer
from helpers import experiment, layered_acoustic
ctx = _lib.Context(0)
mk = lambda k: experiment(layered_acoustic(nz=48 + 4 * k, nx=52, seed=k), nt=200, nr=9, npml=8, ctx=ctx, seed=k)
plain = [mk(k).run_shot(0)[0] for k in range(3)]
with JobServer(workers=2) as js:
    tickets = []
    for k in range(3):
        pa = mk(k)
        tickets.append(js.submit(pa, 0))
        del pa          # the caller lets go of the experiment right after submitting
        gc.collect()
        junk = [np.full(200000, np.nan, np.float32) for _ in range(8)]   # reuse of freed host memory
    outs = [js.wait(t)["records"] for t in tickets]
for a, b in zip(plain, outs):
    for f in a:
        assert np.array_equal(a[f], b[f]), f
print("jobs keep their inputs")
"""


def test_job_queue_keeps_its_inputs(built):
    """Regression: a job reads its inputs (CPML vectors among them) when a worker starts it,
    which may be after the caller has dropped the experiment; the host wrapper keeps every
    array it hands to the queue alive until the job is waited for."""
    code = JOBS.format(pkg=os.path.join(ROOT, "geophyinv.jl_amd"), tests=os.path.join(ROOT, "tests"), lib=built)
    r = run([sys.executable, "-c", code])
    assert r.returncode == 0 and "jobs keep their inputs" in r.stdout, (r.returncode, r.stderr[-3000:])


def test_serving_stream_equals_plain(built):
    r = run([sys.executable, os.path.join(EMU, "serving_sample.py"), "--lib", built, "--shots", "6", "--side", "64",
             "160", "--steps", "600", "900", "--plain", "--time-scale", "100", "--chunk-steps", "16"])
    assert r.returncode == 0 and "records equal: 6/6" in r.stdout, (r.returncode, r.stdout[-2000:], r.stderr[-3000:])
