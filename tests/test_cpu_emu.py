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
