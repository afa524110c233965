"""Serving: many shots of different sizes on one GPU (``gpi_server_*`` of libgpi).

``Server.submit(pa, iss)`` hands supersource ``iss`` of a 2-D ``SeisForwExpt`` to the
library's scheduler: the unpadded medium (coefficients are derived on the device), the
CPML vectors, the spray/interpolation weights and the wavelets prepared by the
experiment, exactly the inputs of ``update!(pa)`` for that supersource.  Shots are
admitted by free HBM, stay resident and advance together (continuous batching); each
result equals ``pa.run_shot(iss)`` bit for bit.  No reference counterpart: the reference
runs supersources one after another (src/fdtd/propagate.jl:100-106).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .fields import FIELD_ID

F32 = np.float32


class Server:
    def __init__(self, physics: int = 0, order: int = 2, *, max_shots: int = 64, chunk_steps: int = 32,
                 autotune: bool = True, tile: int = 0, hbm_budget_bytes: int = 0, device: int = 0,
                 graphs: bool = True):
        self.lib = _lib.load()
        o = _lib.gpi_server_opts()
        o.physics, o.order, o.max_shots, o.chunk_steps = physics, order, max_shots, chunk_steps
        o.autotune, o.tile, o.hbm_budget_bytes = int(bool(autotune)), tile, hbm_budget_bytes
        o.direct_launch = 0 if graphs else 1
        h = C.c_void_p()
        rc = self.lib.gpi_server_create(device, C.byref(o), C.byref(h))
        if rc != _lib.GPI_OK:
            raise _lib.GpiError(rc, "gpi_server_create failed (no HIP device or bad options)")
        self.h = h
        self._out = {}

    def _check(self, rc: int):
        if rc != _lib.GPI_OK:
            raise _lib.GpiError(rc, self.lib.gpi_server_error(self.h).decode())

    def submit(self, pa, iss: int = 0) -> int:
        """Queue supersource ``iss`` of ``pa``; returns a ticket."""
        if pa.ndim != 2:
            raise ValueError("serving takes 2-D experiments")
        nt = len(pa.tgrid)
        keep = []
        src_arr, wav_arr, nsrc = pa._points(pa.spray[iss], pa.wavelets[iss], keep)
        recs, outs, names = [], [], []
        for f, (idx, w) in pa.interp[iss].items():
            idx = np.ascontiguousarray(idx)
            w = np.ascontiguousarray(w)
            out = np.zeros((nt, idx.shape[0]), F32, order="F")
            keep += [idx, w]
            p = _lib.gpi_points()
            p.field, p.npts, p.nnz = FIELD_ID[f], idx.shape[0], idx.shape[1]
            p.idx, p.w = _lib.iptr(idx), _lib.fptr(w)
            recs.append(p)
            outs.append(out)
            names.append(f)
        pmls = []
        for df, (a, b, k) in pa.pml.items():
            p = _lib.gpi_pml()
            p.field = FIELD_ID[df]
            p.len = a.size
            p.a, p.b, p.kI = _lib.fptr(a), _lib.fptr(b), _lib.fptr(k)
            pmls.append(p)
        med, mkeep = pa.medium_struct()
        keep += mkeep
        rq = _lib.gpi_request()
        rq.grid = pa.grid_struct()
        rq.medium = med
        rq.pml = (_lib.gpi_pml * max(1, len(pmls)))(*pmls)
        rq.npml_arrays = len(pmls)
        sh = rq.shot
        sh.nt = nt
        sh.nsrc = nsrc
        sh.nrec = len(recs)
        sh.src = src_arr
        sh.wavelets = wav_arr
        sh.rec = (_lib.gpi_points * max(1, len(recs)))(*recs)
        sh.records = (C.POINTER(C.c_float) * max(1, len(outs)))(*[_lib.fptr(o) for o in outs])
        t = C.c_uint64()
        self._check(self.lib.gpi_submit(self.h, C.byref(rq), C.byref(t)))
        self._out[t.value] = dict(zip(names, outs))   # filled by the scheduler
        return t.value

    def poll(self, ticket: int) -> int:
        st = C.c_int32()
        self._check(self.lib.gpi_poll(self.h, ticket, C.byref(st)))
        return st.value

    def wait(self, ticket: int, timeout: float = 0.0) -> dict:
        """Block until the shot is done; returns its records (nt x nr per field)."""
        st = C.c_int32()
        self._check(self.lib.gpi_wait(self.h, ticket, float(timeout), C.byref(st)))
        if st.value == 1:
            raise TimeoutError(f"shot {ticket} still pending")
        if st.value < 0:
            raise _lib.GpiError(st.value, self.lib.gpi_server_error(self.h).decode())
        return self._out.pop(ticket)

    def latency_ms(self, ticket: int) -> float:
        """Submission to records-on-host time of a finished shot."""
        ms = C.c_double()
        self._check(self.lib.gpi_latency(self.h, ticket, C.byref(ms)))
        return ms.value

    def stats(self) -> dict:
        s = _lib.gpi_server_stats()
        self._check(self.lib.gpi_server_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _lib.gpi_server_stats._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.lib.gpi_server_destroy(self.h)   # waits for every submitted shot
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class JobServer:
    """Forward and gradient jobs of 2-D or 3-D experiments on one GPU (``gpi_jobs_*``).

    ``submit(pa, iss)`` queues the forward modelling of supersource ``iss``;
    ``submit(pa, iss, dobs=recs)`` queues its gradient (``pa.checkpoint`` > 0): misfit and
    the per-supersource gradients of ``pa.gradients`` (acoustic invK, rho; elastic
    invlambda, invmu, rho), on the padded grid in FP64.  Jobs start in arrival order as
    workers and the HBM budget allow; each result equals the plain calls' (``run_shot``,
    or the :forward_save and :adjoint runs of ``gradient!``)."""

    def __init__(self, workers: int = 2, hbm_budget_bytes: int = 0, device: int = 0):
        self.lib = _lib.load()
        o = _lib.gpi_jobs_opts()
        o.workers, o.hbm_budget_bytes = workers, hbm_budget_bytes
        h = C.c_void_p()
        rc = self.lib.gpi_jobs_create(device, C.byref(o), C.byref(h))
        if rc != _lib.GPI_OK:
            raise _lib.GpiError(rc, "gpi_jobs_create failed (no HIP device or bad options)")
        self.h = h
        self._jobs = {}

    def _check(self, rc: int):
        if rc != _lib.GPI_OK:
            raise _lib.GpiError(rc, self.lib.gpi_jobs_error(self.h).decode())

    def submit(self, pa, iss: int = 0, dobs=None) -> int:
        """Queue a forward job, or with ``dobs`` (the observed Recs of supersource ``iss``)
        a gradient job; returns a ticket."""
        keep = []
        nt = len(pa.tgrid)
        src_arr, wav_arr, nsrc = pa._points(pa.spray[iss], pa.wavelets[iss], keep)
        recs, outs, names = [], [], []
        for f, (idx, w) in pa.interp[iss].items():
            idx = np.ascontiguousarray(idx)
            w = np.ascontiguousarray(w)
            out = np.zeros((nt, idx.shape[0]), F32, order="F")
            keep += [idx, w, out]
            p = _lib.gpi_points()
            p.field, p.npts, p.nnz = FIELD_ID[f], idx.shape[0], idx.shape[1]
            p.idx, p.w = _lib.iptr(idx), _lib.fptr(w)
            recs.append(p)
            outs.append(out)
            names.append(f)
        pmls = []
        for df, (a, b, k) in pa.pml.items():
            p = _lib.gpi_pml()
            p.field = FIELD_ID[df]
            p.len = a.size
            p.a, p.b, p.kI = _lib.fptr(a), _lib.fptr(b), _lib.fptr(k)
            pmls.append(p)
            keep += [a, b, k]   # the queue reads them when the job starts, after pa may be gone
        ckeep, arrs = pa.coef_arrays()
        keep += ckeep
        j = _lib.gpi_job()
        j.grid = pa.grid_struct()
        j.coefs = arrs
        j.ncoefs = len(arrs)
        j.pml = (_lib.gpi_pml * max(1, len(pmls)))(*pmls)
        j.npml_arrays = len(pmls)
        sh = j.shot
        sh.nt = nt
        sh.nsrc, sh.src, sh.wavelets = nsrc, src_arr, wav_arr
        sh.nrec = len(recs)
        sh.rec = (_lib.gpi_points * max(1, len(recs)))(*recs)
        sh.records = (C.POINTER(C.c_float) * max(1, len(outs)))(*[_lib.fptr(o) for o in outs])
        job = {"records": dict(zip(names, outs)),
               "keep": [keep, arrs, j, j.pml, sh.rec, sh.records, src_arr, wav_arr]}
        if dobs is not None:
            if not pa.checkpoint:
                raise ValueError("gradient jobs need an experiment with checkpoint > 0")
            obs = [np.asfortranarray(dobs.d[f], dtype=F32) for f in names]
            if any(o.shape != r.shape for o, r in zip(obs, outs)):
                raise ValueError("observed records must have the modelled records' shapes")
            if pa.physics == 1:
                order = ["invlambda", "invmu", "rho"]
            else:
                order = ["invK", "rho" if pa.grad_rho else None, None]
            grads = {k: np.zeros(pa.ext_shape, np.float64, order="F") for k in order if k}
            mis = C.c_double()
            j.checkpoint = pa.checkpoint
            j.dtI = float(F32(1.0 / pa.tgrid.step))
            j.dobs = (C.POINTER(C.c_float) * len(obs))(*[_lib.fptr(o) for o in obs])
            j.misfit = C.pointer(mis)
            for q, k in enumerate(order):
                j.grad[q] = _lib.dptr(grads[k]) if k else C.POINTER(C.c_double)()
            job.update(gradients=grads, misfit=mis)
            job["keep"] += [obs, j.dobs, j.misfit]
        t = C.c_uint64()
        self._check(self.lib.gpi_jobs_submit(self.h, C.byref(j), C.byref(t)))
        self._jobs[t.value] = job
        return t.value

    def wait(self, ticket: int, timeout: float = 0.0) -> dict:
        """Block until the job is done; returns {"records": ...} (+ "misfit", "gradients")."""
        st = C.c_int32()
        self._check(self.lib.gpi_jobs_wait(self.h, ticket, float(timeout), C.byref(st)))
        if st.value == 1:
            raise TimeoutError(f"job {ticket} still pending")
        job = self._jobs.pop(ticket)
        if st.value < 0:
            raise _lib.GpiError(st.value, self.lib.gpi_jobs_error(self.h).decode())
        out = {"records": job["records"]}
        if "gradients" in job:
            out["gradients"] = job["gradients"]
            out["misfit"] = job["misfit"].value
        return out

    def stats(self) -> dict:
        s = _lib.gpi_jobs_stats()
        self._check(self.lib.gpi_jobs_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _lib.gpi_jobs_stats._fields_}

    def close(self):
        if getattr(self, "h", None):
            self.lib.gpi_jobs_destroy(self.h)   # waits for every submitted job
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
