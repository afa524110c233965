"""ctypes binding of libgpi (include/gpi.h).

The library is loaded from this package directory.  There is no CPU fallback: when the
shared library is missing or no HIP device is present, ``GpiError`` is raised.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpi.so")

GPI_OK = 0
ERRORS = {-1: "GPI_EINVAL", -2: "GPI_ENOMEM", -3: "GPI_EHIP", -4: "GPI_ENONFINITE", -5: "GPI_ENOTFOUND"}


class GpiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class gpi_grid(C.Structure):
    _fields_ = [("ndim", C.c_int32), ("order", C.c_int32), ("physics", C.c_int32),
                ("npml", C.c_int32), ("n", C.c_int64 * 3), ("pml_mask", C.c_uint32),
                ("rigid_mask", C.c_uint32), ("free_mask", C.c_uint32), ("reserved", C.c_uint32),
                ("dt", C.c_float), ("dsI", C.c_float * 3)]


class gpi_array(C.Structure):
    _fields_ = [("id", C.c_int32), ("reserved", C.c_int32), ("count", C.c_int64),
                ("data", C.POINTER(C.c_float))]


class gpi_pml(C.Structure):
    _fields_ = [("field", C.c_int32), ("len", C.c_int32), ("a", C.POINTER(C.c_float)),
                ("b", C.POINTER(C.c_float)), ("kI", C.POINTER(C.c_float))]


class gpi_points(C.Structure):
    _fields_ = [("field", C.c_int32), ("npts", C.c_int32), ("nnz", C.c_int32),
                ("reserved", C.c_int32), ("idx", C.POINTER(C.c_int64)), ("w", C.POINTER(C.c_float))]


class gpi_shot(C.Structure):
    _fields_ = [("nt", C.c_int64), ("nsrc", C.c_int32), ("nrec", C.c_int32),
                ("src", C.POINTER(gpi_points)), ("wavelets", C.POINTER(C.POINTER(C.c_float))),
                ("rec", C.POINTER(gpi_points)), ("records", C.POINTER(C.POINTER(C.c_float))),
                ("nfinal", C.c_int32), ("nsnap", C.c_int32),
                ("final_fields", C.POINTER(C.c_int32)),
                ("final_out", C.POINTER(C.POINTER(C.c_float))),
                ("snap_field", C.c_int32), ("reserved", C.c_int32),
                ("snap_steps", C.POINTER(C.c_int64)),
                ("snap_out", C.POINTER(C.POINTER(C.c_float))),
                ("illum", C.POINTER(C.c_double))]


class gpi_adjoint_shot(C.Structure):
    _fields_ = [("nt", C.c_int64), ("nsrc1", C.c_int32), ("nsrc2", C.c_int32),
                ("src1", C.POINTER(gpi_points)), ("wav1", C.POINTER(C.POINTER(C.c_float))),
                ("src2", C.POINTER(gpi_points)), ("wav2", C.POINTER(C.POINTER(C.c_float)))]


class gpi_medium(C.Structure):
    _fields_ = [("n", C.c_int64 * 3), ("vp", C.POINTER(C.c_float)), ("vs", C.POINTER(C.c_float)),
                ("rho", C.POINTER(C.c_float))]


class gpi_server_opts(C.Structure):
    _fields_ = [("physics", C.c_int32), ("order", C.c_int32), ("max_shots", C.c_int32),
                ("chunk_steps", C.c_int32), ("autotune", C.c_int32), ("tile", C.c_int32),
                ("direct_launch", C.c_int32), ("reserved", C.c_int32), ("hbm_budget_bytes", C.c_uint64)]


class gpi_request(C.Structure):
    _fields_ = [("grid", gpi_grid), ("medium", gpi_medium), ("pml", C.POINTER(gpi_pml)),
                ("npml_arrays", C.c_int32), ("reserved", C.c_int32), ("shot", gpi_shot)]


class gpi_server_stats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("active", C.c_int64), ("queued", C.c_int64), ("steps", C.c_int64), ("launches", C.c_int64),
                ("busy_ms", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_batch", C.c_double), ("tile_ty", C.c_int32), ("tile_rx", C.c_int32),
                ("hbm_used", C.c_uint64), ("hbm_budget", C.c_uint64), ("graph_chunks", C.c_int64),
                ("graph_captures", C.c_int64), ("host_admit_ms", C.c_double), ("host_launch_ms", C.c_double),
                ("host_retire_ms", C.c_double), ("host_finalize_ms", C.c_double), ("host_collect_ms", C.c_double),
                ("pool_bytes", C.c_uint64)]


class gpi_jobs_opts(C.Structure):
    _fields_ = [("workers", C.c_int32), ("reserved", C.c_int32), ("hbm_budget_bytes", C.c_uint64)]


class gpi_job(C.Structure):
    _fields_ = [("grid", gpi_grid), ("coefs", C.POINTER(gpi_array)), ("ncoefs", C.c_int32),
                ("npml_arrays", C.c_int32), ("pml", C.POINTER(gpi_pml)), ("shot", gpi_shot),
                ("checkpoint", C.c_int32), ("dtI", C.c_float), ("dobs", C.POINTER(C.POINTER(C.c_float))),
                ("misfit", C.POINTER(C.c_double)), ("grad", C.POINTER(C.c_double) * 3)]


class gpi_jobs_stats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64), ("queued", C.c_int64),
                ("running", C.c_int64), ("forward_done", C.c_int64), ("gradient_done", C.c_int64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("forward_p50_ms", C.c_double),
                ("gradient_p50_ms", C.c_double), ("hbm_used", C.c_uint64), ("hbm_budget", C.c_uint64)]


class gpi_timing(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("velocity_avg_ms", C.c_double), ("stress_avg_ms", C.c_double),
                ("steps", C.c_int64), ("sampled", C.c_int64), ("launches", C.c_int64)]


# every symbol include/gpi.h declares, with its ctypes signature
SIGNATURES = {
    "gpi_abi_version": (C.c_int, []),
    "gpi_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_uint64]),
    "gpi_destroy": (None, [C.c_void_p]),
    "gpi_last_error": (C.c_char_p, [C.c_void_p]),
    "gpi_field_shape": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int64), C.c_int32,
                                  C.POINTER(C.c_int64)]),
    "gpi_model_create": (C.c_int, [C.c_void_p, C.POINTER(gpi_grid), C.POINTER(gpi_array), C.c_int32,
                                   C.POINTER(gpi_pml), C.c_int32, C.POINTER(C.c_uint64)]),
    "gpi_model_update": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(gpi_array), C.c_int32]),
    "gpi_model_free": (C.c_int, [C.c_void_p, C.c_uint64]),
    "gpi_model_create_medium": (C.c_int, [C.c_void_p, C.POINTER(gpi_grid), C.POINTER(gpi_medium),
                                          C.POINTER(gpi_pml), C.c_int32, C.POINTER(C.c_uint64)]),
    "gpi_model_get_coef": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.POINTER(C.c_float), C.c_int64]),
    "gpi_server_create": (C.c_int, [C.c_int, C.POINTER(gpi_server_opts), C.POINTER(C.c_void_p)]),
    "gpi_server_destroy": (None, [C.c_void_p]),
    "gpi_server_error": (C.c_char_p, [C.c_void_p]),
    "gpi_submit": (C.c_int, [C.c_void_p, C.POINTER(gpi_request), C.POINTER(C.c_uint64)]),
    "gpi_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int32)]),
    "gpi_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, C.POINTER(C.c_int32)]),
    "gpi_server_get_stats": (C.c_int, [C.c_void_p, C.POINTER(gpi_server_stats)]),
    "gpi_jobs_create": (C.c_int, [C.c_int, C.POINTER(gpi_jobs_opts), C.POINTER(C.c_void_p)]),
    "gpi_jobs_destroy": (None, [C.c_void_p]),
    "gpi_jobs_error": (C.c_char_p, [C.c_void_p]),
    "gpi_jobs_submit": (C.c_int, [C.c_void_p, C.POINTER(gpi_job), C.POINTER(C.c_uint64)]),
    "gpi_jobs_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, C.POINTER(C.c_int32)]),
    "gpi_jobs_get_stats": (C.c_int, [C.c_void_p, C.POINTER(gpi_jobs_stats)]),
    "gpi_latency": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_double)]),
    "gpi_forward": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(gpi_shot)]),
    "gpi_forward_static": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(gpi_shot)]),
    "gpi_forward_save": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(gpi_shot), C.POINTER(C.c_uint64)]),
    "gpi_adjoint": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(gpi_adjoint_shot), C.c_float,
                              C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "gpi_store_free": (C.c_int, [C.c_void_p, C.c_uint64]),
    "gpi_born": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(gpi_shot), C.POINTER(gpi_array), C.c_int32]),
    "gpi_forward_checkpoint": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(gpi_shot), C.c_int32,
                                         C.POINTER(C.c_uint64)]),
    "gpi_adjoint_checkpoint": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(gpi_adjoint_shot), C.c_float,
                                         C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "gpi_adjoint_checkpoint_elastic": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(gpi_adjoint_shot),
                                                 C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "gpi_interp_apply": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                   C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "gpi_lbfgs_create": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_void_p)]),
    "gpi_lbfgs_destroy": (None, [C.c_void_p]),
    "gpi_lbfgs_reset": (C.c_int, [C.c_void_p]),
    "gpi_lbfgs_direction": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    "gpi_conv_apply": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.POINTER(C.c_double), C.c_int32,
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "gpi_conv_lsmr": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                C.c_int32, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double),
                                C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "gpi_last_timing": (C.c_int, [C.c_void_p, C.POINTER(gpi_timing)]),
    "gpi_set_pass_sampling": (C.c_int, [C.c_void_p, C.c_int32]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load libgpi and bind every entry point; raises when the library is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise GpiError(-3, f"{path} is missing: build it with __graft_entry__.build() "
                           "or make -C geophyinv.jl_amd/csrc")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(C.POINTER(C.c_double))


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(C.POINTER(C.c_float))


def iptr(a: np.ndarray):
    assert a.dtype == np.int64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(C.POINTER(C.c_int64))


class Context:
    """One HIP device with its stream; owns resident models."""

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.gpi_create(C.byref(h), device, hbm_budget_bytes)
        if rc != GPI_OK:
            raise GpiError(rc, "gpi_create failed (no HIP device?)")
        self.h = h

    def check(self, rc: int):
        if rc != GPI_OK:
            raise GpiError(rc, self.lib.gpi_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.gpi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_pass_sampling(self, nsamples: int):
        self.check(self.lib.gpi_set_pass_sampling(self.h, nsamples))

    def timing(self) -> dict:
        t = gpi_timing()
        self.check(self.lib.gpi_last_timing(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in gpi_timing._fields_}


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx
