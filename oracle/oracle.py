"""ctypes driver of the CPU oracle (libfdtd_oracle.so).  TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
``run_oracle`` takes the raw experiment (unpadded medium, coordinates, wavelets) and lets
the oracle do its own preparation, so comparisons against the product also check the
host-side preparation of the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libfdtd_oracle.so")

FIELDS = [
    "p", "vx", "vy", "vz", "tauxx", "tauyy", "tauzz", "tauxy", "tauxz", "tauyz",
    "dpdx", "dpdy", "dpdz",
    "dvxdx", "dvydy", "dvzdz", "dvxdy", "dvxdz", "dvydx", "dvydz", "dvzdx", "dvzdy",
    "dtauxxdx", "dtauyydy", "dtauzzdz", "dtauxydx", "dtauxydy", "dtauxzdx", "dtauxzdz",
    "dtauyzdy", "dtauyzdz",
]
FID = {f: i for i, f in enumerate(FIELDS)}
COEFS = ["dtinvavxirho", "dtinvavyirho", "dtinvavzirho", "dtK", "dtM", "dtlambda", "dtavmu",
         "dtavxzimu", "dtavxyimu", "dtavyzimu", "rho", "invK", "invlambda", "invmu"]
CID = {c: i for i, c in enumerate(COEFS)}
FACES = {"zmin": 0, "zmax": 1, "ymin": 2, "ymax": 3, "xmin": 4, "xmax": 5}


class oracle_expt(C.Structure):
    _fields_ = [
        ("ndim", C.c_int), ("physics", C.c_int), ("order", C.c_int),
        ("nm", C.c_int64 * 3), ("o", C.c_double * 3), ("d", C.c_double * 3),
        ("vp", C.POINTER(C.c_float)), ("vs", C.POINTER(C.c_float)), ("rho", C.POINTER(C.c_float)),
        ("npml", C.c_int), ("pml_mask", C.c_uint), ("rigid_mask", C.c_uint), ("free_mask", C.c_uint),
        ("compat", C.c_uint), ("nt", C.c_int64), ("dt", C.c_double), ("freqpeak", C.c_double),
        ("ns", C.c_int), ("spos", C.POINTER(C.c_double)), ("nsf", C.c_int),
        ("sfields", C.POINTER(C.c_int)), ("wav", C.POINTER(C.c_float)),
        ("nr", C.c_int), ("rpos", C.POINTER(C.c_double)), ("nrf", C.c_int),
        ("rfields", C.POINTER(C.c_int)),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        lib.oracle_create.restype = C.c_void_p
        lib.oracle_create.argtypes = [C.POINTER(oracle_expt)]
        lib.oracle_run.argtypes = [C.c_void_p, C.c_int64]
        for name in ("oracle_get_field", "oracle_get_memory", "oracle_get_coef", "oracle_get_records"):
            getattr(lib, name).argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_int64]
        lib.oracle_run_save.argtypes = [C.c_void_p]
        lib.oracle_step.argtypes = [C.c_void_p, C.c_int64]
        lib.oracle_trailing.argtypes = [C.c_void_p]
        lib.oracle_step_velocity.argtypes = [C.c_void_p, C.c_int64]
        lib.oracle_step_stress.argtypes = [C.c_void_p, C.c_int64]
        lib.oracle_adjoint.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        lib.oracle_set_wavelets.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_int64]
        lib.oracle_get_store.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int64]
        lib.oracle_get_snap.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_int64]
        lib.oracle_set_field.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_int64]
        lib.oracle_set_coef.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_int64]
        lib.oracle_derivative.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.oracle_coef_shape.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64)]
        lib.oracle_get_pml.argtypes = [C.c_void_p, C.c_int] + [C.POINTER(C.c_float)] * 3
        lib.oracle_get_weights.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int64),
                                           C.POINTER(C.c_float)]
        lib.oracle_scalars.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
        lib.oracle_ext_grid.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        lib.oracle_field_shape.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64), C.c_int,
                                           C.POINTER(C.c_int64)]
        lib.oracle_destroy.argtypes = [C.c_void_p]
        _lib = lib
    return _lib


def _mask(faces, ndim):
    m = 0
    for f in faces:
        if f in FACES and not (ndim == 2 and f.startswith("y")):
            m |= 1 << FACES[f]
    return m


def field_shape(physics, ndim, order, n3, field):
    lib = load()
    n = (C.c_int64 * 3)(*n3)
    out = (C.c_int64 * 3)()
    if lib.oracle_field_shape(physics, ndim, order, n, FID[field], out) != 0:
        return None
    return tuple(out)


class Oracle:
    """One supersource of a forward experiment on the CPU."""

    def __init__(self, *, physics, order, grid, vp, rho, vs=None, npml, pml_faces, rigid_faces,
                 stressfree_faces=(), nt, dt, src_pos, src_fields, wavelets, rec_pos, rec_fields,
                 compat=1, freqpeak=0.0):
        lib = load()
        ndim = len(grid)
        e = oracle_expt()
        e.ndim, e.physics, e.order = ndim, physics, order
        axes = [0, 2] if ndim == 2 else [0, 1, 2]
        for k in range(3):
            e.nm[k], e.o[k], e.d[k] = 1, 0.0, 1.0
        for g, a in zip(grid, axes):
            e.nm[a], e.o[a], e.d[a] = len(g), g.first, g.step
        self._keep = []

        def fp(a):
            a = np.asfortranarray(a, dtype=np.float32)
            self._keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_float))

        def dp(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_double))

        def ip(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            self._keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_int))

        e.vp, e.rho = fp(vp), fp(rho)
        e.vs = fp(vs) if vs is not None else None
        e.npml = npml
        rigid = list(dict.fromkeys(list(rigid_faces) + list(pml_faces)))
        e.pml_mask, e.rigid_mask = _mask(pml_faces, ndim), _mask(rigid, ndim)
        e.free_mask = _mask(stressfree_faces, ndim)
        e.compat = compat
        e.nt, e.dt, e.freqpeak = nt, dt, freqpeak
        src_pos = np.asarray(src_pos, np.float64).reshape(-1, ndim)
        rec_pos = np.asarray(rec_pos, np.float64).reshape(-1, ndim)
        e.ns, e.spos = src_pos.shape[0], dp(src_pos)
        e.nsf, e.sfields = len(src_fields), ip([FID[f] for f in src_fields])
        ws = []
        for w in wavelets:
            w = np.asarray(w, np.float32).reshape(-1, src_pos.shape[0], order="F")
            full = np.zeros((nt, src_pos.shape[0]), np.float32, order="F")
            k = min(nt, w.shape[0])
            full[:k] = w[:k]
            ws.append(full.ravel(order="F"))
        wav = np.concatenate(ws)
        e.wav = fp(wav)
        e.nr, e.rpos = rec_pos.shape[0], dp(rec_pos)
        e.nrf, e.rfields = len(rec_fields), ip([FID[f] for f in rec_fields])
        self.e = e
        self.ndim, self.physics, self.order = ndim, physics, order
        self.nt, self.ns, self.nr = nt, e.ns, e.nr
        self.src_fields, self.rec_fields = list(src_fields), list(rec_fields)
        self.h = lib.oracle_create(C.byref(e))
        if not self.h:
            raise RuntimeError("oracle_create failed")
        n = (C.c_int64 * 3)()
        m0 = (C.c_double * 3)()
        lib.oracle_ext_grid(self.h, n, m0)
        self.n3 = tuple(n)
        self.m0 = tuple(m0)

    def __del__(self):
        if getattr(self, "h", None):
            load().oracle_destroy(self.h)
            self.h = None

    def run(self, nsteps=-1):
        load().oracle_run(self.h, nsteps)

    def _shape(self, f):
        s = field_shape(self.physics, self.ndim, self.order, self.n3, f)
        return (s[0], s[2]) if self.ndim == 2 else s

    def field(self, f):
        s = self._shape(f)
        out = np.zeros(s, np.float32, order="F")
        rc = load().oracle_get_field(self.h, FID[f], out.ctypes.data_as(C.POINTER(C.c_float)), out.size)
        assert rc == 0, f
        return out

    def step(self, it):
        """Time step ``it`` of the loop (record p, velocity pass, velocity sources, record v,
        stress pass, stress sources) on the current state."""
        assert load().oracle_step(self.h, it) == 0

    def step_velocity(self, it):
        """First half of step ``it``: record p, velocity pass, velocity sources."""
        assert load().oracle_step_velocity(self.h, it) == 0

    def step_stress(self, it):
        """Second half of step ``it``: record v, stress pass, stress sources."""
        assert load().oracle_step_stress(self.h, it) == 0

    def trailing(self):
        """The trailing velocity half step (src/fdtd/propagate.jl:254-255)."""
        assert load().oracle_trailing(self.h) == 0

    def run_save(self):
        """:forward_save: forward run keeping the time-reversal store and snapshots."""
        assert load().oracle_run_save(self.h) == 0

    def set_wavelets(self, q, w):
        w = np.asfortranarray(np.asarray(w, np.float32).reshape(self.nt, self.ns, order="F"))
        assert load().oracle_set_wavelets(self.h, q, w.ctypes.data_as(C.POINTER(C.c_float)), w.size) == 0

    def adjoint(self, adj: "Oracle"):
        """:adjoint with npw = 2: returns (g_invK, g_rho) on the stress grid (FP32)."""
        s = self._shape("p")
        gk = np.zeros(s, np.float32, order="F")
        gr = np.zeros(s, np.float32, order="F")
        rc = load().oracle_adjoint(self.h, adj.h, gk.ctypes.data_as(C.POINTER(C.c_float)),
                                   gr.ctypes.data_as(C.POINTER(C.c_float)))
        assert rc == 0
        return gk, gr

    def snap(self, f):
        out = np.zeros(self._shape(f), np.float32, order="F")
        assert load().oracle_get_snap(self.h, FID[f], out.ctypes.data_as(C.POINTER(C.c_float)), out.size) == 0
        return out

    def store(self, q, axis, shape):
        out = np.zeros(shape, np.float32, order="F")
        assert load().oracle_get_store(self.h, q, axis, out.ctypes.data_as(C.POINTER(C.c_float)), out.size) == 0
        return out

    def set_field(self, f, a):
        a = np.asfortranarray(a, dtype=np.float32)
        assert a.shape == self._shape(f), (a.shape, self._shape(f))
        rc = load().oracle_set_field(self.h, FID[f], a.ctypes.data_as(C.POINTER(C.c_float)), a.size)
        assert rc == 0

    def set_coef(self, c, a):
        """Overwrite coefficient ``c`` (``"rho"`` is the padded density of the sources)."""
        a = np.asfortranarray(a, dtype=np.float32)
        rc = load().oracle_set_coef(self.h, CID[c], a.ctypes.data_as(C.POINTER(C.c_float)), a.size)
        assert rc == 0, c

    def derivative(self, df, with_cpml=False):
        assert load().oracle_derivative(self.h, FID[df], int(with_cpml)) == 0
        return self.field(df)

    def coef(self, c):
        n = (C.c_int64 * 3)()
        if load().oracle_coef_shape(self.h, CID[c], n) != 0:
            return None
        s = (n[0], n[2]) if self.ndim == 2 else tuple(n)
        out = np.zeros(s, np.float32, order="F")
        load().oracle_get_coef(self.h, CID[c], out.ctypes.data_as(C.POINTER(C.c_float)), out.size)
        return out

    def pml(self, f):
        L = 2 * self.e.npml
        a, b, k = (np.zeros(L, np.float32) for _ in range(3))
        p = [x.ctypes.data_as(C.POINTER(C.c_float)) for x in (a, b, k)]
        if load().oracle_get_pml(self.h, FID[f], *p) != 0:
            return None
        return a, b, k

    def weights(self, src, q):
        npts = self.ns if src else self.nr
        nnz = 4 if self.ndim == 2 else 8
        idx = np.zeros((npts, nnz), np.int64)
        w = np.zeros((npts, nnz), np.float32)
        load().oracle_get_weights(self.h, 1 if src else 0, q, idx.ctypes.data_as(C.POINTER(C.c_int64)),
                                  w.ctypes.data_as(C.POINTER(C.c_float)))
        return idx, w

    def records(self, q):
        out = np.zeros((self.nt, self.nr), np.float32, order="F")
        load().oracle_get_records(self.h, q, out.ctypes.data_as(C.POINTER(C.c_float)), out.size)
        return out

    def scalars(self):
        s = np.zeros(6, np.float32)
        load().oracle_scalars(self.h, s.ctypes.data_as(C.POINTER(C.c_float)))
        return dict(dt=s[0], dzI=s[1], dyI=s[2], dxI=s[3], freqpeak=s[4], velavg=s[5])


def from_experiment(pa, iss=0, **kw) -> Oracle:
    """Oracle of supersource ``iss`` of a host ``SeisForwExpt`` built from the same raw inputs."""
    med = pa.medium
    a = pa.ageom[iss]
    s = pa.srcwav[iss]
    return Oracle(physics=pa.physics, order=pa.order, grid=med.grid, vp=med.vp.m, rho=med.rho.m,
                  vs=getattr(med, "vs", None).m if hasattr(med, "vs") else None,
                  npml=pa.npml, pml_faces=pa.pml_faces, rigid_faces=pa.rigid_faces,
                  stressfree_faces=pa.stressfree_faces, nt=len(pa.tgrid), dt=pa.tgrid.step,
                  src_pos=a.src_points(), src_fields=s.fields, wavelets=[s.d[f] for f in s.fields],
                  rec_pos=a.rec_points(), rec_fields=pa.rfields, compat=pa.compat, **kw)


def threads() -> int:
    return load().oracle_threads()
