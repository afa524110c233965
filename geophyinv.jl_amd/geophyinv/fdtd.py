"""``SeisForwExpt`` / ``update!`` host layer over libgpi.

Mirrors src/fdtd/fdtd.jl:43-295 (construction), src/fdtd/propagate.jl:74-135
(``update!(pa)``), src/fdtd/medium.jl:128-140 (``update!(pa, medium)``),
src/fdtd/source.jl:192-246 (``update!(pa, srcwav, src_types)``),
src/fdtd/ageom.jl:58-98 (``update!(pa, ageom)``) and src/fdtd/getprop.jl:10-31
(``pa[:data]``).  All host-side preparation happens here exactly as in the reference
(FP32 coefficients, CPML vectors, spray/interpolation weights); the time loop of every
supersource runs in libgpi on the GPU.  Julia's ``update!`` is spelled ``update`` here.
"""
from __future__ import annotations

import ctypes as C
import math
import warnings

import numpy as np

from . import _lib
from .cpml import pml_vectors
from .database import Recs, Srcs
from .fields import (COEF_ID, FIELD_ID, coef_grid, derivative_fields, field_shape, get_mgrid,
                     has_field, wavefields_of)
from .grid import Grid
from .media import (MPARAMS, AcousticMedium, ElasticMedium, Medium, MediumParameter, dependent, dtk_of,
                    independent, padarray)
from .projmat import get_proj_matrix
from .wavelets import findfreq

F32 = np.float32
FACES = {"zmin": 0, "zmax": 1, "ymin": 2, "ymax": 3, "xmin": 4, "xmax": 5}


class _Attrib:
    physics = -1
    born = False

    def __init__(self, mode: str = "forward", npw: int | None = None):
        if mode not in ("forward", "forward_save", "adjoint"):
            raise ValueError(f"unknown mode {mode!r}")
        self.mode = mode
        self.npw = npw if npw is not None else (1 if mode == "forward" else 2)

    def __class_getitem__(cls, kind):
        """``FdtdAcoustic[Born]()`` is ``FdtdAcoustic{Born}()`` (physics_types.jl:38-40):
        :forward mode with two wavefields."""
        from .born import Born
        if kind is not Born:
            raise TypeError("only Born is a modelling kind")

        def make(mode: str = "forward"):
            a = cls(mode, 2)
            a.born = True
            return a
        return make

    def __repr__(self):
        kind = "[Born]" if self.born else ""
        return f"{type(self).__name__}{kind}(mode={self.mode!r}, npw={self.npw})"


class FdtdAcoustic(_Attrib):
    """Acoustic staggered-grid modelling (src/physics_types.jl:25-49)."""
    physics = 0


class FdtdElastic(_Attrib):
    """Isotropic elastic staggered-grid modelling (src/physics_types.jl:17-49)."""
    physics = 1


def face_mask(faces, ndim: int) -> int:
    m = 0
    for f in faces:
        if f not in FACES:
            continue
        if ndim == 2 and f.startswith("y"):
            continue
        m |= 1 << FACES[f]
    return m


def get_source(w: np.ndarray, field: str, src_type: int) -> np.ndarray:
    """Source wavelet per source type (src/fdtd/source.jl:3-19)."""
    if src_type == 0:
        return np.zeros_like(w)
    if src_type == 1:
        return w.copy()
    if src_type == -1:
        if field not in ("vx", "vy", "vz"):
            raise ValueError("source type -1 exists for velocity sources only")
        ww = w * F32(-1)
        ww = np.roll(ww, -1, axis=0)
        ww = ww[::-1].copy()
        ww[0, :] = 0
        return ww
    raise ValueError(f"source type {src_type} is not implemented")


def check_stability(medium: Medium, dt: float, freqmax: float, order: int, verbose=False):
    """Courant and sampling checks, warnings only (src/fdtd/stability.jl:8-60)."""
    H = 20 // order
    eps = 1.0 / math.sqrt(medium.ndim)
    ds = [g.step for g in medium.grid]
    if isinstance(medium, ElasticMedium):
        nz = medium.vs.m[medium.vs.m != 0]
        vmin = float(nz.min()) if nz.size else 0.0
        vmax = math.sqrt(float(medium.vp.bounds[1]) ** 2 + float(medium.vs.bounds[1]) ** 2)
    else:
        vmin = float(medium.vp.bounds[0])
        vmax = float(medium.vp.bounds[1])
    if freqmax > 0 and math.isfinite(freqmax):
        ds_temp = round(vmin / H / freqmax, 2)
        if max(ds) > ds_temp and f"{max(ds):0.2e}" != f"{ds_temp:0.2e}":
            warnings.warn(f"decrease maximum spatial sampling ({max(ds):0.2e}) below {ds_temp:0.2e}")
    dt_temp = eps * min(ds) / vmax
    if dt > dt_temp and f"{dt:0.2e}" != f"{dt_temp:0.2e}":
        warnings.warn(f"decrease time sampling ({dt:0.2e}) below {dt_temp:0.2e}")


class SeisForwExpt:
    """Forward modelling experiment (``PFdtd``)."""

    def __init__(self, attrib_mod: _Attrib, *, medium: Medium, tgrid: Grid, ageom, srcwav,
                 rfields=("vz",), pml_faces=None, rigid_faces=None, stressfree_faces=(),
                 order: int = 2, npml: int | None = None, compat: int = 1, verbose: bool = False,
                 context: "_lib.Context | None" = None, coefficients: str = "host", tsnaps=None,
                 snaps_field=None, checkpoint: int = 0, grad_rho: bool = True, illum_flag: bool = False):
        ndim = medium.ndim
        if order not in (2, 4, 6, 8):
            raise ValueError("order must be 2, 4, 6 or 8")
        if attrib_mod.npw not in (1, 2) or (attrib_mod.mode == "adjoint" and attrib_mod.npw != 2):
            raise ValueError("npw must be 1 or 2, and the :adjoint mode needs npw = 2")
        if isinstance(medium, AcousticMedium) != isinstance(attrib_mod, FdtdAcoustic):
            raise ValueError("medium type does not match the modelling attribute")
        physics = attrib_mod.physics
        if pml_faces is None:
            pml_faces = ["zmin", "zmax", "ymin", "ymax", "xmax", "xmin"]
        if rigid_faces is None:
            rigid_faces = list(pml_faces)
        if npml is None:
            npml = 40 + (order - 1)   # _fd_npml (src/GeoPhyInv.jl:90)
        ageom = list(ageom) if isinstance(ageom, (list, tuple)) else [ageom]
        srcwav = list(srcwav) if isinstance(srcwav, (list, tuple)) else [srcwav]
        # ---- checks (src/fdtd/fdtd.jl:88-131)
        rfields = list(rfields)
        for rf in rfields:
            if not has_field(physics, ndim, rf) or rf.startswith("d"):
                raise ValueError(f"receiver field {rf!r} is not a field of {attrib_mod}")
        for s in srcwav:
            for f in s.fields:
                if not has_field(physics, ndim, f) or f.startswith("d"):
                    raise ValueError(f"source field {f!r} is not a field of {attrib_mod}")
        if len(ageom) != len(srcwav):
            raise ValueError("ageom and srcwav mismatch")
        if tgrid.last < srcwav[0].grid.last:
            raise ValueError("modeling time is less than source time")
        for a, s in zip(ageom, srcwav):
            if a.ns != s.n:
                raise ValueError("ageom and srcwav mismatch")
            if not a.inside(medium.grid):
                raise ValueError("sources or receivers not inside medium")
        # checkpoint > 0: :forward_save keeps the whole state every `checkpoint` steps and
        # :adjoint recomputes the forward wavefield exactly from it (gradients in FP64, 2-D
        # and 3-D acoustic); 0: the reference's boundary strips and time reversal
        if checkpoint < 0:
            raise ValueError("checkpoint must be >= 0")
        if checkpoint and physics == 1 and any(f not in ("vx", "vy", "vz") for f in rfields):
            raise ValueError("elastic gradients take velocity records only (rfields vx, vy, vz)")
        self.checkpoint = int(checkpoint)
        self.grad_rho = bool(grad_rho)
        # source illumination (illum_flag, src/fdtd/fdtd.jl:59,73,177): sum over the time steps
        # of p^2 of the first wavefield, stacked over the supersources on the unpadded grid
        if illum_flag and physics != 0:
            raise ValueError("illumination is of the acoustic pressure field")
        self.illum_flag = bool(illum_flag)
        self.illum_stack = np.zeros(tuple(len(g) for g in medium.grid), np.float64, order="F") if illum_flag else None
        self.illums = [None] * len(ageom)
        self.born = bool(attrib_mod.born)
        if self.born and (physics != 0 or attrib_mod.npw != 2):
            raise ValueError("Born modelling is acoustic, with two wavefields")
        if self.born and illum_flag:
            raise ValueError("illumination is produced by the :forward and :forward_save runs, not Born")
        self.dmod = None          # Born: perturbation of mod (invK, rho) on the padded grid
        self.born_dcoefs = None   # Born: linearised coefficient perturbations
        self.attrib_mod = attrib_mod
        self.physics = physics
        self.ndim = ndim
        self.order = order
        self.npml = int(npml)
        self.compat = int(compat)
        self.tgrid = tgrid
        self.pml_faces = list(pml_faces)
        self.rigid_faces = list(dict.fromkeys(list(rigid_faces) + list(pml_faces)))
        self.stressfree_faces = list(stressfree_faces)
        self.rfields = rfields
        self.verbose = verbose
        # snapshots (src/fdtd/fdtd.jl:74-75,179-188): the step nearest to each time in tsnaps
        if snaps_field is not None and (not has_field(physics, ndim, snaps_field) or snaps_field.startswith("d")):
            raise ValueError(f"invalid snaps field {snaps_field!r}")
        self.snaps_field = snaps_field
        if tsnaps is None:
            tsnaps = [0.5 * (tgrid.values()[-1] + tgrid.values()[0])]
        self.tsnaps = [float(t) for t in np.atleast_1d(tsnaps)]
        tv = tgrid.values()
        self.itsnaps = [int(np.argmin(np.abs(tv - t))) for t in self.tsnaps]
        self.snaps = [None] * len(ageom)
        self.medium = medium.copy()
        self.ageom = ageom
        self.srcwav = [s.copy() for s in srcwav]
        self.exmedium = padarray(self.medium, self.npml, self.pml_faces)
        # ---- get_fc (src/fdtd/fdtd.jl:313-332)
        div = {2: 1.0, 4: 24.0, 6: 1920.0, 8: 107520.0}[order]
        self.dt32 = F32(tgrid.step)
        self.dsI = [F32(1.0 / (g.step * div)) for g in self.exmedium.grid]
        # ---- pac.mod, its reference values and pac.dmod (src/fdtd/fdtd.jl:170-176, medium.jl:131-186).
        # coefficients="device" leaves the arrays to gpi_model_create_medium (2-D) and builds
        # the host copies only if a model-vector method asks for them.
        if coefficients not in ("host", "device"):
            raise ValueError("coefficients must be 'host' or 'device'")
        if coefficients == "device" and ndim != 2:
            raise ValueError("device coefficient preparation is 2-D only")
        self.coef_mode = coefficients
        self.mparams = list(MPARAMS["acoustic" if physics == 0 else "elastic"])
        self._exmedium0 = self.exmedium
        self._mod = self._ref_mod = self._coefs = None
        self._stale = set()   # mod entries changed since dmod was derived
        if coefficients == "host":
            self.coefs
        # ---- wavelets and frequency bounds (src/fdtd/source.jl:24-58, 192-246); with npw = 2
        # the second wavefield's sources sit at the receivers (fdtd.jl:90-95)
        self.npw = attrib_mod.npw
        self.srcwav2 = [Srcs(a.nr, tgrid, rfields) for a in ageom] if self.npw == 2 else []
        self.src_types = [1] * self.npw
        self._fill_wavelets()
        # ---- CPML (src/fdtd/cpml.jl:144-155): velavg = mean(exmedium[:vp].bounds)
        b = self.exmedium.vp.bounds
        self.velavg = F32(F32(b[0] + b[1]) / F32(2))
        self.pml = pml_vectors(physics, self.exmedium.grid, self.medium.grid, self.pml_faces,
                               order, self.npml, self.dt32, self.velavg, self.freqpeak)
        self.pml_freqpeak = self.freqpeak   # the CPML profile keeps the construction-time peak
        check_stability(self.medium, float(self.dt32), self.freqmax, order, verbose)
        self._update_weights()
        self.data = [Recs(a.nr, tgrid, rfields) for a in ageom]
        gdt = np.float64 if self.checkpoint else F32
        self.gradients = {k: np.zeros(self.ext_shape, gdt, order="F") for k in self.mparams}
        self._ctx = context
        self._model = None
        self._stores = {}
        self.timing = None
        self.generation = 0   # bumped by every run and every change of model, sources or geometry

    # ------------------------------------------------------------------ host preparation
    @property
    def mod(self) -> dict:
        """``pac.mod``: independent FP32 parameters on the padded grid."""
        if self._mod is None:
            self._mod = independent(self.exmedium)
        return self._mod

    @property
    def ref_mod(self) -> dict:
        """``pac.ref_mod``: reference values of the construction-time parameters."""
        if self._ref_mod is None:
            base = self._mod if self.exmedium is self._exmedium0 and self._mod is not None else independent(self._exmedium0)
            self._ref_mod = {k: MediumParameter(v).ref for k, v in base.items()}
        return self._ref_mod

    @property
    def coefs(self) -> dict:
        """``pac.mod`` and ``pac.dmod`` keyed by the reference's names."""
        if self._coefs is None:
            mod = self.mod
            self.ref_mod
            self._coefs = dependent(mod, self.dt32, self.order, self.ndim)
        return self._coefs

    @property
    def ext_shape(self) -> tuple:
        return tuple(g.n for g in self.exmedium.grid)

    def _n3(self):
        s = self.ext_shape
        return (s[0], 1, s[1]) if self.ndim == 2 else s

    def _fill_wavelets(self):
        nt = len(self.tgrid)
        self.wavelets = []
        fmins, fmaxs, fpeaks = [], [], []
        for s in self.srcwav:
            per = {}
            peaks = []
            for f in s.fields:
                w = get_source(s.d[f], f, self.src_types[0])
                full = np.zeros((nt, s.n), F32, order="F")
                snt = min(len(s.grid), nt)
                full[:snt] = w[:snt]
                per[f] = full
                if np.allclose(w, 0.0):
                    continue
                fmaxs.append(findfreq(w, s.grid, "max"))
                fmins.append(findfreq(w, s.grid, "min"))
                peaks.append(findfreq(w, s.grid, "peak"))
            self.wavelets.append(per)
            if peaks:
                fpeaks.append(float(np.mean(peaks)))
        self.freqmin = F32(max(fmins) if fmins else 0.0)
        self.freqmax = F32(min(fmaxs) if fmaxs else math.inf)
        self.freqpeak = F32(np.mean(fpeaks) if fpeaks else 0.0)
        # second wavefield: velocity sources only are ever injected (source.jl:141-156)
        self.wavelets2 = []
        for s in self.srcwav2:
            per = {}
            for f in s.fields:
                if f not in ("vx", "vy", "vz"):
                    continue
                full = np.zeros((nt, s.n), F32, order="F")
                snt = min(len(s.grid), nt)
                if self.src_types[1] == 1:     # get_source's copy, straight into place
                    full[:snt] = s.d[f][:snt]
                elif self.src_types[1] != 0:   # type 0: zeros
                    full[:snt] = get_source(s.d[f], f, self.src_types[1])[:snt]
                per[f] = full
            self.wavelets2.append(per)

    def _update_weights(self):
        grid = self.exmedium.grid
        self.spray = []
        self.interp = []
        for a, s in zip(self.ageom, self.srcwav):
            sp = {}
            for f in s.fields:
                sp[f] = get_proj_matrix(a.src_points(), get_mgrid(f, self.physics, grid, self.order), self.compat)
            self.spray.append(sp)
            ip = {}
            for f in self.rfields:
                ip[f] = get_proj_matrix(a.rec_points(), get_mgrid(f, self.physics, grid, self.order), self.compat)
            self.interp.append(ip)

    def grid_struct(self) -> "_lib.gpi_grid":
        g = _lib.gpi_grid()
        g.ndim = self.ndim
        g.order = self.order
        g.physics = self.physics
        g.npml = self.npml
        n3 = self._n3()
        for a in range(3):
            g.n[a] = n3[a]
        g.pml_mask = face_mask(self.pml_faces, self.ndim)
        g.rigid_mask = face_mask(self.rigid_faces, self.ndim)
        g.free_mask = face_mask(self.stressfree_faces, self.ndim)
        g.dt = float(self.dt32)
        ds = [self.dsI[0], self.dsI[1] if self.ndim == 3 else F32(0), self.dsI[-1]]
        for a in range(3):
            g.dsI[a] = float(ds[a])
        return g

    def coef_arrays(self, names=None):
        keep, arrs = [], []
        for name, a in self.coefs.items():
            if coef_grid(self.physics, self.ndim, name) is None or (names is not None and name not in names):
                continue
            a = np.asfortranarray(a, dtype=F32)
            keep.append(a)
            x = _lib.gpi_array()
            x.id = COEF_ID[name]
            x.count = a.size
            x.data = _lib.fptr(a)
            arrs.append(x)
        return keep, (_lib.gpi_array * len(arrs))(*arrs)

    # ------------------------------------------------------------------ device side
    def _ensure_model(self):
        if self._model is not None:
            return
        if self._ctx is None:
            self._ctx = _lib.default_context()
        ctx = self._ctx
        pmls = []
        for df, (a, b, k) in self.pml.items():
            p = _lib.gpi_pml()
            p.field = FIELD_ID[df]
            p.len = a.size
            p.a, p.b, p.kI = _lib.fptr(a), _lib.fptr(b), _lib.fptr(k)
            pmls.append(p)
        parr = (_lib.gpi_pml * len(pmls))(*pmls)
        h = C.c_uint64()
        g = self.grid_struct()
        if self.coef_mode == "device" and self._coefs is None:
            med, keep = self.medium_struct()
            ctx.check(ctx.lib.gpi_model_create_medium(ctx.h, C.byref(g), C.byref(med), parr, len(pmls), C.byref(h)))
        else:
            keep, arrs = self.coef_arrays()
            ctx.check(ctx.lib.gpi_model_create(ctx.h, C.byref(g), arrs, len(arrs), parr, len(pmls), C.byref(h)))
        self._model = h.value

    def medium_struct(self):
        """``gpi_medium`` of the unpadded medium (arrays kept alive by the second value)."""
        keep = [np.asfortranarray(getattr(self.medium, k).m, dtype=F32) for k in ("vp", "vs", "rho")
                if hasattr(self.medium, k)]
        med = _lib.gpi_medium()
        shp = self.medium.vp.m.shape
        med.n[0], med.n[1], med.n[2] = shp[0], 1, shp[-1]
        med.vp = _lib.fptr(keep[0])
        med.rho = _lib.fptr(keep[-1])
        med.vs = _lib.fptr(keep[1]) if len(keep) == 3 else C.POINTER(C.c_float)()
        return med, keep

    def device_coef(self, name: str) -> np.ndarray:
        """A coefficient array as held on the device (``gpi_model_get_coef``)."""
        self._ensure_model()
        from .fields import field_shape as fshape
        grid = {"invK": "p", "invlambda": "tauxx", "invmu": "tauxx", "rho": "p" if self.physics == 0 else "tauxx"}
        f = grid.get(name) or coef_grid(self.physics, self.ndim, name)
        out = np.zeros(fshape(self.physics, self.ndim, self.order, self.ext_shape, f), F32, order="F")
        self._ctx.check(self._ctx.lib.gpi_model_get_coef(self._ctx.h, self._model, COEF_ID[name], _lib.fptr(out),
                                                          out.size))
        return out

    def free(self):
        if self._model is not None and self._ctx is not None and self._ctx.h:
            self._ctx.check(self._ctx.lib.gpi_model_free(self._ctx.h, self._model))   # frees its stores
        self._model = None
        self._stores = {}

    def _free_store(self, iss):
        h = self._stores.pop(iss, None)
        if h is not None and self._ctx is not None and self._ctx.h:
            self._ctx.check(self._ctx.lib.gpi_store_free(self._ctx.h, h))

    @staticmethod
    def _points(spray: dict, wavelets: dict, keep: list):
        pts, wavs = [], []
        for f, (idx, w) in spray.items():
            if f not in wavelets:
                continue
            idx = np.ascontiguousarray(idx)
            w = np.ascontiguousarray(w)
            wv = np.asfortranarray(wavelets[f], dtype=F32)
            keep += [idx, w, wv]
            p = _lib.gpi_points()
            p.field, p.npts, p.nnz = FIELD_ID[f], idx.shape[0], idx.shape[1]
            p.idx, p.w = _lib.iptr(idx), _lib.fptr(w)
            pts.append(p)
            wavs.append(_lib.fptr(wv))
        return ((_lib.gpi_points * max(1, len(pts)))(*pts), (C.POINTER(C.c_float) * max(1, len(wavs)))(*wavs),
                len(pts))

    def run_shot(self, iss: int, final_fields=(), save: bool = False, snaps: bool = True, into=None,
                 static: bool = False):
        """Propagate supersource ``iss`` on the GPU; returns (records dict, finals dict).

        ``save`` runs the :forward_save variant, which keeps the boundary strips and the
        final snapshots of this supersource resident for a later :adjoint run.  ``into``
        (field -> FP32 Fortran-ordered nt x nr array) receives the records in place.
        ``static`` runs the shot through ``gpi_forward_static``, which takes its per-shot
        device buffers from the library's static scratch array instead of allocating."""
        if static and save:
            raise ValueError("the static-scratch entry point runs :forward shots only")
        self._ensure_model()
        ctx = self._ctx
        nt = len(self.tgrid)
        keep = []
        src_arr, wav_arr, nsrc = self._points(self.spray[iss], self.wavelets[iss], keep)
        recs, outs, names = [], [], []
        for f, (idx, w) in self.interp[iss].items():
            idx = np.ascontiguousarray(idx)
            w = np.ascontiguousarray(w)
            out = None if into is None else into.get(f)
            if (out is None or out.shape != (nt, idx.shape[0]) or out.dtype != F32
                    or not out.flags.f_contiguous or not out.flags.writeable):
                out = np.zeros((nt, idx.shape[0]), F32, order="F")
            keep += [idx, w, out]
            p = _lib.gpi_points()
            p.field, p.npts, p.nnz = FIELD_ID[f], idx.shape[0], idx.shape[1]
            p.idx, p.w = _lib.iptr(idx), _lib.fptr(w)
            recs.append(p)
            outs.append(out)
            names.append(f)
        finals = {}
        fptrs = []
        for f in final_fields:
            a = np.zeros(field_shape(self.physics, self.ndim, self.order, self.ext_shape, f), F32, order="F")
            finals[f] = a
            fptrs.append(_lib.fptr(a))
        shot = _lib.gpi_shot()
        shot.nt = nt
        shot.nsrc = nsrc
        shot.nrec = len(recs)
        shot.src = src_arr
        shot.wavelets = wav_arr
        shot.rec = (_lib.gpi_points * max(1, len(recs)))(*recs)
        shot.records = (C.POINTER(C.c_float) * max(1, len(outs)))(*[_lib.fptr(o) for o in outs])
        fids = np.array([FIELD_ID[f] for f in final_fields], np.int32)
        shot.nfinal = len(final_fields)
        shot.final_fields = fids.ctypes.data_as(C.POINTER(C.c_int32))
        shot.final_out = (C.POINTER(C.c_float) * max(1, len(fptrs)))(*fptrs)
        snap_out = []
        if snaps and self.snaps_field is not None:
            shp = field_shape(self.physics, self.ndim, self.order, self.ext_shape, self.snaps_field)
            snap_out = [np.zeros(shp, F32, order="F") for _ in self.itsnaps]
            steps = np.array(self.itsnaps, np.int64)
            keep.append(steps)
            shot.nsnap = len(snap_out)
            shot.snap_field = FIELD_ID[self.snaps_field]
            shot.snap_steps = steps.ctypes.data_as(C.POINTER(C.c_int64))
            shot.snap_out = (C.POINTER(C.c_float) * len(snap_out))(*[_lib.fptr(a) for a in snap_out])
            self.snaps[iss] = snap_out
        if self.illum_flag:
            ill = np.zeros(field_shape(self.physics, self.ndim, self.order, self.ext_shape, "p"), np.float64, order="F")
            shot.illum = ill.ctypes.data_as(C.POINTER(C.c_double))
            self.illums[iss] = ill
        if save:
            self._free_store(iss)
            h = C.c_uint64()
            if self.checkpoint:
                ctx.check(ctx.lib.gpi_forward_checkpoint(ctx.h, self._model, C.byref(shot), self.checkpoint,
                                                         C.byref(h)))
            else:
                ctx.check(ctx.lib.gpi_forward_save(ctx.h, self._model, C.byref(shot), C.byref(h)))
            self._stores[iss] = h.value
        elif static:
            ctx.check(ctx.lib.gpi_forward_static(ctx.h, self._model, C.byref(shot)))
        else:
            ctx.check(ctx.lib.gpi_forward(ctx.h, self._model, C.byref(shot)))
        self.timing = ctx.timing()
        return dict(zip(names, outs)), finals

    def run_adjoint(self, iss: int):
        """:adjoint run of supersource ``iss`` (npw = 2): returns its gradients
        (``pap[1].ss[iss].gradients``) on the padded grid, w.r.t. invK and rho (acoustic) or
        invlambda, invmu and rho (elastic, checkpoint > 0)."""
        if (self.physics != 0 or self.ndim != 2) and not self.checkpoint:
            raise NotImplementedError("the reference gradient exists for 2-D acoustic only (gradient.jl:17-64); "
                                      "3-D and elastic gradients need checkpoint > 0")
        if iss not in self._stores:
            raise RuntimeError("run the :forward_save mode before the :adjoint mode")
        ctx = self._ctx
        keep = []
        s1, w1, n1 = self._points(self.spray[iss], self.wavelets[iss], keep)
        # adjoint sources are sprayed with the first wavefield's receiver weights (source.jl:150)
        s2, w2, n2 = self._points(self.interp[iss], self.wavelets2[iss], keep)
        sh = _lib.gpi_adjoint_shot()
        sh.nt = len(self.tgrid)
        sh.nsrc1, sh.src1, sh.wav1 = n1, s1, w1
        sh.nsrc2, sh.src2, sh.wav2 = n2, s2, w2
        dtI = F32(1.0 / self.tgrid.step)   # fc[:dtI] (fdtd.jl:326)
        if self.physics == 1:
            out = {k: np.zeros(self.ext_shape, np.float64, order="F") for k in ("invlambda", "invmu", "rho")}
            ctx.check(ctx.lib.gpi_adjoint_checkpoint_elastic(ctx.h, self._model, self._stores[iss], C.byref(sh),
                                                             *[_lib.dptr(out[k]) for k in ("invlambda", "invmu",
                                                                                           "rho")]))
            self.timing = ctx.timing()
            return out
        if self.checkpoint:
            gk = np.zeros(self.ext_shape, np.float64, order="F")
            gr = np.zeros(self.ext_shape, np.float64, order="F") if self.grad_rho else None
            ctx.check(ctx.lib.gpi_adjoint_checkpoint(ctx.h, self._model, self._stores[iss], C.byref(sh), float(dtI),
                                                     _lib.dptr(gk), _lib.dptr(gr) if gr is not None else None))
            self.timing = ctx.timing()
            out = {"invK": gk}
            if gr is not None:
                out["rho"] = gr
            return out
        gk = np.zeros(self.ext_shape, F32, order="F")
        gr = np.zeros(self.ext_shape, F32, order="F")
        ctx.check(ctx.lib.gpi_adjoint(ctx.h, self._model, self._stores[iss], C.byref(sh), float(dtI),
                                      _lib.fptr(gk), _lib.fptr(gr)))
        self.timing = ctx.timing()
        return {"invK": gk, "rho": gr}

    def update(self):
        """``update!(pa)`` in the experiment's mode (src/fdtd/propagate.jl:74-135).

        :forward and :forward_save model every supersource and fill ``pa["data"]``
        (:forward_save also keeps the boundary values); :adjoint leaves the data zeroed, as
        ``initialize!(pa.c)`` does, and stacks the per-supersource gradients into
        ``pa.gradients`` in supersource order (sum_grads!, gradient.jl:2-12)."""
        mode = self.attrib_mod.mode
        self.generation += 1
        for d in self.data:
            d.fill(0.0)
        for g in self.gradients.values():
            g.fill(0.0)
        if self.born and mode == "forward":   # records of the scattered wavefield
            from .born import run_born
            for iss in range(len(self.ageom)):
                for f, r in run_born(self, iss).items():
                    self.data[iss].d[f][:, :] = r
            return self.timing
        if mode == "adjoint":
            self._ensure_model()
            for iss in range(len(self.ageom)):
                gi = self.run_adjoint(iss)
                for k, g in gi.items():
                    if k in self.gradients:
                        np.add(self.gradients[k], g, out=self.gradients[k])
            return self.timing
        if self.illum_flag:   # initialize! zeroes the stack (src/fdtd/types.jl:171)
            self.illum_stack.fill(0.0)
        for iss in range(len(self.ageom)):
            recs, _ = self.run_shot(iss, save=(mode == "forward_save"), into=self.data[iss].d)
            for f, r in recs.items():
                if r is not self.data[iss].d[f]:
                    self.data[iss].d[f][:, :] = r
            if self.illum_flag:   # stack_illums! (src/fdtd/fdtd.jl:556-566)
                self.illum_stack += self.illums[iss][self.view_inner(self.illums[iss])]
        return self.timing

    # ------------------------------------------------------------------ model vectors
    def view_inner(self, a: np.ndarray) -> tuple:
        """Slices of the unpadded region (``view_inner``, src/fdtd/medium.jl:13-26)."""
        axes = ["z", "x"] if self.ndim == 2 else ["z", "y", "x"]
        out = []
        for ax, n in zip(axes, a.shape):
            lo = self.npml if f"{ax}min" in self.pml_faces else 0
            hi = n - self.npml if f"{ax}max" in self.pml_faces else n
            out.append(slice(lo, hi))
        return tuple(out)

    def _upload_coefs(self, changed=None):
        """Re-derive ``dmod`` and refresh the device copies.  With ``changed == {"invK"}``
        (acoustic) only dtK depends on it, so only dtK is derived and uploaded again."""
        only_k = (changed is not None and set(changed) <= {"invK"} and self.physics == 0
                  and self._coefs is not None)
        if only_k:
            self._coefs["invK"] = self.mod["invK"]
            self._coefs["dtK"] = dtk_of(self.mod["invK"], self.dt32)
        else:
            self._coefs = dependent(self.mod, self.dt32, self.order, self.ndim)
        if self._model is not None:
            keep, arrs = self.coef_arrays(["dtK"] if only_k else None)
            self._ctx.check(self._ctx.lib.gpi_model_update(self._ctx.h, self._model, arrs, len(arrs)))

    def update_model(self, m: np.ndarray, mparams=None):
        """``update!(pa, m, mparams)`` (src/fdtd/medium.jl:31-52): the inner part of each
        parameter becomes ``exp(m) * ref``; the padded part keeps its values.  Coefficients
        of unchanged parameters are neither re-derived nor uploaded again."""
        mparams = list(mparams or self.mparams)
        chunks = np.split(np.asarray(m, dtype=np.float64), len(mparams))
        changed = set()
        for x, name in zip(chunks, mparams):
            sl = self.view_inner(self.mod[name])
            shape = self.mod[name][sl].shape
            x32 = x.reshape(shape, order="F").astype(F32)
            new = np.exp(x32) * self.ref_mod[name]
            if not np.array_equal(self.mod[name][sl], new):
                self.mod[name][sl] = new
                changed.add(name)
        if changed:
            self.generation += 1
        changed |= self._stale
        if changed or self._coefs is None:
            self._upload_coefs(changed)
        self._stale = set()

    def model_vector(self, mparams=None) -> np.ndarray:
        """``update!(m, pa, mparams)`` (src/fdtd/medium.jl:55-76): ``log(mod / ref)`` of the
        inner parts, concatenated; ``mod`` goes through the same FP32 round trip."""
        mparams = list(mparams or self.mparams)
        out = []
        for name in mparams:
            r = self.ref_mod[name]
            lg = np.log(self.mod[name] * (F32(1) / r))
            out.append(np.asarray(lg[self.view_inner(lg)], np.float64).ravel(order="F"))
            rt = np.exp(lg) * r
            if not np.array_equal(self.mod[name], rt):
                self.mod[name][...] = rt
                self._stale.add(name)   # dmod and the device copies follow at update_model
        return np.concatenate(out)

    def update_medium(self, medium: Medium):
        """``update!(pa, medium)`` without reallocation (src/fdtd/medium.jl:128-140)."""
        if type(medium) is not type(self.medium) or [g.n for g in medium.grid] != [g.n for g in self.medium.grid]:
            raise ValueError("medium must match the experiment's medium type and grid")
        self.generation += 1
        self.medium = medium.copy()
        self.exmedium = padarray(self.medium, self.npml, self.pml_faces)
        self._mod = None
        if self.coef_mode == "device" and self._coefs is None:
            self.free()   # the next run derives the coefficients on the device again
            return
        self._upload_coefs()

    def set_perturbation(self, dmod: dict):
        """Born: the perturbation of ``mod`` (``invK``, ``rho``; FP64 on the padded grid,
        zero outside the unpadded region)."""
        from .born import dcoefs
        self.dmod = dmod
        self.born_dcoefs = dcoefs(self, dmod)

    def update_perturbation(self, medium_pert: Medium):
        """``update!(pa, medium, medium_pert)`` after the background update (medium.jl:106-116)."""
        from .born import perturbation
        self.set_perturbation(perturbation(self, medium_pert))

    def migrate(self, iss: int, adj: dict) -> dict:
        """The adjoint run of supersource ``iss`` with adjoint sources ``adj`` (field -> nt x
        nr, already time reversed) at the receivers, after a forward run that keeps its
        store: the imaging of the background wavefield with the adjoint one (FP64 arrays
        ``invK``, ``rho`` on the padded grid)."""
        if self.npw != 2:
            raise ValueError("migration needs two wavefields")
        if not self.checkpoint and any(f not in ("vx", "vy", "vz") for f in self.srcwav[iss].fields):
            raise ValueError("the reference's time reversal needs velocity sources; use checkpoint > 0")
        saved_w = self.wavelets[iss]
        self.run_shot(iss, save=True)
        if not self.checkpoint:   # the reference flow re-propagates with the type -1 sink
            s = self.srcwav[iss]
            nt = len(self.tgrid)
            self.wavelets[iss] = {}
            for f in s.fields:
                full = np.zeros((nt, s.n), F32, order="F")
                w = get_source(s.d[f], f, -1)
                full[:min(len(s.grid), nt)] = w[:min(len(s.grid), nt)]
                self.wavelets[iss][f] = full
        saved_a = self.wavelets2[iss] if iss < len(self.wavelets2) else None
        self.wavelets2[iss] = {f: np.asfortranarray(a, dtype=F32) for f, a in adj.items()}
        grad_rho = self.grad_rho
        self.grad_rho = True
        try:
            g = self.run_adjoint(iss)
        finally:
            self.grad_rho = grad_rho
            self.wavelets[iss] = saved_w
            self.wavelets2[iss] = saved_a
            self._free_store(iss)
        return {k: np.asarray(v, np.float64) for k, v in g.items()}

    def update_srcwav(self, srcwav, src_types=None):
        """``update!(pa, srcwav, src_types)`` (src/fdtd/source.jl:192-246).

        ``srcwav`` is a SrcWav (one Srcs per supersource) for the first wavefield, or, with
        npw = 2, a list [SrcWav, SrcWav] for both; ``src_types`` has one entry per wavefield."""
        self.generation += 1
        srcwav = list(srcwav) if isinstance(srcwav, (list, tuple)) else [srcwav]
        per_pw = bool(srcwav) and isinstance(srcwav[0], (list, tuple))
        sw = [list(x) for x in srcwav] if per_pw else [srcwav]
        if len(sw) > self.npw or any(len(x) != len(self.srcwav) for x in sw):
            raise ValueError("srcwav must have one entry per supersource (and per wavefield)")
        # the experiment's own source sets (gradient!'s adjoint sources) are taken as they
        # are; anything else is copied, as update! copies into the experiment
        own = lambda s, mine: any(s is x for x in mine)
        self.srcwav = [s if own(s, self.srcwav) else s.copy() for s in sw[0]]
        if len(sw) == 2:
            self.srcwav2 = [s if own(s, self.srcwav2) else s.copy() for s in sw[1]]
        if src_types is not None:
            st = list(src_types) if isinstance(src_types, (list, tuple)) else [src_types]
            self.src_types[:len(st)] = st
        self._fill_wavelets()
        self._update_weights()

    def update_ageom(self, ageom):
        """``update!(pa, ageom)`` (src/fdtd/ageom.jl:58-98)."""
        self.generation += 1
        ageom = list(ageom) if isinstance(ageom, (list, tuple)) else [ageom]
        for a, b in zip(ageom, self.ageom):
            if a.ns != b.ns or a.nr != b.nr:
                raise ValueError("ageom must keep the number of sources and receivers")
            if not a.inside(self.medium.grid):
                raise ValueError("sources or receivers not inside medium")
        self.ageom = ageom
        self._update_weights()

    def __getitem__(self, key):
        """``pa[:data]`` and friends (src/fdtd/getprop.jl:10-31)."""
        if isinstance(key, tuple):
            key, iss = key
        else:
            iss = None
        if key == "data":
            return self.data if iss is None else self.data[iss]
        if key == "snaps":   # pa[:snaps, i] (src/fdtd/getprop.jl:13-24)
            return self.snaps[0 if iss is None else iss]
        if key == "illum_stack":   # pac.illum_stack (src/fdtd/types.jl:155)
            return self.illum_stack
        if key in ("ageom", "srcwav", "medium", "exmedium", "rfields", "pml_faces", "rigid_faces", "gradients",
                   "mod", "ref_mod"):
            return getattr(self, key)
        raise KeyError(key)


def update(obj, *args, **kw):
    """Dispatch of the reference's ``update!`` methods on a ``SeisForwExpt``;
    ``update(m, pa, mparams)`` fills the model vector ``m`` from ``pa``."""
    if isinstance(obj, np.ndarray) and args and isinstance(args[0], SeisForwExpt):
        m = args[0].model_vector(*args[1:])
        obj[...] = m
        return obj
    if not isinstance(obj, SeisForwExpt):
        raise TypeError("update expects a SeisForwExpt")
    if not args:
        return obj.update()
    a = args[0]
    if isinstance(a, np.ndarray):
        return obj.update_model(a, *args[1:])
    if isinstance(a, Medium):
        obj.update_medium(a)
        if len(args) > 1 and isinstance(args[1], Medium):   # update!(pa, medium, medium_pert)
            obj.update_perturbation(args[1])
        return None
    if isinstance(a, list) and a and hasattr(a[0], "ns") and hasattr(a[0], "s"):
        return obj.update_ageom(a)
    return obj.update_srcwav(a, *args[1:], **kw)
