"""Shared builders for the tests: small experiments and comparison metrics."""
from __future__ import annotations

import numpy as np

import geophyinv as G
from geophyinv.grid import Grid


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def wavelet(nt: int, dt: float, f: float = 20.0, delay: float | None = None) -> np.ndarray:
    """Ricker-shaped pulse that needs no minimum record length (any nt)."""
    t = np.arange(nt) * dt
    t0 = 1.2 / f if delay is None else delay
    a = (np.pi * f * (t - t0)) ** 2
    return (1 - 2 * a) * np.exp(-a)


def layered_acoustic(nz=48, nx=56, ds=5.0, seed=7):
    rng = np.random.Generator(np.random.PCG64(seed))
    grid = [Grid(0.0, ds, nz), Grid(0.0, ds, nx)]
    vp = 1800.0 + 600.0 * (np.arange(nz)[:, None] / nz) + 80.0 * rng.standard_normal((nz, nx))
    rho = 1900.0 + 300.0 * rng.random((nz, nx))
    return G.AcousticMedium(grid, vp.astype(np.float32), rho.astype(np.float32))


def layered_elastic(nz=48, nx=56, ds=5.0, seed=11):
    rng = np.random.Generator(np.random.PCG64(seed))
    grid = [Grid(0.0, ds, nz), Grid(0.0, ds, nx)]
    vp = 2000.0 + 900.0 * (np.arange(nz)[:, None] / nz) + 60.0 * rng.standard_normal((nz, nx))
    vs = vp / (1.7 + 0.2 * rng.random((nz, nx)))
    rho = 310.0 * vp ** 0.25
    return G.ElasticMedium(grid, vp.astype(np.float32), vs.astype(np.float32), rho.astype(np.float32))


def experiment(medium, *, nt=240, sfields=("p",), rfields=("p", "vx", "vz"), ns=1, nr=9, order=2,
               npml=10, pml_faces=None, stressfree_faces=(), seed=3, ctx=None, src_depth=None,
               rec_depth=0.0, jitter=True):
    """A single-supersource experiment with random (seeded) off-node positions."""
    rng = np.random.Generator(np.random.PCG64(seed))
    gz, gx = medium.grid
    Lz, Lx = gz.last, gx.last
    dt = 0.4 * gz.step / (1.6 * float(medium.vp.m.max()))
    tgrid = Grid(0.0, dt, nt)
    sz = np.full(ns, src_depth if src_depth is not None else 0.3 * Lz)
    sz = sz + (rng.uniform(-1, 1, ns) * gz.step if jitter else 0.0)
    sx = np.linspace(0.35 * Lx, 0.65 * Lx, ns) + (rng.uniform(-1, 1, ns) * gx.step if jitter else 0)
    rz = np.full(nr, rec_depth) + (rng.uniform(0, 1, nr) * gz.step if jitter else 0)
    rx = np.linspace(0.0, Lx, nr)
    ag = G.AGeomss({"z": sz, "x": sx}, {"z": rz, "x": rx})
    src = G.Srcs(ns, tgrid, list(sfields))
    w = np.stack([wavelet(nt, dt, 25.0, 0.03 + 0.004 * k) for k in range(ns)], axis=1)
    src.update(list(sfields), w)
    attrib = G.FdtdAcoustic() if isinstance(medium, G.AcousticMedium) else G.FdtdElastic()
    kw = {}
    if pml_faces is not None:
        kw["pml_faces"] = pml_faces
    return G.SeisForwExpt(attrib, medium=medium, tgrid=tgrid, ageom=[ag], srcwav=[src],
                          rfields=list(rfields), order=order, npml=npml,
                          stressfree_faces=list(stressfree_faces), context=ctx, **kw)


def medium3d(physics, n=(22, 18, 24), ds=10.0, seed=5):
    rng = np.random.Generator(np.random.PCG64(seed))
    grid = [Grid(0.0, ds, n[0]), Grid(0.0, ds, n[1]), Grid(0.0, ds, n[2])]
    vp = (2000.0 + 500.0 * (np.arange(n[0])[:, None, None] / n[0]) + 60.0 * rng.standard_normal(n)).astype(np.float32)
    rho = (1800.0 + 300.0 * rng.random(n)).astype(np.float32)
    if physics == 0:
        return G.AcousticMedium(grid, vp, rho)
    vs = (vp / (1.7 + 0.2 * rng.random(n))).astype(np.float32)
    return G.ElasticMedium(grid, vp, vs, rho)


def experiment3d(medium, *, nt=120, sfields=("p",), rfields=("p", "vx", "vy", "vz"), ns=2, nr=6,
                 order=2, npml=5, pml_faces=None, stressfree_faces=(), compat=1, seed=9, ctx=None):
    rng = np.random.Generator(np.random.PCG64(seed))
    gz, gy, gx = medium.grid
    dt = 0.3 * gz.step / (1.7 * float(medium.vp.m.max()))
    tgrid = Grid(0.0, dt, nt)
    s = {"z": gz.last * rng.uniform(0.3, 0.6, ns), "y": gy.last * rng.uniform(0.3, 0.7, ns),
         "x": gx.last * rng.uniform(0.3, 0.7, ns)}
    r = {"z": gz.step * rng.uniform(0, 1.5, nr), "y": gy.last * rng.uniform(0, 1, nr),
         "x": gx.last * rng.uniform(0, 1, nr)}
    ag = G.AGeomss(s, r)
    src = G.Srcs(ns, tgrid, list(sfields))
    w = np.stack([wavelet(nt, dt, 18.0, 0.05 + 0.005 * k) for k in range(ns)], axis=1)
    src.update(list(sfields), w)
    attrib = G.FdtdAcoustic() if isinstance(medium, G.AcousticMedium) else G.FdtdElastic()
    kw = {}
    if pml_faces is not None:
        kw["pml_faces"] = pml_faces
    return G.SeisForwExpt(attrib, medium=medium, tgrid=tgrid, ageom=[ag], srcwav=[src],
                          rfields=list(rfields), order=order, npml=npml, compat=compat,
                          stressfree_faces=list(stressfree_faces), context=ctx, **kw)


# ---------------------------------------------------------------------------------------
# The reference gradient flow (src/fdtd/func_grad.jl:11-49) on the oracle
# ---------------------------------------------------------------------------------------

ORACLE_COEFS = ("dtinvavxirho", "dtinvavzirho", "dtK", "rho")


def _oracle_coefs(pa):
    return ORACLE_COEFS + (("dtinvavyirho",) if pa.ndim == 3 else ())


def oracle_pair(pa, iss, adj_wavelets=None):
    """Oracles of the two wavefields of supersource ``iss`` of a 2-D acoustic npw = 2
    experiment, on ``pa``'s coefficient arrays; the second sprays at the receivers."""
    import oracle as O
    from geophyinv.fdtd import get_source

    a = pa.ageom[iss]
    s1 = O.from_experiment(pa, iss, freqpeak=float(pa.pml_freqpeak))
    for c in _oracle_coefs(pa):
        s1.set_coef(c, pa.coefs[c])
    vfields = [f for f in pa.rfields if f in ("vx", "vy", "vz")]
    nt = len(pa.tgrid)
    if adj_wavelets is None:
        adj_wavelets = {f: np.zeros((nt, a.nr), np.float32) for f in vfields}
    med = pa.medium
    s2 = O.Oracle(physics=0, order=pa.order, grid=med.grid, vp=med.vp.m, rho=med.rho.m, npml=pa.npml,
                  pml_faces=pa.pml_faces, rigid_faces=pa.rigid_faces, nt=nt, dt=pa.tgrid.step,
                  src_pos=a.rec_points(), src_fields=vfields, wavelets=[adj_wavelets[f] for f in vfields],
                  rec_pos=a.rec_points(), rec_fields=vfields, freqpeak=float(pa.pml_freqpeak), compat=pa.compat)
    for c in _oracle_coefs(pa):
        s2.set_coef(c, pa.coefs[c])
    return s1, s2, get_source


def oracle_gradient(pa, dobs):
    """(J, g_invK, g_rho, dcal) of the reference flow for every supersource, on the oracle:
    forward_save, adjoint sources 2 (dobs - dcal) reversed in time, type -1 wavelets for the
    first wavefield, :adjoint run, FP32 stacking over supersources."""
    g_k = np.zeros(pa.ext_shape, np.float32, order="F")
    g_r = np.zeros(pa.ext_shape, np.float32, order="F")
    J = 0.0
    dcal = []
    for iss, a in enumerate(pa.ageom):
        s1, _, get_source = oracle_pair(pa, iss)
        s1.run_save()
        rec = {f: s1.records(q) for q, f in enumerate(pa.rfields)}
        dcal.append(rec)
        adj = {}
        for f in pa.rfields:
            d = dobs[iss].d[f].astype(np.float64)
            J += float(np.sum((d - rec[f].astype(np.float64)) ** 2))
            if f in ("vx", "vy", "vz"):
                adj[f] = (2.0 * (d - rec[f].astype(np.float64))).astype(np.float32)[::-1].copy()
        _, s2, _ = oracle_pair(pa, iss, adj)
        for q, f in enumerate(pa.srcwav[iss].fields):
            s1.set_wavelets(q, get_source(pa.srcwav[iss].d[f], f, -1))
        gk, gr = s1.adjoint(s2)
        g_k = g_k + gk
        g_r = g_r + gr
    return J, g_k, g_r, dcal


def oracle_gradient_exact(pa, dobs, grad_rho=True, lag=0):
    """(J, g_invK, g_rho, dcal) with the exact forward wavefield, on the oracle: the forward
    oracle stepped to keep p_n and v_n after every step (and v at nt+1/2), the adjoint
    oracle driven by the time-reversed 2 (dobs - dcal) as in the reference flow, and after
    adjoint step it (n = nt-1-it)
        g_invK += p2_prev * (p_{n-1} - p_n) * dtI
        g_rho[inn] -= @av_xi(vx2_prev (vx_n - vx_{n+1}) dtI) (+ @av_yi) + @av_zi(...)
    with each term in FP32 and the sums in FP64, stacked over supersources in FP64.  ``lag``
    pairs adjoint step it with forward step n - lag instead (a planted misalignment for
    tests that check their own sensitivity)."""
    nt = len(pa.tgrid)
    h = pa.order - 1
    dtI = np.float32(1.0 / pa.tgrid.step)
    shape = pa.ext_shape
    g_k = np.zeros(shape, np.float64, order="F")
    g_r = np.zeros(shape, np.float64, order="F")
    vnames = ["vx", "vz"] if pa.ndim == 2 else ["vx", "vy", "vz"]
    J = 0.0
    dcal = []
    for iss in range(len(pa.ageom)):
        s1, _, _ = oracle_pair(pa, iss)
        P, V = [np.zeros(shape, np.float32, order="F")], {f: [] for f in vnames}
        for n in range(nt):
            s1.step(n)
            P.append(s1.field("p"))
            if grad_rho:
                for f in vnames:
                    V[f].append(s1.field(f))
        s1.trailing()
        if grad_rho:
            for f in vnames:
                V[f].append(s1.field(f))
        rec = {f: s1.records(q) for q, f in enumerate(pa.rfields)}
        dcal.append(rec)
        adj = {}
        for f in pa.rfields:
            d = dobs[iss].d[f].astype(np.float64)
            J += float(np.sum((d - rec[f].astype(np.float64)) ** 2))
            if f in ("vx", "vy", "vz"):
                adj[f] = (2.0 * (d - rec[f].astype(np.float64))).astype(np.float32)[::-1].copy()
        _, s2, _ = oracle_pair(pa, iss, adj)
        p2 = np.zeros(shape, np.float32, order="F")
        v2 = {f: np.zeros_like(V[f][0]) for f in vnames} if grad_rho else {}
        for it in range(nt):
            n = nt - 1 - it
            s2.step(it)
            q = min(max(n - lag, 0), nt - 1)
            g_k += ((p2 * (P[q] - P[q + 1])) * dtI).astype(np.float64)
            if grad_rho:
                b = {f: (v2[f] * (V[f][n] - V[f][n + 1])) * dtI for f in vnames}
                if pa.ndim == 2:
                    nz, nx = shape
                    I, K = slice(0, nz - 2 * h), slice(0, nx - 2 * h)
                    bx, bz = b["vx"], b["vz"]
                    avx = (bx[h:nz - h, 0:nx - 2 * h] + bx[h:nz - h, 1:nx - 2 * h + 1]) * np.float32(0.5)
                    avz = (bz[0:nz - 2 * h, h:nx - h] + bz[1:nz - 2 * h + 1, h:nx - h]) * np.float32(0.5)
                    inn = (slice(h, nz - h), slice(h, nx - h))
                    g_r[inn] = (g_r[inn] - avx.astype(np.float64)) - avz.astype(np.float64)
                else:
                    nz, ny, nx = shape
                    bx, by, bz = b["vx"], b["vy"], b["vz"]
                    zi, yi, xi = slice(h, nz - h), slice(h, ny - h), slice(h, nx - h)
                    avx = (bx[zi, yi, 0:nx - 2 * h] + bx[zi, yi, 1:nx - 2 * h + 1]) * np.float32(0.5)
                    avy = (by[zi, 0:ny - 2 * h, xi] + by[zi, 1:ny - 2 * h + 1, xi]) * np.float32(0.5)
                    avz = (bz[0:nz - 2 * h, yi, xi] + bz[1:nz - 2 * h + 1, yi, xi]) * np.float32(0.5)
                    inn = (zi, yi, xi)
                    g_r[inn] = ((g_r[inn] - avx.astype(np.float64)) - avy.astype(np.float64)) - avz.astype(np.float64)
                v2 = {f: s2.field(f) for f in vnames}
            p2 = s2.field("p")
    return J, g_k, g_r, dcal


def oracle_born(pa):
    """Born records of every supersource on the oracle: the background oracle (sources)
    and the scattered oracle (no sources, the receivers) stepped half by half; after both
    velocity halves the scattered velocities receive d(dtinvav*irho) * D(p1), after both
    stress halves the scattered pressure receives (D(vx1) + D(vz1) [+ D(vy1)]) * d(dtK),
    in FP32 with the derivatives of the oracle (no CPML memory)."""
    import oracle as O

    h = pa.order - 1
    dc = pa.born_dcoefs
    nt = len(pa.tgrid)
    out = []
    d3 = pa.ndim == 3
    vnames = ["vx", "vy", "vz"] if d3 else ["vx", "vz"]
    for iss, a in enumerate(pa.ageom):
        s1 = O.from_experiment(pa, iss, freqpeak=float(pa.pml_freqpeak))
        med = pa.medium
        zero = [np.zeros((nt, a.ns), np.float32) for _ in pa.srcwav[iss].fields]
        s2 = O.Oracle(physics=0, order=pa.order, grid=med.grid, vp=med.vp.m, rho=med.rho.m, npml=pa.npml,
                      pml_faces=pa.pml_faces, rigid_faces=pa.rigid_faces, nt=nt, dt=pa.tgrid.step,
                      src_pos=a.src_points(), src_fields=pa.srcwav[iss].fields, wavelets=zero,
                      rec_pos=a.rec_points(), rec_fields=pa.rfields, freqpeak=float(pa.pml_freqpeak),
                      compat=pa.compat)
        for s in (s1, s2):
            for c in _oracle_coefs(pa):
                s.set_coef(c, pa.coefs[c])
        for it in range(nt):
            s1.step_velocity(it)
            s2.step_velocity(it)
            for v in vnames:
                ax = v[1]
                d = s1.derivative(f"dpd{ax}")
                f = s2.field(v)
                inn = tuple(slice(h, h + n) for n in d.shape)
                f[inn] = f[inn] + dc[f"dtinvav{ax}irho"] * d
                s2.set_field(v, f)
            s1.step_stress(it)
            s2.step_stress(it)
            dv = s1.derivative("dvxdx") + s1.derivative("dvzdz")
            if d3:
                dv = dv + s1.derivative("dvydy")
            s2.set_field("p", s2.field("p") + dv * dc["dtK"])
        out.append({f: s2.records(q) for q, f in enumerate(pa.rfields)})
    return out


def fwi_experiment(medium, *, nss=2, nt=300, sfields=("vz",), rfields=("vz",), nr=9, order=2, npml=10,
                   pml_faces=None, mode="forward_save", ctx=None, seed=5, **extra):
    """A 2-D acoustic npw = 2 experiment with ``nss`` supersources (one source each) and a
    line of receivers below them."""
    rng = np.random.Generator(np.random.PCG64(seed))
    gz, gx = medium.grid
    Lz, Lx = gz.last, gx.last
    dt = 0.4 * gz.step / (1.6 * float(medium.vp.m.max()))
    tgrid = Grid(0.0, dt, nt)
    ageom, srcwav = [], []
    for k in range(nss):
        sx = Lx * (0.3 + 0.4 * k / max(nss - 1, 1)) + rng.uniform(-1, 1) * gx.step
        sz = 0.2 * Lz + rng.uniform(-1, 1) * gz.step
        ag = G.AGeomss({"z": [sz], "x": [sx]},
                       {"z": np.full(nr, 0.75 * Lz) + rng.uniform(0, 1, nr) * gz.step,
                        "x": np.linspace(0.1 * Lx, 0.9 * Lx, nr)})
        s = G.Srcs(1, tgrid, list(sfields))
        s.update(list(sfields), wavelet(nt, dt, 25.0, 0.03 + 0.002 * k) * 1e3)
        ageom.append(ag)
        srcwav.append(s)
    kw = {} if pml_faces is None else {"pml_faces": pml_faces}
    return G.SeisForwExpt(G.FdtdAcoustic(mode, 2), medium=medium, tgrid=tgrid, ageom=ageom, srcwav=srcwav,
                          rfields=list(rfields), order=order, npml=npml, context=ctx, **kw, **extra)


def fwi_experiment3d(medium, *, nss=1, nt=160, sfields=("vz",), rfields=("vz",), nr=12, order=2, npml=6,
                     ctx=None, seed=13, **extra):
    """A 3-D acoustic npw = 2 experiment: ``nss`` supersources (one source each) in the
    upper part, a patch of receivers below them."""
    rng = np.random.Generator(np.random.PCG64(seed))
    gz, gy, gx = medium.grid
    Lz, Ly, Lx = gz.last, gy.last, gx.last
    dt = 0.3 * gz.step / (1.7 * float(medium.vp.m.max()))
    tgrid = Grid(0.0, dt, nt)
    ageom, srcwav = [], []
    for k in range(nss):
        s = {"z": [0.25 * Lz + rng.uniform(-1, 1) * gz.step], "y": [Ly * rng.uniform(0.35, 0.65)],
             "x": [Lx * rng.uniform(0.35, 0.65)]}
        r = {"z": np.full(nr, 0.7 * Lz) + rng.uniform(0, 1, nr) * gz.step, "y": Ly * rng.uniform(0.15, 0.85, nr),
             "x": Lx * rng.uniform(0.15, 0.85, nr)}
        ag = G.AGeomss(s, r)
        src = G.Srcs(1, tgrid, list(sfields))
        src.update(list(sfields), wavelet(nt, dt, 18.0, 0.05 + 0.004 * k) * 1e3)
        ageom.append(ag)
        srcwav.append(src)
    return G.SeisForwExpt(G.FdtdAcoustic("forward_save", 2), medium=medium, tgrid=tgrid, ageom=ageom,
                          srcwav=srcwav, rfields=list(rfields), order=order, npml=npml, context=ctx, **extra)


def oracle_records(medium, pa_like):
    """Oracle records (one Recs per supersource) of ``pa_like``'s geometry and wavelets in
    another medium: the observed data of gradient tests."""
    import oracle as O
    out = []
    for iss in range(len(pa_like.ageom)):
        a, s = pa_like.ageom[iss], pa_like.srcwav[iss]
        o = O.Oracle(physics=0, order=pa_like.order, grid=medium.grid, vp=medium.vp.m, rho=medium.rho.m,
                     npml=pa_like.npml, pml_faces=pa_like.pml_faces, rigid_faces=pa_like.rigid_faces,
                     nt=len(pa_like.tgrid), dt=pa_like.tgrid.step, src_pos=a.src_points(),
                     src_fields=s.fields, wavelets=[s.d[f] for f in s.fields], rec_pos=a.rec_points(),
                     rec_fields=pa_like.rfields, compat=pa_like.compat)
        o.run()
        out.append(G.Recs(a.nr, pa_like.tgrid, pa_like.rfields,
                          {f: o.records(q) for q, f in enumerate(pa_like.rfields)}))
    return out


# ---------------------------------------------------------------------------------------
# Host restatements the device FWI pieces are checked against
# ---------------------------------------------------------------------------------------

def apply_proj_host(m, mats):
    """apply_proj_matrix! (src/proj_mat.jl:208-226) with numpy: m contracted with mats[0]
    along axis 0, then mats[1] along axis 1 (...), FP64 sums, stored FP32."""
    out = np.asarray(m, np.float64)
    for ax, P in enumerate(mats):
        out = np.moveaxis(np.tensordot(np.asarray(P, np.float64), out, axes=([0], [ax])), 0, ax)
    return out.astype(np.float32)


def lbfgs_two_loop(pairs, g):
    """-H g of L-BFGS (Nocedal & Wright, Algorithm 7.4) for the (s, y) pairs, oldest first,
    with H0 = gamma I, gamma = s'y / y'y of the newest pair (H0 = I without pairs)."""
    q = np.array(g, np.float64)
    alphas = []
    for s, y in reversed(pairs):
        a = float(np.dot(s, q)) / float(np.dot(s, y))
        alphas.append(a)
        q -= a * y
    if pairs:
        s, y = pairs[-1]
        q *= float(np.dot(s, y)) / float(np.dot(y, y))
    for (s, y), a in zip(pairs, reversed(alphas)):
        b = float(np.dot(y, q)) / float(np.dot(s, y))
        q += (a - b) * s
    return -q
