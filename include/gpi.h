/*
 * libgpi — MI355X (gfx950) staggered-grid velocity-stress FDTD engine, C ABI.
 *
 * This is the drop-in boundary for the time-stepping body of GeoPhyInv.jl's
 * `mod_x_proc!` (src/fdtd/propagate.jl:138-261).  The host (Julia `ccall`, Python
 * `ctypes`, C/C++) keeps building the FP32 coefficient arrays (src/fdtd/medium.jl:143-221),
 * the CPML vectors (src/fdtd/cpml.jl:8-142) and the spray/interpolation weights
 * (src/proj_mat.jl:208-247, src/fdtd/ageom.jl:29-54) exactly as the reference does, and
 * hands them to this library, which keeps them resident in HBM and runs the time loop.
 *
 * Conventions
 *  - All arrays are FP32, column-major with z fastest (Julia layout): a field of size
 *    (nz, ny, nx) has element (iz, iy, ix) at iz + nz*(iy + ny*ix), zero-based.  2-D grids
 *    use ny = 1 and omit the y axis everywhere.
 *  - Every entry point returns 0 on success or a negative GPI_E* code; a message is then
 *    available from gpi_last_error().  No entry point aborts or throws across the ABI.
 *  - A context is not re-entrant: use one per host thread or serialise calls.
 *  - The caller owns every host pointer for the duration of a call; the library copies
 *    what it keeps.
 */
#ifndef GPI_H
#define GPI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPI_ABI_VERSION 5

/* error codes */
#define GPI_OK 0
#define GPI_EINVAL (-1)      /* bad argument or inconsistent sizes */
#define GPI_ENOMEM (-2)      /* device allocation failed or over the context's HBM budget */
#define GPI_EHIP (-3)        /* HIP runtime error */
#define GPI_ENONFINITE (-4)  /* non-finite values found in the records */
#define GPI_ENOTFOUND (-5)   /* unknown handle */

/* Field ids.  Names and staggered grids follow src/fields.jl:9-42 (types) and
 * src/fields.jl:92-671 (grids).  The derivative fields exist for the CPML coefficient
 * vectors (src/fdtd/cpml.jl:109-120); the kernels never store them. */
enum gpi_field {
    GPI_P = 0, GPI_VX, GPI_VY, GPI_VZ,
    GPI_TAUXX, GPI_TAUYY, GPI_TAUZZ, GPI_TAUXY, GPI_TAUXZ, GPI_TAUYZ,
    GPI_DPDX, GPI_DPDY, GPI_DPDZ,
    GPI_DVXDX, GPI_DVYDY, GPI_DVZDZ, GPI_DVXDY, GPI_DVXDZ, GPI_DVYDX, GPI_DVYDZ, GPI_DVZDX,
    GPI_DVZDY,
    GPI_DTAUXXDX, GPI_DTAUYYDY, GPI_DTAUZZDZ, GPI_DTAUXYDX, GPI_DTAUXYDY, GPI_DTAUXZDX,
    GPI_DTAUXZDZ, GPI_DTAUYZDY, GPI_DTAUYZDZ,
    GPI_NFIELDS
};

/* Coefficient ids: the reference's `dmod` (dependent) and `mod` (independent) arrays,
 * src/fdtd/medium.jl:81-95.  Each lives on the grid of the field named in that table:
 *   DTINVAV{X,Y,Z}IRHO on dpdx/dpdy/dpdz (acoustic) or dtauxxdx/dtauyydy/dtauzzdz (elastic);
 *   DTK, DTM, DTLAMBDA, RHO, INVK, INVLAMBDA, INVMU on the stress grid;
 *   DTAVMU (2-D) on dvxdz; DTAVXZIMU on dvxdz, DTAVXYIMU on dvxdy, DTAVYZIMU on dvydz. */
enum gpi_coef {
    GPI_C_DTINVAVXIRHO = 0, GPI_C_DTINVAVYIRHO, GPI_C_DTINVAVZIRHO, GPI_C_DTK, GPI_C_DTM,
    GPI_C_DTLAMBDA, GPI_C_DTAVMU, GPI_C_DTAVXZIMU, GPI_C_DTAVXYIMU, GPI_C_DTAVYZIMU,
    GPI_C_RHO, GPI_C_INVK, GPI_C_INVLAMBDA, GPI_C_INVMU,
    GPI_NCOEF
};

/* Face bits: bit (2*axis + side), axis 0 = z, 1 = y, 2 = x; side 0 = min, 1 = max.
 * Mirrors the reference's pml_faces / rigid_faces / stressfree_faces symbol lists
 * (src/fdtd/fdtd.jl:61-78, :215). */
#define GPI_FACE(axis, side) (1u << (2 * (axis) + (side)))

typedef struct gpi_ctx gpi_ctx;

/* Grid and scheme of one modelling problem (P_common's ic/fc, src/fdtd/fdtd.jl:201-202,
 * src/fdtd/fdtd.jl:301-332). */
typedef struct gpi_grid {
    int32_t ndim;        /* 2 or 3 (LocalPreferences `ndims`) */
    int32_t order;       /* 2, 4, 6 or 8 (LocalPreferences `order`) */
    int32_t physics;     /* 0 FdtdAcoustic, 1 FdtdElastic */
    int32_t npml;        /* CPML points per face; the reference uses 40 + order - 1 */
    int64_t n[3];        /* extended stress grid nz, ny, nx (ny = 1 in 2-D) */
    uint32_t pml_mask;   /* faces carrying CPML memory updates */
    uint32_t rigid_mask; /* faces with Dirichlet velocity conditions (pml ∪ rigid) */
    uint32_t free_mask;  /* stress-free faces (only zmin is implemented by the reference) */
    uint32_t reserved;
    float dt;            /* fc[:dt] as Float32 */
    float dsI[3];        /* fc[:dzI], fc[:dyI], fc[:dxI] as Float32, order divisor included */
} gpi_grid;

/* One coefficient array handed over at model creation. */
typedef struct gpi_array {
    int32_t id;          /* gpi_coef */
    int32_t reserved;
    int64_t count;       /* number of floats; must equal the size of its grid */
    const float* data;   /* host pointer, column-major */
} gpi_array;

/* CPML vectors of one derivative field (pml[df][:a], [:b], [:kI], src/fdtd/cpml.jl:88-102):
 * length 2*npml, the first npml for the min face and the last npml for the max face. */
typedef struct gpi_pml {
    int32_t field;       /* derivative field id */
    int32_t len;         /* 2*npml */
    const float* a;
    const float* b;
    const float* kI;
} gpi_pml;

/* Sparse spray or interpolation weights of one field (one CSC matrix of
 * src/proj_mat.jl:229-247): npts columns with nnz entries each, entries of a column in
 * ascending row order, row = zero-based column-major index into the field's array. */
typedef struct gpi_points {
    int32_t field;       /* field the points live on */
    int32_t npts;        /* sources (spray) or receivers (interpolation) */
    int32_t nnz;         /* 4 in 2-D, 8 in 3-D */
    int32_t reserved;
    const int64_t* idx;  /* npts*nnz */
    const float* w;      /* npts*nnz */
} gpi_points;

/* One supersource ("shot") of a forward run, src/fdtd/propagate.jl:138-261 with
 * npw = 1 and mode :forward. */
typedef struct gpi_shot {
    int64_t nt;                      /* time steps (length(tgrid)) */
    int32_t nsrc;                    /* source fields */
    int32_t nrec;                    /* receiver fields */
    const gpi_points* src;           /* nsrc spray matrices (pap.ss[i].ssprayw) */
    const float* const* wavelets;    /* nsrc arrays of nt x npts, time fastest (Srcs.d) */
    const gpi_points* rec;           /* nrec interpolation matrices (rinterpolatew) */
    float* const* records;           /* nrec host outputs of nt x npts, time fastest (Recs.d) */
    int32_t nfinal;                  /* optional final wavefields to copy back */
    int32_t nsnap;                   /* snapshots (snaps_field at tsnaps, fdtd.jl:179-188) */
    const int32_t* final_fields;     /* field ids */
    float* const* final_out;         /* host outputs, each of its field's full size */
    int32_t snap_field;              /* field id of the snapshots */
    int32_t reserved;
    const int64_t* snap_steps;       /* nsnap zero-based steps: the field after that step */
    float* const* snap_out;          /* nsnap host outputs of the field's full size */
    double* illum;                   /* optional (acoustic): sum over the steps of p^2 after each
                                        step, FP32 square and FP64 sum (compute_illum!,
                                        src/fdtd/fdtd.jl:570-581), host output on the stress grid */
} gpi_shot;

/* Timing of the last gpi_forward call on a context (device time, HIP events on the
 * context's stream). */
typedef struct gpi_timing {
    double total_ms;        /* whole time loop, inputs already resident */
    double velocity_avg_ms; /* mean duration of one velocity-pass launch (sampled steps) */
    double stress_avg_ms;   /* mean duration of one stress-pass launch (sampled steps) */
    int64_t steps;          /* time steps run */
    int64_t sampled;        /* steps whose passes were bracketed by events (0 = sampling off) */
    int64_t launches;       /* kernel launches issued by the time loop */
} gpi_timing;

/* ---- context ------------------------------------------------------------------------ */
/* No reference counterpart: the reference selects its backend at load time
 * (src/GeoPhyInv.jl:95-100).  hbm_budget_bytes = 0 means "free HBM at creation". */
int gpi_create(gpi_ctx** out, int device, uint64_t hbm_budget_bytes);
void gpi_destroy(gpi_ctx* ctx);
const char* gpi_last_error(const gpi_ctx* ctx);
int gpi_abi_version(void);

/* Size of a field's array for a grid (src/fields.jl:48-86 `zeros(::field, ...)`).
 * Returns GPI_EINVAL when the field does not exist for that physics/ndim. */
int gpi_field_shape(int32_t ndim, int32_t order, int32_t physics, const int64_t n[3],
                    int32_t field, int64_t out[3]);

/* ---- model (medium + scheme resident in HBM) ----------------------------------------- */
/* Replaces the device-side state built by PFdtd (src/fdtd/fdtd.jl:135-160) and refreshed
 * by update!(pa, medium) (src/fdtd/medium.jl:128-140): coefficient arrays and CPML
 * vectors.  Required coefficients: acoustic DTINVAV{X,(Y),Z}IRHO, DTK, RHO; elastic
 * DTINVAV{X,(Y),Z}IRHO, DTM, DTLAMBDA, DTAVMU (2-D) or DTAV{XZ,XY,YZ}IMU (3-D), RHO,
 * INVLAMBDA and INVMU (the elastic passes form their dt-scaled coefficients from rho,
 * invlambda and invmu).  That is every pac.dmod and pac.mod array of MediumParameters
 * (src/fdtd/medium.jl:81-95); INVK is accepted for acoustic models and kept for
 * gpi_model_get_coef.  One gpi_pml per derivative field of the physics when
 * pml_mask != 0. */
int gpi_model_create(gpi_ctx* ctx, const gpi_grid* grid, const gpi_array* coefs,
                     int32_t ncoefs, const gpi_pml* pml, int32_t npml_arrays,
                     uint64_t* model_out);
/* update!(pa, medium) without reallocation: overwrite coefficient arrays in place. */
int gpi_model_update(gpi_ctx* ctx, uint64_t model, const gpi_array* coefs, int32_t ncoefs);
int gpi_model_free(gpi_ctx* ctx, uint64_t model);

/* A medium on its unpadded grid (column-major, z fastest; vs for elastic only). */
typedef struct gpi_medium {
    int64_t n[3];        /* nz, 1, nx of the unpadded medium */
    const float* vp;
    const float* vs;
    const float* rho;
} gpi_medium;

/* gpi_model_create with the coefficient arrays derived on the device: the medium side of
 * update!(pac, medium) (src/fdtd/medium.jl:128-186) after padarray (src/media/media.jl:
 * 260-304): replicate padding by npml on the PML faces, invK (or invlambda, invmu) and
 * every dt-scaled array in the same FP32 expressions as the host preparation.  2-D only;
 * grid->n must be the medium size plus the padding. */
int gpi_model_create_medium(gpi_ctx* ctx, const gpi_grid* grid, const gpi_medium* medium,
                            const gpi_pml* pml, int32_t npml_arrays, uint64_t* model_out);

/* Copy a coefficient array of a model to the host (count = its field grid's size; the
 * independent invK/invlambda/invmu live on the stress grid).  No reference counterpart:
 * the reference reads pac.dmod directly. */
int gpi_model_get_coef(gpi_ctx* ctx, uint64_t model, int32_t coef, float* out, int64_t count);

/* ---- forward modelling ---------------------------------------------------------------- */
/* One supersource of update!(pa) in :forward mode (src/fdtd/propagate.jl:74-135 with the
 * body of mod_x_proc!, :138-261): zero the wavefields and CPML memory, run nt steps
 *   record p; velocity pass (stress derivatives + CPML + velocity update + rigid faces);
 *   velocity sources; record v; stress pass (velocity derivatives + CPML + stress update +
 *   free surface); stress sources,
 * then the trailing velocity half step of :254-255, and copy records (and the optional
 * final wavefields) to the host. */
int gpi_forward(gpi_ctx* ctx, uint64_t model, const gpi_shot* shot);

/* gpi_forward without device allocation: the same shot, same records, but every per-shot
 * device buffer (source and receiver tables, device records, transposed records, the
 * non-finite flag, snapshot/illumination/packing slabs) is carved out of one static
 * device array of GPI_STATIC_SCRATCH_BYTES that each device holds until the process
 * ends; no hipMalloc or hipFree is called.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  A shot whose buffers exceed the array fails with GPI_ENOMEM before any
 * step runs.  Every call on a device shares the one array, so calls that overlap on one
 * device (several contexts, several host threads) race; the call drains its stream
 * before returning.  The model itself is created as for gpi_forward (gpi_model_create,
 * which does allocate). */
#define GPI_STATIC_SCRATCH_BYTES (64u << 20)
int gpi_forward_static(gpi_ctx* ctx, uint64_t model, const gpi_shot* shot);

/* ---- time reversal and the reference gradient ---------------------------------------- */
/* One supersource in :forward_save mode (src/fdtd/propagate.jl:138-261 with
 * boundary_save! at :229 and the snapshots of :251-258, src/fdtd/boundary.jl:217-306):
 * the forward run of gpi_forward that also keeps, on the device, -stress on the
 * nbound = 3 planes just inside the PML of every face for every step, -stress at nt and
 * the velocities at nt+1/2.  The store stays resident until gpi_store_free (or until its
 * model is freed). */
int gpi_forward_save(gpi_ctx* ctx, uint64_t model, const gpi_shot* shot, uint64_t* store_out);

/* Inputs of the :adjoint run with npw = 2 (src/fdtd/fdtd.jl:90-99, source.jl:124-157). */
typedef struct gpi_adjoint_shot {
    int64_t nt;
    int32_t nsrc1;                 /* first wavefield: the forward sources (velocity fields) */
    int32_t nsrc2;                 /* second wavefield: adjoint sources at the receivers */
    const gpi_points* src1;        /* the forward spray weights (ssprayw of pw1) */
    const float* const* wav1;      /* their wavelets after get_source(type -1) (source.jl:8-19) */
    const gpi_points* src2;        /* the receiver interpolation weights (rinterpolatew of pw1) */
    const float* const* wav2;      /* time-reversed adjoint sources, nt x nr, time fastest */
} gpi_adjoint_shot;

/* The :adjoint run of one supersource for 2-D acoustic physics (the only case with a
 * gradient in the reference, src/fdtd/gradient.jl:17-64): the first wavefield restarts
 * from the store's snapshots and re-propagates backwards with its strips forced from the
 * store (boundary.jl:115-212); the second propagates the adjoint sources; after each step
 *   g_invK += p2_prev * (p1_prev - p1) * dtI
 *   g_rho[i+h, j+h] -= @av_xi(vx2_prev (vx1 - vx1_prev) dtI) + @av_zi(vz2_prev (vz1 - vz1_prev) dtI)
 * g_invK and g_rho are host outputs on the stress grid (FP32, per supersource, as the
 * reference's pap.ss[i].gradients); dtI = Float32(1/dt) of the time grid. */
int gpi_adjoint(gpi_ctx* ctx, uint64_t model, uint64_t store, const gpi_adjoint_shot* shot, float dtI,
                float* g_invK, float* g_rho);
int gpi_store_free(gpi_ctx* ctx, uint64_t store);

/* ---- checkpointed gradient (2-D and 3-D; acoustic and elastic) ----------------------- */
/* The :forward_save run of gpi_forward_save with checkpoints in place of boundary strips
 * (no reference counterpart: the reference keeps strips, src/fdtd/boundary.jl:217-306, and
 * has no 3-D gradient).  The store keeps the whole wavefield and CPML state after every
 * `every`-th time step, the velocities at nt+1/2 and the forward sources, so that
 * gpi_adjoint_checkpoint can recompute the forward wavefield exactly, one segment of
 * `every` steps at a time.  Same records as gpi_forward.  Memory: (ceil(nt/every) - 1)
 * wavefield states. */
int gpi_forward_checkpoint(gpi_ctx* ctx, uint64_t model, const gpi_shot* shot, int32_t every,
                           uint64_t* store_out);

/* The gradient of one supersource from a checkpoint store.  The adjoint wavefield is
 * gpi_adjoint's second wavefield (src2/wav2; src1/wav1 are not used: the forward sources are
 * the stored run's).  The forward wavefield is recomputed from the checkpoints, segment by
 * segment, instead of being re-propagated backwards, so it is exact up to the PML.  After
 * adjoint step it (forward step n = nt-1-it), with p_n, v_n the forward fields after step n
 * and v_nt the trailing half step:
 *   g_invK += p2_prev * (p_{n-1} - p_n) * dtI                    (compute_gmodKI!)
 *   g_rho[inn] -= @av_xi(vx2_prev (vx_n - vx_{n+1}) dtI) [+ @av_yi(vy ...)] + @av_zi(vz ...)
 * (gradient.jl:17-64 with the first wavefield's sign convention; in 3-D @av_yi joins the
 * other two as in diff3D.jl).  Each step's term is formed in FP32 as in the reference and
 * accumulated in FP64.  g_invK, g_rho: host outputs on the stress grid, FP64; g_rho may be
 * NULL (the velocities are then not kept: 1 instead of 1 + ndim slabs per segment step). */
int gpi_adjoint_checkpoint(gpi_ctx* ctx, uint64_t model, uint64_t store, const gpi_adjoint_shot* shot,
                           float dtI, double* g_invK, double* g_rho);

/* Elastic gradient of one supersource from a checkpoint store (2-D P-SV and 3-D; no
 * reference counterpart: the reference's gradient is acoustic only, src/fdtd/gradient.jl).
 * The store comes from gpi_forward_checkpoint on an elastic model whose receivers record
 * velocities only; the adjoint sources (src2/wav2) are the time-reversed velocity
 * residuals, as in gpi_adjoint_checkpoint.  The adjoint wavefield is the forward scheme run
 * backwards, which is the exact adjoint of the discrete update away from the CPML; per
 * forward step n the adjoint stresses tau~^{n+1} and velocities v~^{n+1/2} meet the forward
 * increments tau^{n+1} - tau^n and v^{n+1/2} - v^{n-1/2}, and the chain rule through the
 * dt-scaled coefficients of update_dmod! (src/fdtd/medium.jl:143-186) gives, in FP64, the
 * gradients w.r.t. the independent parameters invlambda, invmu and rho (mod, medium.jl:81-95)
 * on the stress grid.  Memory: per segment step, every stress and velocity slab. */
int gpi_adjoint_checkpoint_elastic(gpi_ctx* ctx, uint64_t model, uint64_t store, const gpi_adjoint_shot* shot,
                                   double* g_invlambda, double* g_invmu, double* g_rho);

/* ---- Born modelling (acoustic, 2-D and 3-D) ------------------------------------------ */
/* One supersource of FdtdAcoustic{Born} in :forward mode (src/fdtd/born.jl,
 * src/fdtd/propagate.jl:53-60: both wavefields propagate, the sources drive the background
 * one, the receivers sample the scattered one).  Per step, after both velocity passes
 *   v2[inn] += d(dtinvav{x,y,z}irho) * D_{x,y,z}i(p1)
 * and after both stress passes
 *   p2 += (D_xa(vx1) + D_za(vz1) [+ D_ya(vy1)]) * d(dtK)
 * with plain stencil derivatives.  These secondary sources are the linearisation of
 * compute_v!/compute_p! (advance_acou.jl) in their coefficients; the reference's own Born
 * calls do not match those kernels' arguments, so there is no reference result to follow.
 * dcoefs: the coefficient perturbations GPI_C_DTK and GPI_C_DTINVAV{X,(Y,)Z}IRHO on their
 * grids (host, FP32), zero wherever a CPML memory variable acts.  The shot's nfinal and
 * nsnap must be 0. */
int gpi_born(gpi_ctx* ctx, uint64_t model, const gpi_shot* shot, const gpi_array* dcoefs, int32_t ndcoefs);

/* ---- serving: continuous batching of many shots on one GPU ------------------------- */
/* No reference counterpart (the reference runs supersources one after another per worker,
 * src/fdtd/propagate.jl:100-106).  A server owns a device and a scheduler thread.  Shots
 * (2-D, one physics and stencil order per server) are admitted as they arrive while
 * HBM and slots allow, stay resident (medium, wavefields, CPML memory, sources,
 * receivers) and advance together: every time step is one velocity and one stress launch
 * covering all resident shots, each at its own step.  Finished shots leave and waiting
 * ones enter between chunks of steps, so requests that arrive while others run overlap
 * with them on the device.  Each shot's records equal those of gpi_forward on the same
 * inputs bit for bit (same kernels, same order of operations). */
typedef struct gpi_server gpi_server;

typedef struct gpi_server_opts {
    int32_t physics;             /* 0 acoustic, 1 elastic */
    int32_t order;               /* 2, 4, 6, 8 */
    int32_t max_shots;           /* resident shots at most (default 64) */
    int32_t chunk_steps;         /* time steps between admissions (default 32, at most 4096) */
    int32_t autotune;            /* 1: pick the workgroup shape per batch size by timing */
    int32_t tile;                /* fixed shape index when autotune = 0 (0..3) */
    int32_t direct_launch;       /* 0: a chunk whose batch and shape are those of the chunk
                                    before it replays a captured HIP graph of its 2 x
                                    chunk_steps launches; 1: every pass launched directly */
    int32_t reserved;            /* 0 */
    uint64_t hbm_budget_bytes;   /* 0: 90 % of the free HBM at creation */
} gpi_server_opts;

/* One shot: the medium on its unpadded grid (coefficients are derived on the device as in
 * gpi_model_create_medium), the CPML vectors, and a gpi_shot whose record buffers must
 * stay valid until the shot is done.  Everything else is copied at submission. */
typedef struct gpi_request {
    gpi_grid grid;
    gpi_medium medium;
    const gpi_pml* pml;
    int32_t npml_arrays;
    int32_t reserved;
    gpi_shot shot;               /* nfinal and nsnap must be 0 */
} gpi_request;

typedef struct gpi_server_stats {
    int64_t submitted, completed, failed;
    int64_t active, queued;
    int64_t steps;               /* batched time steps run */
    int64_t launches;
    double busy_ms;              /* device time of the batched steps */
    double p50_ms, p99_ms;       /* latency (submission to records in the caller's buffers) */
    double mean_batch;           /* mean resident shots per step */
    int32_t tile_ty, tile_rx;    /* workgroup shape of the last chunk */
    uint64_t hbm_used, hbm_budget;
    int64_t graph_chunks;        /* chunks run as a graph replay */
    int64_t graph_captures;      /* graphs captured and instantiated */
    /* host time of the scheduler thread (ms): admitting shots (device state, uploads),
     * launching chunks (table upload, graph capture and launch), retiring shots (record
     * copies queued), handing records to the caller, draining for memory release */
    double host_admit_ms, host_launch_ms, host_retire_ms, host_finalize_ms, host_collect_ms;
    uint64_t pool_bytes;         /* device memory held by the scheduler's pool */
} gpi_server_stats;

int gpi_server_create(int device, const gpi_server_opts* opts, gpi_server** out);
/* Waits for every submitted shot, then stops the scheduler and frees the device state. */
void gpi_server_destroy(gpi_server* srv);
const char* gpi_server_error(const gpi_server* srv);
int gpi_submit(gpi_server* srv, const gpi_request* req, uint64_t* ticket);
/* status: 0 done, 1 pending, < 0 the shot's error code (message in gpi_server_error). */
int gpi_poll(gpi_server* srv, uint64_t ticket, int32_t* status);
int gpi_wait(gpi_server* srv, uint64_t ticket, double timeout_s, int32_t* status);
int gpi_server_get_stats(gpi_server* srv, gpi_server_stats* out);
/* Milliseconds from gpi_submit to the records being in the caller's buffers, for a
 * finished shot (GPI_ENOTFOUND until then). */
int gpi_latency(gpi_server* srv, uint64_t ticket, double* ms);

/* ---- job queue: mixed forward and gradient jobs, 2-D or 3-D -------------------------- */
/* No reference counterpart.  Jobs wait in arrival order and start when a worker is free
 * and their estimated device memory fits the HBM budget; each worker owns a context and
 * stream, so the device work, the host work (coefficient upload, residuals) and the
 * copies of several jobs overlap.  A forward job is gpi_forward; a gradient job is
 * gpi_forward_checkpoint, the misfit sum (dcal - dobs)^2 in FP64 and the adjoint sources
 * 2 (dobs - dcal) in FP32 reversed in time (database.jl:357-380), then
 * gpi_adjoint_checkpoint (acoustic) or gpi_adjoint_checkpoint_elastic; its results equal
 * those calls made one after another (GPU test).  Every buffer a job names (coefficients,
 * CPML vectors, points, wavelets, observed records and the outputs) must stay valid until
 * the job is done. */
typedef struct gpi_jobs gpi_jobs;

typedef struct gpi_jobs_opts {
    int32_t workers;             /* jobs in flight at most (default 2) */
    int32_t reserved;            /* 0 */
    uint64_t hbm_budget_bytes;   /* 0: 90 % of the free HBM at creation */
} gpi_jobs_opts;

typedef struct gpi_job {
    gpi_grid grid;
    const gpi_array* coefs;      /* as gpi_model_create */
    int32_t ncoefs;
    int32_t npml_arrays;
    const gpi_pml* pml;
    gpi_shot shot;               /* records into shot.records; nfinal = nsnap = 0 */
    int32_t checkpoint;          /* 0: forward job; > 0: gradient job, checkpoint interval */
    float dtI;                   /* gradient jobs, acoustic: Float32(1/dt) of the time grid */
    const float* const* dobs;    /* gradient jobs: observed records per receiver set, nt x npts */
    double* misfit;              /* gradient jobs: output */
    double* grad[3];             /* gradient jobs: stress-grid FP64 outputs, invK and rho (grad[1]
                                    may be NULL) or invlambda, invmu and rho */
} gpi_job;

typedef struct gpi_jobs_stats {
    int64_t submitted, completed, failed, queued, running;
    int64_t forward_done, gradient_done;
    double p50_ms, p99_ms;              /* submission to outputs written, all jobs */
    double forward_p50_ms, gradient_p50_ms;
    uint64_t hbm_used, hbm_budget;      /* estimated bytes of the running jobs */
} gpi_jobs_stats;

int gpi_jobs_create(int device, const gpi_jobs_opts* opts, gpi_jobs** out);
/* Waits for every submitted job, then stops the workers. */
void gpi_jobs_destroy(gpi_jobs* q);
const char* gpi_jobs_error(const gpi_jobs* q);
int gpi_jobs_submit(gpi_jobs* q, const gpi_job* job, uint64_t* ticket);
/* status: 0 done, 1 pending, < 0 the job's error code (message in gpi_jobs_error). */
int gpi_jobs_wait(gpi_jobs* q, uint64_t ticket, double timeout_s, int32_t* status);
int gpi_jobs_get_stats(gpi_jobs* q, gpi_jobs_stats* out);

/* ---- the FWI outer loop on the device ------------------------------------------------- */
/* Separable linear map of apply_proj_matrix! (src/proj_mat.jl:208-226): y is x contracted
 * with w[0] along axis 0, then w[1] along axis 1 (then w[2] along axis 2), where w[a] is the
 * dense column-major FP32 (nin[a] x nout[a]) matrix of 1-D interpolation weights of
 * get_proj_matrix (:176-200) or its transpose; FP64 sums, x of nin[0] x .. and y of
 * nout[0] x .. elements (z fastest), host arrays.  The inversion-grid model to the
 * modelling grid (P) and the modelling-grid gradient back to the inversion grid (P'). */
int gpi_interp_apply(gpi_ctx* ctx, int32_t ndim, const int64_t nin[3], const int64_t nout[3],
                     const float* const w[3], const double* x, double* y);

/* L-BFGS direction with the pair history resident in HBM (the optimizer the reference
 * drives from src/fwi/fwi.jl:66-75, Optim.jl LBFGS with memory 10).  Each call takes the
 * current model m and gradient g (n doubles each, host); from the second call on it first
 * folds the pair (m - m_prev, g - g_prev) into the history when s'y > 0 (at most `memory`
 * pairs, oldest dropped), then returns d = -H g by the two-loop recursion with H0 = gamma I,
 * gamma = s'y / y'y of the newest pair (d = -g while the history is empty), and the number
 * of pairs used.  No reference counterpart for the split: the reference keeps the history
 * on the host. */
typedef struct gpi_lbfgs gpi_lbfgs;
int gpi_lbfgs_create(gpi_ctx* ctx, int64_t n, int32_t memory, gpi_lbfgs** out);
void gpi_lbfgs_destroy(gpi_lbfgs* lbfgs);
int gpi_lbfgs_reset(gpi_lbfgs* lbfgs);   /* forget the history and the previous point */
int gpi_lbfgs_direction(gpi_lbfgs* lbfgs, const double* m, const double* g, double* d, int32_t* pairs);

/* Source-signature estimation (update!(paw, pac), src/fwi/fwi.jl:37-64, with the
 * convolution operator of src/Conv.jl:30-480).  g and d are nt x ntr (time fastest, FP64,
 * host); the filter s has nt samples: nnl = nt - 1 - (nt - 1) / 2 negative lags (most
 * negative first), lag 0 at index nnl, then (nt - 1) / 2 positive lags.  d = g * s
 * convolves every trace of g with the shared s (lags 0 .. nt-1 kept).
 * gpi_conv_apply: mode 0 out (nt x ntr) = g * in (nt); mode 1 out (nt) = the adjoint, the
 * sum over traces of in (nt x ntr) correlated with g (Conv.mod! with D() and S()).
 * gpi_conv_lsmr: s = argmin ||g * s - d|| by LSMR from zero (IterativeSolvers.lsmr! on
 * Conv.operator(pa, G()), maxiter, atol, btol, conlim as there), optionally gs = g * s;
 * *iters and *istop report the iterations and the stopping rule (1-7 as in LSMR). */
int gpi_conv_apply(gpi_ctx* ctx, int64_t nt, int64_t ntr, const double* g, int32_t mode, const double* in,
                   double* out);
int gpi_conv_lsmr(gpi_ctx* ctx, int64_t nt, int64_t ntr, const double* g, const double* d, int32_t maxiter,
                  double atol, double btol, double conlim, double* s, double* gs, int32_t* iters, int32_t* istop);

/* Device timing of the last gpi_forward on this context. */
int gpi_last_timing(const gpi_ctx* ctx, gpi_timing* out);

/* Bracket the velocity and stress passes of up to `nsamples` evenly spaced steps of each
 * following gpi_forward with HIP events (0 turns sampling off).  No reference counterpart:
 * it replaces the TimerOutputs sections of src/fdtd/propagate.jl:176-232 with device
 * time. */
int gpi_set_pass_sampling(gpi_ctx* ctx, int32_t nsamples);

#ifdef __cplusplus
}
#endif
#endif
