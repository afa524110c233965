/*
 * CPU oracle for the staggered-grid velocity-stress FDTD path of GeoPhyInv.jl.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, unfused and on the CPU, the reference's
 * forward-modelling algorithm so that the HIP product path can be checked against it and
 * so that bench.py can time a CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product library never links it.
 *
 * Every function follows the reference file:line named in its comment (paths relative to
 * the GeoPhyInv.jl repository root).  The pass structure of the reference is kept on
 * purpose: derivative arrays are materialised, CPML memory updates run as separate strip
 * loops, sources are sprayed into full-grid buffers and receivers are sampled through the
 * bilinear weights in CSC order.  Arithmetic is FP32 with the reference's evaluation order;
 * build with -ffp-contract=off so no fused multiply-adds are introduced.
 *
 * Unpinnable details (no Julia toolchain and no ParallelStencil source is available):
 *  - float literals inside ParallelStencil kernels (27.0, 0.5, 0.25, ...) are taken as
 *    FP32 (ParallelStencil converts kernel literals to the number type of
 *    @init_parallel_stencil); host-side broadcasts (medium.jl:165-167) keep FP64 promotion.
 *  - Statistics.mean over FP32 arrays (media.jl:26) is taken as the FP64 mean rounded to
 *    FP32; Julia's pairwise FP32 sum can differ in the last bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "fdtd_oracle.h"

/* ------------------------------------------------------------------------------------ */
/* Field tables: src/fields.jl:92-671 (staggered grids) and src/fdtd/diff{2,3}D.jl        */
/* ------------------------------------------------------------------------------------ */

/* per-axis grid code of a field, relative to the extended stress grid n:
 * N: n points at m[1];  A: n+h at m[1]-h/2 d;  H: n-h at m[1]+h/2 d;  I: n-2h at m[1]+h d   */
enum { GN = 1, GA, GH, GI };

/* grid codes (z, y, x) per field for [physics][ndim-2]; 0 = field absent */
static int field_code(int physics, int ndim, int f, int axis)
{
    /* acoustic: src/fields.jl:92-216 */
    static const char* ac2[O_NFIELDS] = {
        [O_P] = "N.N", [O_VX] = "N.A", [O_VZ] = "A.N", [O_DPDX] = "I.H", [O_DPDZ] = "H.I",
        [O_DVXDX] = "N.N", [O_DVZDZ] = "N.N"};
    static const char* ac3[O_NFIELDS] = {
        [O_P] = "NNN", [O_VX] = "NNA", [O_VY] = "NAN", [O_VZ] = "ANN", [O_DPDX] = "IIH",
        [O_DPDY] = "IHI", [O_DPDZ] = "HII", [O_DVXDX] = "NNN", [O_DVYDY] = "NNN",
        [O_DVZDZ] = "NNN"};
    /* elastic: src/fields.jl:222-671 */
    static const char* el2[O_NFIELDS] = {
        [O_TXX] = "N.N", [O_TZZ] = "N.N", [O_TXZ] = "H.H", [O_VX] = "N.A", [O_VZ] = "A.N",
        [O_DTXXDX] = "I.H", [O_DTZZDZ] = "H.I", [O_DTXZDX] = "H.I", [O_DTXZDZ] = "I.H",
        [O_DVXDX] = "N.N", [O_DVZDZ] = "N.N", [O_DVXDZ] = "H.H", [O_DVZDX] = "H.H"};
    static const char* el3[O_NFIELDS] = {
        [O_TXX] = "NNN", [O_TYY] = "NNN", [O_TZZ] = "NNN", [O_TXY] = "IHH", [O_TXZ] = "HIH",
        [O_TYZ] = "HHI", [O_VX] = "NNA", [O_VY] = "NAN", [O_VZ] = "ANN",
        [O_DTXXDX] = "IIH", [O_DTYYDY] = "IHI", [O_DTZZDZ] = "HII", [O_DTXYDX] = "IHI",
        [O_DTXYDY] = "IIH", [O_DTXZDX] = "HII", [O_DTXZDZ] = "IIH", [O_DTYZDY] = "HII",
        [O_DTYZDZ] = "IHI", [O_DVXDX] = "NNN", [O_DVYDY] = "NNN", [O_DVZDZ] = "NNN",
        [O_DVXDY] = "IHH", [O_DVXDZ] = "HIH", [O_DVYDX] = "IHH", [O_DVYDZ] = "HHI",
        [O_DVZDX] = "HIH", [O_DVZDY] = "HHI"};
    const char* s;
    if (f < 0 || f >= O_NFIELDS) return 0;
    if (physics == 0) s = (ndim == 2) ? ac2[f] : ac3[f];
    else s = (ndim == 2) ? el2[f] : el3[f];
    if (!s) return 0;
    if (ndim == 2 && axis == 1) return GN; /* absent y axis: size 1, offset 0 */
    switch (s[axis]) {
    case 'N': return GN;
    case 'A': return GA;
    case 'H': return GH;
    case 'I': return GI;
    }
    return 0;
}

int oracle_field_shape(int physics, int ndim, int order, const int64_t n[3], int f, int64_t out[3])
{
    int h = order - 1;
    for (int a = 0; a < 3; ++a) {
        int c = field_code(physics, ndim, f, a);
        if (!c) return -1;
        if (ndim == 2 && a == 1) { out[a] = 1; continue; }
        out[a] = n[a] + (c == GA ? h : c == GH ? -h : c == GI ? -2 * h : 0);
    }
    return 0;
}

/* origin of a field's grid along an axis, in units of the spacing (src/fields.jl get_mgrid) */
static double field_origin_units(int code, int h)
{
    switch (code) {
    case GA: return -0.5 * h;
    case GH: return 0.5 * h;
    case GI: return (double)h;
    default: return 0.0;
    }
}

/* derivative definitions: source field, axis (0 z, 1 y, 2 x), kind 'a' (no shift of the
 * other axes) or 'i' (other axes shifted by h).  src/fdtd/advance_acou.jl:258-299,
 * src/fdtd/advance_elastic.jl:10-153 */
typedef struct { int src, axis; char kind; } dspec_t;
static dspec_t dspec(int f)
{
    switch (f) {
    case O_DPDX: return (dspec_t){O_P, 2, 'i'};
    case O_DPDY: return (dspec_t){O_P, 1, 'i'};
    case O_DPDZ: return (dspec_t){O_P, 0, 'i'};
    case O_DVXDX: return (dspec_t){O_VX, 2, 'a'};
    case O_DVYDY: return (dspec_t){O_VY, 1, 'a'};
    case O_DVZDZ: return (dspec_t){O_VZ, 0, 'a'};
    case O_DVXDY: return (dspec_t){O_VX, 1, 'i'};
    case O_DVXDZ: return (dspec_t){O_VX, 0, 'i'};
    case O_DVYDX: return (dspec_t){O_VY, 2, 'i'};
    case O_DVYDZ: return (dspec_t){O_VY, 0, 'i'};
    case O_DVZDX: return (dspec_t){O_VZ, 2, 'i'};
    case O_DVZDY: return (dspec_t){O_VZ, 1, 'i'};
    case O_DTXXDX: return (dspec_t){O_TXX, 2, 'i'};
    case O_DTYYDY: return (dspec_t){O_TYY, 1, 'i'};
    case O_DTZZDZ: return (dspec_t){O_TZZ, 0, 'i'};
    case O_DTXYDX: return (dspec_t){O_TXY, 2, 'a'};
    case O_DTXYDY: return (dspec_t){O_TXY, 1, 'a'};
    case O_DTXZDX: return (dspec_t){O_TXZ, 2, 'a'};
    case O_DTXZDZ: return (dspec_t){O_TXZ, 0, 'a'};
    case O_DTYZDY: return (dspec_t){O_TYZ, 1, 'a'};
    case O_DTYZDZ: return (dspec_t){O_TYZ, 0, 'a'};
    }
    return (dspec_t){-1, -1, 0};
}

static int is_deriv(int f) { return f >= O_DPDX && f < O_NFIELDS; }

/* ------------------------------------------------------------------------------------ */
/* state                                                                                 */
/* ------------------------------------------------------------------------------------ */

typedef struct { int64_t n[3]; float* a; } arr_t;

static inline int64_t IX(const arr_t* A, int64_t i, int64_t j, int64_t k)
{
    return i + A->n[0] * (j + A->n[1] * k);
}
static inline int64_t numel(const arr_t* A) { return A->n[0] * A->n[1] * A->n[2]; }

typedef struct { int npts, nnz; int64_t* idx; float* w; } sparse_t; /* CSC columns */

struct oracle {
    oracle_expt e;          /* copy of the description (pointers borrowed during create) */
    int h, o2;
    int64_t n[3];           /* extended grid */
    double m0[3];           /* extended grid origin */
    double d[3];
    float dt, dsI[3];
    float freqpeak, velavg;
    arr_t f[O_NFIELDS];     /* wavefields and derivative arrays */
    arr_t mem[O_NFIELDS];   /* CPML memory of derivative fields */
    float* pml_a[O_NFIELDS];
    float* pml_b[O_NFIELDS];
    float* pml_kI[O_NFIELDS];
    arr_t coef[O_NCOEF];
    arr_t vp, vs, rho;      /* padded medium */
    sparse_t spray[O_MAXSF];
    sparse_t gath[O_MAXRF];
    float* wav[O_MAXSF];    /* nt x ns per source field, it fastest (Srcs.d layout) */
    float* rec[O_MAXRF];    /* nt x nr per receiver field */
    float* buf;             /* spray buffer (largest field) */
    int64_t it_done;
    /* time-reversal store of :forward_save (boundary.jl): per stored stress field and
     * axis, nt strips of 2*nbound planes; snapshots of the stress and velocity fields */
    int nstore;
    int store_f[3];
    arr_t store[3][3];      /* [field][axis]: dims of the field with the axis = 2*nbound, x nt */
    arr_t snap[O_NFIELDS];
    arr_t tp[O_NFIELDS];    /* save_tp! copies (adjoint) */
};

static int present(const oracle_t* S, int f)
{
    return field_code(S->e.physics, S->e.ndim, f, 0) != 0;
}

static void alloc_field(oracle_t* S, arr_t* A, int f)
{
    oracle_field_shape(S->e.physics, S->e.ndim, S->e.order, S->n, f, A->n);
    A->a = calloc((size_t)numel(A), sizeof(float));
}

/* ------------------------------------------------------------------------------------ */
/* host-side preparation                                                                 */
/* ------------------------------------------------------------------------------------ */

/* replicate padding by npml on the faces that carry a PML: src/media/media.jl:260-304 */
static void pad_medium(const oracle_t* S, const float* m, arr_t* out)
{
    const oracle_expt* e = &S->e;
    int64_t lo[3], nm[3];
    for (int a = 0; a < 3; ++a) {
        nm[a] = e->nm[a];
        lo[a] = (e->ndim == 2 && a == 1) ? 0 : ((e->pml_mask >> (2 * a)) & 1) ? e->npml : 0;
    }
    out->n[0] = S->n[0]; out->n[1] = S->n[1]; out->n[2] = S->n[2];
    out->a = malloc(sizeof(float) * (size_t)numel(out));
    for (int64_t k = 0; k < out->n[2]; ++k)
        for (int64_t j = 0; j < out->n[1]; ++j)
            for (int64_t i = 0; i < out->n[0]; ++i) {
                int64_t s[3] = {i - lo[0], j - lo[1], k - lo[2]};
                for (int a = 0; a < 3; ++a) {
                    if (s[a] < 0) s[a] = 0;
                    if (s[a] > nm[a] - 1) s[a] = nm[a] - 1;
                }
                out->a[IX(out, i, j, k)] = m[s[0] + nm[0] * (s[1] + nm[1] * s[2])];
            }
}

/* Statistics.mean / bounds of a medium parameter: src/media/media.jl:24-32 */
static void ref_bounds(const arr_t* A, float* ref, float b[2])
{
    double s = 0.0;
    float mn = A->a[0], mx = A->a[0];
    int64_t N = numel(A);
    for (int64_t i = 0; i < N; ++i) {
        s += A->a[i];
        if (A->a[i] < mn) mn = A->a[i];
        if (A->a[i] > mx) mx = A->a[i];
    }
    *ref = (float)(s / (double)N);
    float lo = mn - (float)0.1 * *ref;   /* convert(eltype(m), frac) * ref */
    b[0] = lo > 0.0f ? lo : 0.0f;
    b[1] = mx + (float)0.1 * *ref;
}

/* finite-difference coefficient prep: src/fdtd/medium.jl:143-221, media.jl:111-130 */
static void prep_coefficients(oracle_t* S)
{
    const int h = S->h, nd = S->e.ndim;
    const float dt = S->dt;
    arr_t* rho = &S->rho;
    int64_t N = numel(rho);
    /* independent parameters (mod), FP32 as the reference's broadcasts on Float32 arrays */
    if (S->e.physics == 0) {
        arr_t* invK = &S->coef[O_C_INVK];
        *invK = (arr_t){{rho->n[0], rho->n[1], rho->n[2]}, malloc(sizeof(float) * N)};
        for (int64_t i = 0; i < N; ++i) {
            float vp = S->vp.a[i];
            invK->a[i] = 1.0f / ((vp * vp) * rho->a[i]); /* invK(vp,rho)=inv(abs2(vp)*rho) */
        }
        arr_t* dtK = &S->coef[O_C_DTK];
        alloc_field(S, dtK, O_P);
        for (int64_t i = 0; i < N; ++i) dtK->a[i] = (1.0f / invK->a[i]) * dt; /* :147-148 */
    } else {
        arr_t* il = &S->coef[O_C_INVLAMBDA];
        arr_t* im = &S->coef[O_C_INVMU];
        *il = (arr_t){{rho->n[0], rho->n[1], rho->n[2]}, malloc(sizeof(float) * N)};
        *im = (arr_t){{rho->n[0], rho->n[1], rho->n[2]}, malloc(sizeof(float) * N)};
        for (int64_t i = 0; i < N; ++i) {
            float vp = S->vp.a[i], vs = S->vs.a[i], r = S->rho.a[i];
            il->a[i] = 1.0f / (((vp * vp) - 2.0f * (vs * vs)) * r);
            im->a[i] = 1.0f / ((vs * vs) * r);
        }
        arr_t* dtl = &S->coef[O_C_DTLAMBDA];
        arr_t* dtM = &S->coef[O_C_DTM];
        alloc_field(S, dtl, O_TXX);
        alloc_field(S, dtM, O_TXX);
        for (int64_t i = 0; i < N; ++i) {
            dtl->a[i] = (1.0f / il->a[i]) * dt;                       /* :163-164 */
            /* inv(invλ) + 2.0*inv(invμ) is promoted to Float64 (literal 2.0), stored F32 */
            double m = (double)(1.0f / il->a[i]) + 2.0 * (double)(1.0f / im->a[i]);
            dtM->a[i] = ((float)m) * dt;                              /* :165-168 */
        }
    }
    /* dt / @av_{x,y,z}i(rho) on the velocity-update grids: medium.jl:191-204 */
    int cfx = S->e.physics == 0 ? O_DPDX : O_DTXXDX;
    int cfy = S->e.physics == 0 ? O_DPDY : O_DTYYDY;
    int cfz = S->e.physics == 0 ? O_DPDZ : O_DTZZDZ;
    int ids[3] = {O_C_DTINVAVZIRHO, O_C_DTINVAVYIRHO, O_C_DTINVAVXIRHO};
    int grids[3] = {cfz, cfy, cfx};
    for (int ax = 0; ax < 3; ++ax) {
        if (nd == 2 && ax == 1) continue;
        arr_t* C = &S->coef[ids[ax]];
        alloc_field(S, C, grids[ax]);
        for (int64_t k = 0; k < C->n[2]; ++k)
            for (int64_t j = 0; j < C->n[1]; ++j)
                for (int64_t i = 0; i < C->n[0]; ++i) {
                    int64_t b[3] = {i + h, nd == 3 ? j + h : 0, k + h};
                    b[ax] -= h; /* the averaged axis is not shifted */
                    int64_t c[3] = {b[0], b[1], b[2]};
                    c[ax] += 1;
                    float av = (rho->a[IX(rho, b[0], b[1], b[2])] + rho->a[IX(rho, c[0], c[1], c[2])]) * 0.5f;
                    C->a[IX(C, i, j, k)] = dt / av;
                }
    }
    if (S->e.physics == 1) {
        arr_t* im = &S->coef[O_C_INVMU];
        if (nd == 2) {
            /* store_invav!: dt / @av(invmu), diff2D.jl:222-225 */
            arr_t* C = &S->coef[O_C_DTAVMU];
            alloc_field(S, C, O_DVXDZ);
            for (int64_t k = 0; k < C->n[2]; ++k)
                for (int64_t i = 0; i < C->n[0]; ++i) {
                    float av = (((im->a[IX(im, i, 0, k)] + im->a[IX(im, i + 1, 0, k)]) +
                                 im->a[IX(im, i, 0, k + 1)]) + im->a[IX(im, i + 1, 0, k + 1)]) * 0.25f;
                    C->a[IX(C, i, 0, k)] = dt / av;
                }
        } else {
            /* @av_xzi, @av_xyi, @av_yzi: diff3D.jl:331-378 */
            struct { int id, grid, a0, a1; } sh[3] = {
                {O_C_DTAVXZIMU, O_DVXDZ, 0, 2}, {O_C_DTAVXYIMU, O_DVXDY, 1, 2},
                {O_C_DTAVYZIMU, O_DVYDZ, 0, 1}};
            for (int s = 0; s < 3; ++s) {
                arr_t* C = &S->coef[sh[s].id];
                alloc_field(S, C, sh[s].grid);
                int a0 = sh[s].a0, a1 = sh[s].a1;
                for (int64_t k = 0; k < C->n[2]; ++k)
                    for (int64_t j = 0; j < C->n[1]; ++j)
                        for (int64_t i = 0; i < C->n[0]; ++i) {
                            int64_t b[3] = {i + h, j + h, k + h};
                            b[a0] -= h; b[a1] -= h;
                            int64_t q[4][3];
                            for (int t = 0; t < 4; ++t) {
                                q[t][0] = b[0]; q[t][1] = b[1]; q[t][2] = b[2];
                            }
                            q[1][a0] += 1;                       /* (a0+1, a1)   */
                            q[2][a1] += 1;                       /* (a0, a1+1)   */
                            q[3][a0] += 1; q[3][a1] += 1;        /* (a0+1, a1+1) */
                            float av = (((im->a[IX(im, q[0][0], q[0][1], q[0][2])] +
                                          im->a[IX(im, q[1][0], q[1][1], q[1][2])]) +
                                         im->a[IX(im, q[2][0], q[2][1], q[2][2])]) +
                                        im->a[IX(im, q[3][0], q[3][1], q[3][2])]) * 0.25f;
                            C->a[IX(C, i, j, k)] = dt / av;
                        }
            }
        }
    }
}

/* CPML profile on one derivative field's own 1-D grid: src/fdtd/cpml.jl:8-103 */
static void pml_profile(const oracle_t* S, int f, int axis, float* pa, float* pb, float* pk)
{
    const int npml = S->e.npml, h = S->h;
    const double ds = S->d[axis];
    int code = field_code(S->e.physics, S->e.ndim, f, axis);
    int64_t nf = S->n[axis] + (code == GA ? h : code == GH ? -h : code == GI ? -2 * h : 0);
    double x0 = S->m0[axis] + field_origin_units(code, h) * ds;
    double mg0 = S->e.o[axis], mgend = S->e.o[axis] + (double)(S->e.nm[axis] - 1) * ds;
    int fmin = (S->e.pml_mask >> (2 * axis)) & 1, fmax = (S->e.pml_mask >> (2 * axis + 1)) & 1;
    double xl = mg0 - 2.0 * ds, xr = mgend + 2.0 * ds;
    const double NPOWER = 2.0, K_MAX = 1.0, ALPHA_MAX = M_PI * (double)S->freqpeak;
    double thick = (double)(npml - 1) * ds;
    const double R = 0.001;
    double d0 = -(NPOWER + 1.0) * (double)S->velavg * log(R) / (2.0 * thick);
    double* k = malloc(sizeof(double) * nf);
    double* a = calloc(nf, sizeof(double));
    double* b = calloc(nf, sizeof(double));
    for (int64_t ix = 0; ix < nf; ++ix) {
        double x = x0 + (double)ix * ds, d = 0.0, al = 0.0;
        k[ix] = 1.0;
        if (fmin) {
            double ab = xl - x;
            if (ab >= 0.0) {
                double an = ab / thick;
                d = d0 * pow(an, NPOWER);
                k[ix] = 1.0 + (K_MAX - 1.0) * pow(an, NPOWER);
                al = ALPHA_MAX * (1.0 - an) + 0.01 * ALPHA_MAX;
            }
        }
        if (fmax) {
            double ab = x - xr;
            if (ab >= 0.0) {
                double an = ab / thick;
                d = d0 * pow(an, NPOWER);
                k[ix] = 1.0 + (K_MAX - 1.0) * pow(an, NPOWER);
                al = ALPHA_MAX * (1.0 - an) + 0.01 * ALPHA_MAX;
            }
        }
        if (al < 0.0) al = 0.0;
        b[ix] = exp(-(d / k[ix] + al) * (double)S->dt);
        if (fabs(d) > 1e-6) a[ix] = d * (b[ix] - 1.0) / (k[ix] * (d + k[ix] * al));
    }
    for (int i = 0; i < npml; ++i) {
        int j = i + npml;
        pk[i] = (float)(1.0 / k[i]);
        pk[j] = (float)(1.0 / k[nf - npml + i]);
        pa[i] = (float)a[i];
        pa[j] = (float)a[nf - npml + i];
        pb[i] = (float)b[i];
        pb[j] = (float)b[nf - npml + i];
    }
    free(k); free(a); free(b);
}

/* searchsortedfirst clamp of src/proj_mat.jl:69-79 on a uniform grid g0 + i*dg */
static void neighbours(double g0, double dg, int64_t len, double v, int64_t* i1, int64_t* i2)
{
    int64_t idx = 0; /* 0-based first index with g[idx] >= v */
    while (idx < len && g0 + (double)idx * dg < v) ++idx;
    if (idx == 0) { *i1 = 0; *i2 = 1; }
    else if (idx >= len - 1) { *i1 = len - 2; *i2 = len - 1; }
    else { *i1 = idx - 1; *i2 = idx; }
}

/* bilinear / trilinear weights of one point on field f: src/proj_mat.jl:85-177, 229-247,
 * src/fdtd/ageom.jl:19-54.  Emits (index, weight) sorted by index (CSC column order).
 * In 3-D the reference calls bilinear_interp(mz,my,mx, Pz,Py,Px) whose signature is
 * (x,y,z, zi,yi,xi): the z coordinate is searched in the x grid and vice versa.  compat
 * bit O_COMPAT_SWAP3D reproduces that; otherwise each coordinate uses its own axis. */
static void point_weights(const oracle_t* S, int f, const double* pos, int64_t* idx, float* w)
{
    const int h = S->h, nd = S->e.ndim;
    int64_t nf[3], lo[3], hi[3];
    double g0[3], frac[3];
    oracle_field_shape(S->e.physics, nd, S->e.order, S->n, f, nf);
    for (int a = 0; a < 3; ++a) {
        int c = field_code(S->e.physics, nd, f, a);
        g0[a] = S->m0[a] + field_origin_units(c, h) * S->d[a];
    }
    double p[3] = {pos[0], nd == 3 ? pos[1] : 0.0, nd == 3 ? pos[2] : pos[1]};
    int axes[3] = {0, 1, 2};
    if (nd == 3 && (S->e.compat & O_COMPAT_SWAP3D)) { double t = p[0]; p[0] = p[2]; p[2] = t; }
    for (int a = 0; a < 3; ++a) {
        int ax = axes[a];
        if (nd == 2 && ax == 1) { lo[ax] = hi[ax] = 0; frac[ax] = 0.0; continue; }
        neighbours(g0[ax], S->d[ax], nf[ax], p[ax], &lo[ax], &hi[ax]);
        double gmin = g0[ax], gmax = g0[ax] + (double)(nf[ax] - 1) * S->d[ax];
        double x1 = g0[ax] + (double)lo[ax] * S->d[ax], x2 = g0[ax] + (double)hi[ax] * S->d[ax];
        if (p[ax] < gmin) frac[ax] = 0.0;
        else if (p[ax] > gmax) frac[ax] = 1.0;
        else frac[ax] = (p[ax] - x1) / (x2 - x1);
    }
    if (nd == 2) {
        /* weights (1-dz)(1-dx), (1-dz)dx, dz(1-dx), dz dx on (z1,x1),(z1,x2),(z2,x1),(z2,x2) */
        double dz = frac[0], dx = frac[2];
        int64_t ii[4] = {lo[0] + nf[0] * lo[2], lo[0] + nf[0] * hi[2], hi[0] + nf[0] * lo[2],
                         hi[0] + nf[0] * hi[2]};
        double ww[4] = {(1 - dz) * (1 - dx), (1 - dz) * dx, dz * (1 - dx), dz * dx};
        /* sparsevec sorts by index: z1x1, z2x1, z1x2, z2x2 */
        int ord[4] = {0, 2, 1, 3};
        for (int t = 0; t < 4; ++t) { idx[t] = ii[ord[t]]; w[t] = (float)ww[ord[t]]; }
    } else {
        double fz = frac[0], fy = frac[1], fx = frac[2];
        int t = 0;
        /* ascending linear index: z fastest, then y, then x */
        for (int bx = 0; bx < 2; ++bx)
            for (int by = 0; by < 2; ++by)
                for (int bz = 0; bz < 2; ++bz) {
                    int64_t iz = bz ? hi[0] : lo[0], iy = by ? hi[1] : lo[1], ix = bx ? hi[2] : lo[2];
                    idx[t] = iz + nf[0] * (iy + nf[1] * ix);
                    /* weight product order of proj_mat.jl:121-128: (x-term)*(y-term)*(z-term)
                     * where the reference's x is the first grid argument (our z) */
                    double wz = bz ? fz : (1 - fz), wy = by ? fy : (1 - fy), wx = bx ? fx : (1 - fx);
                    w[t] = (float)((wz * wy) * wx);
                    ++t;
                }
    }
}

/* direct DFT power-spectrum argmax of findfreq(:peak): src/Utils/freq.jl:28-57 */
static double findfreq_peak(const float* w, int64_t nt, int ns, double dt)
{
    int64_t nf = nt / 2 + 1;
    double best = -1.0;
    int64_t kbest = 0;
    int any = 0;
    for (int s = 0; s < ns; ++s)
        for (int64_t k = 0; k < nf; ++k) {
            double re = 0.0, im = 0.0;
            for (int64_t t = 0; t < nt; ++t) {
                double ang = -2.0 * M_PI * (double)((k * t) % nt) / (double)nt;
                re += (double)w[t + nt * s] * cos(ang);
                im += (double)w[t + nt * s] * sin(ang);
            }
            double p = re * re + im * im;
            if (p != 0.0) any = 1;
            if (p > best) { best = p; kbest = k; }
        }
    if (!any) return NAN;
    return (double)kbest * ((1.0 / dt) / (double)nt);
}

/* ------------------------------------------------------------------------------------ */
/* time-step kernels                                                                     */
/* ------------------------------------------------------------------------------------ */

/* staggered difference with the reference's literal order: diff2D.jl:15-212 */
static inline float stencil(int order, const float* v)
{
    switch (order) {
    case 2: return v[1] - v[0];
    case 4: return ((v[2] * 27.0f - v[1] * 27.0f) + v[0]) - v[3];
    case 6: return ((((v[3] * 2250.0f - v[2] * 2250.0f) + v[1] * 125.0f) - v[4] * 125.0f) + v[5] * 9.0f) - v[0] * 9.0f;
    default:
        return ((((((v[4] * 128625.0f - v[3] * 128625.0f) + v[2] * 8575.0f) - v[5] * 8575.0f) +
                  v[6] * 1029.0f) - v[1] * 1029.0f) + v[0] * 75.0f) - v[7] * 75.0f;
    }
}

/* @all(D) = @d_<axis><kind>(A) * dsI : advance_acou.jl:258-299, advance_elastic.jl:10-153 */
static void compute_derivative(oracle_t* S, int df)
{
    dspec_t sp = dspec(df);
    arr_t* D = &S->f[df];
    const arr_t* A = &S->f[sp.src];
    const int h = S->h, o = S->e.order, nd = S->e.ndim;
    const float sI = S->dsI[sp.axis];
    int64_t sh[3] = {0, 0, 0};
    if (sp.kind == 'i')
        for (int a = 0; a < 3; ++a)
            if (a != sp.axis && !(nd == 2 && a == 1)) sh[a] = h;
    int64_t st = sp.axis == 0 ? 1 : sp.axis == 1 ? A->n[0] : A->n[0] * A->n[1];
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < D->n[2]; ++k)
        for (int64_t j = 0; j < D->n[1]; ++j)
            for (int64_t i = 0; i < D->n[0]; ++i) {
                float v[8];
                int64_t base = IX(A, i + sh[0], j + sh[1], k + sh[2]);
                for (int t = 0; t < o; ++t) v[t] = A->a[base + t * st];
                D->a[IX(D, i, j, k)] = stencil(o, v) * sI;
            }
}

/* memory{x,y,z}! on the first/last npml points of a derivative: cpml.jl:170-212 */
static void cpml_update(oracle_t* S, int df)
{
    dspec_t sp = dspec(df);
    const int axis = sp.axis, npml = S->e.npml;
    arr_t* D = &S->f[df];
    arr_t* M = &S->mem[df];
    const float *a = S->pml_a[df], *b = S->pml_b[df], *kI = S->pml_kI[df];
    for (int face = 0; face < 2; ++face) {
        if (!((S->e.pml_mask >> (2 * axis + face)) & 1)) continue;
        int64_t moff = face ? npml : 0, doff = face ? D->n[axis] - npml : 0;
        int64_t rng[3] = {M->n[0], M->n[1], M->n[2]};
        rng[axis] = npml;
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < rng[2]; ++k)
            for (int64_t j = 0; j < rng[1]; ++j)
                for (int64_t i = 0; i < rng[0]; ++i) {
                    int64_t mi[3] = {i, j, k}, di[3] = {i, j, k};
                    mi[axis] += moff;
                    di[axis] += doff;
                    int64_t c = (axis == 0 ? i : axis == 1 ? j : k) + moff;
                    int64_t pm = IX(M, mi[0], mi[1], mi[2]), pd = IX(D, di[0], di[1], di[2]);
                    M->a[pm] = b[c] * M->a[pm] + a[c] * D->a[pd];
                    D->a[pd] = D->a[pd] * kI[c] + M->a[pm];
                }
    }
}

/* v[inn] (+|-)= coef * (sum of derivatives): advance_acou.jl:273-284, advance_elastic.jl:69-110 */
static void update_velocity(oracle_t* S)
{
    const int h = S->h, nd = S->e.ndim;
    const int vel[3] = {O_VZ, O_VY, O_VX};
    const int cid[3] = {O_C_DTINVAVZIRHO, O_C_DTINVAVYIRHO, O_C_DTINVAVXIRHO};
    for (int ax = 0; ax < 3; ++ax) {
        if (nd == 2 && ax == 1) continue;
        arr_t* V = &S->f[vel[ax]];
        const arr_t* C = &S->coef[cid[ax]];
        int d1, d2 = -1, d3 = -1;
        if (S->e.physics == 0) {
            d1 = ax == 0 ? O_DPDZ : ax == 1 ? O_DPDY : O_DPDX;
        } else if (nd == 2) {
            /* vx -= c*(dtauxxdx + dtauxzdz); vz -= c*(dtauxzdx + dtauzzdz) */
            if (ax == 2) { d1 = O_DTXXDX; d2 = O_DTXZDZ; }
            else { d1 = O_DTXZDX; d2 = O_DTZZDZ; }
        } else {
            if (ax == 2) { d1 = O_DTXXDX; d2 = O_DTXYDY; d3 = O_DTXZDZ; }
            else if (ax == 1) { d1 = O_DTXYDX; d2 = O_DTYYDY; d3 = O_DTYZDZ; }
            else { d1 = O_DTXZDX; d2 = O_DTYZDY; d3 = O_DTZZDZ; }
        }
        const arr_t* D1 = &S->f[d1];
        int64_t shy = nd == 3 ? h : 0;
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < C->n[2]; ++k)
            for (int64_t j = 0; j < C->n[1]; ++j)
                for (int64_t i = 0; i < C->n[0]; ++i) {
                    int64_t pv = IX(V, i + h, j + shy, k + h), pc = IX(C, i, j, k);
                    float s;
                    if (S->e.physics == 0) {
                        V->a[pv] = V->a[pv] + C->a[pc] * D1->a[pc];
                        continue;
                    }
                    if (d3 < 0) s = D1->a[pc] + S->f[d2].a[pc];
                    else s = (D1->a[pc] + S->f[d2].a[pc]) + S->f[d3].a[pc];
                    V->a[pv] = V->a[pv] - C->a[pc] * s;
                }
    }
}

/* rigid faces, applied in the order xmin, xmax, (ymin, ymax,) zmin, zmax:
 * advance_acou.jl:50-57,155-192, dirichlet.jl:3-78 */
static void dirichlet(oracle_t* S)
{
    const int nd = S->e.ndim, o = S->e.order, fdh = o / 2;
    const int vel[3] = {O_VZ, O_VY, O_VX};
    const int axes[3] = {2, 1, 0};
    for (int q = 0; q < 3; ++q) {
        int ax = axes[q];
        if (nd == 2 && ax == 1) continue;
        for (int face = 0; face < 2; ++face) {
            if (!((S->e.rigid_mask >> (2 * ax + face)) & 1)) continue;
            arr_t* V = &S->f[vel[ax]];
            int64_t n = S->n[ax];
            /* irest runs over the stress-grid sizes of the other axes */
            int o1 = -1, o2 = -1;
            for (int a = 0; a < 3; ++a)
                if (a != ax && !(nd == 2 && a == 1)) { if (o1 < 0) o1 = a; else o2 = a; }
            int64_t r1 = S->n[o1], r2 = o2 >= 0 ? S->n[o2] : 1;
            for (int64_t p2 = 0; p2 < r2; ++p2)
                for (int64_t p1 = 0; p1 < r1; ++p1) {
                    int64_t id[3] = {0, 0, 0};
                    id[o1] = p1;
                    if (o2 >= 0) id[o2] = p2;
                    /* tangential velocities on the face */
                    for (int b = 0; b < 3; ++b) {
                        if (b == ax || (nd == 2 && b == 1)) continue;
                        arr_t* T = &S->f[vel[b]];
                        int64_t t[3] = {id[0], id[1], id[2]};
                        t[ax] = face ? n - 1 : 0;
                        T->a[IX(T, t[0], t[1], t[2])] = 0.0f;
                    }
                    /* ghost mirror of the normal velocity */
                    for (int ifd = 1; ifd <= fdh; ++ifd) {
                        int64_t g[3] = {id[0], id[1], id[2]}, s[3] = {id[0], id[1], id[2]};
                        if (!face) { g[ax] = ifd - 1; s[ax] = o - ifd; }
                        else { g[ax] = n + o - ifd - 1; s[ax] = n + ifd - 2; }
                        V->a[IX(V, g[0], g[1], g[2])] = -V->a[IX(V, s[0], s[1], s[2])];
                    }
                }
        }
    }
}

/* stress updates: advance_acou.jl:303-313, advance_elastic.jl:155-230 */
static void update_stress(oracle_t* S)
{
    const int nd = S->e.ndim;
    if (S->e.physics == 0) {
        arr_t* P = &S->f[O_P];
        const float* dtK = S->coef[O_C_DTK].a;
        const float *dx = S->f[O_DVXDX].a, *dz = S->f[O_DVZDZ].a;
        const float* dy = nd == 3 ? S->f[O_DVYDY].a : NULL;
        int64_t N = numel(P);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            if (nd == 2) P->a[i] = P->a[i] + (dx[i] + dz[i]) * dtK[i];
            else P->a[i] = P->a[i] + ((dx[i] + dz[i]) + dy[i]) * dtK[i];
        }
        return;
    }
    const float* dtM = S->coef[O_C_DTM].a;
    const float* dtl = S->coef[O_C_DTLAMBDA].a;
    float *txx = S->f[O_TXX].a, *tzz = S->f[O_TZZ].a;
    const float *vxx = S->f[O_DVXDX].a, *vzz = S->f[O_DVZDZ].a;
    int64_t N = numel(&S->f[O_TXX]);
    if (nd == 2) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            txx[i] = (txx[i] - (dtM[i] * vxx[i])) - (dtl[i] * vzz[i]);
            tzz[i] = (tzz[i] - (dtM[i] * vzz[i])) - (dtl[i] * vxx[i]);
        }
        float* txz = S->f[O_TXZ].a;
        const float *c = S->coef[O_C_DTAVMU].a, *a1 = S->f[O_DVXDZ].a, *a2 = S->f[O_DVZDX].a;
        int64_t M = numel(&S->f[O_TXZ]);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < M; ++i) txz[i] = txz[i] - c[i] * (a1[i] + a2[i]);
    } else {
        float* tyy = S->f[O_TYY].a;
        const float* vyy = S->f[O_DVYDY].a;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            float x = txx[i], y = tyy[i], z = tzz[i];
            txx[i] = (x - (dtM[i] * vxx[i])) - (dtl[i] * (vyy[i] + vzz[i]));
            tyy[i] = (y - (dtM[i] * vyy[i])) - (dtl[i] * (vxx[i] + vzz[i]));
            tzz[i] = (z - (dtM[i] * vzz[i])) - (dtl[i] * (vyy[i] + vxx[i]));
        }
        struct { int t, c, a, b; } sh[3] = {{O_TXZ, O_C_DTAVXZIMU, O_DVXDZ, O_DVZDX},
                                            {O_TXY, O_C_DTAVXYIMU, O_DVXDY, O_DVYDX},
                                            {O_TYZ, O_C_DTAVYZIMU, O_DVYDZ, O_DVZDY}};
        for (int s = 0; s < 3; ++s) {
            float* t = S->f[sh[s].t].a;
            const float *c = S->coef[sh[s].c].a, *a1 = S->f[sh[s].a].a, *a2 = S->f[sh[s].b].a;
            int64_t M = numel(&S->f[sh[s].t]);
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < M; ++i) t[i] = t[i] - c[i] * (a1[i] + a2[i]);
        }
    }
    if (S->e.free_mask & 1) { /* zmin only: advance_elastic.jl:519-529, 677-680 */
        arr_t* Z = &S->f[O_TZZ];
        for (int64_t k = 0; k < Z->n[2]; ++k)
            for (int64_t j = 0; j < Z->n[1]; ++j)
                Z->a[IX(Z, 0, j, k)] = -Z->a[IX(Z, 1, j, k)];
        int sf[2] = {O_TXZ, O_TYZ};
        for (int s = 0; s < (nd == 3 ? 2 : 1); ++s) {
            arr_t* T = &S->f[sf[s]];
            for (int64_t k = 0; k < T->n[2]; ++k)
                for (int64_t j = 0; j < T->n[1]; ++j) T->a[IX(T, 0, j, k)] = 0.0f;
        }
    }
}

/* mul!(buf, W, w): CSC spray in column order (SparseArrays), into a zeroed full grid */
static void spray(const sparse_t* W, const float* wt, float* buf, int64_t n)
{
    memset(buf, 0, sizeof(float) * (size_t)n);
    for (int s = 0; s < W->npts; ++s) {
        float x = wt[s];
        for (int t = 0; t < W->nnz; ++t) {
            int64_t r = W->idx[s * W->nnz + t];
            buf[r] += W->w[s * W->nnz + t] * x;
        }
    }
}

/* add_velocity_source!: source.jl:124-177 (ipw = 1 only) */
static void add_velocity_source(oracle_t* S, int64_t it)
{
    const int h = S->h, nd = S->e.ndim;
    const arr_t* rho = &S->rho;
    for (int q = 0; q < S->e.nsf; ++q) {
        int sf = S->e.sfields[q];
        if (sf != O_VX && sf != O_VY && sf != O_VZ) continue;
        arr_t* V = &S->f[sf];
        int64_t N = numel(V);
        float wt[O_MAXPTS_STACK];
        float* wv = S->e.ns <= O_MAXPTS_STACK ? wt : malloc(sizeof(float) * S->e.ns);
        for (int s = 0; s < S->e.ns; ++s) wv[s] = S->wav[q][it + S->e.nt * s];
        spray(&S->spray[q], wv, S->buf, N);
        if (wv != wt) free(wv);
        int ax = sf == O_VZ ? 0 : sf == O_VY ? 1 : 2;
        int64_t m[3] = {V->n[0] - 2 * h, nd == 3 ? V->n[1] - 2 * h : 1, V->n[2] - 2 * h};
        int64_t shy = nd == 3 ? h : 0;
        for (int64_t k = 0; k < m[2]; ++k)
            for (int64_t j = 0; j < m[1]; ++j)
                for (int64_t i = 0; i < m[0]; ++i) {
                    int64_t b[3] = {i + h, j + shy, k + h};
                    b[ax] -= h;
                    int64_t c[3] = {b[0], b[1], b[2]};
                    c[ax] += 1;
                    float av = (rho->a[IX(rho, b[0], b[1], b[2])] + rho->a[IX(rho, c[0], c[1], c[2])]) * 0.5f;
                    int64_t p = IX(V, i + h, j + shy, k + h);
                    V->a[p] = V->a[p] + ((S->buf[p] / av) * S->dt);
                }
    }
}

/* add_stress_source!: source.jl:61-119 (ipw = 1 only) */
static void add_stress_source(oracle_t* S, int64_t it)
{
    for (int q = 0; q < S->e.nsf; ++q) {
        int sf = S->e.sfields[q];
        if (!(sf == O_P || sf == O_TXX || sf == O_TYY || sf == O_TZZ)) continue;
        int tgt = S->e.physics == 0 ? O_P : O_TXX;
        int64_t N = numel(&S->f[tgt]);
        float wt[O_MAXPTS_STACK];
        float* wv = S->e.ns <= O_MAXPTS_STACK ? wt : malloc(sizeof(float) * S->e.ns);
        for (int s = 0; s < S->e.ns; ++s) wv[s] = S->wav[q][it + S->e.nt * s];
        spray(&S->spray[q], wv, S->buf, N);
        if (wv != wt) free(wv);
        if (S->e.physics == 0) {
            float* p = S->f[O_P].a;
            const float* c = S->coef[O_C_DTK].a;
            for (int64_t i = 0; i < N; ++i) p[i] = p[i] + (S->buf[i] * c[i]);
        } else {
            const float* c = S->coef[O_C_DTM].a;
            int tg[3] = {O_TXX, O_TYY, O_TZZ};
            for (int t = 0; t < 3; ++t) {
                if (S->e.ndim == 2 && tg[t] == O_TYY) continue;
                float* p = S->f[tg[t]].a;
                for (int64_t i = 0; i < N; ++i) p[i] = p[i] + (S->buf[i] * c[i]);
            }
        }
    }
}

/* record!: rec = transpose(W) * field, SparseArrays column-wise dot (receiver.jl:3-14) */
static void record(oracle_t* S, int64_t it, int stress)
{
    for (int q = 0; q < S->e.nrf; ++q) {
        int rf = S->e.rfields[q];
        int is_stress = (rf == O_P || rf == O_TXX || rf == O_TYY || rf == O_TZZ ||
                         rf == O_TXY || rf == O_TXZ || rf == O_TYZ);
        /* propagate.jl:177 records [:p] only; :208 records [:vx,:vy,:vz] */
        if (stress && rf != O_P) continue;
        if (!stress && is_stress) continue;
        const sparse_t* W = &S->gath[q];
        const float* F = S->f[rf].a;
        for (int r = 0; r < W->npts; ++r) {
            float tmp = 0.0f;
            for (int t = 0; t < W->nnz; ++t) tmp += W->w[r * W->nnz + t] * F[W->idx[r * W->nnz + t]];
            S->rec[q][it + S->e.nt * r] = tmp;
        }
    }
}

static const int derivs_ac2[] = {O_DPDX, O_DPDZ, -1};
static const int derivs_ac3[] = {O_DPDX, O_DPDY, O_DPDZ, -1};
static const int derivv_ac2[] = {O_DVXDX, O_DVZDZ, -1};
static const int derivv_ac3[] = {O_DVXDX, O_DVYDY, O_DVZDZ, -1};
static const int derivs_el2[] = {O_DTXXDX, O_DTXZDX, O_DTZZDZ, O_DTXZDZ, -1};
static const int derivs_el3[] = {O_DTXXDX, O_DTXYDX, O_DTXZDX, O_DTYYDY, O_DTXYDY, O_DTYZDY,
                                 O_DTZZDZ, O_DTYZDZ, O_DTXZDZ, -1};
static const int derivv_el2[] = {O_DVXDX, O_DVZDZ, O_DVXDZ, O_DVZDX, -1};
static const int derivv_el3[] = {O_DVXDX, O_DVYDY, O_DVZDZ, O_DVXDY, O_DVXDZ, O_DVYDX,
                                 O_DVYDZ, O_DVZDX, O_DVZDY, -1};

static const int* stress_derivs(const oracle_t* S)
{
    if (S->e.physics == 0) return S->e.ndim == 2 ? derivs_ac2 : derivs_ac3;
    return S->e.ndim == 2 ? derivs_el2 : derivs_el3;
}
static const int* velocity_derivs(const oracle_t* S)
{
    if (S->e.physics == 0) return S->e.ndim == 2 ? derivv_ac2 : derivv_ac3;
    return S->e.ndim == 2 ? derivv_el2 : derivv_el3;
}

/* update_dstress! + update_v! (advance_acou.jl:11-58, advance_elastic.jl:233-394,534-608) */
static void velocity_half_step(oracle_t* S)
{
    for (const int* d = stress_derivs(S); *d >= 0; ++d) compute_derivative(S, *d);
    for (const int* d = stress_derivs(S); *d >= 0; ++d) cpml_update(S, *d);
    update_velocity(S);
    dirichlet(S);
}

/* update_dv! + update_stress! */
static void stress_half_step(oracle_t* S)
{
    for (const int* d = velocity_derivs(S); *d >= 0; ++d) compute_derivative(S, *d);
    for (const int* d = velocity_derivs(S); *d >= 0; ++d) cpml_update(S, *d);
    update_stress(S);
}

/* ------------------------------------------------------------------------------------ */
/* public API                                                                            */
/* ------------------------------------------------------------------------------------ */

static int bad(const char* msg)
{
    fprintf(stderr, "fdtd_oracle: %s\n", msg);
    return -1;
}

oracle_t* oracle_create(const oracle_expt* e)
{
    if (e->ndim != 2 && e->ndim != 3) { bad("ndim"); return NULL; }
    if (e->order != 2 && e->order != 4 && e->order != 6 && e->order != 8) { bad("order"); return NULL; }
    if (e->nsf > O_MAXSF || e->nrf > O_MAXRF) { bad("too many fields"); return NULL; }
    oracle_t* S = calloc(1, sizeof(oracle_t));
    S->e = *e;
    S->h = e->order - 1;
    S->o2 = e->order / 2;
    for (int a = 0; a < 3; ++a) {
        S->d[a] = e->d[a];
        if (e->ndim == 2 && a == 1) { S->n[a] = 1; S->m0[a] = 0.0; continue; }
        int lo = (e->pml_mask >> (2 * a)) & 1, hi = (e->pml_mask >> (2 * a + 1)) & 1;
        S->n[a] = e->nm[a] + (lo ? e->npml : 0) + (hi ? e->npml : 0);
        S->m0[a] = e->o[a] - (lo ? e->npml : 0) * e->d[a]; /* padmgrid, media.jl:278-296 */
    }
    /* get_fc: fdtd.jl:313-332 -> Data.Number */
    double div = e->order == 4 ? 24.0 : e->order == 6 ? 1920.0 : e->order == 8 ? 107520.0 : 1.0;
    for (int a = 0; a < 3; ++a) S->dsI[a] = (float)(1.0 / (e->d[a] * div));
    S->dt = (float)e->dt;
    /* padded medium, its reference values and bounds */
    pad_medium(S, e->vp, &S->vp);
    pad_medium(S, e->rho, &S->rho);
    if (e->physics == 1) pad_medium(S, e->vs, &S->vs);
    float ref, bnd[2];
    ref_bounds(&S->vp, &ref, bnd);
    S->velavg = (float)(((double)bnd[0] + (double)bnd[1]) / 2.0);
    /* wavelets (type 1 copy: source.jl:6) */
    for (int q = 0; q < e->nsf; ++q) {
        S->wav[q] = malloc(sizeof(float) * (size_t)(e->nt * e->ns));
        memcpy(S->wav[q], e->wav + (size_t)q * e->nt * e->ns, sizeof(float) * (size_t)(e->nt * e->ns));
    }
    /* freqpeak: mean over source fields of findfreq peaks, stored as Data.Number */
    if (e->freqpeak > 0.0) S->freqpeak = (float)e->freqpeak;
    else {
        double s = 0.0;
        int c = 0;
        for (int q = 0; q < e->nsf; ++q) {
            double f = findfreq_peak(S->wav[q], e->nt, e->ns, e->dt);
            if (!isnan(f)) { s += f; ++c; }
        }
        S->freqpeak = (float)(c ? s / c : 0.0);
    }
    /* fields, derivatives, CPML memory */
    for (int f = 0; f < O_NFIELDS; ++f) {
        if (!present(S, f)) continue;
        alloc_field(S, &S->f[f], f);
        if (is_deriv(f)) {
            dspec_t sp = dspec(f);
            arr_t* M = &S->mem[f];
            M->n[0] = S->f[f].n[0]; M->n[1] = S->f[f].n[1]; M->n[2] = S->f[f].n[2];
            M->n[sp.axis] = 2 * e->npml;
            M->a = calloc((size_t)numel(M), sizeof(float));
            S->pml_a[f] = calloc(2 * e->npml, sizeof(float));
            S->pml_b[f] = calloc(2 * e->npml, sizeof(float));
            S->pml_kI[f] = calloc(2 * e->npml, sizeof(float));
            pml_profile(S, f, sp.axis, S->pml_a[f], S->pml_b[f], S->pml_kI[f]);
        }
    }
    prep_coefficients(S);
    /* spray / gather weights */
    int nnz = e->ndim == 2 ? 4 : 8;
    int64_t maxn = 0;
    for (int f = 0; f < O_NFIELDS; ++f)
        if (S->f[f].a && numel(&S->f[f]) > maxn) maxn = numel(&S->f[f]);
    S->buf = calloc((size_t)maxn, sizeof(float));
    for (int q = 0; q < e->nsf; ++q) {
        sparse_t* W = &S->spray[q];
        W->npts = e->ns; W->nnz = nnz;
        W->idx = malloc(sizeof(int64_t) * nnz * e->ns);
        W->w = malloc(sizeof(float) * nnz * e->ns);
        for (int s = 0; s < e->ns; ++s)
            point_weights(S, e->sfields[q], e->spos + (size_t)s * e->ndim, W->idx + s * nnz, W->w + s * nnz);
    }
    for (int q = 0; q < e->nrf; ++q) {
        sparse_t* W = &S->gath[q];
        W->npts = e->nr; W->nnz = nnz;
        W->idx = malloc(sizeof(int64_t) * nnz * e->nr);
        W->w = malloc(sizeof(float) * nnz * e->nr);
        for (int r = 0; r < e->nr; ++r)
            point_weights(S, e->rfields[q], e->rpos + (size_t)r * e->ndim, W->idx + r * nnz, W->w + r * nnz);
        S->rec[q] = calloc((size_t)(e->nt * e->nr), sizeof(float));
    }
    return S;
}

/* ---------------------------------------------------------------------------------- */
/* time reversal: boundary saving and forcing (src/fdtd/boundary.jl), gradient          */
/* ---------------------------------------------------------------------------------- */

#define NBOUND 3 /* _fd_nbound, src/GeoPhyInv.jl:92 */

/* stress fields kept by boundary_save! (boundary.jl:217-264): p, or tauxx/tauxz/tauzz */
static int stored_fields(const oracle_t* S, int out[3])
{
    if (S->e.physics == 0) { out[0] = O_P; return 1; }
    out[0] = O_TXX; out[1] = O_TXZ; out[2] = O_TZZ;
    return 3;
}

static int64_t strip_np(const oracle_t* S, int axis)
{
    /* boundary_{save,force}{x,y,z}! use the min face flag for both strips (boundary.jl:24,40) */
    return ((S->e.pml_mask >> (2 * axis)) & 1) ? S->e.npml : 0;
}

/* boundary_half{x,y,z}!: store[.., a] <-> field[.., np + a] (a < nbound) and
 * field[.., n - np - 2 nbound + a] (a >= nbound), all other indices full */
static void strip_copy(const oracle_t* S, arr_t* F, float* st, int axis, int save)
{
    int64_t sd[3] = {F->n[0], F->n[1], F->n[2]};
    sd[axis] = 2 * NBOUND;
    const int64_t np = strip_np(S, axis), n = F->n[axis];
    for (int64_t k = 0; k < sd[2]; ++k)
        for (int64_t j = 0; j < sd[1]; ++j)
            for (int64_t i = 0; i < sd[0]; ++i) {
                int64_t a = axis == 0 ? i : axis == 1 ? j : k;
                int64_t fi[3] = {i, j, k};
                fi[axis] = a < NBOUND ? np + a : n - np - 2 * NBOUND + a;
                int64_t so = i + sd[0] * (j + sd[1] * k);
                int64_t fo = IX(F, fi[0], fi[1], fi[2]);
                if (save) st[so] = F->a[fo];
                else F->a[fo] = st[so];
            }
}

static int64_t strip_size(const arr_t* F, int axis)
{
    int64_t sd[3] = {F->n[0], F->n[1], F->n[2]};
    sd[axis] = 2 * NBOUND;
    return sd[0] * sd[1] * sd[2];
}

static int store_alloc(oracle_t* S)
{
    if (S->nstore) return 0;
    if (S->e.physics == 1 && S->e.ndim == 3) return bad("boundary saving of 3-D elastic is not in the reference");
    S->nstore = stored_fields(S, S->store_f);
    for (int q = 0; q < S->nstore; ++q)
        for (int ax = 0; ax < 3; ++ax) {
            if (S->e.ndim == 2 && ax == 1) continue;
            arr_t* F = &S->f[S->store_f[q]];
            arr_t* B = &S->store[q][ax];
            B->n[0] = strip_size(F, ax); B->n[1] = S->e.nt; B->n[2] = 1;
            B->a = calloc((size_t)(B->n[0] * B->n[1]), sizeof(float));
        }
    return 0;
}

/* boundary_save!(it): store -stress on the strips (boundary.jl:217-264) */
static void boundary_save(oracle_t* S, int64_t it)
{
    for (int q = 0; q < S->nstore; ++q)
        for (int ax = 2; ax >= 0; --ax) {   /* x, (y,) z as in the reference */
            if (S->e.ndim == 2 && ax == 1) continue;
            arr_t* F = &S->f[S->store_f[q]];
            float* st = S->store[q][ax].a + S->store[q][ax].n[0] * it;
            strip_copy(S, F, st, ax, 1);
            for (int64_t m = 0; m < S->store[q][ax].n[0]; ++m) st[m] = st[m] * -1.0f;
        }
}

/* boundary_force!(it): overwrite the strips with the store (boundary.jl:115-169) */
static void boundary_force(oracle_t* S, int64_t it)
{
    for (int q = 0; q < S->nstore; ++q)
        for (int ax = 2; ax >= 0; --ax) {
            if (S->e.ndim == 2 && ax == 1) continue;
            arr_t* F = &S->f[S->store_f[q]];
            strip_copy(S, F, S->store[q][ax].a + S->store[q][ax].n[0] * it, ax, 0);
        }
}

static void copy_field(arr_t* dst, const arr_t* src)
{
    if (!dst->a) {
        dst->n[0] = src->n[0]; dst->n[1] = src->n[1]; dst->n[2] = src->n[2];
        dst->a = malloc(sizeof(float) * (size_t)numel(src));
    }
    memcpy(dst->a, src->a, sizeof(float) * (size_t)numel(src));
}

static void reset_wavefields(oracle_t* S)
{
    for (int f = 0; f < O_NFIELDS; ++f) {
        if (S->f[f].a) memset(S->f[f].a, 0, sizeof(float) * (size_t)numel(&S->f[f]));
        if (S->mem[f].a) memset(S->mem[f].a, 0, sizeof(float) * (size_t)numel(&S->mem[f]));
    }
}

static const int vel_fields[3] = {O_VX, O_VY, O_VZ};

/* mod_x_proc! in :forward_save mode (propagate.jl:138-261 with boundary_save! at :229 and
 * the final snapshots of :251-258) */
int oracle_run_save(oracle_t* S)
{
    if (store_alloc(S)) return -1;
    reset_wavefields(S);
    for (int64_t it = 0; it < S->e.nt; ++it) {
        record(S, it, 1);
        velocity_half_step(S);
        add_velocity_source(S, it);
        record(S, it, 0);
        stress_half_step(S);
        add_stress_source(S, it);
        boundary_save(S, it);
    }
    for (int q = 0; q < S->nstore; ++q) {   /* boundary_save_snap_tau!: -stress */
        arr_t* D = &S->snap[S->store_f[q]];
        copy_field(D, &S->f[S->store_f[q]]);
        for (int64_t m = 0; m < numel(D); ++m) D->a[m] = D->a[m] * -1.0f;
    }
    velocity_half_step(S);
    for (int v = 0; v < 3; ++v)             /* boundary_save_snap_v! */
        if (S->f[vel_fields[v]].a) copy_field(&S->snap[vel_fields[v]], &S->f[vel_fields[v]]);
    S->it_done = S->e.nt;
    return 0;
}

/* save_tp!: copy the wavefields (no derivatives) before the step (save_tp.jl:5-11) */
static void save_tp(oracle_t* S)
{
    for (int f = O_P; f <= O_TYZ; ++f)
        if (S->f[f].a) copy_field(&S->tp[f], &S->f[f]);
}

/* mod_x_proc! in :adjoint mode with npw = 2 (propagate.jl:138-261): S1 re-propagates the
 * forward wavefield backwards from its snapshots with the strips forced from the store and
 * its (type -1) sources; S2 propagates the adjoint sources.  compute_gradient! after each
 * step accumulates (gradient.jl:17-64, 2-D acoustic only):
 *   g_invK += p2_prev * (p1_prev - p1) * dtI
 *   g_rho[inn] -= @av_xi(vx2_prev (vx1 - vx1_prev) dtI) + @av_zi(vz2_prev (vz1 - vz1_prev) dtI) */
int oracle_adjoint(oracle_t* S1, oracle_t* S2, float* g_invK, float* g_rho)
{
    if (S1->e.physics != 0 || S1->e.ndim != 2) return bad("the reference gradient exists for 2-D acoustic only");
    if (!S1->nstore || !S1->snap[O_P].a) return bad("run oracle_run_save first");
    const int h = S1->h;
    const float dtI = (float)(1.0 / S1->e.dt);   /* fc[:dtI] = Data.Number(inv(dt)) */
    reset_wavefields(S1);
    reset_wavefields(S2);
    copy_field(&S1->f[O_P], &S1->snap[O_P]);     /* boundary_force_snap_tau! */
    copy_field(&S1->f[O_VX], &S1->snap[O_VX]);   /* boundary_force_snap_v! */
    copy_field(&S1->f[O_VZ], &S1->snap[O_VZ]);
    const arr_t* P = &S1->f[O_P];
    const int64_t N = numel(P);
    memset(g_invK, 0, sizeof(float) * (size_t)N);
    memset(g_rho, 0, sizeof(float) * (size_t)N);
    const int64_t nt = S1->e.nt;
    for (int64_t it = 0; it < nt; ++it) {
        record(S1, it, 1);            /* computed but not returned by the reference (rec_flags) */
        save_tp(S1);
        save_tp(S2);
        boundary_force(S1, nt - 1 - it);
        velocity_half_step(S1);
        velocity_half_step(S2);
        add_velocity_source(S1, it);
        add_velocity_source(S2, it);
        record(S1, it, 0);
        stress_half_step(S1);
        stress_half_step(S2);
        add_stress_source(S1, it);
        /* gradlame!: compute_gmodKI! */
        const float *p1 = S1->f[O_P].a, *p1p = S1->tp[O_P].a, *p2p = S2->tp[O_P].a;
        for (int64_t m = 0; m < N; ++m) g_invK[m] = g_invK[m] + (p2p[m] * (p1p[m] - p1[m])) * dtI;
        /* gradrho!: compute_gmodrho! into buffers, then combine_gmodrho! */
        const arr_t *VX = &S1->f[O_VX], *VZ = &S1->f[O_VZ];
        const float *vx1 = VX->a, *vx1p = S1->tp[O_VX].a, *vx2p = S2->tp[O_VX].a;
        const float *vz1 = VZ->a, *vz1p = S1->tp[O_VZ].a, *vz2p = S2->tp[O_VZ].a;
        const int64_t nz = P->n[0], nx = P->n[2];
        for (int64_t j = 0; j < nx - 2 * h; ++j)
            for (int64_t i = 0; i < nz - 2 * h; ++i) {
                int64_t a0 = IX(VX, i + h, 0, j), a1 = IX(VX, i + h, 0, j + 1);
                int64_t b0 = IX(VZ, i, 0, j + h), b1 = IX(VZ, i + 1, 0, j + h);
                float bx0 = (vx2p[a0] * (vx1[a0] - vx1p[a0])) * dtI;
                float bx1 = (vx2p[a1] * (vx1[a1] - vx1p[a1])) * dtI;
                float bz0 = (vz2p[b0] * (vz1[b0] - vz1p[b0])) * dtI;
                float bz1 = (vz2p[b1] * (vz1[b1] - vz1p[b1])) * dtI;
                int64_t g = IX(P, i + h, 0, j + h);
                g_rho[g] = (g_rho[g] - (bx0 + bx1) * 0.5f) - (bz0 + bz1) * 0.5f;
            }
    }
    velocity_half_step(S1);   /* :254-255 (pw1 only) */
    return 0;
}

int oracle_set_wavelets(oracle_t* S, int q, const float* w, int64_t n)
{
    if (q < 0 || q >= S->e.nsf || n != S->e.nt * S->e.ns) return -1;
    memcpy(S->wav[q], w, sizeof(float) * (size_t)n);
    return 0;
}

int oracle_get_store(const oracle_t* S, int q, int axis, float* out, int64_t n)
{
    if (q < 0 || q >= S->nstore || !S->store[q][axis].a || n != S->store[q][axis].n[0] * S->store[q][axis].n[1])
        return -1;
    memcpy(out, S->store[q][axis].a, sizeof(float) * (size_t)n);
    return 0;
}

int oracle_get_snap(const oracle_t* S, int f, float* out, int64_t n)
{
    if (f < 0 || f >= O_NFIELDS || !S->snap[f].a || numel(&S->snap[f]) != n) return -1;
    memcpy(out, S->snap[f].a, sizeof(float) * (size_t)n);
    return 0;
}

/* mod_x_proc! time loop for one supersource, forward mode: propagate.jl:138-261.
 * nsteps < 0 runs all nt steps plus the trailing velocity half step (:254-255). */
int oracle_run(oracle_t* S, int64_t nsteps)
{
    int64_t nt = S->e.nt, stop = nsteps < 0 ? nt : (nsteps < nt ? nsteps : nt);
    for (int64_t it = 0; it < stop; ++it) {
        record(S, it, 1);            /* :177  p at [it-1] */
        velocity_half_step(S);       /* :192, :197 */
        add_velocity_source(S, it);  /* :202 */
        record(S, it, 0);            /* :208 */
        stress_half_step(S);         /* :213, :218 */
        add_stress_source(S, it);    /* :223 */
    }
    S->it_done = stop;
    if (nsteps < 0) velocity_half_step(S); /* :254-255 */
    return 0;
}

/* Step `it` of the loop of oracle_run, on the current state (test hook: the exact-wavefield
 * gradient composes per-step fields of two oracles), and the trailing half step. */
int oracle_step(oracle_t* S, int64_t it)
{
    if (it < 0 || it >= S->e.nt) return -1;
    record(S, it, 1);
    velocity_half_step(S);
    add_velocity_source(S, it);
    record(S, it, 0);
    stress_half_step(S);
    add_stress_source(S, it);
    S->it_done = it + 1;
    return 0;
}

/* The two halves of oracle_step (test hook: Born modelling adds secondary sources between
 * them): record p, velocity half step, velocity sources; record v, stress half step,
 * stress sources. */
int oracle_step_velocity(oracle_t* S, int64_t it)
{
    if (it < 0 || it >= S->e.nt) return -1;
    record(S, it, 1);
    velocity_half_step(S);
    add_velocity_source(S, it);
    return 0;
}

int oracle_step_stress(oracle_t* S, int64_t it)
{
    if (it < 0 || it >= S->e.nt) return -1;
    record(S, it, 0);
    stress_half_step(S);
    add_stress_source(S, it);
    S->it_done = it + 1;
    return 0;
}

int oracle_trailing(oracle_t* S)
{
    velocity_half_step(S);
    return 0;
}

int oracle_get_field(const oracle_t* S, int f, float* out, int64_t n)
{
    if (f < 0 || f >= O_NFIELDS || !S->f[f].a || numel(&S->f[f]) != n) return -1;
    memcpy(out, S->f[f].a, sizeof(float) * (size_t)n);
    return 0;
}

/* test hooks: load a wavefield and evaluate one derivative (optionally with its CPML) */
int oracle_set_field(oracle_t* S, int f, const float* in, int64_t n)
{
    if (f < 0 || f >= O_NFIELDS || !S->f[f].a || numel(&S->f[f]) != n) return -1;
    memcpy(S->f[f].a, in, sizeof(float) * (size_t)n);
    return 0;
}

int oracle_derivative(oracle_t* S, int df, int with_cpml)
{
    if (df < O_DPDX || df >= O_NFIELDS || !S->f[df].a) return -1;
    compute_derivative(S, df);
    if (with_cpml) cpml_update(S, df);
    return 0;
}

int oracle_get_memory(const oracle_t* S, int f, float* out, int64_t n)
{
    if (f < 0 || f >= O_NFIELDS || !S->mem[f].a || numel(&S->mem[f]) != n) return -1;
    memcpy(out, S->mem[f].a, sizeof(float) * (size_t)n);
    return 0;
}

/* coefficient array c; O_C_RHO is the padded density (mod[:rho]) the velocity sources read */
static const arr_t* coef_array(const oracle_t* S, int c)
{
    if (c == O_C_RHO && S->rho.a) return &S->rho;
    return (c >= 0 && c < O_NCOEF && S->coef[c].a) ? &S->coef[c] : NULL;
}

int oracle_get_coef(const oracle_t* S, int c, float* out, int64_t n)
{
    const arr_t* A = coef_array(S, c);
    if (!A || numel(A) != n) return -1;
    memcpy(out, A->a, sizeof(float) * (size_t)n);
    return 0;
}

/* test hook: overwrite a coefficient array (O_C_RHO is the padded density used by the
 * velocity sources), for runs on coefficients prepared elsewhere */
int oracle_set_coef(oracle_t* S, int c, const float* in, int64_t n)
{
    arr_t* A = c == O_C_RHO ? &S->rho : (c >= 0 && c < O_NCOEF ? &S->coef[c] : NULL);
    if (!A || !A->a || numel(A) != n) return -1;
    memcpy(A->a, in, sizeof(float) * (size_t)n);
    return 0;
}

int oracle_coef_shape(const oracle_t* S, int c, int64_t out[3])
{
    const arr_t* A = coef_array(S, c);
    if (!A) return -1;
    for (int a = 0; a < 3; ++a) out[a] = A->n[a];
    return 0;
}

int oracle_get_pml(const oracle_t* S, int f, float* a, float* b, float* kI)
{
    if (f < 0 || f >= O_NFIELDS || !S->pml_a[f]) return -1;
    int n = 2 * S->e.npml;
    memcpy(a, S->pml_a[f], sizeof(float) * n);
    memcpy(b, S->pml_b[f], sizeof(float) * n);
    memcpy(kI, S->pml_kI[f], sizeof(float) * n);
    return 0;
}

int oracle_get_weights(const oracle_t* S, int src, int q, int64_t* idx, float* w)
{
    const sparse_t* W = src ? &S->spray[q] : &S->gath[q];
    if ((src && q >= S->e.nsf) || (!src && q >= S->e.nrf) || !W->idx) return -1;
    memcpy(idx, W->idx, sizeof(int64_t) * W->npts * W->nnz);
    memcpy(w, W->w, sizeof(float) * W->npts * W->nnz);
    return 0;
}

int oracle_get_records(const oracle_t* S, int q, float* out, int64_t n)
{
    if (q < 0 || q >= S->e.nrf || n != S->e.nt * S->e.nr) return -1;
    memcpy(out, S->rec[q], sizeof(float) * (size_t)n);
    return 0;
}

void oracle_scalars(const oracle_t* S, float* out)
{
    out[0] = S->dt; out[1] = S->dsI[0]; out[2] = S->dsI[1]; out[3] = S->dsI[2];
    out[4] = S->freqpeak; out[5] = S->velavg;
}

void oracle_ext_grid(const oracle_t* S, int64_t n[3], double m0[3])
{
    for (int a = 0; a < 3; ++a) { n[a] = S->n[a]; m0[a] = S->m0[a]; }
}

int oracle_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_destroy(oracle_t* S)
{
    if (!S) return;
    for (int f = 0; f < O_NFIELDS; ++f) {
        free(S->f[f].a); free(S->mem[f].a);
        free(S->pml_a[f]); free(S->pml_b[f]); free(S->pml_kI[f]);
    }
    for (int c = 0; c < O_NCOEF; ++c) free(S->coef[c].a);
    free(S->vp.a); free(S->vs.a); free(S->rho.a); free(S->buf);
    for (int q = 0; q < O_MAXSF; ++q) { free(S->spray[q].idx); free(S->spray[q].w); free(S->wav[q]); }
    for (int q = 0; q < O_MAXRF; ++q) { free(S->gath[q].idx); free(S->gath[q].w); free(S->rec[q]); }
    for (int q = 0; q < 3; ++q)
        for (int a = 0; a < 3; ++a) free(S->store[q][a].a);
    for (int f = 0; f < O_NFIELDS; ++f) { free(S->snap[f].a); free(S->tp[f].a); }
    free(S);
}
