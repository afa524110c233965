/*
 * CPU oracle of the GeoPhyInv.jl staggered-grid FDTD path.  TEST INFRASTRUCTURE ONLY:
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
 *
 * Field and coefficient numbering matches the product C-ABI (include/gpi.h) so that test
 * code can pass ids straight through; the tests check that both sides agree on shapes.
 */
#ifndef FDTD_ORACLE_H
#define FDTD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    O_P = 0, O_VX, O_VY, O_VZ, O_TXX, O_TYY, O_TZZ, O_TXY, O_TXZ, O_TYZ,
    O_DPDX, O_DPDY, O_DPDZ,
    O_DVXDX, O_DVYDY, O_DVZDZ, O_DVXDY, O_DVXDZ, O_DVYDX, O_DVYDZ, O_DVZDX, O_DVZDY,
    O_DTXXDX, O_DTYYDY, O_DTZZDZ, O_DTXYDX, O_DTXYDY, O_DTXZDX, O_DTXZDZ, O_DTYZDY, O_DTYZDZ,
    O_NFIELDS
};

enum {
    O_C_DTINVAVXIRHO = 0, O_C_DTINVAVYIRHO, O_C_DTINVAVZIRHO, O_C_DTK, O_C_DTM, O_C_DTLAMBDA,
    O_C_DTAVMU, O_C_DTAVXZIMU, O_C_DTAVXYIMU, O_C_DTAVYZIMU, O_C_RHO, O_C_INVK, O_C_INVLAMBDA,
    O_C_INVMU, O_NCOEF
};

/* faces: bit 2*axis + (0 min | 1 max), axis 0 z, 1 y, 2 x */
#define O_COMPAT_SWAP3D 1u
#define O_MAXSF 4
#define O_MAXRF 8
#define O_MAXPTS_STACK 64

typedef struct {
    int ndim, physics, order;       /* physics 0 acoustic, 1 elastic */
    int64_t nm[3];                  /* medium (unpadded) sizes z, y, x; y = 1 in 2-D */
    double o[3], d[3];              /* medium grid origin and spacing per axis */
    const float *vp, *vs, *rho;     /* column-major, z fastest; vs NULL for acoustic */
    int npml;
    unsigned pml_mask, rigid_mask, free_mask, compat;
    int64_t nt;
    double dt;
    double freqpeak;                /* <= 0: computed from the wavelets (findfreq :peak) */
    int ns;                         /* sources of the supersource */
    const double* spos;             /* ns x ndim, (z,[y],x) per source */
    int nsf;
    const int* sfields;
    const float* wav;               /* nsf blocks of nt x ns, time fastest (Srcs.d layout) */
    int nr;
    const double* rpos;             /* nr x ndim */
    int nrf;
    const int* rfields;
} oracle_expt;

typedef struct oracle oracle_t;

int oracle_field_shape(int physics, int ndim, int order, const int64_t n[3], int f, int64_t out[3]);
oracle_t* oracle_create(const oracle_expt* e);
int oracle_run(oracle_t* S, int64_t nsteps);
int oracle_run_save(oracle_t* S);
int oracle_step(oracle_t* S, int64_t it);
int oracle_trailing(oracle_t* S);
int oracle_step_velocity(oracle_t* S, int64_t it);
int oracle_step_stress(oracle_t* S, int64_t it);
int oracle_adjoint(oracle_t* S1, oracle_t* S2, float* g_invK, float* g_rho);
int oracle_set_wavelets(oracle_t* S, int q, const float* w, int64_t n);
int oracle_get_store(const oracle_t* S, int q, int axis, float* out, int64_t n);
int oracle_get_snap(const oracle_t* S, int f, float* out, int64_t n);
int oracle_get_field(const oracle_t* S, int f, float* out, int64_t n);
int oracle_set_field(oracle_t* S, int f, const float* in, int64_t n);
int oracle_derivative(oracle_t* S, int df, int with_cpml);
int oracle_get_memory(const oracle_t* S, int f, float* out, int64_t n);
int oracle_get_coef(const oracle_t* S, int c, float* out, int64_t n);
int oracle_coef_shape(const oracle_t* S, int c, int64_t out[3]);
int oracle_set_coef(oracle_t* S, int c, const float* in, int64_t n);
int oracle_get_pml(const oracle_t* S, int f, float* a, float* b, float* kI);
int oracle_get_weights(const oracle_t* S, int src, int q, int64_t* idx, float* w);
int oracle_get_records(const oracle_t* S, int q, float* out, int64_t n);
void oracle_scalars(const oracle_t* S, float* out);
void oracle_ext_grid(const oracle_t* S, int64_t n[3], double m0[3]);
int oracle_threads(void);
void oracle_destroy(oracle_t* S);

#ifdef __cplusplus
}
#endif
#endif
