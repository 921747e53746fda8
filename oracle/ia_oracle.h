/*
 * ia_oracle.h -- CPU oracle for the two-phase time step of InterfaceAdvection.jl.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP library is
 * compared against; it is never linked into, loaded by, or called from the
 * product path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it.
 *
 * The same source (ia_oracle.c) is compiled twice, once with REAL=float and
 * once with REAL=double, into libia_oracle_f32.so and libia_oracle_f64.so.
 * `real_t` below is that type.  Every array is laid out exactly as the
 * reference's Julia arrays: column-major (x fastest), one ghost layer per
 * side, vector components last:  idx = i + Ngx*(j + Ngy*(k + Ngz*c)).
 */
#ifndef IA_ORACLE_H
#define IA_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifndef REAL
#define REAL double
#endif
typedef REAL real_t;

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oc_params {
    int D;              /* 2 or 3 */
    int N[3];           /* interior cells per axis */
    int perdir;         /* bit d set -> axis d (0-based) periodic */
    double u_bc[3];     /* constant boundary velocity tuple u_BC */
    double g[3];        /* constant gravity g(i,x,t) = g[i] */
    int has_g;          /* 0 -> g === nothing */
    double eta;         /* surface tension coefficient η; 0 -> η === nothing */
    double nu;          /* flow.ν (= μ of the dark fluid, ρ_dark = 1); 0 -> no viscosity */
    double lambda_rho;  /* λρ */
    double lambda_mu;   /* λμ */
    int normal_scheme;  /* 0 WH (default), 1 Column, 2 WY, 3 MYC, 4 PCD, 5 Youngs, 6 SLIC */
    int limiter;        /* 0 Koren (default), 1 minmod, 2 upwind, 3 vanAlbada1, 4 superbee, 5 TVDcen, 6 TVDdown, 7 Sweby */
    double dt0;         /* initial Δt of the Simulation (WaterLily default 0.25) */
    int mg_itmx;        /* solver! itmx (reference: 200) */
    double mg_tol;      /* solver! tol on Σr² (reference: 1e-4) */
    int poisson_mode;   /* 0: WaterLily solver! (reference); 1: FP64 flexible CG + V-cycle to poisson_rtol;
                           2: psolver! on the single-level Poisson (flow.jl:300-341) */
    double poisson_rtol;
    int smooth_it;      /* fine-level pcg! iterations per preconditioning V-cycle (mode 1) */
    int exit_bc;        /* WaterLily exitBC: BC! keeps the exit (upper x) face of u_x */
} oc_params;

/* field ids */
enum {
    OC_F = 0, OC_F0, OC_ALPHA, OC_NHAT, OC_FFLUX, OC_CBAR, OC_RHOU, OC_RHOUF, OC_DRHO,
    OC_U, OC_U0, OC_FORCE, OC_P, OC_SIGMA, OC_MU0, OC_MG_D, OC_MG_ID, OC_MG_R, OC_MG_EPS,
    OC_NFIELDS
};

/* status codes (same meaning as the HIP library's) */
enum { OC_OK = 0, OC_NAN = 1, OC_EXPLODE = 2, OC_FILLWARN = 3, OC_BADARG = 5 };

typedef struct oc_ctx oc_ctx;

oc_ctx *oc_create(const oc_params *p);
void oc_destroy(oc_ctx *c);
int oc_field_len(oc_ctx *c, int id);           /* number of elements */
int oc_set_field(oc_ctx *c, int id, const void *src);
int oc_get_field(oc_ctx *c, int id, void *dst);
int oc_mg_nlevels(oc_ctx *c);
int oc_mg_level_dims(oc_ctx *c, int l, int *Ng);
int oc_mg_level_get(oc_ctx *c, int l, int which, real_t *out);
int oc_mg_transfer(oc_ctx *c, int l, int dir, const real_t *in);
int oc_mg_level_set(oc_ctx *c, int l, int which, const real_t *in);
int oc_mg_pcg(oc_ctx *c, int l, int it, int single);

/* construction helpers */
int oc_apply_vof(oc_ctx *c, const real_t *sdf);   /* applyVOF! + cleanWisp! + BCf! + f0=f */
int oc_set_initial_velocity(oc_ctx *c, const real_t *u); /* apply!, BC!, u0=u */
double oc_cfl(oc_ctx *c);                          /* MPCFL */

/* the path */
int oc_mom_step(oc_ctx *c, double dt, int step_index, double *dt_next);
int oc_advect(oc_ctx *c, double dt, int step_index); /* advect!(flow, intf, f, u0, u, dt) */

/* stage entry points (for parity tests of single stages) */
int oc_reconstruct(oc_ctx *c);
int oc_vof_sweep(oc_ctx *c, int d, double dt);   /* advectVOF1d! on f with u0/u; fill tol = 10eps */
int oc_visc(oc_ctx *c, int use_f0);               /* viscSurfTenρu! into FORCE with f or f0 */
int oc_update_L(oc_ctx *c, int use_f0);
int oc_mg_update(oc_ctx *c);
int oc_project(oc_ctx *c, double dt_w, int *iters, double *r2);
int oc_mg_solve(oc_ctx *c, double tol, int itmx, int *iters, double *r2);

/* iterative refinement variant: FP64 outer residual/accumulation around the
 * MultiLevelPoisson V-cycles; stops at ||z-Ax||2/||z||2 <= rtol */
int oc_project_refined(oc_ctx *c, double dt_w, double rtol, int max_outer, int *outer, double *rel);
/* FP64 flexible CG (FCG(1)) preconditioned by the FP32 V-cycle + smooth_it fine pcg! iterations */
int oc_project_fcg(oc_ctx *c, double dt_w, double rtol, int max_it, int smooth_it, int *iters, double *rel);

/* point functions (PLIC.jl, normalEstimation.jl, VOFutil.jl) for the KATs */
real_t oc_getIntercept2(real_t n1, real_t n2, real_t g);
real_t oc_getIntercept3(real_t n1, real_t n2, real_t n3, real_t g);
real_t oc_getVolumeFraction2(real_t n1, real_t n2, real_t b);
real_t oc_getVolumeFraction3(real_t n1, real_t n2, real_t n3, real_t b);
real_t oc_f2alpha3(real_t m1, real_t m2, real_t m3, real_t v);
real_t oc_alpha2f3(real_t m1, real_t m2, real_t m3, real_t a);
/* normal of cell (i,j,k) (1-based) of a small standalone f block of size Ng */
int oc_normal_block(int D, const int *Ng, const real_t *f, int scheme, const int *I, real_t *n_out);
real_t oc_getmu(int D, const int *Ng, const real_t *fface, int i, int j, const int *I,
                real_t lm, real_t mu, real_t lr);
int oc_bc_scalar(int D, const int *Ng, real_t *f, int perdir, int d);  /* BCf! (d=0) / BCf!(d,f) */
int oc_bc_facevec(int D, const int *Ng, real_t *v, int perdir);        /* BCv! */
int oc_bc_vof(int D, const int *Ng, real_t *f, real_t *alpha, real_t *nhat, int perdir);
int oc_f2face(int D, const int *Ng, const real_t *fc, real_t *ff, int perdir);
int oc_vof_flux_face(int D, const int *Ng, real_t *ff, const real_t *f, const real_t *alpha,
                     const real_t *nhat, real_t dl, int d, const int *IFace, real_t *rhouf, real_t lr);
real_t oc_koren(real_t u, real_t c, real_t d);
/* flow.jl:5-15: 0 Koren, 1 minmod, 2 upwind, 3 vanAlbada1, 4 superbee, 5 TVDcen, 6 TVDdown, 7 Sweby(γ=1.5) */
real_t oc_limiter(int lim, real_t u, real_t c, real_t d);
/* ϕq's face value (flow.jl:43-56) for a given Ψ and donor face fraction fd */
real_t oc_phiq_value(real_t Psi, real_t fd, real_t dt, real_t lr, int lim, real_t u, real_t c, real_t d);
/* surfaceTension.jl: height (getPopinetHeight) and curvature (getCurvature) about
 * cell I (1-based) of a standalone block, along the signed axis i (1-based) */
real_t oc_popinet_height(int D, const int *Ng, const real_t *f, const int *I, int i);
/* metrics.jl:15-50 per cell (I 1-based, U / StatWL / g as doubles) and the domain integrals
 * out = {Σρke, Σρu_1, Σρu_2, Σρu_3, Σρgh, Σenstrophy, Σf} over the interior */
double oc_rhoke(int D, const int *Ng, const real_t *u, const real_t *f, real_t lr, const int *I, const double *U);
double oc_rhou(int D, const int *Ng, int i, const real_t *u, const real_t *f, real_t lr, const int *I, const double *U);
double oc_rhogh(int D, const int *Ng, const double *gv, const real_t *f, real_t lr, const int *I, const double *wl);
double oc_ens(int D, const int *Ng, const real_t *w, const int *I);
int oc_metrics(oc_ctx *c, const double *U, const double *wl, double *out);
real_t oc_curvature(int D, const int *Ng, const real_t *f, const int *I, int i);
/* redistaning.jl on standalone arrays of Ng points per axis (at least 2 interior
 * cells each): the ENO contribution 𝛁ϕᵢ², computeL!, _redistaningStage! (in place,
 * L as scratch) and redistaning!(ϕ; d, dτ, perdir) with ϕ⁰ and L as scratch */
real_t oc_ls_grad2(real_t a, real_t b, real_t c, real_t d, real_t e, real_t s);
int oc_ls_computeL(int D, const int *Ng, real_t *L, const real_t *phi, const real_t *ini, int perdir);
int oc_ls_stage(int D, const int *Ng, real_t *phi, const real_t *phi0, const real_t *ini, real_t *L, real_t dtau,
                real_t alpha, int perdir);
int oc_redistance(int D, const int *Ng, real_t *phi, real_t *phi0, const real_t *ini, real_t *L, double d, double dtau,
                  int perdir);
void oc_set_threads(int n);
const char *oc_last_error(oc_ctx *c);

/* WaterLily BC!(a, A, saveexit, perdir) on a staggered vector field (the step's exitBC path) */
int oc_bc_vel(int D, const int *Ng, real_t *a, const real_t *A, int perdir, int saveexit);

#ifdef __cplusplus
}
#endif
#endif
