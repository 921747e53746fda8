/*
 * ia.h -- C ABI of the MI355X (gfx950) implementation of the two-phase time
 * step of InterfaceAdvection.jl.
 *
 * A Julia host binds these with `ccall` from a package extension whose
 * methods of MPFMomStep! and advect! dispatch on device-array fields; a
 * Python host binds them with ctypes (interfaceadvection.jl_amd/__init__.py).
 * Plain pointers and sizes only.
 *
 * Layout of every field buffer handed across this boundary is exactly the
 * reference's Julia array: column-major, x fastest, one ghost layer per side
 * (Ng = N + 2), vector components last:
 *     idx = i + Ngx*(j + Ngy*(k + Ngz*c)),  0 <= i < Ngx, ...
 * Element type is float (IA_F32) or double (IA_F64) as chosen at creation;
 * IA_FIELD_CBAR is int8.
 *
 * Threading: one context <-> one device <-> one HIP stream; a context is not
 * re-entrant.  Every call that returns scalars to the host (dt, flags,
 * solver counts) synchronises its stream before returning.
 */
#ifndef IA_H
#define IA_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ia_ctx ia_ctx;

enum ia_dtype { IA_F32 = 0, IA_F64 = 1 };

/* normalScheme plug-ins (normalEstimation.jl): WH is the reference default
 * (InterfaceAdvection.jl:70); Column is its first stage (:75-97); WY is
 * :35-68; MYC is Youngs + CCi (:185-257); PCD :161-165; Youngs :232-257;
 * SLIC :172-176. */
enum ia_normal_scheme {
    IA_NORMAL_WH = 0, IA_NORMAL_COLUMN = 1, IA_NORMAL_WY = 2, IA_NORMAL_MYC = 3, IA_NORMAL_PCD = 4,
    IA_NORMAL_YOUNGS = 5, IA_NORMAL_SLIC = 6
};

/* λ plug-ins (flow.jl:5-15); Koren is the reference default (InterfaceAdvection.jl:70). */
enum ia_limiter {
    IA_LIMITER_KOREN = 0, IA_LIMITER_MINMOD = 1, IA_LIMITER_UPWIND = 2, IA_LIMITER_VANALBADA1 = 3,
    IA_LIMITER_SUPERBEE = 4, IA_LIMITER_TVDCEN = 5, IA_LIMITER_TVDDOWN = 6, IA_LIMITER_SWEBY = 7
};

/* Pressure solve.
 * REFERENCE: WaterLily solver!(ml; tol, itmx) exactly as the reference calls it
 *            (flow.jl:346): stop when Σr² <= tol (absolute) or after itmx V-cycles.
 * PLAIN:     plain multigrid to a relative residual: FP32 MultiLevelPoisson
 *            V-cycles (+ the fine pcg!, i.e. solver!'s loop body) inside an FP64
 *            iterative-refinement loop; stop when ||z - A x||2 / ||z||2 <= rtol.
 * REFINED:   FP64 flexible CG (FCG(1)) preconditioned by one FP32 V-cycle (+ two
 *            Chebyshev smoothing steps on the finest level) per iteration;
 *            same stopping test, FP64 residual and pressure accumulation.  No
 *            reference counterpart.  Read at ia_create, for tuning sweeps:
 *            IA_FCG_CIT=a,b,... (coarse iterations on levels 1, 2, ...),
 *            IA_FCG_SMOOTH=n (n fine pcg! iterations instead of Chebyshev),
 *            IA_CG1_NB=a,b,... (coarse CG workgroup sizes on levels 1, 2, ...).
 * PCG:       inproject!(a, b::Poisson, dt) (flow.jl:337-341): psolver! (flow.jl:300-326),
 *            Jacobi-preconditioned CG on the single (finest) level until the
 *            absolute Σr² <= 50eps(T) or 2000 iterations, as the reference runs it
 *            when the simulation's Poisson is a single-level `Poisson`.
 * LEARNED:   REFINED, whose warm start is first corrected by the learned network
 *            (ia_learned_*; no reference counterpart).  The correction is kept
 *            only when it lowers the FP64 residual.  3D, FP32, N[0] % 16 == 0,
 *            N[1] and N[2] even. */
enum ia_poisson_mode {
    IA_POISSON_REFERENCE = 0, IA_POISSON_REFINED = 1, IA_POISSON_LEARNED = 2, IA_POISSON_PLAIN = 3, IA_POISSON_PCG = 4
};

/* Status codes (advection.jl:145-189 reportFillError and the glue's errors). */
enum ia_status {
    IA_OK = 0,
    IA_ERR_NAN = 1,        /* error("NaN!")                         advection.jl:148 */
    IA_ERR_EXPLODE = 2,    /* error("divergence ... is exploding!")  advection.jl:160,:180 */
    IA_WARN_FILL = 3,      /* fill error printed, not thrown          advection.jl:161-166 */
    IA_ERR_HIP = 4,        /* HIP runtime error */
    IA_ERR_BADARG = 5      /* bad argument / unsupported plug-in */
};

/* Field ids (the cVOF and Flow arrays of cVOF.jl:18-43 and WaterLily's Flow). */
enum ia_field {
    IA_FIELD_F = 0,      /* cVOF.f   (s) */
    IA_FIELD_F0 = 1,     /* cVOF.f⁰  (s) */
    IA_FIELD_ALPHA = 2,  /* cVOF.α   (s) -- materialised by ia_reconstruct */
    IA_FIELD_NHAT = 3,   /* cVOF.n̂   (v) -- materialised by ia_reconstruct */
    IA_FIELD_CBAR = 5,   /* cVOF.c̄   (s, int8) */
    IA_FIELD_RHOU = 6,   /* cVOF.ρu  (v) */
    IA_FIELD_RHOUF = 7,  /* cVOF.ρuf (v) */
    IA_FIELD_U = 9,      /* Flow.u   (v) */
    IA_FIELD_U0 = 10,    /* Flow.u⁰  (v) */
    IA_FIELD_FORCE = 11, /* Flow.f   (v) -- written by ia_visc */
    IA_FIELD_P = 12,     /* Flow.p   (s) */
    IA_FIELD_SIGMA = 13, /* Flow.σ   (s) -- Poisson source z */
    IA_FIELD_MU0 = 14    /* Flow.μ₀  (v) -- Poisson coefficients L */
};

typedef struct ia_params {
    int D;                 /* 2 or 3 */
    int N[3];              /* interior cells per axis; MultiLevelPoisson needs N+2 = a*2^n, n>2 */
    int dtype;             /* enum ia_dtype */
    int perdir;            /* bit d set -> axis d (0-based) periodic */
    double u_bc[3];        /* constant u_BC tuple */
    double g[3];           /* constant gravity, g(i,x,t) = g[i] */
    int has_g;             /* 0 -> g === nothing */
    double nu;             /* Flow.ν; cVOF.μ = ν (0 -> inviscid, cVOF.jl:78) */
    double lambda_rho;     /* λρ light/dark density ratio */
    double lambda_mu;      /* λμ light/dark viscosity ratio */
    double eta;            /* surface tension coefficient η (surfaceTension.jl); 0 -> η === nothing */
    int normal_scheme;     /* enum ia_normal_scheme */
    int limiter;           /* enum ia_limiter */
    int poisson_mode;      /* enum ia_poisson_mode */
    double poisson_tol;    /* REFERENCE: tol on Σr² (reference 1e-4) */
    int poisson_itmx;      /* REFERENCE: itmx (reference 200) */
    double poisson_rtol;   /* REFINED/LEARNED: relative FP64 residual (e.g. 1e-6) */
    int poisson_max_outer; /* REFINED/LEARNED: outer iteration cap */
    int device;            /* HIP device ordinal */
    int exit_bc;           /* WaterLily exitBC: the step's BC! leaves u_x on the exit face (the
                              upper x boundary) alone (flow.jl:69, :79, :82, :197, :207); needs
                              no gravity along x (g[0] = 0) */
} ia_params;

typedef struct ia_solve_stats {
    int iters;             /* V-cycles (REFERENCE) or outer refinement steps (REFINED) */
    double r2;             /* final Σr² of the FP32/T solve (REFERENCE) */
    double rel;            /* final FP64 relative residual (REFINED/LEARNED), else 0 */
    double ms;             /* device time of the solve, milliseconds */
} ia_solve_stats;

typedef struct ia_step_stats {
    double dt_next;        /* min(MPCFL, 1.2 dt)  (flow.jl:108) */
    int status;            /* worst status over the step's sweeps */
    int fill_warnings;     /* number of sweeps that reported a fill error */
    ia_solve_stats solve[2]; /* predictor, corrector projections */
    double vol;            /* Σf over the interior after the call, accumulated in FP64 on the device */
    double fmin, fmax;     /* min and max of f over the interior after the call (of the final
                              field; reportFillError's findmax/findmin, advection.jl:146-147,
                              see each sweep's field before cleanWisp!) */
} ia_step_stats;

/* TwoPhaseSimulation / cVOF construction (InterfaceAdvection.jl:67-85,
 * cVOF.jl:45-89, WaterLily Flow/MultiLevelPoisson): allocates every field
 * with the reference's initial values (u = u_BC, f = 1, dρ = 1, μ₀ = 1 + BC). */
int ia_create(const ia_params *p, ia_ctx **out);
void ia_destroy(ia_ctx *c);

/* Copy a whole field in/out (nbytes must equal the field's size); the other
 * buffer may be host memory or device memory (unified addressing). */
int ia_set_field(ia_ctx *c, int field, const void *host, size_t nbytes);
int ia_get_field(ia_ctx *c, int field, void *host, size_t nbytes);
/* Zero-copy access to the current device buffer of a field. */
int ia_field_device_ptr(ia_ctx *c, int field, void **dptr, size_t *nbytes);
/* Zero-copy adoption of a caller's device buffer (on the context's device,
 * nbytes = the field's size) as the storage of cVOF.f, cVOF.f⁰, Flow.u,
 * Flow.u⁰ or Flow.p from now on: a host whose arrays already live on the GPU
 * (the reference's ROCArrays, InterfaceAdvection.jl:59-60 `mem`) hands them
 * over instead of copying.  The buffer's contents become the field's value;
 * every later call returns with the field's current value in it (the sweeps'
 * f/f⁰ rotation ends with device copies back into attached buffers).  The
 * library never frees it: the caller keeps it alive until ia_destroy.  No
 * reference counterpart (the reference's arrays are its own). */
int ia_attach_field(ia_ctx *c, int field, void *dptr, size_t nbytes);
/* Sizes of MultiLevelPoisson level l (ghost-inclusive); returns the level count. */
int ia_mg_levels(ia_ctx *c, int l, int *Ng);

/* applyVOF!(f,α,n̂,sdf) + cleanWisp! + BCf! + f⁰ = copy(f)  (VOFutil.jl:8-37,
 * cVOF.jl:64-65).  `sdf` holds, per interior cell in x-fastest order, the
 * 1+2D samples [sdf(x), sdf(x+Δx e1), sdf(x-Δx e1), ...] at x = I .- 1.5,
 * Δx = 0.01 (the SDF is a host closure, evaluated by the caller). */
int ia_apply_vof(ia_ctx *c, const void *sdf, size_t nbytes);
/* WaterLily Flow: apply!(uλ,u); BC!(u); u⁰ = copy(u).  `u` is the full
 * ghost-inclusive field of uλ(i, loc(i,I)) samples. */
int ia_set_initial_velocity(ia_ctx *c, const void *u, size_t nbytes);

/* MPCFL(flow, intf)  (flow.jl:262-281). */
int ia_cfl(ia_ctx *c, double *dt);

/* sim_step!(::TwoPhaseSimulation) -> MPFMomStep!(flow,pois,intf,body; δt=dt)
 * (InterfaceAdvection.jl:100-103, flow.jl:60-109).  step_index = length(Δt)
 * selects the sweep order (flow.jl:163).  stats->dt_next is the Δt to push. */
int ia_mom_step(ia_ctx *c, double dt, int step_index, ia_step_stats *stats);

/* advect!(a::Flow, c::cVOF, f=c.f, u¹=a.u⁰, u²=a.u, dt=last(a.Δt))
 * (advection.jl:17-23, :34-78): pure cVOF advection of the field `f` with
 * the average of `u1` and `u2`.  f is IA_FIELD_F or IA_FIELD_F0 (advected in
 * place; the other one is untouched); u1 and u2 are any of IA_FIELD_U,
 * IA_FIELD_U0, IA_FIELD_FORCE, IA_FIELD_RHOU (a caller that advects with its
 * own velocity uploads it into one of them, or writes it through
 * ia_field_device_ptr).  The reference's defaults are
 * ia_advect(c, IA_FIELD_F, IA_FIELD_U0, IA_FIELD_U, dt, length(Δt), &st).
 * stats->vol/fmin/fmax describe the advected field; dt_next = dt. */
int ia_advect(ia_ctx *c, int f, int u1, int u2, double dt, int step_index, ia_step_stats *stats);

/* Stage entry points (each one reference routine, for callers that drive the
 * step themselves and for stage-level parity tests). */
/* reconstructInterface!(f,α,n̂,normalScheme)  normalEstimation.jl:10-28 */
int ia_reconstruct(ia_ctx *c);
/* c̄ = f≥½; ρuf = 0; advectVOF1d!(f,…,u⁰,u,dt,d)  advection.jl:80-89 (d is 1-based) */
int ia_vof_sweep(ia_ctx *c, int d, double dt, ia_step_stats *stats);
/* viscSurfTenρu!(Flow.f, u, …, f or f⁰)  flow.jl:113-139 */
int ia_visc(ia_ctx *c, int use_f0);
/* fill!(μ₀,1); updateL!(μ₀, f or f⁰)  flow.jl:254-259 */
int ia_update_L(ia_ctx *c, int use_f0);
/* update!(::MultiLevelPoisson)  (WaterLily; flow.jl:81) */
int ia_mg_update(ia_ctx *c);
/* myproject!(flow,pois,w) with dt_w = w*Δt, then BC!(u)  flow.jl:328-347, :82 */
int ia_project(ia_ctx *c, double dt_w, ia_solve_stats *stats);

/* Diagnostics of metrics.jl:15-50 integrated over the interior, in FP64:
 * out[0] = Σ ρkeI(I,u,f,λρ,U), out[1..3] = Σ ρuI(i,I,u,f,λρ,U), out[4] = Σ ρgh(I,g,f,λρ,StatWL),
 * out[5] = Σ EnsI(I,ω) with ω_i = ∂_j u_k - ∂_k u_j by backward differences at
 * the cell's lower edges (2D: ω = ∂_x v - ∂_y u), out[6] = Σ f.  U and statwl may
 * be NULL (zero). */
int ia_metrics(ia_ctx *c, const double *U, const double *statwl, double *out);

/* ia_metrics without any allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight time-stepping or serving loops, harnesses that forbid hipMalloc).
 * Same integrals as ia_metrics, summed in another fixed order (equal to
 * round-off).  Its partial sums live in a static device array of the library
 * (1024 block partials x 7 + 8 doubles, 57 KiB per GPU, held until the
 * process ends); any grid size folds into those 1024 partials.  Every call on
 * one GPU shares that array: calls that overlap on one GPU (two threads, two
 * contexts on one device) race; serialise them. */
int ia_metrics_noalloc(ia_ctx *c, const double *U, const double *statwl, double *out);

/* LevelSet(sim)  redistaning.jl:14-25: ϕ = 2f - 1 on every entry of cVOF.f⁰
 * (IA_FIELD_F0 then reads ϕ) and ϕini = ϕ in a scratch buffer of the context
 * (the reference's cVOF.fᶠ), valid until f is next advected or re-applied. */
int ia_levelset(ia_ctx *c);
/* redistaning!(ls; d, dτ, perdir)  redistaning.jl:44-56 (computeL! :67-108):
 * round(d/dτ) SSP-RK3 pseudo-time steps of ∂τϕ = ϕini (1 - |∇ϕ|) on ϕ (f⁰),
 * BCf!(ϕ; perdir) after every stage.  perdir is a bit mask (bit a-1: axis a
 * periodic), passed separately as in the reference.  cVOF.α and Flow.σ
 * (the reference's ϕ⁰ and L) are left holding scratch.  Needs ia_levelset
 * since the last advection and at least 2 interior cells per axis. */
int ia_redistance(ia_ctx *c, double d, double dtau, int perdir);

/* Learned Poisson correction (no counterpart in the reference).  A five-layer
 * 3x3x3 convolutional network (3 -> 16 -> 16 -> 16 -> 16 -> 2 channels, ReLU
 * between layers), evaluated with BF16 matrix-core products and FP32
 * accumulation.  Inputs per cell: f, z/rms(z), r0/rms(r0) with r0 = z - A x0
 * the residual of the warm start; the warm start becomes
 * x0 + iD·(2/3 r0 + rms(r0)·(out0 + 2x2x2-block-average(out1))), a damped
 * Jacobi step plus the network's correction in the same scaled units.  Trained on the device
 * from the solver's own operator: loss ‖z - A x‖² / ‖r0‖², one Adam step per
 * projection while training is on.
 * Parameters: IA_LEARNED_NPAR floats, W[layer][tap][ci][co] (tap =
 * (dz+1)*9 + (dy+1)*3 + dx+1, 16x16 channels per tap) for the 5 layers, then
 * the 5 bias vectors of 16.  Input channels >= 3 of layer 0 and output
 * channels >= 2 of layer 4 are unused and stay zero. */
#define IA_LEARNED_NPAR (5 * (27 * 16 * 16 + 16))
typedef struct ia_learned_stats {
    int projections;       /* projections that evaluated the network */
    int accepted;          /* ... whose correction lowered the residual */
    int train_steps;       /* Adam steps taken */
    double last_ratio;     /* ‖r‖² after / before the last correction */
} ia_learned_info;
/* He-normal initialisation from `seed` (3D FP32 contexts). */
int ia_learned_init(ia_ctx *c, unsigned seed);
int ia_learned_set_weights(ia_ctx *c, const float *w, size_t n);
int ia_learned_get_weights(ia_ctx *c, float *w, size_t n);
/* Train (one Adam step per LEARNED projection) while on != 0. */
int ia_learned_train(ia_ctx *c, int on, double lr);
int ia_learned_stats(ia_ctx *c, ia_learned_info *s);
/* The network alone on host features (3 planes of the interior, x fastest)
 * -> 2 output planes; and the parameter gradient for a given output
 * gradient (IA_LEARNED_NPAR floats).  For numerical checks. */
int ia_learned_forward(ia_ctx *c, const float *feat, float *out);
int ia_learned_grad(ia_ctx *c, const float *feat, const float *dout, float *gw);

/* Kernel-class timing with HIP events on the context's stream (no reference
 * counterpart; used by the benchmark to measure one kernel's launches inside
 * a timed region).  Enabling resets the record; MG classes time level 0 only,
 * the CG1 classes the coarse level they name, with one event pair around each
 * run of back-to-back step launches (launches counts the launches, total_ms
 * the bracketed time, gaps between those launches included). */
enum ia_prof_class {
    IA_PROF_VOF = 0,          /* cVOF sweep (face pass + cell pass, one record per sweep) */
    IA_PROF_MOMENTUM = 1,     /* CMOM/SynDRoM sweep (face pass + cell pass, one record per sweep) */
    IA_PROF_VISC_UPDATE = 2,  /* fused viscosity + RK update + coefficients */
    IA_PROF_PCG_AE = 3,       /* pcg z = Aϵ, finest level */
    IA_PROF_PCG_UPDATE = 4,   /* pcg x,r,z update, finest level */
    IA_PROF_INCREMENT = 5,    /* increment! r -= Aϵ, x += ϵ, finest level (unfused) */
    IA_PROF_LOAD_INC = 6,     /* FCG residual load + Jacobi! + increment!, finest level */
    IA_PROF_PROLONG_INC = 7,  /* prolongate! + increment! + pcg! ϵ = r·iD, finest level */
    IA_PROF_FCG_AP = 8,       /* FCG p = w + βp, q = A p (FP64) */
    IA_PROF_FCG_UPDATE = 9,   /* FCG x += αp, r -= αq (FP64) */
    IA_PROF_PCG_LAST = 10,    /* the flexible CG's fine smoother fused with its dots: two Chebyshev steps
                                 (default), or the last finest-level pcg! update (IA_FCG_SMOOTH set) */
    IA_PROF_REDIST = 11,      /* one level-set reinitialisation stage (computeL! + update + BCf!) */
    IA_PROF_CG1_L1 = 12,      /* single-reduction CG step of the FCG preconditioner, coarse level 1 */
    IA_PROF_CG1_L2 = 13,      /* ... level 2 */
    IA_PROF_CG1_L3 = 14,      /* ... level 3 */
    IA_PROF_NCLASS = 15
};
int ia_profile_enable(ia_ctx *c, int kernel, int on);
int ia_profile_read(ia_ctx *c, int kernel, int *launches, double *total_ms);

int ia_synchronize(ia_ctx *c);
const char *ia_last_error(ia_ctx *c);

#ifdef __cplusplus
}
#endif
#endif
