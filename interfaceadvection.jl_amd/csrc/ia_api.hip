// ia_api.hip -- context, step orchestration and the C ABI (include/ia.h).
//
// The host side only sequences kernels on the context's stream; every field
// lives in HBM for the life of the context.  Scalars cross to the host once
// per pressure V-cycle (the solver's stopping test) and once per step (CFL
// and fill flags).
#include "ia.h"
#include "ia_poisson.h"
#include "ia_learned.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

using namespace ia;

// ---------------------------------------------------------------------------
// Scratch of the allocation-free diagnostics entry point (ia_metrics_noalloc):
// the per-block partial sums of the metrics integrals, then the 7 totals.
// Every GPU the library runs on holds one copy until the process ends; in
// host code `scratch` is only a symbol, and its device address comes from
// hipGetSymbolAddress, once per device, at context creation (noalloc_ptr).
// ---------------------------------------------------------------------------
#define IA_NOALLOC_BLOCKS 1024  // partial sums: at most this many blocks of NT threads
#define IA_NOALLOC_DOUBLES (IA_NOALLOC_BLOCKS * 7 + 8)
static __device__ __align__(256) double scratch[IA_NOALLOC_DOUBLES];
namespace {
constexpr int NOALLOC_MAXDEV = 64;
double *noalloc_dptr[NOALLOC_MAXDEV] = {};
std::mutex noalloc_mu;
// the current device's address of `scratch` (the caller has made c->dev current)
int noalloc_setup(int dev, const char **why) {
    if (dev < 0 || dev >= NOALLOC_MAXDEV) {
        *why = "device ordinal out of range for the no-allocation scratch";
        return IA_ERR_BADARG;
    }
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (noalloc_dptr[dev]) return IA_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        *why = hipGetErrorString(e);
        return IA_ERR_HIP;
    }
    noalloc_dptr[dev] = (double *)ptr;
    return IA_OK;
}
}  // namespace

namespace {

struct Level {
    Grid g;
    Tiling t;
    CTiling ct;
    void *L = nullptr, *Dg = nullptr, *iD = nullptr, *x = nullptr, *e = nullptr, *r = nullptr, *z = nullptr;
    void *e2 = nullptr, *z2 = nullptr;  // second ϵ/z pair of the fused pcg! iterations (levels >= 1)
    void *r2 = nullptr, *w = nullptr, *w2 = nullptr;  // single-reduction CG: r, w pairs (levels >= 1)
    CTiling wt;      // the single-reduction CG step's tiling on one-cell-per-thread levels (wnb threads)
    int wnb = NT;
    bool own = false;
};

}  // namespace

struct ia_ctx {
    ia_params P;
    int es = 4;
    Grid g;
    Tiling t;
    int dev = 0;
    hipStream_t s = nullptr;
    // fields
    void *fb[4] = {};   // pool of scalar buffers for f, f⁰ and sweep outputs
    int fi = 0, f0i = 1;
    void *u = nullptr, *u0 = nullptr, *ru = nullptr, *ru2 = nullptr, *ruf = nullptr, *F = nullptr;
    void *p = nullptr, *sig = nullptr, *L = nullptr, *al = nullptr, *nh = nullptr;
    void *vface = nullptr;  // swept-volume face fluxes of the current VOF sweep
    void *mface = nullptr;  // momentum face fluxes of the current sweep (components last)
    int8_t *cb = nullptr;
    Level lev[12];
    int nlev = 0;
    // reductions and device-side control
    FillPart *fpart = nullptr;
    double *part = nullptr;
    long part_n = 0;
    StepState *st = nullptr;
    PcgState *pst = nullptr;   // pcg! state slots
    FcgState *fst = nullptr;   // flexible-CG scalars
    double *pp = nullptr;      // partial-sum buffers (PP_N of pp_stride doubles)
    long pp_stride = 0;
    std::vector<double> hpart; // host copy of a partial-sum buffer
    double *dsc = nullptr;   // device scalars
    double *x64 = nullptr, *r64 = nullptr, *z64 = nullptr, *pa64 = nullptr, *pb64 = nullptr, *q64 = nullptr;
    int fine_cheb = 1;   // the finest level's smoother after each preconditioning V-cycle: two Chebyshev steps
    int smooth_it = 2;   // ... or, with IA_FCG_SMOOTH=n, n pcg! iterations
    // pcg! iterations on each coarse level inside Vcycle!: WaterLily's 6 (REFERENCE,
    // PLAIN), and fcg_coarse_its(level cells) in the flexible CG's preconditioner
    // (IA_FCG_CIT=a,b,... overrides levels 1, 2, ...; the last value repeats)
    int cit_wl[MAXLEV], cit_fcg[MAXLEV];
    int cg1_nb[MAXLEV];  // workgroup size of the coarse CG step per level (one cell per thread)
    struct Cnn {
        bool ready = false;
        void *act[CNN_L + 1] = {};   // layer inputs/outputs, BF16 channels-last
        void *grad[2] = {};          // output-gradient ping-pong
        double *xkeep = nullptr;
        float *gtmp = nullptr, *io = nullptr;
        float *w = nullptr, *m = nullptr, *v = nullptr, *gw = nullptr, *zbias = nullptr, *wpart = nullptr;
        void *ff = nullptr, *fb = nullptr;  // MFMA operand order, [layer][27][64] x 4 bf16
        long wnblk = 0;
        int train = 0;
        double lr = 1e-3;
        long step = 0;
        int projections = 0, accepted = 0;
        double last_ratio = 0;
    } cnn;
    void *st_buf = nullptr, *st_fb = nullptr;  // surface tension forcing (vector) and face-averaged f
    int ls_ini = -1;          // pool buffer holding ϕini of the LevelSet (valid until f is next advected)
    bool ls_valid = false;
    const int *halt = nullptr;   // the flexible CG's stop flag while a pipelined solve runs
    FcgState *pin = nullptr;     // pinned host copies of the solver state (two slots)
    hipEvent_t it_ev[2] = {nullptr, nullptr};
    // caller-owned device buffers adopted by ia_attach_field (nullptr: the
    // library's own); att_fi/att_f0i: the f-pool slots holding attached f, f⁰
    void *att_u = nullptr, *att_u0 = nullptr, *att_p = nullptr;
    int att_fi = -1, att_f0i = -1;
    int tail_level = 99; // first level handled by the one-workgroup tail kernel
    size_t tail_lds = 0; // its dynamic LDS bytes
    hipEvent_t e0 = nullptr, e1 = nullptr;
    struct ProfClass {
        bool on = false;
        int n = 0;         // event pairs
        long launches = 0; // kernel launches inside them
        std::vector<hipEvent_t> ev;
    } prof[IA_PROF_NCLASS];
    char msg[512] = {0};
};

namespace {

int fail(ia_ctx *c, int code, const char *fmt, ...) {
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->msg, sizeof c->msg, fmt, ap);
        va_end(ap);
    }
    return code;
}
#define HIPCHK(c, x)                                                                        \
    do {                                                                                    \
        hipError_t _e = (x);                                                                \
        if (_e != hipSuccess) return fail(c, IA_ERR_HIP, "%s: %s", #x, hipGetErrorString(_e)); \
    } while (0)
#define LAUNCHCHK(c)                                                                              \
    do {                                                                                          \
        hipError_t _e = hipGetLastError();                                                        \
        if (_e != hipSuccess) return fail(c, IA_ERR_HIP, "launch: %s", hipGetErrorString(_e));    \
    } while (0)

// HIP events around the launches of one kernel class (ia_profile_enable)
void prof_begin(ia_ctx *c, int k) {
    auto &p = c->prof[k];
    if (!p.on) return;
    while ((int)p.ev.size() < 2 * (p.n + 1)) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) { p.on = false; return; }
        p.ev.push_back(e);
    }
    (void)hipEventRecord(p.ev[2 * p.n], c->s);
}
// nlaunch: kernel launches since the matching prof_begin (a bracket around a
// run of back-to-back launches costs two events instead of two per launch)
void prof_end(ia_ctx *c, int k, int nlaunch = 1) {
    auto &p = c->prof[k];
    if (!p.on) return;
    (void)hipEventRecord(p.ev[2 * p.n + 1], c->s);
    p.n++;
    p.launches += nlaunch;
}

void grid_init(Grid &g, int D, const int *Ng, int perdir) {
    g.D = D;
    g.exitx = 0;
    for (int a = 0; a < 3; ++a) {
        g.Ng[a] = a < D ? Ng[a] : 1;
        g.per[a] = a < D ? ((perdir >> a) & 1) : 0;
    }
    g.st[0] = 1;
    g.st[1] = g.Ng[0];
    g.st[2] = (long)g.Ng[0] * g.Ng[1];
    g.ns = (long)g.Ng[0] * g.Ng[1] * g.Ng[2];
}
Tiling tiling(const Grid &g) {
    Tiling t;
    int n0 = g.Ng[0] - 2, n1 = g.D > 1 ? g.Ng[1] - 2 : 1, n2 = g.D > 2 ? g.Ng[2] - 2 : 1;
    t.nt[0] = (n0 + TX - 1) / TX;
    t.nt[1] = (n1 + TY - 1) / TY;
    t.nt[2] = n2;
    t.tz = 1;
    t.nblk = (long)t.nt[0] * t.nt[1] * t.nt[2];
    return t;
}
// narrow: a coarse level may take 32-wide rows (only the one-cell-per-thread
// kernels run there)
CTiling ctiling(const Grid &g, bool narrow = false) {
    CTiling t;
    int n0 = g.Ng[0] - 2, n1 = g.D > 1 ? g.Ng[1] - 2 : 1, n2 = g.D > 2 ? g.Ng[2] - 2 : 1;
    t.nx = TX;
    t.nt[0] = (n0 + TX - 1) / TX;
    t.nt[1] = (n1 + TY - 1) / TY;
    long base = (long)t.nt[0] * t.nt[1];
    // cells per thread column: at most 32, and no fewer workgroups than about
    // 1024 (128³: 8-cell columns, 1024 workgroups; measured faster than 4- or
    // 16-cell columns on the coarse CG there)
    int zb = 1;
    if (g.D > 2) zb = (int)std::max(1L, std::min(32L, base * n2 / 1024));
    t.zb = zb;
    t.nt[2] = g.D > 2 ? (n2 + zb - 1) / zb : 1;
    if (narrow && g.D == 3 && zb == 1 && n0 <= 32) {
        t.nx = 32;
        t.nt[0] = 1;
        t.nt[1] = (n1 + NT / 32 - 1) / (NT / 32);
        t.nt[2] = n2;
    }
    t.nblk = (long)t.nt[0] * t.nt[1] * t.nt[2];
    return t;
}
long interior(const Grid &g) {
    long n = 1;
    for (int a = 0; a < g.D; ++a) n *= g.Ng[a] - 2;
    return n;
}
inline dim3 blocks(const Tiling &t) { return dim3((unsigned)t.nblk); }
inline dim3 blocks1(long n) { return dim3((unsigned)((n + 255) / 256)); }

size_t field_elems(ia_ctx *c, int id) {
    long ns = c->g.ns, D = c->g.D;
    switch (id) {
    case IA_FIELD_F: case IA_FIELD_F0: case IA_FIELD_ALPHA: case IA_FIELD_CBAR: case IA_FIELD_P: case IA_FIELD_SIGMA:
        return ns;
    case IA_FIELD_NHAT: case IA_FIELD_RHOU: case IA_FIELD_RHOUF: case IA_FIELD_U: case IA_FIELD_U0:
    case IA_FIELD_FORCE: case IA_FIELD_MU0:
        return ns * D;
    default:
        return 0;
    }
}
void *field_buf(ia_ctx *c, int id) {
    switch (id) {
    case IA_FIELD_F: return c->fb[c->fi];
    case IA_FIELD_F0: return c->fb[c->f0i];
    case IA_FIELD_ALPHA: return c->al;
    case IA_FIELD_NHAT: return c->nh;
    case IA_FIELD_CBAR: return c->cb;
    case IA_FIELD_RHOU: return c->ru;
    case IA_FIELD_RHOUF: return c->ruf;
    case IA_FIELD_U: return c->u;
    case IA_FIELD_U0: return c->u0;
    case IA_FIELD_FORCE: return c->F;
    case IA_FIELD_P: return c->p;
    case IA_FIELD_SIGMA: return c->sig;
    case IA_FIELD_MU0: return c->L;
    default: return nullptr;
    }
}
int free_fbuf(ia_ctx *c, int a, int b, int d2) {
    for (int q = 0; q < 4; ++q)
        if (q != a && q != b && q != d2) return q;
    return -1;
}
bool is_attached(const ia_ctx *c, const void *b) {
    if (!b) return false;
    if (b == c->att_u || b == c->att_u0 || b == c->att_p) return true;
    if (c->att_fi >= 0 && b == c->fb[c->att_fi]) return true;
    if (c->att_f0i >= 0 && b == c->fb[c->att_f0i]) return true;
    return false;
}
// the sweeps rotate f and f⁰ through the four pool buffers; when the caller
// attached its own f or f⁰ buffers, move the current values back into them
// (device copies on the context's stream) so every call returns with
// cVOF.f and cVOF.f⁰ in the caller's arrays
int restore_attached(ia_ctx *c) {
    int a = c->att_fi, b = c->att_f0i, x = c->fi, y = c->f0i;
    if ((a < 0 || x == a) && (b < 0 || y == b)) return IA_OK;
    size_t bytes = (size_t)c->g.ns * c->es;
    auto cp = [&](int dst, int src) -> hipError_t {
        return hipMemcpyAsync(c->fb[dst], c->fb[src], bytes, hipMemcpyDeviceToDevice, c->s);
    };
    hipError_t e = hipSuccess;
    if (a >= 0 && b >= 0) {
        if (x == b && y == a) {  // exchanged: through a third buffer
            int t = free_fbuf(c, a, b, -1);
            if (e == hipSuccess) e = cp(t, a);
            if (e == hipSuccess) e = cp(a, b);
            if (e == hipSuccess) e = cp(b, t);
        } else if (x == b) {     // f sits in f⁰'s slot, f⁰ elsewhere
            if (e == hipSuccess) e = cp(a, b);
            if (e == hipSuccess) e = cp(b, y);
        } else if (y == a) {     // f⁰ sits in f's slot, f elsewhere
            if (e == hipSuccess) e = cp(b, a);
            if (e == hipSuccess) e = cp(a, x);
        } else {
            if (x != a && e == hipSuccess) e = cp(a, x);
            if (y != b && e == hipSuccess) e = cp(b, y);
        }
        c->fi = a;
        c->f0i = b;
    } else if (a >= 0) {
        if (y == a) {            // f⁰ moves to a free slot first
            int t = free_fbuf(c, a, x, -1);
            if (e == hipSuccess) e = cp(t, a);
            c->f0i = t;
        }
        if (e == hipSuccess) e = cp(a, x);
        c->fi = a;
    } else {
        if (x == b) {
            int t = free_fbuf(c, b, y, -1);
            if (e == hipSuccess) e = cp(t, b);
            c->fi = t;
        }
        if (e == hipSuccess) e = cp(b, y);
        c->f0i = b;
    }
    if (e != hipSuccess) return fail(c, IA_ERR_HIP, "restoring attached f/f0: %s", hipGetErrorString(e));
    return IA_OK;
}

// partial-sum buffers
enum { PP_A = 0, PP_B, PP_C, PP_MR, PP_ML2, PP_FI, PP_FZ, PP_FR, PP_FRR, PP_FD, PP_FAP, PP_FU, PP_N };

template <class T> T *P(void *v) { return (T *)v; }
template <class T> const T *P(const void *v) { return (const T *)v; }

// WaterLily BC!(a, A, saveexit) on a staggered vector field whose interior is
// current (kind 1: saveexit false; kind 2: the velocity's exitBC, the exit
// face keeping its own value)
template <class T>
__global__ void __launch_bounds__(NT) k_bcvel(Grid g, Tiling t, T *a, T d0, T d1, T d2, int kind) {
    int p[3];
    if (!thread_cell(g, t, p)) return;
    long I = lin(g, p);
    T dv[3] = {d0, d1, d2};
    for (int i = 0; i < g.D; ++i) {
        T xv = kind == 2 && i == 0 && g.exitx ? a[I + g.st[0]] : T(0);
        write_images<T>(g, a + g.ns * i, p, a[g.ns * i + I], kind, i, dv[i], xv);
    }
}

// metrics integrands, grid-stride over the tiles: every block folds the tiles
// blockIdx.x, blockIdx.x + gridDim.x, ... in that order, then writes its 7
// block sums (ia_metrics_noalloc; part = the scratch array's partials)
template <int D, class T>
__global__ void __launch_bounds__(NT) k_metrics_gs(Grid g, Tiling t, const T *u, const T *f, T lr, double U0,
                                                   double U1, double U2, double g0, double g1, double g2, double w0,
                                                   double w1, double w2, double *part) {
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    double U[3] = {U0, U1, U2}, gv[3] = {g0, g1, g2}, wl[3] = {w0, w1, w2};
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    for (long b = blockIdx.x; b < t.nblk; b += gridDim.x) {
        long bx = b % t.nt[0], r = b / t.nt[0], by = r % t.nt[1], bz = r / t.nt[1];
        int p[3];
        p[0] = 1 + (int)bx * TX + tx;
        p[1] = g.D > 1 ? 1 + (int)by * TY + ty : 0;
        p[2] = g.D > 2 ? 1 + (int)bz : 0;
        bool ok = p[0] <= g.Ng[0] - 2;
        if (g.D > 1) ok = ok && p[1] <= g.Ng[1] - 2;
        if (g.D > 2) ok = ok && p[2] <= g.Ng[2] - 2;
        if (ok) metrics_cell<D, T>(g, u, f, lr, U, gv, wl, p, acc);
    }
    for (int q = 0; q < 7; ++q) {
        double s = block_sum(acc[q]);
        if (threadIdx.x == 0) part[(long)blockIdx.x * 7 + q] = s;
    }
}
// the 7 totals of nb block partials (each thread adds its blocks in order,
// then a fixed tree)
__global__ void __launch_bounds__(NT) k_metrics_fin(const double *part, long nb, double *out) {
    for (int q = 0; q < 7; ++q) {
        double s = 0;
        for (long b = threadIdx.x; b < nb; b += NT) s += part[b * 7 + q];
        s = block_sum(s);
        if (threadIdx.x == 0) out[q] = s;
    }
}

// ---------------------------------------------------------------------------
// typed implementation
// ---------------------------------------------------------------------------
// the stop flag that view() hands to the kernels of the pipelined solve in progress
static thread_local const int *cur_halt = nullptr;
template <class T> struct Impl {
    static double epsT() { return sizeof(T) == 4 ? 1.1920928955078125e-07 : 2.220446049250313e-16; }

    static int alloc_zero(ia_ctx *c, void **p, size_t bytes) {
        HIPCHK(c, hipMalloc(p, bytes));
        HIPCHK(c, hipMemsetAsync(*p, 0, bytes, c->s));
        return IA_OK;
    }
    static int fill(ia_ctx *c, void *p, long n, T v) {
        hipLaunchKernelGGL(k_fill<T>, blocks1(n), dim3(256), 0, c->s, P<T>(p), n, v);
        LAUNCHCHK(c);
        return IA_OK;
    }
    static int bc_vel(ia_ctx *c, void *a, const T *dv, int kind = 2) {
        hipLaunchKernelGGL(k_bcvel<T>, blocks(c->t), dim3(NT), 0, c->s, c->g, c->t, P<T>(a), dv[0], dv[1], dv[2],
                           kind);
        LAUNCHCHK(c);
        return IA_OK;
    }

    static int create(ia_ctx *c) {
        Grid &g = c->g;
        size_t ns = g.ns, nv = ns * g.D, es = sizeof(T);
        for (int q = 0; q < 4; ++q)
            if (int r = alloc_zero(c, &c->fb[q], ns * es)) return r;
        void **vec[] = {&c->u, &c->u0, &c->ru, &c->ru2, &c->ruf, &c->F, &c->L, &c->nh, &c->mface};
        for (void **v : vec)
            if (int r = alloc_zero(c, v, nv * es)) return r;
        void **sca[] = {&c->p, &c->sig, &c->al, &c->vface};
        for (void **v : sca)
            if (int r = alloc_zero(c, v, ns * es)) return r;
        HIPCHK(c, hipMalloc(&c->cb, ns));
        HIPCHK(c, hipMemsetAsync(c->cb, 0, ns, c->s));
        // Flow: u = u_BC, BC!, u⁰ = u ; μ₀ = 1, BC!(μ₀, 0)
        T ub[3], z3[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a) ub[a] = (T)c->P.u_bc[a];
        for (int d = 0; d < g.D; ++d) fill(c, P<T>(c->u) + ns * d, ns, ub[d]);
        bc_vel(c, c->u, ub);
        HIPCHK(c, hipMemcpyAsync(c->u0, c->u, nv * es, hipMemcpyDeviceToDevice, c->s));
        fill(c, c->L, nv, T(1));
        bc_vel(c, c->L, z3, 1);
        // cVOF: f = 1, f⁰ = 1
        fill(c, c->fb[0], ns, T(1));
        fill(c, c->fb[1], ns, T(1));
        c->fi = 0;
        c->f0i = 1;
        // MultiLevelPoisson(p, μ₀, σ)
        Level &l0 = c->lev[0];
        l0.g = g;
        l0.t = c->t;
        l0.ct = ctiling(g);
        l0.L = c->L;
        l0.x = c->p;
        l0.z = c->sig;
        void **l0b[] = {&l0.Dg, &l0.iD, &l0.e, &l0.r};
        for (void **v : l0b)
            if (int r = alloc_zero(c, v, ns * es)) return r;
        c->nlev = 1;
        long maxblk = c->t.nblk;
        for (;;) {
            Grid &gp = c->lev[c->nlev - 1].g;
            bool ok = true;
            for (int a = 0; a < g.D; ++a) ok = ok && gp.Ng[a] % 2 == 0 && gp.Ng[a] > 4;
            if (!(ok && c->nlev <= 10)) break;
            int Na[3];
            for (int a = 0; a < 3; ++a) Na[a] = a < g.D ? 1 + gp.Ng[a] / 2 : 1;
            Level &lv = c->lev[c->nlev];
            grid_init(lv.g, g.D, Na, c->P.perdir);
            lv.t = tiling(lv.g);
            lv.ct = ctiling(lv.g, true);
            lv.wt = lv.ct;
            lv.wnb = c->cg1_nb[c->nlev];
            if (lv.ct.zb == 1 && lv.g.D == 3) {
                int n1 = lv.g.D > 1 ? lv.g.Ng[1] - 2 : 1, rows = lv.wnb / lv.wt.nx;
                lv.wt.nt[1] = (n1 + rows - 1) / rows;
                lv.wt.nblk = (long)lv.wt.nt[0] * lv.wt.nt[1] * lv.wt.nt[2];
            }
            size_t n = lv.g.ns;
            if (int r = alloc_zero(c, &lv.L, n * g.D * es)) return r;
            void **lb[] = {&lv.Dg, &lv.iD, &lv.e, &lv.r, &lv.x, &lv.z, &lv.e2, &lv.z2, &lv.r2, &lv.w, &lv.w2};
            for (void **v : lb)
                if (int r = alloc_zero(c, v, n * es)) return r;
            lv.own = true;
            maxblk = std::max(maxblk, lv.t.nblk);
            c->nlev++;
        }
        if (c->nlev <= 2) return fail(c, IA_ERR_BADARG, "MultiLevelPoisson requires size=a2^n, where n>2");
        // the coarsest levels run in one workgroup with every array in LDS:
        // start at the first level of at most 4096 cells whose levels fit
        {
            Grid gs[MAXLEV];
            for (int l = 0; l < c->nlev && l < MAXLEV; ++l) gs[l] = c->lev[l].g;
            c->tail_level = c->nlev;
            for (int l = 1; l < c->nlev && l < MAXLEV; ++l)
                if (interior(c->lev[l].g) <= TAIL_KC * NT1 && tail_lds_bytes<T>(gs, l, c->nlev) <= TAIL_LDS_MAX) {
                    c->tail_level = l;
                    break;
                }
            c->tail_lds = c->tail_level < c->nlev ? tail_lds_bytes<T>(gs, c->tail_level, c->nlev) : 0;
            if (c->tail_lds > 0)
                HIPCHK(c, hipFuncSetAttribute((const void *)k_mg_tail<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)c->tail_lds));
        }
        c->part_n = 5 * maxblk + 64;
        HIPCHK(c, hipMalloc(&c->part, c->part_n * sizeof(double)));
        HIPCHK(c, hipMalloc(&c->fpart, (maxblk + FILL_R2) * sizeof(FillPart)));
        HIPCHK(c, hipMalloc(&c->st, sizeof(StepState)));
        HIPCHK(c, hipMalloc(&c->pst, PCG_SLOTS * sizeof(PcgState)));
        HIPCHK(c, hipMemsetAsync(c->pst, 0, PCG_SLOTS * sizeof(PcgState), c->s));
        HIPCHK(c, hipMalloc(&c->fst, sizeof(FcgState)));
        HIPCHK(c, hipMemsetAsync(c->fst, 0, sizeof(FcgState), c->s));
        HIPCHK(c, hipHostMalloc((void **)&c->pin, 2 * sizeof(FcgState)));
        for (auto &ev : c->it_ev) HIPCHK(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        {
            long mb = 0;
            for (int l = 0; l < c->nlev; ++l) mb = std::max(mb, c->lev[l].ct.nblk);
            c->pp_stride = 2 * mb + 64;
            HIPCHK(c, hipMalloc(&c->pp, PP_N * c->pp_stride * sizeof(double)));
            HIPCHK(c, hipMemsetAsync(c->pp, 0, PP_N * c->pp_stride * sizeof(double), c->s));
        }
        HIPCHK(c, hipMalloc(&c->dsc, 64 * sizeof(double)));
        HIPCHK(c, hipMemsetAsync(c->dsc, 0, 64 * sizeof(double), c->s));
        if (int r = mg_update(c)) return r;
        return IA_OK;
    }

    // ----------------------------------------------------------------- MG
    static int mg_update(ia_ctx *c) {
        for (int l = 0; l < c->nlev; ++l) {
            Level &lv = c->lev[l];
            if (l > 0) {
                Level &lf = c->lev[l - 1];
                hipLaunchKernelGGL(k_restrictL<T>, blocks(lv.t), dim3(NT), 0, c->s, lv.g, lv.t, lf.g, P<T>(lf.L),
                                   P<T>(lv.L));
            }
            hipLaunchKernelGGL(k_set_diag<T>, blocks(lv.t), dim3(NT), 0, c->s, lv.g, lv.t, P<T>(lv.L), P<T>(lv.Dg),
                               P<T>(lv.iD));
        }
        LAUNCHCHK(c);
        return IA_OK;
    }
    static LevV<T> view(Level &l) {
        LevV<T> v;
        v.halt = cur_halt;
        v.g = l.g;
        v.t = l.ct;
        v.L = P<T>(l.L);
        v.Dg = P<T>(l.Dg);
        v.iD = P<T>(l.iD);
        v.x = P<T>(l.x);
        v.e = P<T>(l.e);
        v.r = P<T>(l.r);
        v.z = P<T>(l.z);
        return v;
    }
    static dim3 cb(const Level &l) { return dim3((unsigned)l.ct.nblk); }
    static void jacobi(ia_ctx *c, Level &l) {
        hipLaunchKernelGGL(k_mg_jacobi<T>, cb(l), dim3(NT), 0, c->s, view(l));
        increment(c, l);
    }
    static void increment(ia_ctx *c, Level &l) {
        bool f0 = &l == &c->lev[0];
        if (f0) prof_begin(c, IA_PROF_INCREMENT);
        hipLaunchKernelGGL(k_mg_increment<T>, cb(l), dim3(NT), 0, c->s, view(l));
        if (f0) prof_end(c, IA_PROF_INCREMENT);
    }
    static double *pp(ia_ctx *c, int k) { return c->pp + (long)k * c->pp_stride; }
    // WaterLily pcg!(p; it).  init_done: ϵ = r·iD and the partials of ρ₀ (PP_A)
    // were produced by the preceding fused prolongate!/increment!.  last: the
    // final update is fused with the FCG dots of the preconditioned residual.
    static void pcg(ia_ctx *c, Level &l, int it = 6, bool init_done = false, const FcgV<T> *last = nullptr,
                    int first = 0) {
        LevV<T> v = view(l);
        bool f0 = &l == &c->lev[0];
        dim3 B = cb(l);
        if (!init_done) hipLaunchKernelGGL(k_pcg_init<T>, B, dim3(NT), 0, c->s, v, pp(c, PP_A));
        // levels >= 1 fuse each ϵ update into the next z = Aϵ, alternating
        // between the (e, z) and (e2, z2) buffer pairs
        bool pair = l.e2 != nullptr && it > 1;
        for (int i = 1; i <= it; ++i) {
            int sa = 3 * (i - 1), su = sa + 1, se = sa + 2;
            int sin = i == 1 ? -1 : sa - 1;
            if (i > 1 && pair) {
                LevV<T> w = v;
                w.e = (i % 2 == 0) ? P<T>(l.e2) : P<T>(l.e);
                w.z = (i % 2 == 0) ? P<T>(l.z2) : P<T>(l.z);
                hipLaunchKernelGGL(k_pcg_epsAe<T>, B, dim3(NT), 0, c->s, v, w.e, w.z, c->pst, sa - 2, sa, pp(c, PP_C),
                                   pp(c, PP_B));
                v = w;
            } else {
                if (f0) prof_begin(c, IA_PROF_PCG_AE);
                hipLaunchKernelGGL(k_pcg_Ae<T>, B, dim3(NT), 0, c->s, v, c->pst, sin, sa, pp(c, PP_A), pp(c, PP_B));
                if (f0) prof_end(c, IA_PROF_PCG_AE);
            }
            if (i == it && last) {
                if (f0) prof_begin(c, IA_PROF_PCG_LAST);
                hipLaunchKernelGGL(k_pcg_last_dots<T>, B, dim3(NT), 0, c->s, v, *last, c->pst, sa, su, pp(c, PP_B),
                                   pp(c, PP_FD), first);
                if (f0) prof_end(c, IA_PROF_PCG_LAST);
                continue;
            }
            if (f0) prof_begin(c, IA_PROF_PCG_UPDATE);
            hipLaunchKernelGGL(k_pcg_update<T>, B, dim3(NT), 0, c->s, v, c->pst, sa, su, pp(c, PP_B), pp(c, PP_C), i, it);
            if (f0) prof_end(c, IA_PROF_PCG_UPDATE);
            if (i < it && !pair) hipLaunchKernelGGL(k_pcg_eps<T>, B, dim3(NT), 0, c->s, v, c->pst, su, se, pp(c, PP_C));
        }
    }
    // Vcycle!(ml; l); the levels from c->tail_level down run in one workgroup.
    // pre_done: Jacobi! + increment! of level l were fused into the caller's
    // residual load.  then_pcg: a pcg! on level l follows, so its first
    // statement is fused into the final prolongate! + increment!.
    // the flexible CG preconditioner's coarse solve: single-reduction CG, one
    // launch per iteration (k_cg1_init, k_cg1_step).  The final r may sit in
    // r2; nothing reads a coarse r after its solve (restrict! rewrites it).
    static void pcg1(ia_ctx *c, Level &l, int it) {
        LevV<T> v = view(l);
        dim3 B = cb(l);
        T *r[2] = {P<T>(l.r), P<T>(l.r2)}, *w[2] = {P<T>(l.w), P<T>(l.w2)}, *s[2] = {P<T>(l.z), P<T>(l.z2)};
        hipLaunchKernelGGL(k_cg1_init<T>, B, dim3(NT), 0, c->s, v, w[0], c->pst, pp(c, PP_A));
        int lv = (int)(&l - c->lev), pc = lv >= 1 && lv <= 3 ? IA_PROF_CG1_L1 + lv - 1 : -1;
        if (pc >= 0) prof_begin(c, pc);
        for (int q = 1; q <= it; ++q) {
            int a = (q - 1) & 1;
            Cg1V<T> b{r[a], w[a], s[a], r[1 - a], w[1 - a], s[1 - a]};
            double *pin = pp(c, (q & 1) ? PP_A : PP_B), *pout = pp(c, (q & 1) ? PP_B : PP_A);
            if (l.ct.zb == 1) {
                LevV<T> vw = v;
                vw.t = l.wt;
                long nin = q == 1 ? l.ct.nblk : l.wt.nblk;
                dim3 W((unsigned)l.wt.nblk);
                if (l.g.D == 2)
                    hipLaunchKernelGGL((k_cg1_step1<2, T>), B, dim3(NT), 0, c->s, v, b, c->pst, q, it, pin, nin, pout);
                else if (l.wnb == 1024)
                    hipLaunchKernelGGL((k_cg1_step1<3, T, 1024>), W, dim3(1024), 0, c->s, vw, b, c->pst, q, it, pin,
                                       nin, pout);
                else if (l.wnb == 128)
                    hipLaunchKernelGGL((k_cg1_step1<3, T, 128>), W, dim3(128), 0, c->s, vw, b, c->pst, q, it, pin, nin,
                                       pout);
                else if (l.wnb == 512)
                    hipLaunchKernelGGL((k_cg1_step1<3, T, 512>), W, dim3(512), 0, c->s, vw, b, c->pst, q, it, pin, nin,
                                       pout);
                else
                    hipLaunchKernelGGL((k_cg1_step1<3, T>), W, dim3(NT), 0, c->s, vw, b, c->pst, q, it, pin, nin,
                                       pout);
            } else {
                hipLaunchKernelGGL(k_cg1_step<T>, B, dim3(NT), 0, c->s, v, b, c->pst, q, it, pin, pout);
            }
        }
        if (pc >= 0) prof_end(c, pc, it);
    }
    // e_done: Jacobi!'s ϵ = r·iD on level l was written by the restrict! that
    // produced its r, so only increment! remains.
    static void vcycle(ia_ctx *c, int l, const int *cit, bool pre_done = false, bool then_pcg = false,
                       bool e_done = false) {
        Level &fl = c->lev[l], &co = c->lev[l + 1];
        if (!pre_done) {
            if (e_done) increment(c, fl);
            else jacobi(c, fl);
        }
        bool rec = l + 2 < c->nlev, co_jacobi = l + 1 < c->tail_level && rec;
        hipLaunchKernelGGL(k_mg_restrict<T>, cb(co), dim3(NT), 0, c->s, view(co), fl.g, P<T>(fl.r),
                           co_jacobi ? 1 : 0);
        if (l + 1 >= c->tail_level) {
            TailArgs<T> a;
            for (int q = 0; q < c->nlev && q < MAXLEV; ++q) {
                Level &lv = c->lev[q];
                a.lv[q].g = lv.g;
                a.lv[q].L = P<T>(lv.L);
                a.lv[q].Dg = P<T>(lv.Dg);
                a.lv[q].iD = P<T>(lv.iD);
                a.lv[q].x = P<T>(lv.x);
                a.lv[q].e = P<T>(lv.e);
                a.lv[q].r = P<T>(lv.r);
                a.lv[q].z = P<T>(lv.z);
                a.lv[q].n = (int)interior(lv.g);
            }
            a.l0 = l + 1;
            a.nlev = c->nlev;
            for (int q = 0; q < MAXLEV; ++q) a.cit[q] = cit[q];
            a.halt = cur_halt;
            hipLaunchKernelGGL(k_mg_tail<T>, dim3(1), dim3(NT1), c->tail_lds, c->s, a);
        } else {
            if (rec) vcycle(c, l + 1, cit, false, true, true);
            if (cit == c->cit_fcg) pcg1(c, co, cit[l + 1]);
            else pcg(c, co, cit[l + 1], rec);
        }
        if (then_pcg) {
            bool f0 = l == 0;
            if (f0) prof_begin(c, IA_PROF_PROLONG_INC);
            hipLaunchKernelGGL(k_mg_prolong_inc_init<T>, cb(fl), dim3(NT), 0, c->s, view(fl), co.g, P<T>(co.x),
                               pp(c, PP_A));
            if (f0) prof_end(c, IA_PROF_PROLONG_INC);
        } else {
            hipLaunchKernelGGL(k_mg_prolong<T>, cb(fl), dim3(NT), 0, c->s, view(fl), co.g, P<T>(co.x));
            increment(c, fl);
        }
    }
    // fixed-order host sum of one partial-sum buffer (K values per block)
    static int host_total(ia_ctx *c, int buf, long nblk, int K, double *out) {
        c->hpart.resize(K * nblk);
        HIPCHK(c, hipMemcpyAsync(c->hpart.data(), pp(c, buf), K * nblk * sizeof(double), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        for (int k = 0; k < K; ++k) {
            double s = 0;
            for (long q = 0; q < nblk; ++q) s += c->hpart[K * q + k];
            out[k] = s;
        }
        return IA_OK;
    }
    // WaterLily solver!(ml; tol, itmx)
    static int solver(ia_ctx *c, double tol, int itmx, int *iters, double *r2o) {
        Level &l0 = c->lev[0];
        long ncell = interior(l0.g), nb = l0.ct.nblk;
        hipLaunchKernelGGL(k_mg_resid<T>, cb(l0), dim3(NT), 0, c->s, view(l0), pp(c, PP_MR));
        hipLaunchKernelGGL(k_mg_center_l2<T>, cb(l0), dim3(NT), 0, c->s, view(l0), pp(c, PP_MR), ncell, 1,
                           pp(c, PP_ML2));
        LAUNCHCHK(c);
        double r2;
        if (int e = host_total(c, PP_ML2, nb, 1, &r2)) return e;
        r2 = (double)(T)r2;
        int n = 0;
        while (r2 > tol && n < itmx) {
            vcycle(c, 0, c->cit_wl, false, true);
            pcg(c, l0, 6, true);
            hipLaunchKernelGGL(k_mg_center_l2<T>, cb(l0), dim3(NT), 0, c->s, view(l0), pp(c, PP_MR), ncell, 0,
                               pp(c, PP_ML2));
            LAUNCHCHK(c);
            if (int e = host_total(c, PP_ML2, nb, 1, &r2)) return e;
            r2 = (double)(T)r2;
            n++;
        }
        hipLaunchKernelGGL(k_mg_perbc<T>, cb(l0), dim3(NT), 0, c->s, view(l0), P<T>(l0.x));
        LAUNCHCHK(c);
        *iters = n;
        *r2o = r2;
        return IA_OK;
    }
    // myproject! (flow.jl:328-347) + BC!(u), REFERENCE mode
    static int project_ref(ia_ctx *c, void *u, double dtw_d, ia_solve_stats *ss) {
        Grid &g = c->g;
        T dtw = (T)dtw_d;
        hipLaunchKernelGGL(k_div<T>, blocks(c->t), dim3(NT), 0, c->s, g, c->t, P<T>(u), P<T>(c->sig));
        hipLaunchKernelGGL(k_scale<T>, blocks1(g.ns), dim3(256), 0, c->s, P<T>(c->p), g.ns, dtw);
        int it = 0;
        double r2 = 0;
        if (int e = solver(c, c->P.poisson_tol, c->P.poisson_itmx, &it, &r2)) return e;
        hipLaunchKernelGGL(k_project<T>, blocks(c->t), dim3(NT), 0, c->s, g, c->t, P<T>(u), P<T>(c->L), P<T>(c->p),
                           (T)c->P.u_bc[0], (T)c->P.u_bc[1], (T)c->P.u_bc[2]);
        hipLaunchKernelGGL(k_scale<T>, blocks1(g.ns), dim3(256), 0, c->s, P<T>(c->p), g.ns, T(1) / dtw);
        LAUNCHCHK(c);
        if (ss) {
            ss->iters = it;
            ss->r2 = r2;
            ss->rel = 0;
        }
        return IA_OK;
    }
    // inproject!(a, b::Poisson, dt) + myproject! + BC!(u) (flow.jl:300-341, :82):
    // psolver!(b; tol = 50eps(T), itmx = 2000) on the single-level Poisson
    static int project_pcg(ia_ctx *c, void *u, double dtw_d, ia_solve_stats *ss) {
        Grid &g = c->g;
        Level &l0 = c->lev[0];
        T dtw = (T)dtw_d;
        long ncell = interior(g), nb = l0.ct.nblk;
        // fill!(b.z, 0); fill!(b.ϵ, 0); fill!(b.r, 0); z = div u; x *= dt
        fill(c, c->sig, g.ns, T(0));
        fill(c, l0.e, g.ns, T(0));
        fill(c, l0.r, g.ns, T(0));
        hipLaunchKernelGGL(k_div<T>, blocks(c->t), dim3(NT), 0, c->s, g, c->t, P<T>(u), P<T>(c->sig));
        hipLaunchKernelGGL(k_scale<T>, blocks1(g.ns), dim3(256), 0, c->s, P<T>(c->p), g.ns, dtw);
        // residual!(p); r₂ = L₂(p)   (x's periodic images are read through the wrapped offsets)
        hipLaunchKernelGGL(k_mg_resid<T>, cb(l0), dim3(NT), 0, c->s, view(l0), pp(c, PP_MR));
        hipLaunchKernelGGL(k_mg_center_l2<T>, cb(l0), dim3(NT), 0, c->s, view(l0), pp(c, PP_MR), ncell, 1,
                           pp(c, PP_ML2));
        // z = ϵ = r·iD, ρ = r·z
        hipLaunchKernelGGL(k_pcg_init<T>, cb(l0), dim3(NT), 0, c->s, view(l0), pp(c, PP_A));
        double *rho = c->dsc + 20;  // two slots: 20, 21
        hipLaunchKernelGGL(k_ps_rho<T>, dim3(1), dim3(NT), 0, c->s, pp(c, PP_A), nb, rho);
        LAUNCHCHK(c);
        double r2;
        if (int e = host_total(c, PP_ML2, nb, 1, &r2)) return e;
        r2 = (double)(T)r2;
        const double tol = (double)(T)(50 * epsT());
        const int itmx = 2000;
        int n = 0;
        while ((r2 > tol || (r2 > tol / 4 && n == 0)) && n < itmx) {
            LevV<T> v = view(l0);
            hipLaunchKernelGGL(k_ps_Ae<T>, cb(l0), dim3(NT), 0, c->s, v, pp(c, PP_B));
            hipLaunchKernelGGL(k_ps_update<T>, cb(l0), dim3(NT), 0, c->s, v, rho + (n & 1), pp(c, PP_B), pp(c, PP_C));
            hipLaunchKernelGGL(k_ps_eps<T>, cb(l0), dim3(NT), 0, c->s, v, rho + (n & 1), rho + ((n + 1) & 1),
                               pp(c, PP_C));
            LAUNCHCHK(c);
            double tot[2];
            if (int e = host_total(c, PP_C, nb, 2, tot)) return e;
            r2 = (double)(T)tot[1];
            n++;
        }
        hipLaunchKernelGGL(k_mg_perbc<T>, cb(l0), dim3(NT), 0, c->s, view(l0), P<T>(l0.x));
        hipLaunchKernelGGL(k_project<T>, blocks(c->t), dim3(NT), 0, c->s, g, c->t, P<T>(u), P<T>(c->L), P<T>(c->p),
                           (T)c->P.u_bc[0], (T)c->P.u_bc[1], (T)c->P.u_bc[2]);
        hipLaunchKernelGGL(k_scale<T>, blocks1(g.ns), dim3(256), 0, c->s, P<T>(c->p), g.ns, T(1) / dtw);
        LAUNCHCHK(c);
        if (ss) {
            ss->iters = n;
            ss->r2 = r2;
            ss->rel = 0;
        }
        return IA_OK;
    }
    // FP64 flexible CG preconditioned by the FP32 V-cycle (REFINED / LEARNED)
    static FcgV<T> fview(ia_ctx *c, int par = 0) {
        Level &l0 = c->lev[0];
        FcgV<T> f;
        f.g = l0.g;
        f.t = l0.ct;
        f.L = P<T>(l0.L);
        f.iD = P<T>(l0.iD);
        f.z = c->z64;
        f.x = c->x64;
        f.r = c->r64;
        f.pa = par ? c->pb64 : c->pa64;
        f.pb = par ? c->pa64 : c->pb64;
        f.q = c->q64;
        f.w = P<T>(l0.x);
        f.wr = P<T>(l0.r);
        f.S = c->fst;
        f.ncell = interior(l0.g);
        return f;
    }
    static int true_resid(ia_ctx *c, double *rr) {
        Level &l0 = c->lev[0];
        FcgV<T> f = fview(c);
        hipLaunchKernelGGL(k_fcg_resid<T>, cb(l0), dim3(NT), 0, c->s, f, pp(c, PP_FR));
        hipLaunchKernelGGL(k_fcg_center_r<T>, cb(l0), dim3(NT), 0, c->s, f, pp(c, PP_FR), pp(c, PP_FRR));
        LAUNCHCHK(c);
        return host_total(c, PP_FRR, l0.ct.nblk, 1, rr);
    }
    // one flexible-CG iteration: preconditioner (V-cycle + fine smoothing) and
    // the FP64 Krylov update.  The input residual's partial sums are in
    // buffer `rbuf` (the true residual after a restart, else the last update).
    static void fcg_body(ia_ctx *c, int first, int par, int rbuf) {
        Level &l0 = c->lev[0];
        FcgV<T> f = fview(c, par);
        dim3 B = cb(l0);
        prof_begin(c, IA_PROF_LOAD_INC);
        hipLaunchKernelGGL(k_fcg_load_inc<T>, B, dim3(NT), 0, c->s, f, view(l0), pp(c, rbuf));
        prof_end(c, IA_PROF_LOAD_INC);
        if (c->fine_cheb) {
            vcycle(c, 0, c->cit_fcg, true, true);
            double c0, ca, cb;
            cheb2_coeffs(c0, ca, cb);
            prof_begin(c, IA_PROF_PCG_LAST);
            hipLaunchKernelGGL(k_cheb2_dots<T>, B, dim3(NT), 0, c->s, view(l0), f, (T)c0, (T)ca, (T)cb, pp(c, PP_FD),
                               first);
            prof_end(c, IA_PROF_PCG_LAST);
        } else if (c->smooth_it > 0) {
            vcycle(c, 0, c->cit_fcg, true, true);
            pcg(c, l0, c->smooth_it, true, &f, first);
        } else {
            vcycle(c, 0, c->cit_fcg, true, false);
            hipLaunchKernelGGL(k_fcg_dots<T>, B, dim3(NT), 0, c->s, f, pp(c, PP_FD), first);
        }
        prof_begin(c, IA_PROF_FCG_AP);
        hipLaunchKernelGGL(k_fcg_Ap<T>, B, dim3(NT), 0, c->s, f, pp(c, PP_FD), pp(c, PP_FAP), first);
        prof_end(c, IA_PROF_FCG_AP);
        prof_begin(c, IA_PROF_FCG_UPDATE);
        hipLaunchKernelGGL(k_fcg_update<T>, B, dim3(NT), 0, c->s, f, pp(c, PP_FAP), pp(c, PP_FU));
        prof_end(c, IA_PROF_FCG_UPDATE);
    }
    static int fcg_iteration(ia_ctx *c, int first, int par, int rbuf) {
        fcg_body(c, first, par, rbuf);
        LAUNCHCHK(c);
        return IA_OK;
    }
    static int project_refined(ia_ctx *c, void *u, double dtw, ia_solve_stats *ss) {
        Grid &g = c->g;
        Level &l0 = c->lev[0];
        long nb = l0.ct.nblk;
        if (!c->x64) {
            double **bufs[] = {&c->x64, &c->r64, &c->z64, &c->pa64, &c->pb64, &c->q64};
            for (double **b : bufs) {
                HIPCHK(c, hipMalloc((void **)b, g.ns * sizeof(double)));
                HIPCHK(c, hipMemsetAsync(*b, 0, g.ns * sizeof(double), c->s));
            }
        }
        FcgV<T> f = fview(c);
        dim3 B = cb(l0);
        hipLaunchKernelGGL(k_fcg_halt, dim3(1), dim3(64), 0, c->s, c->fst, 0);
        hipLaunchKernelGGL(k_fcg_init<T>, B, dim3(NT), 0, c->s, f, P<T>(u), P<T>(c->p), dtw, pp(c, PP_FI));
        hipLaunchKernelGGL(k_fcg_center_z<T>, B, dim3(NT), 0, c->s, f, pp(c, PP_FI), pp(c, PP_FZ));
        double zn2 = 0, rr = 0;
        if (int e = host_total(c, PP_FZ, nb, 1, &zn2)) return e;
        if (int e = true_resid(c, &rr)) return e;
        if (c->P.poisson_mode == IA_POISSON_LEARNED)
            if (int e = cnn_prelude(c, &rr)) return e;
        int rbuf = PP_FRR;  // buffer holding the partials of ‖r‖²
        double zn = sqrt(zn2);
        double rel = zn > 0 ? sqrt(rr) / zn : 0.0;
        int it = 0;
        int par = 0;
        while (c->P.poisson_mode == IA_POISSON_PLAIN && rel > c->P.poisson_rtol && it < c->P.poisson_max_outer) {
            // plain multigrid: one V-cycle + pcg!(fine, 6) on the unit-RMS residual, x += s·w
            f = fview(c);
            prof_begin(c, IA_PROF_LOAD_INC);
            hipLaunchKernelGGL(k_fcg_load_inc<T>, B, dim3(NT), 0, c->s, f, view(l0), pp(c, rbuf));
            prof_end(c, IA_PROF_LOAD_INC);
            vcycle(c, 0, c->cit_wl, true, true);
            pcg(c, l0, 6, true);
            hipLaunchKernelGGL(k_ir_accum<T>, B, dim3(NT), 0, c->s, f);
            LAUNCHCHK(c);
            if (int e = true_resid(c, &rr)) return e;
            rbuf = PP_FRR;
            rel = sqrt(rr) / zn;
            it++;
        }
        // flexible CG, pipelined: the host queues iteration n+1 before it reads
        // iteration n's residual, so the device never waits for the host; the
        // device-side test sets a stop flag that turns a queued iteration past
        // convergence into immediate returns
        const double rtol = c->P.poisson_rtol;
        const int maxo = c->P.poisson_max_outer;
        if (rel > rtol && it < maxo) {
            struct HaltScope {
                explicit HaltScope(const int *h) { cur_halt = h; }
                ~HaltScope() { cur_halt = nullptr; }
            } hs(&c->fst->halt);
            int queued = 0, qslot = 0, rslot = 0;
            auto enqueue = [&](int first) -> int {
                if (int e = fcg_iteration(c, first, par, rbuf)) return e;
                par ^= 1;
                rbuf = PP_FU;
                hipLaunchKernelGGL(k_fcg_check<T>, dim3(1), dim3(NT), 0, c->s, fview(c), pp(c, PP_FU), zn, rtol);
                HIPCHK(c, hipMemcpyAsync(&c->pin[qslot], c->fst, sizeof(FcgState), hipMemcpyDeviceToHost, c->s));
                HIPCHK(c, hipEventRecord(c->it_ev[qslot], c->s));
                qslot ^= 1;
                queued++;
                return IA_OK;
            };
            if (int e = enqueue(1)) return e;
            for (;;) {
                if (queued < 2 && it + queued < maxo)
                    if (int e = enqueue(0)) return e;
                HIPCHK(c, hipEventSynchronize(c->it_ev[rslot]));
                FcgState st = c->pin[rslot];
                rslot ^= 1;
                queued--;
                it++;
                rr = st.rr;
                rel = sqrt(rr) / zn;
                if (st.halt) {
                    // the queued iteration (if any) returned at once; confirm with
                    // the true FP64 residual and restart from it if it drifted
                    HIPCHK(c, hipStreamSynchronize(c->s));
                    queued = 0;
                    rslot = qslot;
                    if (int e = true_resid(c, &rr)) return e;
                    rel = sqrt(rr) / zn;
                    if (rel <= rtol || it >= maxo) break;
                    hipLaunchKernelGGL(k_fcg_halt, dim3(1), dim3(64), 0, c->s, c->fst, 0);
                    rbuf = PP_FRR;
                    if (int e = enqueue(1)) return e;
                    continue;
                }
                if (queued == 0) break;  // iteration cap reached
            }
        }
        f = fview(c);
        hipLaunchKernelGGL(k_fcg_project<T>, B, dim3(NT), 0, c->s, f, P<T>(u), P<T>(c->p), dtw, (T)c->P.u_bc[0],
                           (T)c->P.u_bc[1], (T)c->P.u_bc[2]);
        hipLaunchKernelGGL(k_mg_perbc<T>, B, dim3(NT), 0, c->s, view(l0), P<T>(c->p));
        LAUNCHCHK(c);
        if (ss) {
            ss->iters = it;
            ss->r2 = rr;
            ss->rel = rel;
        }
        return IA_OK;
    }
    // ----------------------------------------------------------- learned correction
    static CnnTiles cnn_tiles(ia_ctx *c) {
        CnnTiles t;
        t.g = c->lev[0].g;
        t.tx = (t.g.Ng[0] - 2) / 16;
        t.ntile = (long)t.tx * (t.g.Ng[1] - 2) * (t.g.Ng[2] - 2);
        return t;
    }
    static CnnSegs cnn_segs(ia_ctx *c) {
        CnnSegs t;
        t.g = c->lev[0].g;
        t.sx = (t.g.Ng[0] - 2 + WG_SEG - 1) / WG_SEG;
        t.nseg = (long)t.sx * (t.g.Ng[1] - 2) * (t.g.Ng[2] - 2);
        return t;
    }
    static dim3 cnn_blocks(const CnnTiles &t) { return dim3((unsigned)std::min<long>(2048, (t.ntile + 3) / 4)); }
    static int cnn_alloc(ia_ctx *c) {
        auto &n = c->cnn;
        size_t ns = c->g.ns;
        for (auto &p : n.act) {
            HIPCHK(c, hipMalloc(&p, ns * CNN_C * 2));
            HIPCHK(c, hipMemsetAsync(p, 0, ns * CNN_C * 2, c->s));
        }
        for (auto &p : n.grad) {
            HIPCHK(c, hipMalloc(&p, ns * CNN_C * 2));
            HIPCHK(c, hipMemsetAsync(p, 0, ns * CNN_C * 2, c->s));
        }
        HIPCHK(c, hipMalloc((void **)&n.xkeep, ns * sizeof(double)));
        HIPCHK(c, hipMalloc((void **)&n.gtmp, ns * sizeof(float)));
        HIPCHK(c, hipMemsetAsync(n.gtmp, 0, ns * sizeof(float), c->s));
        HIPCHK(c, hipMalloc((void **)&n.io, 3 * ns * sizeof(float)));
        float **par[] = {&n.w, &n.m, &n.v, &n.gw};
        for (float **p : par) {
            HIPCHK(c, hipMalloc((void **)p, CNN_NPAR * sizeof(float)));
            HIPCHK(c, hipMemsetAsync(*p, 0, CNN_NPAR * sizeof(float), c->s));
        }
        HIPCHK(c, hipMalloc((void **)&n.zbias, CNN_C * sizeof(float)));
        HIPCHK(c, hipMemsetAsync(n.zbias, 0, CNN_C * sizeof(float), c->s));
        HIPCHK(c, hipMalloc(&n.ff, CNN_L * CNN_P * 64 * sizeof(bf16x8)));
        HIPCHK(c, hipMalloc(&n.fb, CNN_L * CNN_P * 64 * sizeof(bf16x8)));
        n.wnblk = std::min<long>(1024, cnn_segs(c).nseg);
        HIPCHK(c, hipMalloc((void **)&n.wpart, n.wnblk * (CNN_W + CNN_C) * sizeof(float)));
        return IA_OK;
    }
    static void cnn_frags(ia_ctx *c) {
        auto &n = c->cnn;
        for (int l = 0; l < CNN_L; ++l)
            hipLaunchKernelGGL(k_cnn_frag, dim3(CNN_P), dim3(64), 0, c->s, n.w + (long)l * CNN_W,
                               (bf16x8 *)n.ff + l * CNN_P * 64, (bf16x8 *)n.fb + l * CNN_P * 64);
    }
    static CnnBlocks cnn_lds_blocks(ia_ctx *c) {
        CnnBlocks t;
        t.g = c->lev[0].g;
        t.tx = (t.g.Ng[0] - 2) / 16;
        t.ty = (t.g.Ng[1] - 2) / 4;
        t.tz = (t.g.Ng[2] - 2 + CNN_ZB - 1) / CNN_ZB;
        return t;
    }
    // Y = act(conv(X) + b): the LDS-staged kernel where rows of 4 fit, else the direct one
    static void cnn_conv(ia_ctx *c, const void *X, void *Y, const void *frag, const float *bias, int relu,
                         const void *mask = nullptr) {
        Grid &g = c->lev[0].g;
        if ((g.Ng[1] - 2) % 4 == 0) {
            CnnBlocks t = cnn_lds_blocks(c);
            hipLaunchKernelGGL(k_cnn_conv_lds, dim3((unsigned)((long)t.tx * t.ty * t.tz)), dim3(256), 0, c->s, t,
                               (const unsigned short *)X, (unsigned short *)Y, (const bf16x8 *)frag, bias, relu,
                               (const unsigned short *)mask);
        } else {
            CnnTiles t = cnn_tiles(c);
            hipLaunchKernelGGL(k_cnn_conv, cnn_blocks(t), dim3(256), 0, c->s, t, (const unsigned short *)X,
                               (unsigned short *)Y, (const bf16x8 *)frag, bias, relu, (const unsigned short *)mask);
        }
    }
    static void cnn_forward(ia_ctx *c) {
        auto &n = c->cnn;
        for (int l = 0; l < CNN_L; ++l)
            cnn_conv(c, n.act[l], n.act[l + 1], (const bf16x8 *)n.ff + l * CNN_P * 64,
                     n.w + (long)CNN_L * CNN_W + l * CNN_C, l < CNN_L - 1 ? 1 : 0);
    }
    // gradients of every parameter into gw, from the output gradient in grad[0]
    static void cnn_backward(ia_ctx *c) {
        auto &n = c->cnn;
        CnnSegs sg = cnn_segs(c);
        int cur = 0;
        for (int l = CNN_L - 1; l >= 0; --l) {
            hipLaunchKernelGGL(k_cnn_wgrad, dim3((unsigned)n.wnblk), dim3(256), 0, c->s, sg,
                               (const unsigned short *)n.act[l], (const unsigned short *)n.grad[cur], n.wpart);
            hipLaunchKernelGGL(k_cnn_wgrad_sum, dim3((CNN_W + CNN_C + 15) / 16), dim3(256), 0, c->s, n.wpart, n.wnblk,
                               n.gw + (long)l * CNN_W, n.gw + (long)CNN_L * CNN_W + l * CNN_C);
            if (l == 0) break;
            // the transposed convolution, masked by the ReLU derivative of layer l's input
            cnn_conv(c, n.grad[cur], n.grad[1 - cur], (const bf16x8 *)n.fb + l * CNN_P * 64, n.zbias, 0, n.act[l]);
            cur ^= 1;
        }
    }
    static void cnn_adam(ia_ctx *c) {
        auto &n = c->cnn;
        n.step++;
        const float b1 = 0.9f, b2 = 0.999f;
        float c1 = (float)(1 - pow(b1, (double)n.step)), c2 = (float)(1 - pow(b2, (double)n.step));
        hipLaunchKernelGGL(k_cnn_adam, dim3((CNN_NPAR + 255) / 256), dim3(256), 0, c->s, n.w, n.gw, n.m, n.v, CNN_NPAR,
                           (float)n.lr, b1, b2, c1, c2, 1e-8f);
        cnn_frags(c);
    }
    // the learned warm start: propose x += s_r·net(f, z, r), keep it if the
    // FP64 residual dropped; optionally one training step on this solve
    static int cnn_prelude(ia_ctx *c, double *rr) {
        auto &n = c->cnn;
        if (!n.ready) return fail(c, IA_ERR_BADARG, "learned Poisson mode needs ia_learned_init first");
        Level &l0 = c->lev[0];
        dim3 B = cb(l0);
        CnnSolve<T> a;
        a.f = fview(c);
        a.t = cnn_tiles(c);
        a.fv = P<T>(c->fb[c->fi]);
        double rr0 = *rr, rr1 = 0;
        hipLaunchKernelGGL(k_cnn_features<T>, B, dim3(NT), 0, c->s, a, (unsigned short *)n.act[0], pp(c, PP_FZ),
                           pp(c, PP_FRR));
        cnn_forward(c);
        hipLaunchKernelGGL(k_cnn_apply<T>, B, dim3(NT), 0, c->s, a, (const unsigned short *)n.act[CNN_L], n.xkeep);
        if (int e = true_resid(c, &rr1)) return e;
        if (n.train) {
            hipLaunchKernelGGL(k_cnn_loss_grad<T>, B, dim3(NT), 0, c->s, a, n.gtmp, rr0);
            hipLaunchKernelGGL(k_cnn_loss_grad2<T>, B, dim3(NT), 0, c->s, a, (unsigned short *)n.grad[0], n.gtmp);
            cnn_backward(c);
            cnn_adam(c);
        }
        n.projections++;
        n.last_ratio = rr0 > 0 ? rr1 / rr0 : 0;
        if (rr1 < rr0) {
            n.accepted++;
            *rr = rr1;
        } else {
            hipLaunchKernelGGL(k_cnn_restore<T>, B, dim3(NT), 0, c->s, a, n.xkeep);
            if (int e = true_resid(c, rr)) return e;
        }
        LAUNCHCHK(c);
        return IA_OK;
    }
    static int cnn_init(ia_ctx *c, unsigned seed) {
        auto &n = c->cnn;
        if (!n.act[0])
            if (int e = cnn_alloc(c)) return e;
        std::vector<float> w(CNN_NPAR, 0.f);
        uint64_t s = 0x9E3779B97F4A7C15ull ^ seed;
        auto uni = [&]() {  // splitmix64 -> (0,1)
            s += 0x9E3779B97F4A7C15ull;
            uint64_t z = s;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        };
        auto gauss = [&]() { return sqrt(-2 * log(uni())) * cos(2 * M_PI * uni()); };
        for (int l = 0; l < CNN_L; ++l) {
            int cin = l == 0 ? CNN_IN : CNN_C, cout = l == CNN_L - 1 ? CNN_OUT : CNN_C;
            double sd = sqrt(2.0 / (CNN_T * cin)) * (l == CNN_L - 1 ? 0.1 : 1.0);
            for (int t = 0; t < CNN_T; ++t)
                for (int ci = 0; ci < cin; ++ci)
                    for (int co = 0; co < cout; ++co) w[(long)l * CNN_W + (t * CNN_C + ci) * CNN_C + co] = (float)(sd * gauss());
        }
        return cnn_set(c, w.data());
    }
    static int cnn_set(ia_ctx *c, const float *w) {
        auto &n = c->cnn;
        if (!n.act[0])
            if (int e = cnn_alloc(c)) return e;
        HIPCHK(c, hipMemcpyAsync(n.w, w, CNN_NPAR * sizeof(float), hipMemcpyHostToDevice, c->s));
        HIPCHK(c, hipMemsetAsync(n.m, 0, CNN_NPAR * sizeof(float), c->s));
        HIPCHK(c, hipMemsetAsync(n.v, 0, CNN_NPAR * sizeof(float), c->s));
        n.step = 0;
        cnn_frags(c);
        HIPCHK(c, hipStreamSynchronize(c->s));
        n.ready = true;
        return IA_OK;
    }
    static int cnn_pack(ia_ctx *c, const float *host, int nch, void *X) {
        long n = interior(c->g);
        HIPCHK(c, hipMemcpyAsync(c->cnn.io, host, nch * n * sizeof(float), hipMemcpyHostToDevice, c->s));
        hipLaunchKernelGGL(k_cnn_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->s, c->g, c->cnn.io, nch,
                           (unsigned short *)X);
        return IA_OK;
    }
    static int cnn_test_forward(ia_ctx *c, const float *feat, float *out) {
        auto &n = c->cnn;
        if (!n.ready) return fail(c, IA_ERR_BADARG, "ia_learned_init first");
        long ncell = interior(c->g);
        if (int e = cnn_pack(c, feat, CNN_IN, n.act[0])) return e;
        cnn_forward(c);
        hipLaunchKernelGGL(k_cnn_unpack, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, c->s, c->g,
                           (const unsigned short *)n.act[CNN_L], CNN_OUT, n.io);
        HIPCHK(c, hipMemcpyAsync(out, n.io, CNN_OUT * ncell * sizeof(float), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        return IA_OK;
    }
    static int cnn_test_grad(ia_ctx *c, const float *feat, const float *dout, float *gw) {
        auto &n = c->cnn;
        if (!n.ready) return fail(c, IA_ERR_BADARG, "ia_learned_init first");
        if (int e = cnn_pack(c, feat, CNN_IN, n.act[0])) return e;
        cnn_forward(c);
        if (int e = cnn_pack(c, dout, CNN_OUT, n.grad[0])) return e;
        cnn_backward(c);
        HIPCHK(c, hipMemcpyAsync(gw, n.gw, CNN_NPAR * sizeof(float), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        return IA_OK;
    }
    static int project(ia_ctx *c, void *u, double dtw, ia_solve_stats *ss) {
        HIPCHK(c, hipEventRecord(c->e0, c->s));
        int e = c->P.poisson_mode == IA_POISSON_REFERENCE ? project_ref(c, u, dtw, ss)
                : c->P.poisson_mode == IA_POISSON_PCG     ? project_pcg(c, u, dtw, ss)
                                                          : project_refined(c, u, dtw, ss);
        if (e) return e;
        HIPCHK(c, hipEventRecord(c->e1, c->s));
        if (ss) {
            HIPCHK(c, hipEventSynchronize(c->e1));
            float ms = 0;
            HIPCHK(c, hipEventElapsedTime(&ms, c->e0, c->e1));
            ss->ms = ms;
        }
        return IA_OK;
    }

    // ----------------------------------------------------------------- CFL
    // vol (optional): Σf in FP64, min f and max f over the interior of the current f
    static int cfl(ia_ctx *c, double *out, double *vol = nullptr) {
        Grid &g = c->g;
        const CTiling &ct = c->lev[0].ct;
        hipLaunchKernelGGL(k_cfl<T>, dim3((unsigned)ct.nblk), dim3(NT), 0, c->s, g, ct, P<T>(c->u),
                           P<T>(c->fb[c->fi]), c->part);
        hipLaunchKernelGGL(k_fin_cfl, dim3(1), dim3(NT), 0, c->s, c->part, ct.nblk, c->dsc + 8);
        LAUNCHCHK(c);
        double mx[5];
        HIPCHK(c, hipMemcpyAsync(mx, c->dsc + 8, 5 * sizeof(double), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        if (vol)
            for (int q = 0; q < 3; ++q) vol[q] = mx[2 + q];
        T nu = (T)c->P.nu, lm = (T)c->P.lambda_mu, lr = (T)c->P.lambda_rho;
        T dtAdv = T(1) / ((T)mx[0] + 5 * nu);
        T dtV = T(1) / (2 * (T)mx[1]);
        T dtG = 1;
        if (c->P.has_g) {
            T g2 = 0;
            for (int d = 0; d < g.D; ++d) g2 += (T)c->P.g[d] * (T)c->P.g[d];
            dtG = T(1) / (2 * (T)sqrt(g2));
        }
        T dtVisc = c->P.nu != 0 ? T(3) / ((14 * nu) * std::max(T(1), lm / lr)) : T(1);
        T dtST = c->P.eta != 0 ? (T)std::sqrt((1 + lr) / ((T)(8 * M_PI) * (T)c->P.eta)) : T(1);
        T m = std::min(std::min(std::min(std::min(dtV, dtAdv), dtG), dtVisc), dtST);
        *out = (double)(T(0.8) * m);
        return IA_OK;
    }

    // ----------------------------------------------------------------- level set
    // LevelSet(sim) (redistaning.jl:14-25): ϕ = 2f - 1 in f⁰'s storage, ϕini in a
    // free buffer of the f pool (the reference's cVOF.fᶠ scratch)
    static int levelset(ia_ctx *c) {
        int q = free_fbuf(c, c->fi, c->f0i, -1);
        long ns = c->g.ns;
        hipLaunchKernelGGL(k_ls_init<T>, dim3((unsigned)((ns + NT - 1) / NT)), dim3(NT), 0, c->s, ns,
                           P<T>(c->fb[c->fi]), P<T>(c->fb[c->f0i]), P<T>(c->fb[q]));
        LAUNCHCHK(c);
        c->ls_ini = q;
        c->ls_valid = true;
        return IA_OK;
    }
    // redistaning!(ls; d, dτ, perdir) (redistaning.jl:44-56): round(d/dτ) SSP-RK3
    // steps, each stage one fused launch (computeL!, the stage update, BCf!).
    // The stage outputs ping-pong ϕⁿ (f⁰) -> ϕ¹ (σ) -> ϕ² (α) -> ϕⁿ⁺¹ (f⁰), ϕⁿ
    // staying in f⁰ as ϕ⁰ until the third stage overwrites it in place.
    static int redistance(ia_ctx *c, long itmx, double dtau, int perdir) {
        Grid g = c->g;
        for (int a = 0; a < 3; ++a) g.per[a] = a < g.D ? (perdir >> a) & 1 : 0;
        T dt = (T)dtau;
        T *phi = P<T>(c->fb[c->f0i]), *s1 = P<T>(c->sig), *s2 = P<T>(c->al);
        const T *ini = P<T>(c->fb[c->ls_ini]);
        auto k = g.D == 3 ? k_ls_stage<3, T> : k_ls_stage<2, T>;
        const T *in[3] = {phi, s1, s2};
        T *out[3] = {s1, s2, phi};
        const T al[3] = {T(0), (T)(3.0 / 4), (T)(1.0 / 3)};
        // 3D grids with multi-cell columns take the LDS-tiled stage
        CTiling ct = ctiling(g);
        bool tiled = g.D == 3 && ct.zb > 1;
        for (long it = 0; it < itmx; ++it)
            for (int q = 0; q < 3; ++q) {
                prof_begin(c, IA_PROF_REDIST);
                if (tiled)
                    hipLaunchKernelGGL(k_ls_stage_tiled<T>, dim3((unsigned)ct.nblk), dim3(NT), 0, c->s, g, ct, in[q],
                                       phi, ini, out[q], dt, al[q]);
                else
                    hipLaunchKernelGGL(k, blocks(c->t), dim3(NT), 0, c->s, g, c->t, in[q], phi, ini, out[q], dt, al[q]);
                prof_end(c, IA_PROF_REDIST);
                LAUNCHCHK(c);
            }
        return IA_OK;
    }

    // metrics.jl integrals (Σρke, Σρu_i, Σρgh, Σ enstrophy, Σf)
    static int metrics(ia_ctx *c, const double *U, const double *wl, double *out) {
        Grid &g = c->g;
        long nb = c->t.nblk;
        double *part = nullptr;  // seven per-block partials (a diagnostic: allocated per call)
        HIPCHK(c, hipMalloc((void **)&part, 7 * nb * sizeof(double)));
        double gv[3] = {0, 0, 0};
        if (c->P.has_g)
            for (int d = 0; d < g.D; ++d) gv[d] = (double)(T)c->P.g[d];
        auto km = g.D == 3 ? k_metrics<3, T> : k_metrics<2, T>;
        hipLaunchKernelGGL(km, blocks(c->t), dim3(NT), 0, c->s, g, c->t, P<T>(c->u), P<T>(c->fb[c->fi]),
                           (T)c->P.lambda_rho, U[0], U[1], U[2], gv[0], gv[1], gv[2], wl[0], wl[1], wl[2], part);
        std::vector<double> h(7 * nb);
        hipError_t e = hipMemcpyAsync(h.data(), part, 7 * nb * sizeof(double), hipMemcpyDeviceToHost, c->s);
        if (e == hipSuccess) e = hipStreamSynchronize(c->s);
        (void)hipFree(part);
        HIPCHK(c, e);
        for (int q = 0; q < 7; ++q) {
            double s = 0;
            for (long b = 0; b < nb; ++b) s += h[b * 7 + q];
            out[q] = s;
        }
        return IA_OK;
    }
    // ia_metrics without allocating: block partials and totals in the
    // device's `scratch` (address ws); the totals are copied into `out`
    static int metrics_noalloc(ia_ctx *c, double *ws, const double *U, const double *wl, double *out) {
        Grid &g = c->g;
        long nb = std::min<long>(c->t.nblk, IA_NOALLOC_BLOCKS);
        if (nb * 7 + 7 > IA_NOALLOC_DOUBLES)
            return fail(c, IA_ERR_BADARG, "ia_metrics_noalloc: %ld block partials exceed the scratch of %d doubles",
                        nb, IA_NOALLOC_DOUBLES);
        double gv[3] = {0, 0, 0};
        if (c->P.has_g)
            for (int d = 0; d < g.D; ++d) gv[d] = (double)(T)c->P.g[d];
        double *tot = ws + (long)IA_NOALLOC_BLOCKS * 7;
        auto km = g.D == 3 ? k_metrics_gs<3, T> : k_metrics_gs<2, T>;
        hipLaunchKernelGGL(km, dim3((unsigned)nb), dim3(NT), 0, c->s, g, c->t, P<T>(c->u), P<T>(c->fb[c->fi]),
                           (T)c->P.lambda_rho, U[0], U[1], U[2], gv[0], gv[1], gv[2], wl[0], wl[1], wl[2], ws);
        hipLaunchKernelGGL(k_metrics_fin, dim3(1), dim3(NT), 0, c->s, ws, nb, tot);
        LAUNCHCHK(c);
        HIPCHK(c, hipMemcpyAsync(out, tot, 7 * sizeof(double), hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        return IA_OK;
    }
    // ----------------------------------------------------------------- sweeps
    static void prepare(ia_ctx *c, const void *f, const void *usrc, bool do_ru) {
        hipLaunchKernelGGL(k_prepare<T>, blocks(c->t), dim3(NT), 0, c->s, c->g, c->t, P<T>(f), P<T>(usrc), P<T>(c->ru),
                           c->cb, (T)c->P.lambda_rho, do_ru ? 1 : 0);
    }
    static void vof_sweep(ia_ctx *c, const void *fin, void *fout, const void *u1, const void *u2, int d, T dt,
                          int mode, double rtol, const void *fn, void *fmid) {
        Grid &g = c->g;
        SweepArgs<T> a;
        a.g = g;
        a.t = c->t;
        a.fin = P<T>((void *)fin);
        a.fout = P<T>(fout);
        a.u1 = P<T>((void *)u1) + g.ns * d;
        a.u2 = P<T>((void *)u2) + g.ns * d;
        a.cb = c->cb;
        a.ruf = P<T>(c->ruf) + g.ns * d;
        a.fn = P<T>((void *)fn);
        a.fmid = P<T>(fmid);
        a.part = c->fpart;
        a.d = d;
        a.scheme = c->P.normal_scheme;
        a.mode = mode;
        a.dt = dt;
        a.lr = (T)c->P.lambda_rho;
        a.idt = T(1) / dt;
        a.rtol = (T)rtol;
        a.ctol = 10 * (T)epsT();
        a.ubc_d = (T)c->P.u_bc[d];
        prof_begin(c, IA_PROF_VOF);
        long nfb = c->t.nblk + (d == 0 ? top_face_x_blocks(g) : 0);
        hipLaunchKernelGGL(k_vof_face<T>, dim3((unsigned)nfb), dim3(NT), 0, c->s, a, P<T>(c->vface));
        hipLaunchKernelGGL(k_vof_update<T>, blocks(c->t), dim3(NT), 0, c->s, a, P<T>(c->vface));
        prof_end(c, IA_PROF_VOF);  // one sweep (its two launches) per record
        hipLaunchKernelGGL(k_fill_reduce, dim3(FILL_R2), dim3(NT), 0, c->s, c->fpart, c->t.nblk, c->fpart + c->t.nblk);
        hipLaunchKernelGGL(k_fill_finalize<T>, dim3(1), dim3(NT), 0, c->s, g, c->fpart + c->t.nblk, (long)FILL_R2,
                           P<T>((void *)u1),
                           P<T>((void *)u2), rtol, c->st);
    }
    static void momentum(ia_ctx *c, const void *fpre, const void *u1, const void *u2, const void *uold, int d, T dt) {
        Grid &g = c->g;
        MomArgs<T> a;
        a.g = g;
        a.t = c->t;
        a.ru_in = P<T>(c->ru);
        a.ru_out = P<T>(c->ru2);
        a.fpre = P<T>((void *)fpre);
        a.ruf = P<T>(c->ruf) + g.ns * d;
        a.cb = c->cb;
        a.u1 = P<T>((void *)u1) + g.ns * d;
        a.u2 = P<T>((void *)u2) + g.ns * d;
        a.uold = P<T>((void *)uold);
        a.d = d;
        a.lim = c->P.limiter;
        a.dt = dt;
        a.lr = (T)c->P.lambda_rho;
        for (int q = 0; q < 3; ++q) a.ubc[q] = (T)c->P.u_bc[q];
        a.us = P<T>(c->nh);
        hipLaunchKernelGGL(k_ustar<T>, blocks1(g.ns), dim3(NT), 0, c->s, a);
        prof_begin(c, IA_PROF_MOMENTUM);
        auto kf = k_mom_face<3, 0, T>;
        auto ku = k_mom_update<3, 0, T>;
        if (g.D == 3) {
            kf = d == 0 ? k_mom_face<3, 0, T> : d == 1 ? k_mom_face<3, 1, T> : k_mom_face<3, 2, T>;
            ku = d == 0 ? k_mom_update<3, 0, T> : d == 1 ? k_mom_update<3, 1, T> : k_mom_update<3, 2, T>;
        } else {
            kf = d == 0 ? k_mom_face<2, 0, T> : k_mom_face<2, 1, T>;
            ku = d == 0 ? k_mom_update<2, 0, T> : k_mom_update<2, 1, T>;
        }
        long nfb = c->t.nblk + (d == 0 ? top_face_x_blocks(g) : 0);
        hipLaunchKernelGGL(kf, dim3((unsigned)nfb), dim3(NT), 0, c->s, a, P<T>(c->mface));
        hipLaunchKernelGGL(ku, blocks(c->t), dim3(NT), 0, c->s, a, P<T>(c->mface));
        prof_end(c, IA_PROF_MOMENTUM);  // one sweep (its two launches) per record
        std::swap(c->ru, c->ru2);
    }
    static void dir_order(int D, int step_index, int *dirO) {
        for (int i = 1; i <= D; ++i) dirO[i - 1] = ((step_index + i) % D);  // 0-based (flow.jl:163)
    }
    // advectfq!/advectVOFρuu! (flow.jl:157-210): advect pool buffer `fsrc`;
    // returns the pool index holding the result; the last sweep also writes
    // f_mid = (f* + f^n)/2 into pool index `mid` when mid >= 0.
    static int advect_fq(ia_ctx *c, int fsrc, const void *u1, const void *u2, const void *uold, T dt, int step_index,
                         int keep, int mid) {
        int dirO[3];
        dir_order(c->g.D, step_index, dirO);
        int cur = fsrc;
        int a = free_fbuf(c, fsrc, keep, mid), b = -1;
        for (int q = 0; q < 4; ++q)
            if (q != fsrc && q != keep && q != mid && q != a) { b = q; break; }
        int outs[2] = {a, b};
        double rtol = 100 * epsT();
        for (int op = 0; op < c->g.D; ++op) {
            int d = dirO[op];
            int out = outs[op & 1];
            bool last = op == c->g.D - 1;
            vof_sweep(c, c->fb[cur], c->fb[out], u1, u2, d, dt, 0, rtol, last && mid >= 0 ? c->fb[keep] : nullptr,
                      last && mid >= 0 ? c->fb[mid] : nullptr);
            momentum(c, c->fb[cur], u1, u2, uold, d, dt);
            cur = out;
        }
        return cur;
    }
    // surfTen! (surfaceTension.jl:7-19) on f into `out`: written (acc = 0) or added (acc = 1)
    static int surften(ia_ctx *c, const void *f, void *out, int acc) {
        Grid &g = c->g;
        if (!c->st_fb) {
            HIPCHK(c, hipMalloc(&c->st_fb, g.ns * sizeof(T)));
            HIPCHK(c, hipMemsetAsync(c->st_fb, 0, g.ns * sizeof(T), c->s));
        }
        for (int d = 0; d < g.D; ++d) {
            hipLaunchKernelGGL(k_st_fbuf<T>, blocks(c->t), dim3(NT), 0, c->s, g, c->t, P<T>((void *)f), P<T>(c->st_fb), d);
            for (int j = 0; j < g.D; ++j)
                hipLaunchKernelGGL(k_st_bc<T>, blocks1(g.ns / g.Ng[j]), dim3(NT), 0, c->s, g, P<T>(c->st_fb), d, j);
            auto ks = g.D == 3 ? k_surften<3, T> : k_surften<2, T>;
            hipLaunchKernelGGL(ks, blocks(c->t), dim3(NT), 0, c->s, g, c->t, P<T>(c->st_fb), P<T>((void *)f),
                               (T)c->P.eta, d, P<T>(out), acc);
        }
        LAUNCHCHK(c);
        return IA_OK;
    }
    static int upd(ia_ctx *c, const void *ust, const void *u0, const void *fcur, const void *fupd, void *uout, T dt,
                   T w) {
        Grid &g = c->g;
        UpdArgs<T> a;
        a.v.g = g;
        a.v.t = c->t;
        a.v.u = P<T>((void *)ust);
        a.v.fv = P<T>((void *)fupd);
        a.v.lm = (T)c->P.lambda_mu;
        a.v.mu = (T)c->P.nu;
        a.v.lr = (T)c->P.lambda_rho;
        for (int q = 0; q < 3; ++q) a.v.ubc[q] = (T)c->P.u_bc[q];
        a.v.has_mu = c->P.nu != 0;
        a.u0 = P<T>((void *)u0);
        a.fcur = P<T>((void *)fcur);
        a.fupd = P<T>((void *)fupd);
        a.ru = P<T>(c->ru);
        a.uout = P<T>(uout);
        a.L = P<T>(c->L);
        a.dt = dt;
        a.w = w;
        for (int q = 0; q < 3; ++q) a.g[q] = c->P.has_g ? (T)c->P.g[q] : T(0);
        a.v.ff = P<T>(c->nh);
        if (a.v.has_mu) hipLaunchKernelGGL(k_fface<T>, blocks1(g.ns), dim3(NT), 0, c->s, a.v, P<T>(c->nh));
        a.st = nullptr;
        if (c->P.eta != 0) {
            if (!c->st_buf) HIPCHK(c, hipMalloc(&c->st_buf, g.ns * g.D * sizeof(T)));
            if (int e = surften(c, fupd, c->st_buf, 0)) return e;
            a.st = P<T>(c->st_buf);
        }
        prof_begin(c, IA_PROF_VISC_UPDATE);
        auto kv = g.D == 3 ? k_visc_update<3, T> : k_visc_update<2, T>;
        hipLaunchKernelGGL(kv, blocks(c->t), dim3(NT), 0, c->s, a);
        prof_end(c, IA_PROF_VISC_UPDATE);
        LAUNCHCHK(c);
        return IA_OK;
    }
    static int read_state(ia_ctx *c, ia_step_stats *ss) {
        StepState h;
        HIPCHK(c, hipMemcpyAsync(&h, c->st, sizeof h, hipMemcpyDeviceToHost, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        if (ss) {
            ss->status = h.status;
            ss->fill_warnings = h.nwarn;
        }
        if (h.status == IA_ERR_NAN) fail(c, IA_ERR_NAN, "NaN!");
        else if (h.status == IA_ERR_EXPLODE) fail(c, IA_ERR_EXPLODE, "divergence is exploding! (VOF %.6g / %.6g)", h.maxf, h.minf);
        else if (h.status == IA_WARN_FILL)
            fail(c, IA_WARN_FILL, "VOF not in [0,1]: max %.9g @ %ld, min %.9g @ %ld", h.maxf, h.imax, h.minf, h.imin);
        return h.status;
    }

    // MPFMomStep! (flow.jl:60-109)
    static int mom_step(ia_ctx *c, double dtd, int step_index, ia_step_stats *ss) {
        Grid &g = c->g;
        T dt = (T)dtd;
        HIPCHK(c, hipMemsetAsync(c->st, 0, sizeof(StepState), c->s));
        int fn = c->fi;
        void *U0 = c->u, *U1 = c->u0;  // u⁰ ← u: the predictor reads u^n from U0 and writes u^{n+1/2} to U1
        // predictor (w = ½)
        T w = T(0.5);
        prepare(c, c->fb[fn], U0, true);
        int mid = free_fbuf(c, fn, -1, -1);
        int fstar = advect_fq(c, fn, U0, U0, U0, dt, step_index, fn, mid);
        (void)fstar;
        if (int e = upd(c, U0, U0, c->fb[fn], c->fb[mid], U1, dt, w)) return e;
        if (int e = mg_update(c)) return e;
        if (int e = project(c, U1, (double)(w * dt), ss ? &ss->solve[0] : nullptr)) return e;
        // corrector (w = 1): f⁰ ← f; ρu = u⁰ρ(f); advect f with u^{n+½}
        prepare(c, c->fb[fn], U0, true);
        int fnew = advect_fq(c, fn, U1, U1, U0, dt, step_index, fn, -1);
        if (int e = upd(c, U1, U0, c->fb[fnew], c->fb[fnew], U0, dt, T(1))) return e;
        if (int e = mg_update(c)) return e;
        if (int e = project(c, U0, (double)dt, ss ? &ss->solve[1] : nullptr)) return e;
        c->u = U0;
        c->u0 = U1;
        c->fi = fnew;
        c->f0i = fn;
        if (int e = restore_attached(c)) return e;
        double cf, vol[3];
        if (int e = cfl(c, &cf, vol)) return e;
        int st = read_state(c, ss);
        double nx = std::min(cf, 1.2 * (double)dt);
        if (ss) {
            ss->dt_next = (double)(T)nx;
            ss->vol = vol[0];
            ss->fmin = vol[1];
            ss->fmax = vol[2];
        }
        return st;
    }
    // advect!(a, c, f, u¹, u², dt) (advection.jl:17-78): f is cVOF.f or cVOF.f⁰
    // (pool index fsrc; the other one is kept), u¹ and u² any of the
    // context's vector fields that the sweeps do not write
    static int advect(ia_ctx *c, int fsrc, const void *u1, const void *u2, double dtd, int step_index,
                      ia_step_stats *ss) {
        Grid &g = c->g;
        T dt = (T)dtd;
        int keep = fsrc == c->fi ? c->f0i : c->fi;
        HIPCHK(c, hipMemsetAsync(c->st, 0, sizeof(StepState), c->s));
        HIPCHK(c, hipMemsetAsync(c->ruf, 0, g.ns * g.D * sizeof(T), c->s));
        prepare(c, c->fb[fsrc], u1, false);
        int dirO[3];
        dir_order(g.D, step_index, dirO);
        int cur = fsrc;
        for (int op = 0; op < g.D; ++op) {
            int d = dirO[op];
            int out = free_fbuf(c, cur, keep, -1);
            vof_sweep(c, c->fb[cur], c->fb[out], u1, u2, d, dt, 1, 10 * epsT(), nullptr, nullptr);
            cur = out;
        }
        LAUNCHCHK(c);
        bool adv_f = fsrc == c->fi;
        if (adv_f) c->fi = cur;
        else c->f0i = cur;
        if (int e = restore_attached(c)) return e;
        double cf, vol[3];
        int fi_save = c->fi;
        c->fi = adv_f ? c->fi : c->f0i;  // the statistics describe the advected field
        int e = cfl(c, &cf, vol);
        c->fi = fi_save;
        if (e) return e;
        int st = read_state(c, ss);
        if (ss) {
            ss->dt_next = dtd;
            ss->vol = vol[0];
            ss->fmin = vol[1];
            ss->fmax = vol[2];
        }
        return st;
    }
    static int vof_sweep_stage(ia_ctx *c, int d, double dtd, ia_step_stats *ss) {
        Grid &g = c->g;
        HIPCHK(c, hipMemsetAsync(c->st, 0, sizeof(StepState), c->s));
        HIPCHK(c, hipMemsetAsync(c->ruf, 0, g.ns * g.D * sizeof(T), c->s));
        prepare(c, c->fb[c->fi], c->u0, false);
        int out = free_fbuf(c, c->fi, c->f0i, -1);
        vof_sweep(c, c->fb[c->fi], c->fb[out], c->u0, c->u, d, (T)dtd, 1, 100 * epsT(), nullptr, nullptr);
        LAUNCHCHK(c);
        c->fi = out;
        if (int e = restore_attached(c)) return e;
        return read_state(c, ss);
    }
    static int visc_stage(ia_ctx *c, int use_f0) {
        ViscArgs<T> a;
        a.g = c->g;
        a.t = c->t;
        a.u = P<T>(c->u);
        a.fv = P<T>(c->fb[use_f0 ? c->f0i : c->fi]);
        a.lm = (T)c->P.lambda_mu;
        a.mu = (T)c->P.nu;
        a.lr = (T)c->P.lambda_rho;
        for (int q = 0; q < 3; ++q) a.ubc[q] = (T)c->P.u_bc[q];
        a.has_mu = c->P.nu != 0;
        a.ff = P<T>(c->nh);
        hipLaunchKernelGGL(k_fface<T>, blocks1(c->g.ns), dim3(NT), 0, c->s, a, P<T>(c->nh));
        auto kf = c->g.D == 3 ? k_visc_force<3, T> : k_visc_force<2, T>;
        hipLaunchKernelGGL(kf, blocks(c->t), dim3(NT), 0, c->s, a, P<T>(c->F));
        if (c->P.eta != 0)
            if (int e = surften(c, a.fv, c->F, 1)) return e;
        LAUNCHCHK(c);
        return IA_OK;
    }
    static int update_L_stage(ia_ctx *c, int use_f0) {
        hipLaunchKernelGGL(k_update_L<T>, blocks(c->t), dim3(NT), 0, c->s, c->g, c->t,
                           P<T>(c->fb[use_f0 ? c->f0i : c->fi]), P<T>(c->L), (T)c->P.lambda_rho);
        LAUNCHCHK(c);
        return IA_OK;
    }
    static int reconstruct_stage(ia_ctx *c) {
        hipLaunchKernelGGL(k_reconstruct<T>, blocks(c->t), dim3(NT), 0, c->s, c->g, c->t, P<T>(c->fb[c->fi]),
                           P<T>(c->al), P<T>(c->nh), c->P.normal_scheme);
        LAUNCHCHK(c);
        return IA_OK;
    }
    static int apply_vof(ia_ctx *c, const void *sdf, size_t nbytes) {
        Grid &g = c->g;
        size_t need = (size_t)interior(g) * (1 + 2 * g.D) * sizeof(T);
        if (nbytes != need) return fail(c, IA_ERR_BADARG, "ia_apply_vof: %zu bytes, expected %zu", nbytes, need);
        void *d = nullptr;
        HIPCHK(c, hipMalloc(&d, need));
        HIPCHK(c, hipMemcpyAsync(d, sdf, need, hipMemcpyHostToDevice, c->s));
        hipLaunchKernelGGL(k_apply_vof<T>, blocks(c->t), dim3(NT), 0, c->s, g, c->t, P<T>(d), P<T>(c->fb[c->fi]),
                           P<T>(c->al), P<T>(c->nh), 10 * (T)epsT());
        LAUNCHCHK(c);
        HIPCHK(c, hipMemcpyAsync(c->fb[c->f0i], c->fb[c->fi], g.ns * sizeof(T), hipMemcpyDeviceToDevice, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        HIPCHK(c, hipFree(d));
        return IA_OK;
    }
    static int set_initial_velocity(ia_ctx *c, const void *u, size_t nbytes) {
        Grid &g = c->g;
        size_t need = g.ns * g.D * sizeof(T);
        if (nbytes != need) return fail(c, IA_ERR_BADARG, "ia_set_initial_velocity: %zu bytes, expected %zu", nbytes, need);
        HIPCHK(c, hipMemcpyAsync(c->u, u, need, hipMemcpyHostToDevice, c->s));
        T ub[3];
        for (int a = 0; a < 3; ++a) ub[a] = (T)c->P.u_bc[a];
        if (int e = bc_vel(c, c->u, ub)) return e;
        HIPCHK(c, hipMemcpyAsync(c->u0, c->u, need, hipMemcpyDeviceToDevice, c->s));
        HIPCHK(c, hipStreamSynchronize(c->s));
        return IA_OK;
    }
};

template <class F> int dispatch(ia_ctx *c, F &&f) {
    if (!c) return IA_ERR_BADARG;
    HIPCHK(c, hipSetDevice(c->dev));
    return c->P.dtype == IA_F64 ? f(Impl<double>()) : f(Impl<float>());
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int ia_create(const ia_params *p, ia_ctx **out) {
    if (!p || !out) return IA_ERR_BADARG;
    *out = nullptr;
    if (!(p->D == 2 || p->D == 3) || !(p->dtype == IA_F32 || p->dtype == IA_F64)) return IA_ERR_BADARG;
    for (int a = 0; a < p->D; ++a)
        if (p->N[a] < 2) return IA_ERR_BADARG;
    if (!(p->eta >= 0)) return IA_ERR_BADARG;  // η = 0 means η === nothing
    if (p->normal_scheme < 0 || p->normal_scheme > IA_NORMAL_SLIC || p->limiter < 0 || p->limiter > IA_LIMITER_SWEBY)
        return IA_ERR_BADARG;
    if (p->poisson_mode < 0 || p->poisson_mode > IA_POISSON_PCG) return IA_ERR_BADARG;
    // exitBC: uStar's exit face is the reference's forcing scratch there, 0 only when g_x = 0
    if (p->exit_bc && p->has_g && p->g[0] != 0) return IA_ERR_BADARG;
    if (p->poisson_mode != IA_POISSON_REFERENCE && p->poisson_mode != IA_POISSON_PCG && p->dtype != IA_F32)
        return IA_ERR_BADARG;
    if (p->poisson_mode == IA_POISSON_LEARNED &&
        !(p->D == 3 && p->N[0] % 16 == 0 && p->N[1] % 2 == 0 && p->N[2] % 2 == 0))
        return IA_ERR_BADARG;  // the network runs on 3D grids of 16-cell rows and even sides
    ia_ctx *c = new ia_ctx();
    c->P = *p;
    if (c->P.poisson_tol <= 0) c->P.poisson_tol = 1e-4;
    if (c->P.poisson_itmx <= 0) c->P.poisson_itmx = 200;
    if (c->P.poisson_rtol <= 0) c->P.poisson_rtol = 1e-6;
    if (c->P.poisson_max_outer <= 0) c->P.poisson_max_outer = 100;
    c->dev = p->device;
    c->es = p->dtype == IA_F64 ? 8 : 4;
    if (const char *sv = getenv("IA_FCG_SMOOTH"); sv && *sv) {
        c->smooth_it = std::max(0, std::min(CIT_MAX, atoi(sv)));
        c->fine_cheb = 0;
    }
    for (int l = 0; l < MAXLEV; ++l) {
        long n = 1;
        for (int a = 0; a < p->D; ++a) n *= std::max(1L, (long)(p->N[a] >> l));
        c->cit_wl[l] = 6;
        c->cit_fcg[l] = fcg_coarse_its(n, l);
        c->cg1_nb[l] = p->D == 3 ? cg1_workgroup(n) : NT;
    }
    // IA_CG1_NB=a,b,...: the coarse CG step's workgroup size on levels 1, 2, ...
    // (128, 256, 512 or 1024; used where a level has one cell per thread)
    if (const char *nv = getenv("IA_CG1_NB"); nv && *nv) {
        int l = 1;
        const char *q = nv;
        while (*q && l < MAXLEV) {
            char *end;
            int v = (int)strtol(q, &end, 10);
            if (end == q || (v != 128 && v != 256 && v != 512 && v != 1024)) {
                delete c;
                return IA_ERR_BADARG;
            }
            c->cg1_nb[l++] = v;
            q = *end == ',' ? end + 1 : end;
        }
    }
    if (const char *cv = getenv("IA_FCG_CIT"); cv && *cv) {
        int l = 1, v = 0;
        const char *q = cv;
        while (*q && l < MAXLEV) {
            char *end;
            v = (int)strtol(q, &end, 10);
            if (end == q || v < 1 || v > CIT_MAX) {
                delete c;
                return IA_ERR_BADARG;
            }
            c->cit_fcg[l++] = v;
            q = *end == ',' ? end + 1 : end;
            if (end == q && *q) {
                delete c;
                return IA_ERR_BADARG;
            }
        }
        for (; l < MAXLEV; ++l) c->cit_fcg[l] = v;
    }
    int Ng[3];
    for (int a = 0; a < 3; ++a) Ng[a] = a < p->D ? p->N[a] + 2 : 1;
    grid_init(c->g, p->D, Ng, p->perdir);
    c->g.exitx = p->exit_bc && !(p->perdir & 1);
    c->t = tiling(c->g);
    if (hipSetDevice(c->dev) != hipSuccess || hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->e0) != hipSuccess || hipEventCreate(&c->e1) != hipSuccess) {
        delete c;
        return IA_ERR_HIP;
    }
    const char *why = "";
    int r = noalloc_setup(c->dev, &why);
    if (r != IA_OK) fail(c, r, "no-allocation scratch: %s", why);
    if (r == IA_OK) r = dispatch(c, [&](auto impl) { return decltype(impl)::create(c); });
    if (r == IA_OK) r = hipStreamSynchronize(c->s) == hipSuccess ? IA_OK : IA_ERR_HIP;
    if (r != IA_OK) {
        fprintf(stderr, "ia_create: %s\n", c->msg);
        ia_destroy(c);
        return r;
    }
    *out = c;
    return IA_OK;
}

void ia_destroy(ia_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->dev);
    if (c->s) (void)hipStreamSynchronize(c->s);
    void *bufs[] = {c->fb[0], c->fb[1], c->fb[2], c->fb[3], c->u, c->u0, c->ru, c->ru2, c->ruf, c->F, c->p, c->sig,
                    c->st_buf, c->st_fb,
                    c->L, c->al, c->nh, c->cb, c->fpart, c->part, c->st, c->pst, c->fst, c->pp, c->dsc, c->x64, c->r64, c->z64,
                    c->pa64, c->pb64, c->q64, c->vface, c->mface};
    for (void *b : bufs)
        if (b && !is_attached(c, b)) (void)hipFree(b);
    for (int l = 0; l < c->nlev; ++l) {
        Level &lv = c->lev[l];
        void *lb[] = {lv.Dg, lv.iD, lv.e, lv.r};
        for (void *b : lb)
            if (b) (void)hipFree(b);
        if (lv.own) {
            void *ob[] = {lv.L, lv.x, lv.z, lv.e2, lv.z2, lv.r2, lv.w, lv.w2};
            for (void *b : ob)
                if (b) (void)hipFree(b);
        }
    }
    for (auto &p : c->prof)
        for (hipEvent_t e : p.ev) (void)hipEventDestroy(e);
    {
        auto &n = c->cnn;
        for (void *p : n.act)
            if (p) (void)hipFree(p);
        for (void *p : n.grad)
            if (p) (void)hipFree(p);
        void *cp[] = {n.xkeep, n.gtmp, n.io, n.w, n.m, n.v, n.gw, n.zbias, n.wpart, n.ff, n.fb};
        for (void *p : cp)
            if (p) (void)hipFree(p);
    }
    for (hipEvent_t e : c->it_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->pin) (void)hipHostFree(c->pin);
    if (c->e0) (void)hipEventDestroy(c->e0);
    if (c->e1) (void)hipEventDestroy(c->e1);
    if (c->s) (void)hipStreamDestroy(c->s);
    delete c;
}

int ia_set_field(ia_ctx *c, int field, const void *host, size_t nbytes) {
    if (!c) return IA_ERR_BADARG;
    size_t n = field_elems(c, field) * (field == IA_FIELD_CBAR ? 1 : c->es);
    void *b = field_buf(c, field);
    if (!b || n != nbytes) return fail(c, IA_ERR_BADARG, "ia_set_field(%d): %zu bytes, expected %zu", field, nbytes, n);
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipMemcpyAsync(b, host, n, hipMemcpyDefault, c->s));  // host or device source
    HIPCHK(c, hipStreamSynchronize(c->s));
    return IA_OK;
}
int ia_get_field(ia_ctx *c, int field, void *host, size_t nbytes) {
    if (!c) return IA_ERR_BADARG;
    size_t n = field_elems(c, field) * (field == IA_FIELD_CBAR ? 1 : c->es);
    void *b = field_buf(c, field);
    if (!b || n != nbytes) return fail(c, IA_ERR_BADARG, "ia_get_field(%d): %zu bytes, expected %zu", field, nbytes, n);
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipMemcpyAsync(host, b, n, hipMemcpyDefault, c->s));  // host or device destination
    HIPCHK(c, hipStreamSynchronize(c->s));
    return IA_OK;
}
int ia_field_device_ptr(ia_ctx *c, int field, void **dptr, size_t *nbytes) {
    if (!c || !dptr) return IA_ERR_BADARG;
    void *b = field_buf(c, field);
    if (!b) return fail(c, IA_ERR_BADARG, "unknown field %d", field);
    *dptr = b;
    if (nbytes) *nbytes = field_elems(c, field) * (field == IA_FIELD_CBAR ? 1 : c->es);
    return IA_OK;
}
int ia_attach_field(ia_ctx *c, int field, void *dptr, size_t nbytes) {
    if (!c) return IA_ERR_BADARG;
    if (!dptr) return fail(c, IA_ERR_BADARG, "ia_attach_field: null buffer");
    if (!(field == IA_FIELD_F || field == IA_FIELD_F0 || field == IA_FIELD_U || field == IA_FIELD_U0 ||
          field == IA_FIELD_P))
        return fail(c, IA_ERR_BADARG, "ia_attach_field: field %d cannot be attached (f, f0, u, u0 or p)", field);
    size_t need = field_elems(c, field) * c->es;
    if (nbytes != need) return fail(c, IA_ERR_BADARG, "ia_attach_field: %zu bytes, expected %zu", nbytes, need);
    HIPCHK(c, hipSetDevice(c->dev));
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, dptr) != hipSuccess || at.type != hipMemoryTypeDevice || at.device != c->dev)
        return fail(c, IA_ERR_BADARG, "ia_attach_field: not device memory of device %d", c->dev);
    HIPCHK(c, hipStreamSynchronize(c->s));
    void *old = nullptr;
    switch (field) {
    case IA_FIELD_F:
    case IA_FIELD_F0: {
        int &slot = field == IA_FIELD_F ? c->att_fi : c->att_f0i;
        int q = field == IA_FIELD_F ? c->fi : c->f0i;
        old = c->fb[q];
        c->fb[q] = dptr;
        slot = q;
        c->ls_valid = false;
        break;
    }
    case IA_FIELD_U: old = c->u; c->u = c->att_u = dptr; break;
    case IA_FIELD_U0: old = c->u0; c->u0 = c->att_u0 = dptr; break;
    default: old = c->p; c->p = c->att_p = dptr; c->lev[0].x = dptr; break;
    }
    if (old && old != dptr && !is_attached(c, old)) HIPCHK(c, hipFree(old));
    return IA_OK;
}
int ia_mg_levels(ia_ctx *c, int l, int *Ng) {
    if (!c) return IA_ERR_BADARG;
    if (Ng && l >= 0 && l < c->nlev)
        for (int a = 0; a < 3; ++a) Ng[a] = c->lev[l].g.Ng[a];
    return c->nlev;
}
int ia_apply_vof(ia_ctx *c, const void *sdf, size_t nbytes) {
    if (c) c->ls_valid = false;
    return dispatch(c, [&](auto impl) { return decltype(impl)::apply_vof(c, sdf, nbytes); });
}
int ia_set_initial_velocity(ia_ctx *c, const void *u, size_t nbytes) {
    return dispatch(c, [&](auto impl) { return decltype(impl)::set_initial_velocity(c, u, nbytes); });
}
int ia_cfl(ia_ctx *c, double *dt) {
    if (!dt) return IA_ERR_BADARG;
    return dispatch(c, [&](auto impl) { return decltype(impl)::cfl(c, dt); });
}
int ia_mom_step(ia_ctx *c, double dt, int step_index, ia_step_stats *stats) {
    if (c) c->ls_valid = false;
    if (stats) memset(stats, 0, sizeof *stats);
    return dispatch(c, [&](auto impl) { return decltype(impl)::mom_step(c, dt, step_index, stats); });
}
int ia_advect(ia_ctx *c, int f, int u1, int u2, double dt, int step_index, ia_step_stats *stats) {
    if (!c) return IA_ERR_BADARG;
    if (stats) memset(stats, 0, sizeof *stats);
    if (f != IA_FIELD_F && f != IA_FIELD_F0)
        return fail(c, IA_ERR_BADARG, "ia_advect: f must be IA_FIELD_F or IA_FIELD_F0 (got %d)", f);
    auto vel_ok = [](int id) {
        return id == IA_FIELD_U || id == IA_FIELD_U0 || id == IA_FIELD_FORCE || id == IA_FIELD_RHOU;
    };
    if (!vel_ok(u1) || !vel_ok(u2))
        return fail(c, IA_ERR_BADARG, "ia_advect: u1, u2 must be IA_FIELD_U, _U0, _FORCE or _RHOU (got %d, %d)", u1,
                    u2);
    c->ls_valid = false;
    int fsrc = f == IA_FIELD_F ? c->fi : c->f0i;
    const void *b1 = field_buf(c, u1), *b2 = field_buf(c, u2);
    return dispatch(c, [&](auto impl) { return decltype(impl)::advect(c, fsrc, b1, b2, dt, step_index, stats); });
}
int ia_reconstruct(ia_ctx *c) {
    return dispatch(c, [&](auto impl) { return decltype(impl)::reconstruct_stage(c); });
}
int ia_vof_sweep(ia_ctx *c, int d, double dt, ia_step_stats *stats) {
    if (!c || d < 1 || d > c->g.D) return IA_ERR_BADARG;
    c->ls_valid = false;
    if (stats) memset(stats, 0, sizeof *stats);
    return dispatch(c, [&](auto impl) { return decltype(impl)::vof_sweep_stage(c, d - 1, dt, stats); });
}
int ia_visc(ia_ctx *c, int use_f0) {
    return dispatch(c, [&](auto impl) { return decltype(impl)::visc_stage(c, use_f0); });
}
int ia_update_L(ia_ctx *c, int use_f0) {
    return dispatch(c, [&](auto impl) { return decltype(impl)::update_L_stage(c, use_f0); });
}
int ia_mg_update(ia_ctx *c) {
    return dispatch(c, [&](auto impl) { return decltype(impl)::mg_update(c); });
}
int ia_project(ia_ctx *c, double dt_w, ia_solve_stats *stats) {
    if (stats) memset(stats, 0, sizeof *stats);
    return dispatch(c, [&](auto impl) { return decltype(impl)::project(c, c->u, dt_w, stats); });
}
int ia_synchronize(ia_ctx *c) {
    if (!c) return IA_ERR_BADARG;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return IA_OK;
}
int ia_metrics(ia_ctx *c, const double *U, const double *statwl, double *out) {
    if (!c || !out) return IA_ERR_BADARG;
    double z[3] = {0, 0, 0};
    (void)hipSetDevice(c->dev);
    return dispatch(c, [&](auto impl) { return decltype(impl)::metrics(c, U ? U : z, statwl ? statwl : z, out); });
}
int ia_metrics_noalloc(ia_ctx *c, const double *U, const double *statwl, double *out) {
    if (!c || !out) return IA_ERR_BADARG;
    if (c->dev < 0 || c->dev >= NOALLOC_MAXDEV || !noalloc_dptr[c->dev])
        return fail(c, IA_ERR_BADARG, "ia_metrics_noalloc: no scratch address for device %d", c->dev);
    double *ws = noalloc_dptr[c->dev];
    double z[3] = {0, 0, 0};
    return dispatch(c, [&](auto impl) {
        return decltype(impl)::metrics_noalloc(c, ws, U ? U : z, statwl ? statwl : z, out);
    });
}
int ia_levelset(ia_ctx *c) {
    if (!c) return IA_ERR_BADARG;
    (void)hipSetDevice(c->dev);
    return dispatch(c, [&](auto impl) { return decltype(impl)::levelset(c); });
}
int ia_redistance(ia_ctx *c, double d, double dtau, int perdir) {
    if (!c) return IA_ERR_BADARG;
    if (!c->ls_valid) return fail(c, IA_ERR_BADARG, "ia_redistance: no level set (call ia_levelset after the last advection)");
    if (perdir < 0 || perdir >= (1 << c->g.D)) return fail(c, IA_ERR_BADARG, "ia_redistance: perdir mask %d", perdir);
    double r = d / dtau;
    if (!std::isfinite(r)) return fail(c, IA_ERR_BADARG, "ia_redistance: d/dτ = %g is not finite", r);
    r = std::nearbyint(r);  // round(d/dτ), ties to even
    if (r > 1e8) return fail(c, IA_ERR_BADARG, "ia_redistance: %g pseudo-time steps", r);
    for (int a = 0; a < c->g.D; ++a)
        if (c->g.Ng[a] < 4) return fail(c, IA_ERR_BADARG, "ia_redistance: axis %d has fewer than 2 interior cells", a + 1);
    (void)hipSetDevice(c->dev);
    long itmx = r > 0 ? (long)r : 0;
    return dispatch(c, [&](auto impl) { return decltype(impl)::redistance(c, itmx, dtau, perdir); });
}
int ia_learned_init(ia_ctx *c, unsigned seed) {
    if (!c || c->P.dtype != IA_F32 || c->g.D != 3 || (c->g.Ng[0] - 2) % 16) return IA_ERR_BADARG;
    (void)hipSetDevice(c->dev);
    return Impl<float>::cnn_init(c, seed);
}
int ia_learned_set_weights(ia_ctx *c, const float *w, size_t n) {
    if (!c || !w || c->P.dtype != IA_F32 || c->g.D != 3 || (c->g.Ng[0] - 2) % 16) return IA_ERR_BADARG;
    if (n != (size_t)CNN_NPAR) return fail(c, IA_ERR_BADARG, "ia_learned_set_weights: %zu values, expected %d", n, CNN_NPAR);
    (void)hipSetDevice(c->dev);
    return Impl<float>::cnn_set(c, w);
}
int ia_learned_get_weights(ia_ctx *c, float *w, size_t n) {
    if (!c || !w || !c->cnn.ready) return IA_ERR_BADARG;
    if (n != (size_t)CNN_NPAR) return fail(c, IA_ERR_BADARG, "ia_learned_get_weights: %zu values, expected %d", n, CNN_NPAR);
    (void)hipSetDevice(c->dev);
    HIPCHK(c, hipMemcpyAsync(w, c->cnn.w, CNN_NPAR * sizeof(float), hipMemcpyDeviceToHost, c->s));
    HIPCHK(c, hipStreamSynchronize(c->s));
    return IA_OK;
}
int ia_learned_train(ia_ctx *c, int on, double lr) {
    if (!c || !c->cnn.ready || !(lr > 0)) return IA_ERR_BADARG;
    c->cnn.train = on != 0;
    c->cnn.lr = lr;
    return IA_OK;
}
int ia_learned_stats(ia_ctx *c, ia_learned_info *s) {
    if (!c || !s) return IA_ERR_BADARG;
    s->projections = c->cnn.projections;
    s->accepted = c->cnn.accepted;
    s->train_steps = (int)c->cnn.step;
    s->last_ratio = c->cnn.last_ratio;
    return IA_OK;
}
int ia_learned_forward(ia_ctx *c, const float *feat, float *out) {
    if (!c || !feat || !out || c->P.dtype != IA_F32 || c->g.D != 3 || (c->g.Ng[0] - 2) % 16) return IA_ERR_BADARG;
    (void)hipSetDevice(c->dev);
    return Impl<float>::cnn_test_forward(c, feat, out);
}
int ia_learned_grad(ia_ctx *c, const float *feat, const float *dout, float *gw) {
    if (!c || !feat || !dout || !gw || c->P.dtype != IA_F32 || c->g.D != 3 || (c->g.Ng[0] - 2) % 16)
        return IA_ERR_BADARG;
    (void)hipSetDevice(c->dev);
    return Impl<float>::cnn_test_grad(c, feat, dout, gw);
}
int ia_profile_enable(ia_ctx *c, int kernel, int on) {
    if (!c || kernel < 0 || kernel >= IA_PROF_NCLASS) return IA_ERR_BADARG;
    c->prof[kernel].on = on != 0;
    c->prof[kernel].n = 0;
    c->prof[kernel].launches = 0;
    return IA_OK;
}
int ia_profile_read(ia_ctx *c, int kernel, int *launches, double *total_ms) {
    if (!c || kernel < 0 || kernel >= IA_PROF_NCLASS) return IA_ERR_BADARG;
    auto &p = c->prof[kernel];
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->s));
    double t = 0;
    for (int q = 0; q < p.n; ++q) {
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, p.ev[2 * q], p.ev[2 * q + 1]));
        t += ms;
    }
    if (launches) *launches = (int)p.launches;
    if (total_ms) *total_ms = t;
    return IA_OK;
}
const char *ia_last_error(ia_ctx *c) { return c ? c->msg : "null context"; }

}  // extern "C"
