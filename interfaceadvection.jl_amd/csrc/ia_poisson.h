// ia_poisson.h -- the pressure solve: WaterLily's MultiLevelPoisson
// (Jacobi!/increment!/restrict!/prolongate!/pcg!/residual!/solver!) and the
// FP64 flexible CG that uses its V-cycle as a preconditioner.
//
// Solver kernels walk a column of ZB z-planes per thread (x,y from the
// 64x4 tile), which cuts the block count (so a reduction's partial sums are
// few) and lets each z-neighbour be loaded once.  Every reduction is
// finished inside its own kernel by the last block to arrive (fixed-order
// sum of the block partials, then the scalar update), so no 1-block
// finishing launches are needed.
#pragma once
#include "ia_kernels.h"

namespace ia {

struct CTiling {
    int nt[3];
    int zb;
    long nblk;
};
__device__ inline bool thread_col(const Grid &g, const CTiling &t, int p[3], int &zlo, int &zhi) {
    long b = swizzle_block((long)blockIdx.x, t.nblk);
    long bx = b % t.nt[0];
    long r = b / t.nt[0];
    long by = r % t.nt[1];
    long bz = r / t.nt[1];
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    p[0] = 1 + (int)bx * TX + tx;
    p[1] = g.D > 1 ? 1 + (int)by * TY + ty : 0;
    if (g.D > 2) {
        zlo = 1 + (int)bz * t.zb;
        zhi = zlo + t.zb - 1;
        if (zhi > g.Ng[2] - 2) zhi = g.Ng[2] - 2;
    } else {
        zlo = zhi = 0;
    }
    p[2] = zlo;
    bool ok = p[0] <= g.Ng[0] - 2;
    if (g.D > 1) ok = ok && p[1] <= g.Ng[1] - 2;
    return ok;
}

// scalar state of an inner pcg! step, one slot per launch: each slot is
// written by block 0 of exactly one kernel and read by later kernels, so no
// block ever reads a value that is being written in the same launch
struct PcgState {
    double rho, alpha;
    int done, stop;
};
// scalar state of the flexible CG
struct FcgState {
    double sload, wr, beta, pq, alpha;
};

// every block of a consumer sums the producer's per-block partials itself
// (fixed order: strided per-thread sums, then the block tree), so a
// reduction costs neither an atomic nor an extra launch
__device__ inline double block_total(const double *part, long n, int K = 1, int k = 0) {
    __shared__ double tot;
    double v = 0;
    for (long q = threadIdx.x; q < n; q += NT) v += part[K * q + k];
    double s = block_sum(v);
    if (threadIdx.x == 0) tot = s;
    __syncthreads();
    double r = tot;
    __syncthreads();
    return r;
}
__device__ inline void block_partial(double v, double *part, int K = 1, int k = 0) {
    double s = block_sum(v);
    if (threadIdx.x == 0) part[K * blockIdx.x + k] = s;
}

template <class T> struct LevV {
    Grid g;
    CTiling t;
    const T *L, *Dg, *iD;
    T *x, *e, *r, *z;
};

// neighbour offsets of an interior cell: ±δ on every axis, wrapped on a
// periodic axis at the boundary (perBC! on read); ghosts of non-periodic
// axes are read as stored (their coefficient L is zero)
struct Nb {
    long m[3], p[3];
};
template <int D> __device__ inline Nb neighbours_d(const Grid &g, const int p[3]) {
    Nb n;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        long s = g.st[a];
        n.m[a] = -s;
        n.p[a] = s;
        if (g.per[a]) {
            long w = (long)(g.Ng[a] - 3) * s;
            if (p[a] == 1) n.m[a] = w;
            if (p[a] == g.Ng[a] - 2) n.p[a] = -w;
        }
    }
    return n;
}
// mult(I,L,D,x)  (WaterLily: x[I]D[I] + Σ_i x[I-δi]L[I,i] + Σ_i x[I+δi]L[I+δi,i])
template <int D, class T>
__device__ inline T mgmult_d(const Grid &g, const int p[3], long I, const T *__restrict__ L, const T *__restrict__ Dg,
                             const T *__restrict__ x) {
    Nb n = neighbours_d<D>(g, p);
    T xm[D], xp[D], lm[D], lp[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        xm[i] = x[I + n.m[i]];
        xp[i] = x[I + n.p[i]];
        lm[i] = L[g.ns * i + I];
        lp[i] = L[g.ns * i + I + g.st[i]];
    }
    T sl = 0, su = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) sl += xm[i] * lm[i];
#pragma unroll
    for (int i = 0; i < D; ++i) su += xp[i] * lp[i];
    return (x[I] * Dg[I] + sl) + su;
}
template <class T> __device__ inline T mgmult(const Grid &g, const int p[3], long I, const T *L, const T *Dg, const T *x) {
    return g.D == 3 ? mgmult_d<3>(g, p, I, L, Dg, x) : mgmult_d<2>(g, p, I, L, Dg, x);
}

#define COL_LOOP(lv, ...)                                           \
    {                                                               \
        int p[3], zlo, zhi;                                         \
        if (thread_col(lv.g, lv.t, p, zlo, zhi))                    \
            for (int z_ = zlo; z_ <= zhi; ++z_) {                   \
                p[2] = z_;                                          \
                long I = (long)p[0] + lv.g.st[1] * p[1] + lv.g.st[2] * p[2]; \
                __VA_ARGS__                                         \
            }                                                       \
    }

// residual! part 1: r = z - A x (0 where iD == 0); partial Σr
template <class T>
__global__ void __launch_bounds__(NT) k_mg_resid(LevV<T> l, double *part) {
    double v = 0;
    COL_LOOP(l, {
        T rr = l.iD[I] == T(0) ? T(0) : l.z[I] - mgmult(l.g, p, I, l.L, l.Dg, l.x);
        l.r[I] = rr;
        v += (double)rr;
    })
    block_partial(v, part);
}
// residual! part 2: r -= mean when |mean| > 2eps (sub), fused with partial r·r
template <class T>
__global__ void __launch_bounds__(NT) k_mg_center_l2(LevV<T> l, const double *part_in, long ncell, int sub,
                                                     double *part) {
    bool doit = false;
    T m = 0;
    if (sub) {
        m = (T)block_total(part_in, l.t.nblk) / (T)ncell;
        doit = !(fabs(m) <= 2 * eps_of<T>());
    }
    double v = 0;
    COL_LOOP(l, {
        T rr = l.r[I];
        if (doit) {
            rr = rr - m;
            l.r[I] = rr;
        }
        v += (double)rr * (double)rr;
    })
    block_partial(v, part);
}
// Jacobi!: ϵ = r·iD
template <class T>
__global__ void __launch_bounds__(NT) k_mg_jacobi(LevV<T> l) {
    COL_LOOP(l, { l.e[I] = l.r[I] * l.iD[I]; })
}
// increment!: r -= Aϵ; x += ϵ
template <class T>
__global__ void __launch_bounds__(NT) k_mg_increment(LevV<T> l) {
    COL_LOOP(l, {
        l.r[I] = l.r[I] - mgmult(l.g, p, I, l.L, l.Dg, l.e);
        l.x[I] = l.x[I] + l.e[I];
    })
}
// restrict!: coarse r = Σ children; coarse x = 0
template <class T>
__global__ void __launch_bounds__(NT) k_mg_restrict(LevV<T> c, Grid gf, const T *rf) {
    COL_LOOP(c, {
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = a < c.g.D ? 2 * p[a] - 1 : 0;
            hi[a] = a < c.g.D ? 2 * p[a] : 0;
        }
        T s = 0;
        for (int c2 = lo[2]; c2 <= hi[2]; ++c2)
            for (int c1 = lo[1]; c1 <= hi[1]; ++c1)
                for (int c0 = lo[0]; c0 <= hi[0]; ++c0) s += rf[c0 + gf.st[1] * c1 + gf.st[2] * c2];
        c.r[I] = s;
        c.x[I] = T(0);
    })
}
// prolongate!: fine ϵ = coarse x[down(I)]
template <class T>
__global__ void __launch_bounds__(NT) k_mg_prolong(LevV<T> f, Grid gc, const T *xc) {
    COL_LOOP(f, {
        int q[3];
        for (int a = 0; a < 3; ++a) q[a] = a < f.g.D ? (p[a] + 3) / 2 - 1 : 0;
        f.e[I] = xc[(long)q[0] + gc.st[1] * q[1] + gc.st[2] * q[2]];
    })
}
// pcg! (WaterLily):  z = ϵ = r·iD ; partial r·z  (-> ρ₀)
template <class T>
__global__ void __launch_bounds__(NT) k_pcg_init(LevV<T> l, double *partA) {
    double v = 0;
    COL_LOOP(l, {
        T zz = l.r[I] * l.iD[I];
        l.z[I] = zz;
        l.e[I] = zz;
        v += (double)l.r[I] * (double)zz;
    })
    block_partial(v, partA);
}
// iteration i: ρ (from ρ₀ when i == 1, else from the previous ϵ-update), z = Aϵ, partial z·ϵ
template <class T>
__global__ void __launch_bounds__(NT) k_pcg_Ae(LevV<T> l, PcgState *S, int sin, int sout, const double *partA,
                                               double *partB) {
    PcgState st;
    if (sin < 0) {
        T rho = (T)block_total(partA, l.t.nblk);
        st.rho = (double)rho;
        st.alpha = 0;
        st.done = fabs(rho) < 10 * eps_of<T>() ? 1 : 0;
        st.stop = 0;
    } else {
        st = S[sin];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) S[sout] = st;
    if (st.done) return;
    double v = 0;
    COL_LOOP(l, {
        T zz = mgmult(l.g, p, I, l.L, l.Dg, l.e);
        l.z[I] = zz;
        v += (double)zz * (double)l.e[I];
    })
    block_partial(v, partB);
}
// α = ρ/(z·ϵ); x += αϵ; r -= αz; z = r·iD; partial r·z; stop when i == itmax or |α| < 1e-2
template <class T>
__global__ void __launch_bounds__(NT) k_pcg_update(LevV<T> l, PcgState *S, int sin, int sout, const double *partB,
                                                   double *partC, int i, int itmax) {
    PcgState st = S[sin];
    if (st.done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) S[sout] = st;
        return;
    }
    T alpha = (T)st.rho / (T)block_total(partB, l.t.nblk);
    st.alpha = (double)alpha;
    st.stop = (i == itmax || fabs((double)alpha) < 1e-2) ? 1 : 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) S[sout] = st;
    double v = 0;
    COL_LOOP(l, {
        l.x[I] += alpha * l.e[I];
        T rr = l.r[I] - alpha * l.z[I];
        l.r[I] = rr;
        T zz = rr * l.iD[I];
        l.z[I] = zz;
        v += (double)rr * (double)zz;
    })
    block_partial(v, partC);
}
// ρ₂ = r·z; exit when stop or |ρ₂| < 10eps; β = ρ₂/ρ; ϵ = βϵ + z
template <class T>
__global__ void __launch_bounds__(NT) k_pcg_eps(LevV<T> l, PcgState *S, int sin, int sout, const double *partC) {
    PcgState st = S[sin];
    if (!st.done) {
        if (st.stop) {
            st.done = 1;
        } else {
            T rho2 = (T)block_total(partC, l.t.nblk);
            if (fabs(rho2) < 10 * eps_of<T>()) {
                st.done = 1;
            } else {
                T beta = rho2 / (T)st.rho;
                st.rho = (double)rho2;
                if (blockIdx.x == 0 && threadIdx.x == 0) S[sout] = st;
                COL_LOOP(l, { l.e[I] = beta * l.e[I] + l.z[I]; })
                return;
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) S[sout] = st;
}
// perBC!(x): periodic images of the interior
template <class T>
__global__ void __launch_bounds__(NT) k_mg_perbc(LevV<T> l, T *x) {
    const Grid &g = l.g;
    COL_LOOP(l, {
        bool any = false;
        for (int a = 0; a < g.D; ++a) any |= g.per[a] && (p[a] == 1 || p[a] == g.Ng[a] - 2);
        if (any) {
            int alt[3][2], na[3];
            for (int a = 0; a < 3; ++a) {
                na[a] = 1;
                alt[a][0] = p[a];
                if (a < g.D && g.per[a]) {
                    if (p[a] == 1) { alt[a][1] = g.Ng[a] - 1; na[a] = 2; }
                    else if (p[a] == g.Ng[a] - 2) { alt[a][1] = 0; na[a] = 2; }
                }
            }
            T v = x[I];
            for (int i2 = 0; i2 < na[2]; ++i2)
                for (int i1 = 0; i1 < na[1]; ++i1)
                    for (int i0 = 0; i0 < na[0]; ++i0)
                        x[(long)alt[0][i0] + g.st[1] * alt[1][i1] + g.st[2] * alt[2][i2]] = v;
        }
    })
}

// ---------------------------------------------------------------------------
// FP64 flexible CG (FCG(1)) preconditioned by the FP32 V-cycle
// ---------------------------------------------------------------------------
template <class T> struct FcgV {
    Grid g;
    CTiling t;
    const T *L, *iD;   // level-0 coefficients (values of type T, evaluated in FP64)
    double *z, *x, *r, *pa, *pb, *q;
    T *w;              // preconditioner output (level-0 x)
    T *wr;             // preconditioner input (level-0 r)
    FcgState *S;
    long ncell;
};
template <int D, class T>
__device__ inline double mult64_d(const Grid &g, const int p[3], long I, const T *__restrict__ L,
                                  const double *__restrict__ x) {
    Nb n = neighbours_d<D>(g, p);
    double lm[D], lp[D], xm[D], xp[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        lm[i] = (double)L[g.ns * i + I];
        lp[i] = (double)L[g.ns * i + I + g.st[i]];
        xm[i] = x[I + n.m[i]];
        xp[i] = x[I + n.p[i]];
    }
    double dg = 0, sl = 0, su = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) dg -= (lm[i] + lp[i]);
#pragma unroll
    for (int i = 0; i < D; ++i) sl += xm[i] * lm[i];
#pragma unroll
    for (int i = 0; i < D; ++i) su += xp[i] * lp[i];
    return (x[I] * dg + sl) + su;
}
template <class T> __device__ inline double mult64(const Grid &g, const int p[3], long I, const T *L, const double *x) {
    return g.D == 3 ? mult64_d<3>(g, p, I, L, x) : mult64_d<2>(g, p, I, L, x);
}
// z = div u (FP64), x = p·dt ; partial Σz
template <class T>
__global__ void __launch_bounds__(NT) k_fcg_init(FcgV<T> f, const T *u, const T *pT, double dtw, double *part) {
    double v = 0;
    COL_LOOP(f, {
        double s = 0;
        for (int i = 0; i < f.g.D; ++i) s += (double)u[f.g.ns * i + I + f.g.st[i]] - (double)u[f.g.ns * i + I];
        f.z[I] = s;
        f.x[I] = (double)pT[I] * dtw;
        v += s;
    })
    block_partial(v, part);
}
// z -= mean(z) ; partial ‖z‖²
template <class T>
__global__ void __launch_bounds__(NT) k_fcg_center_z(FcgV<T> f, const double *part_in, double *part) {
    double m = block_total(part_in, f.t.nblk) / (double)f.ncell;
    double v = 0;
    COL_LOOP(f, {
        double q = f.z[I] - m;
        f.z[I] = q;
        v += q * q;
    })
    block_partial(v, part);
}
// r = z - A x ; partial Σr
template <class T>
__global__ void __launch_bounds__(NT) k_fcg_resid(FcgV<T> f, double *part) {
    double v = 0;
    COL_LOOP(f, {
        double rr = f.iD[I] == T(0) ? 0.0 : f.z[I] - mult64(f.g, p, I, f.L, f.x);
        f.r[I] = rr;
        v += rr;
    })
    block_partial(v, part);
}
// r -= mean(r) ; partial ‖r‖²
template <class T>
__global__ void __launch_bounds__(NT) k_fcg_center_r(FcgV<T> f, const double *part_in, double *part) {
    double m = block_total(part_in, f.t.nblk) / (double)f.ncell;
    double v = 0;
    COL_LOOP(f, {
        double q = f.r[I] - m;
        f.r[I] = q;
        v += q * q;
    })
    block_partial(v, part);
}
// preconditioner input r/s (s = ‖r‖/√n, so WaterLily's absolute pcg! exit
// thresholds see a unit-RMS residual), x = 0
template <class T>
__global__ void __launch_bounds__(NT) k_fcg_load(FcgV<T> f, const double *part_rr) {
    double s = sqrt(block_total(part_rr, f.t.nblk) / (double)f.ncell);
    if (blockIdx.x == 0 && threadIdx.x == 0) f.S->sload = s;
    COL_LOOP(f, {
        f.wr[I] = (T)(f.r[I] / s);
        f.w[I] = T(0);
    })
}
// partials of w·r and w·q (w = s·x_T)
template <class T>
__global__ void __launch_bounds__(NT) k_fcg_dots(FcgV<T> f, double *part, int first) {
    double a = 0, b = 0;
    double s = f.S->sload;
    COL_LOOP(f, {
        double w = s * (double)f.w[I];
        a += w * f.r[I];
        if (!first) b += w * f.q[I];
    })
    block_partial(a, part, 2, 0);
    block_partial(b, part, 2, 1);
}
// p' = s·w + βp at the cell and its neighbours; q = A p' (FP64)
template <int D, class T>
__device__ inline void fcg_apply(const FcgV<T> &f, const int p[3], long I, double s, double beta, double &pc,
                                 double &qq) {
    const Grid &g = f.g;
    const T *__restrict__ w = f.w;
    const T *__restrict__ L = f.L;
    const double *__restrict__ pa = f.pa;
    Nb n = neighbours_d<D>(g, p);
    double lm[D], lp[D], pm[D], pp[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        lm[i] = (double)L[g.ns * i + I];
        lp[i] = (double)L[g.ns * i + I + g.st[i]];
        long Km = I + n.m[i], Kp = I + n.p[i];
        pm[i] = s * (double)w[Km] + beta * pa[Km];
        pp[i] = s * (double)w[Kp] + beta * pa[Kp];
    }
    pc = s * (double)w[I] + beta * pa[I];
    double dg = 0, sl = 0, su = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) dg -= (lm[i] + lp[i]);
#pragma unroll
    for (int i = 0; i < D; ++i) sl += pm[i] * lm[i];
#pragma unroll
    for (int i = 0; i < D; ++i) su += pp[i] * lp[i];
    qq = (pc * dg + sl) + su;
}
// ρ = w·r, β = -(w·q)/(p·q); p' = w + βp (on the fly at the stencil points);
// q = A p'; partial p'·q
template <class T>
__global__ void __launch_bounds__(NT) k_fcg_Ap(FcgV<T> f, const double *part_dots, double *part, int first) {
    double wr = block_total(part_dots, f.t.nblk, 2, 0);
    double wq = first ? 0.0 : block_total(part_dots, f.t.nblk, 2, 1);
    FcgState st = *f.S;
    double beta = first ? 0.0 : -wq / st.pq;
    double s = st.sload;
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        f.S->wr = wr;
        f.S->beta = beta;
    }
    const Grid &g = f.g;
    double v = 0;
    COL_LOOP(f, {
        double pc, qq;
        if (g.D == 3) fcg_apply<3>(f, p, I, s, beta, pc, qq);
        else fcg_apply<2>(f, p, I, s, beta, pc, qq);
        f.pb[I] = pc;
        f.q[I] = qq;
        v += pc * qq;
    })
    block_partial(v, part);
}
// α = ρ/(p·q); x += αp ; r -= αq ; partial ‖r‖²
template <class T>
__global__ void __launch_bounds__(NT) k_fcg_update(FcgV<T> f, const double *part_pq, double *part) {
    double pq = block_total(part_pq, f.t.nblk);
    double alpha = f.S->wr / pq;
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        f.S->pq = pq;
        f.S->alpha = alpha;
    }
    double v = 0;
    COL_LOOP(f, {
        f.x[I] += alpha * f.pb[I];
        double rr = f.r[I] - alpha * f.q[I];
        f.r[I] = rr;
        v += rr * rr;
    })
    block_partial(v, part);
}
// plain multigrid refinement: x += s·x_T ; r = z - A x is recomputed by the caller
template <class T>
__global__ void __launch_bounds__(NT) k_ir_accum(FcgV<T> f) {
    double s = f.S->sload;
    COL_LOOP(f, { f.x[I] += s * (double)f.w[I]; })
}
// u -= L ∂x ; p = x/dt ; BC!(u)  (flow.jl:331-334, :82)
template <class T>
__global__ void __launch_bounds__(NT) k_fcg_project(FcgV<T> f, T *u, T *pT, double dtw, T ubc0, T ubc1, T ubc2) {
    const Grid &g = f.g;
    T ub[3] = {ubc0, ubc1, ubc2};
    COL_LOOP(f, {
        for (int i = 0; i < g.D; ++i) {
            int q[3] = {p[0], p[1], p[2]};
            q[i] -= 1;
            double xm = xper(g, (const double *)f.x, q[0], q[1], q[2]);
            T un = (T)((double)u[g.ns * i + I] - (double)f.L[g.ns * i + I] * (f.x[I] - xm));
            write_images<T>(g, u + g.ns * i, p, un, 1, i, ub[i]);
        }
        pT[I] = (T)(f.x[I] / dtw);
    })
}

}  // namespace ia

// ---------------------------------------------------------------------------
// the coarse tail of the V-cycle in one workgroup: Vcycle!(l_t) + pcg!(l_t)
// for every level from l_t down (levels of <= 4096 cells), so the ~20 tiny
// launches per level become one.  Same arithmetic as the multi-block
// kernels; dot products accumulate in FP64 then round to T.
// ---------------------------------------------------------------------------
namespace ia {
constexpr int NT1 = 1024;
constexpr int MAXLEV = 12;
template <class T> struct TailLev {
    Grid g;
    const T *L, *Dg, *iD;
    T *x, *e, *r, *z;
    int n;  // interior cells
};
template <class T> struct TailArgs {
    TailLev<T> lv[MAXLEV];
    int l0, nlev;
};
__device__ inline double bsum1024(double v, double *sh) {
    v = wave_sum(v);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0;
    for (int q = 0; q < NT1 / 64; ++q) s += sh[q];
    __syncthreads();
    return s;  // every thread
}
__device__ inline void tcell(const Grid &g, int q, int p[3], long &I) {
    int n0 = g.Ng[0] - 2, n1 = g.D > 1 ? g.Ng[1] - 2 : 1;
    p[0] = 1 + q % n0;
    int r = q / n0;
    p[1] = g.D > 1 ? 1 + r % n1 : 0;
    p[2] = g.D > 2 ? 1 + r / n1 : 0;
    I = (long)p[0] + g.st[1] * p[1] + g.st[2] * p[2];
}
#define TLOOP(L, ...)                                      \
    for (int q_ = threadIdx.x; q_ < (L).n; q_ += NT1) {    \
        int p[3];                                          \
        long I;                                            \
        tcell((L).g, q_, p, I);                            \
        __VA_ARGS__                                        \
    }                                                      \
    __syncthreads();

template <class T> __device__ void tail_increment(const TailLev<T> &l) {
    const T *__restrict__ e = l.e;
    T *__restrict__ r = l.r;
    T *__restrict__ x = l.x;
    TLOOP(l, {
        r[I] = r[I] - mgmult(l.g, p, I, l.L, l.Dg, e);
        x[I] = x[I] + e[I];
    })
}
template <class T> __device__ void tail_pcg(const TailLev<T> &l, int it, double *sh) {
    T *__restrict__ e = l.e;
    T *__restrict__ z = l.z;
    T *__restrict__ r = l.r;
    T *__restrict__ x = l.x;
    const T *__restrict__ iD = l.iD;
    double v = 0;
    TLOOP(l, {
        T zz = r[I] * iD[I];
        z[I] = zz;
        e[I] = zz;
        v += (double)r[I] * (double)zz;
    })
    T rho = (T)bsum1024(v, sh);
    if (fabs(rho) < 10 * eps_of<T>()) return;
    for (int i = 1; i <= it; ++i) {
        v = 0;
        TLOOP(l, {
            T zz = mgmult(l.g, p, I, l.L, l.Dg, (const T *)e);
            z[I] = zz;
            v += (double)zz * (double)e[I];
        })
        T alpha = rho / (T)bsum1024(v, sh);
        bool last = i == it || fabs((double)alpha) < 1e-2;
        v = 0;
        TLOOP(l, {
            x[I] += alpha * e[I];
            T rr = r[I] - alpha * z[I];
            r[I] = rr;
            T zz = rr * iD[I];
            z[I] = zz;
            v += (double)rr * (double)zz;
        })
        if (last) return;
        T rho2 = (T)bsum1024(v, sh);
        if (fabs(rho2) < 10 * eps_of<T>()) return;
        T beta = rho2 / rho;
        TLOOP(l, { e[I] = beta * e[I] + z[I]; })
        rho = rho2;
    }
}
template <class T> __global__ void __launch_bounds__(NT1) k_mg_tail(TailArgs<T> a) {
    __shared__ double sh[NT1 / 64];
    int last = a.nlev - 1;
    // descend: Jacobi!(l), restrict!(l -> l+1), coarse x = 0
    for (int l = a.l0; l < last; ++l) {
        const TailLev<T> &f = a.lv[l], &c = a.lv[l + 1];
        TLOOP(f, { f.e[I] = f.r[I] * f.iD[I]; })
        tail_increment(f);
        TLOOP(c, {
            T s = 0;
            for (int c2 = (c.g.D > 2 ? 2 * p[2] - 1 : 0); c2 <= (c.g.D > 2 ? 2 * p[2] : 0); ++c2)
                for (int c1 = 2 * p[1] - 1; c1 <= 2 * p[1]; ++c1)
                    for (int c0 = 2 * p[0] - 1; c0 <= 2 * p[0]; ++c0) s += f.r[c0 + f.g.st[1] * c1 + f.g.st[2] * c2];
            c.r[I] = s;
            c.x[I] = T(0);
        })
    }
    // ascend: pcg!(l+1), prolongate!, increment!(l)
    for (int l = last - 1; l >= a.l0; --l) {
        const TailLev<T> &f = a.lv[l], &c = a.lv[l + 1];
        tail_pcg(c, 6, sh);
        TLOOP(f, {
            int q[3];
            for (int d = 0; d < 3; ++d) q[d] = d < f.g.D ? (p[d] + 3) / 2 - 1 : 0;
            f.e[I] = c.x[(long)q[0] + c.g.st[1] * q[1] + c.g.st[2] * q[2]];
        })
        tail_increment(f);
    }
    tail_pcg(a.lv[a.l0], 6, sh);
}
}  // namespace ia
