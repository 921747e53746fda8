/*
 * ia_oracle.c -- CPU oracle: a loop-by-loop C restatement of the two-phase
 * time step of InterfaceAdvection.jl (MPFMomStep!, advect!) and of the
 * WaterLily.jl v1.x routines it calls.
 *
 * TEST INFRASTRUCTURE ONLY.  It is the checker the HIP library is compared
 * against.  Nothing in the product path links, loads or calls it; only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
 *
 * Conventions (identical to the reference's arrays):
 *   - column-major, x fastest, one ghost layer per side (Ng = N + 2),
 *     vector components last; axis and component numbers are 1-based in the
 *     code below, exactly as in the Julia source, so each loop can be read
 *     against the line it restates.
 *   - the reference's scratch aliasing is reproduced literally
 *     (uStar = n̂, dila = α, Φ = flow.σ, r = flow.f, fFace = n̂, ρu⁰ = n̂;
 *     flow.jl:157-159, :75-76, :98-99), because stale values left in
 *     never-written face slots are read by the step (e.g. the upper wall
 *     face of dρ and of the viscous face fraction).
 *   - literal constants that are Float64 in the Julia source (4.5, 0.5,
 *     2.5, 1.2, 0. in flux_out) are double here too, so the FP32 build
 *     promotes exactly where Julia promotes.
 *
 * Deliberate deviations (each documented where it happens):
 *   - dot products and sums accumulate in double and round to real_t
 *     (the reference uses BLAS in T).
 *   - ghost n̂/α on non-periodic sides are zeroed before each
 *     reconstruction (the reference leaves stale scratch there; only read
 *     when a non-periodic wall has a non-zero normal velocity).
 *
 * Parity pins: every function is checked against the reference's own
 * known-answer tests (test/maintests.jl) or against closed forms /
 * invariants in tests/test_oracle_*.py; see the "pinned by" notes.
 */
#include "ia_oracle.h"
#include <tgmath.h>
#undef I /* complex.h (pulled in by tgmath.h) defines I; the code uses I as the cell index */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef real_t T;
#define EPS_T ((T)(sizeof(T) == 4 ? 1.1920928955078125e-07 : 2.220446049250313e-16))

/* ------------------------------------------------------------------ */
/* grid + ranges (WaterLily util.jl: inside, inside_u, slice, CIj, δ)   */
/* ------------------------------------------------------------------ */
typedef struct {
    int D, Ng[3], per[3];
    long st[3];
    size_t ns;
    int exitbc; /* WaterLily exitBC: BC! leaves the exit face (the upper x face of u_x) alone */
} grid_t;

typedef struct { int lo[3], hi[3]; } rng_t;

static void grid_init(grid_t *g, int D, const int *Ng, int perdir) {
    g->D = D;
    g->exitbc = 0;
    for (int d = 0; d < 3; ++d) {
        g->Ng[d] = d < D ? Ng[d] : 1;
        g->per[d] = d < D ? ((perdir >> d) & 1) : 0;
    }
    g->st[0] = 1;
    g->st[1] = g->Ng[0];
    g->st[2] = (long)g->Ng[0] * g->Ng[1];
    g->ns = (size_t)g->Ng[0] * g->Ng[1] * g->Ng[2];
}

#define NG(j) (g->Ng[(j)-1])
#define ST(j) (g->st[(j)-1])
#define PER(j) (g->per[(j)-1])
#define VC(a, c) ((a) + g->ns * (size_t)((c)-1))

static inline size_t IXg(const grid_t *g, int i, int j, int k) {
    return (size_t)(i - 1) + (size_t)g->Ng[0] * ((size_t)(j - 1) + (size_t)g->Ng[1] * (size_t)(k - 1));
}

static rng_t r_all(const grid_t *g) {
    rng_t r;
    for (int d = 0; d < 3; ++d) { r.lo[d] = 1; r.hi[d] = g->Ng[d]; }
    return r;
}
/* inside(a) : 2:end-1 on every axis */
static rng_t r_in(const grid_t *g) {
    rng_t r;
    for (int d = 0; d < 3; ++d) {
        r.lo[d] = d < g->D ? 2 : 1;
        r.hi[d] = d < g->D ? g->Ng[d] - 1 : 1;
    }
    return r;
}
/* WaterLily inside_u(N,j): axis j 3:N-1, others 2:N-1 */
static rng_t r_in_u(const grid_t *g, int j) {
    rng_t r = r_in(g);
    r.lo[j - 1] = 3;
    return r;
}
/* util.jl:47-49 inside_uWB(dims,j): axis j 2:N, others 2:N-1 */
static rng_t r_in_uWB(const grid_t *g, int j) {
    rng_t r = r_in(g);
    r.hi[j - 1] = g->Ng[j - 1];
    return r;
}
/* WaterLily slice(N,s,j,low): axis j s:s, others low:N */
static rng_t r_slice(const grid_t *g, int s, int j, int low) {
    rng_t r;
    for (int d = 0; d < 3; ++d) {
        r.lo[d] = d < g->D ? low : 1;
        r.hi[d] = d < g->D ? g->Ng[d] : 1;
    }
    r.lo[j - 1] = r.hi[j - 1] = s;
    return r;
}
static long rng_count(rng_t r) {
    long n = 1;
    for (int d = 0; d < 3; ++d) n *= (r.hi[d] - r.lo[d] + 1);
    return n;
}

/* serial and OpenMP loops over a range; body sees i,j,k (1-based) */
#define FOR_R(R)                                      \
    for (int k = (R).lo[2]; k <= (R).hi[2]; ++k)      \
        for (int j = (R).lo[1]; j <= (R).hi[1]; ++j)  \
            for (int i = (R).lo[0]; i <= (R).hi[0]; ++i)
#define PFOR_R(R) _Pragma("omp parallel for collapse(2) schedule(static)") FOR_R(R)
#define IX(i, j, k) IXg(g, i, j, k)
#define PCOORD(P) int P[3] = {i, j, k}

/* deterministic reductions: per-row double partials, summed in row order */
static double r_dot(const grid_t *g, rng_t R, const T *a, const T *b) {
    int nj = R.hi[1] - R.lo[1] + 1, nk = R.hi[2] - R.lo[2] + 1;
    double *part = (double *)calloc((size_t)nj * nk, sizeof(double));
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = R.lo[2]; k <= R.hi[2]; ++k)
        for (int j = R.lo[1]; j <= R.hi[1]; ++j) {
            double s = 0;
            for (int i = R.lo[0]; i <= R.hi[0]; ++i) {
                size_t I = IX(i, j, k);
                s += (double)a[I] * (double)(b ? b[I] : 1);
            }
            part[(size_t)(k - R.lo[2]) * nj + (j - R.lo[1])] = s;
        }
    double s = 0;
    for (long q = 0; q < (long)nj * nk; ++q) s += part[q];
    free(part);
    return s;
}
static T r_max(const grid_t *g, rng_t R, const T *a) {
    T m = a[IX(R.lo[0], R.lo[1], R.lo[2])];
    FOR_R(R) {
        T v = a[IX(i, j, k)];
        if (v > m || v != v) m = v;
    }
    return m;
}

/* ------------------------------------------------------------------ */
/* scalar helpers (Julia semantics)                                     */
/* ------------------------------------------------------------------ */
static inline T tmin(T a, T b) { return b < a ? b : a; }
static inline T tmax(T a, T b) { return b > a ? b : a; }
static inline double dmin(double a, double b) { return b < a ? b : a; }
static inline T tsign(T x) { return x > 0 ? (T)1 : (x < 0 ? (T)-1 : x); }
/* VOFutil.jl:166 linInterpProp(f,λ) = 1*(λ+(1-λ)f) */
static inline T lin(T f, T lam) { return lam + ((T)1 - lam) * f; }
/* VOFutil.jl:157 fullorempty */
static inline int fullorempty(T f) { return f == 0 || f == 1; }
/* WaterLily median(a,b,c) */
static inline T median3(T a, T b, T c) { return tmax(tmin(a, b), tmin(tmax(a, b), c)); }

/* ------------------------------------------------------------------ */
/* PLIC.jl (Scardovelli & Zaleski 2000)                                 */
/* pinned by maintests.jl:25-51 KATs + round trips + closed forms        */
/* ------------------------------------------------------------------ */
/* PLIC.jl:178 */
static inline void sort2(T *a, T *b) {
    if (!(*a < *b)) { T t = *a; *a = *b; *b = t; }
}
/* PLIC.jl:186-191 */
static inline void sort3(T *a, T *b, T *c) {
    T t;
    if (*a > *c) { t = *a; *a = *c; *c = t; }
    if (*a > *b) { t = *a; *a = *b; *b = t; }
    if (*b > *c) { t = *b; *b = *c; *c = t; }
}
/* PLIC.jl:165-176 proot */
static T proot(T c0, T c1, T c2, T c3) {
    T a0 = c0 / c3, a1 = c1 / c3, a2 = c2 / c3;
    T p0 = a1 / 3 - (a2 * a2) / 9;
    T q0 = (a1 * a2 - 3 * a0) / 6 - (a2 * a2 * a2) / 27;
    T a = q0 / sqrt(-(p0 * p0 * p0));
    T t = acos((a * a) <= 1 ? a : (T)0) / 3;
    return sqrt(-p0) * (sqrt((T)3) * sin(t) - cos(t)) - a2 / 3;
}
/* PLIC.jl:92 */
static inline T a2f2(T m1, T m2, T a) {
    return a < m1 ? (a * a) / ((2 * m1) * m2) : (a - m1 / 2) / m2;
}
/* PLIC.jl:93-107 */
static T a2f3(T m1, T m2, T m3, T a) {
    T m12 = m1 + m2;
    if (a < m1) return (a * a * a) / (((6 * m1) * m2) * m3);
    if (a < m2) return (a * (a - m1)) / ((2 * m2) * m3) + ((m2 == 0 ? (T)1 : m1 / m2) * m1) / (6 * m3);
    if (a < tmin(m3, m12))
        return ((a * a) * (3 * m12 - a) + (m1 * m1) * (m1 - 3 * a) + (m2 * m2) * (m2 - 3 * a)) / (((6 * m1) * m2) * m3);
    if (m3 < m12)
        return ((a * a) * (3 - 2 * a) + (m1 * m1) * (m1 - 3 * a) + (m2 * m2) * (m2 - 3 * a) + (m3 * m3) * (m3 - 3 * a)) /
               (((6 * m1) * m2) * m3);
    return (2 * a - m12) / (2 * m3);
}
/* PLIC.jl:117 */
static inline T f2a2(T m1, T m2, T v) {
    return v < m1 / (2 * m2) ? sqrt(((2 * m1) * m2) * v) : m2 * v + m1 / 2;
}
/* PLIC.jl:118-149 */
static T f2a3(T m1, T m2, T m3, T v) {
    T m12 = m1 + m2;
    T p = ((6 * m1) * m2) * m3;
    T v1 = ((m2 == 0 ? (T)1 : m1 / m2) * m1) / (6 * m3);
    T v2 = v1 + (m2 - m1) / (2 * m3);
    T v3 = m3 < m12 ? ((m3 * m3) * (3 * m12 - m3) + (m1 * m1) * (m1 - 3 * m3) + (m2 * m2) * (m2 - 3 * m3)) / p
                    : m12 / (2 * m3);
    if (v < v1) return cbrt(p * v);
    if (v < v2) return (m1 + sqrt(m1 * m1 + ((8 * m2) * m3) * (v - v1))) / 2;
    if (v < v3) {
        T c0 = (m1 * m1 * m1 + m2 * m2 * m2) - p * v;
        T c1 = -3 * (m1 * m1 + m2 * m2);
        T c2 = 3 * m12;
        return proot(c0, c1, c2, (T)-1);
    }
    if (m3 < m12) {
        T c0 = ((m1 * m1 * m1 + m2 * m2 * m2) + m3 * m3 * m3) - p * v;
        T c1 = -3 * ((m1 * m1 + m2 * m2) + m3 * m3);
        return proot(c0, c1, (T)3, (T)-2);
    }
    return m3 * v + m12 / 2;
}
/* PLIC.jl:20-29 */
static T getIntercept2(T n1, T n2, T g) {
    T t = fabs(n1) + fabs(n2), a;
    if (g != (T)0.5) {
        T m1 = fabs(n1) / t, m2 = fabs(n2) / t;
        sort2(&m1, &m2);
        a = f2a2(m1, m2, g < (T)0.5 ? g : 1 - g);
    } else
        a = (T)0.5;
    return ((g < (T)0.5 ? a : 1 - a) * t + tmin(n1, 0)) + tmin(n2, 0);
}
/* PLIC.jl:30-39 */
static T getIntercept3(T n1, T n2, T n3, T g) {
    T t = (fabs(n1) + fabs(n2)) + fabs(n3), a;
    if (g != (T)0.5) {
        T m1 = fabs(n1) / t, m2 = fabs(n2) / t, m3 = fabs(n3) / t;
        sort3(&m1, &m2, &m3);
        a = f2a3(m1, m2, m3, g < (T)0.5 ? g : 1 - g);
    } else
        a = (T)0.5;
    return (((g < (T)0.5 ? a : 1 - a) * t + tmin(n1, 0)) + tmin(n2, 0)) + tmin(n3, 0);
}
/* PLIC.jl:59-70 */
static T getVolumeFraction2(T n1, T n2, T b) {
    T t = fabs(n1) + fabs(n2);
    T a = ((b - tmin(n1, 0)) - tmin(n2, 0)) / t;
    if (a <= 0 || a == (T)0.5 || a >= 1) return tmin(tmax(a, 0), 1);
    T m1 = fabs(n1) / t, m2 = fabs(n2) / t;
    sort2(&m1, &m2);
    T v = a2f2(m1, m2, a < (T)0.5 ? a : 1 - a);
    return a < (T)0.5 ? v : 1 - v;
}
/* PLIC.jl:71-82 */
static T getVolumeFraction3(T n1, T n2, T n3, T b) {
    T t = (fabs(n1) + fabs(n2)) + fabs(n3);
    T a = (((b - tmin(n1, 0)) - tmin(n2, 0)) - tmin(n3, 0)) / t;
    if (a <= 0 || a == (T)0.5 || a >= 1) return tmin(tmax(a, 0), 1);
    T m1 = fabs(n1) / t, m2 = fabs(n2) / t, m3 = fabs(n3) / t;
    sort3(&m1, &m2, &m3);
    T v = a2f3(m1, m2, m3, a < (T)0.5 ? a : 1 - a);
    return a < (T)0.5 ? v : 1 - v;
}
static T getVolumeFractionN(int D, const T *n, T b) {
    return D == 2 ? getVolumeFraction2(n[0], n[1], b) : getVolumeFraction3(n[0], n[1], n[2], b);
}
static T getInterceptN(int D, const T *n, T f) {
    return D == 2 ? getIntercept2(n[0], n[1], f) : getIntercept3(n[0], n[1], n[2], f);
}

/* ------------------------------------------------------------------ */
/* normalEstimation.jl + util.jl                                        */
/* ------------------------------------------------------------------ */
/* VOFutil.jl:158 get3CellHeight */
static inline T h3(const grid_t *g, const T *f, size_t I, int s) {
    return (f[I] + f[I - ST(s)]) + f[I + ST(s)];
}
/* util.jl:19-31 myArgAbsMax(vec,I) */
static int argabsmax(const grid_t *g, const T *nh, size_t I) {
    T mx = 0;
    int im = 1;
    for (int i = 1; i <= g->D; ++i) {
        T v = VC(nh, i)[I];
        T cur = v * v;
        if (cur > mx) { mx = cur; im = i; }
    }
    return im;
}
/* normalEstimation.jl:161-165 */
static void normal_pcd(const grid_t *g, const T *f, T *nh, size_t I) {
    for (int d = 1; d <= g->D; ++d) VC(nh, d)[I] = f[I - ST(d)] - f[I + ST(d)];
}
/* normalEstimation.jl:75-97 */
static void normal_column(const grid_t *g, const T *f, T *nh, size_t I) {
    normal_pcd(g, f, nh, I);
    int dom = argabsmax(g, nh, I);
    for (int d = 1; d <= g->D; ++d) {
        if (d == dom) {
            T s = tsign(VC(nh, d)[I]);
            VC(nh, d)[I] = s == 0 ? (T)1 : s;
            continue;
        }
        T hl = h3(g, f, I - ST(d), dom), hr = h3(g, f, I + ST(d), dom);
        VC(nh, d)[I] = (hl - hr) / 2;
    }
}
/* normalEstimation.jl:105-154 (default scheme).  Float64 literals promote
 * exactly as in Julia.  pinned by: planar interfaces from the exact forward
 * PLIC in both regimes, and first-order convergence on circles and spheres
 * (tests/test_oracle_pins.py) -- exact values on generic blocks are parity
 * unpinned. */
static void normal_wh(const grid_t *g, const T *f, T *nh, size_t I) {
    normal_column(g, f, nh, I);
    int im = argabsmax(g, nh, I);
    int domDir = signbit(VC(nh, im)[I]) ? -im : im; /* util.jl:72-75 majorDir */
    int ad = im;
    T an = fabs(VC(nh, ad)[I]);
    an = an == 0 ? (T)1 : an;
    for (int d = 1; d <= g->D; ++d) VC(nh, d)[I] /= an;
    for (int d = 1; d <= g->D; ++d) {
        if (d == ad) continue;
        T slope = fabs(VC(nh, d)[I]);
        int cur = signbit(VC(nh, d)[I]) ? -d : d;
        long off = (cur > 0 ? 1 : -1) * ST(d); /* util.jl:56 δd */
        T hl = h3(g, f, I - off, ad), hc = h3(g, f, I, ad), hr = h3(g, f, I + off, ad);
        T sh = (hl + hc) + hr;
        double s45 = 4.5 * (double)slope;
        if (s45 <= (double)sh && (double)sh <= 9 - s45) continue;
        long offd = (domDir > 0 ? 1 : -1) * ST(ad);
        double mn = dmin((double)((T)1 / (2 * slope)), s45);
        if ((double)sh < s45) {
            T wb = h3(g, f, I - offd, d);
            if ((double)wb > mn) VC(nh, d)[I] = (T)copysign(((double)hl - 0.5) / ((double)wb - 0.5), (double)cur);
        }
        if ((double)sh > 9 - s45) {
            T wt = h3(g, f, I + offd, d);
            if ((double)wt < 3 - mn) VC(nh, d)[I] = (T)copysign((2.5 - (double)hr) / (2.5 - (double)wt), (double)cur);
        }
    }
}
/* normalEstimation.jl:35-68 (WY; a non-default plug-in kept as a self-check) */
static void normal_wy(const grid_t *g, const T *f, T *nh, size_t I) {
    normal_pcd(g, f, nh, I);
    int dom = argabsmax(g, nh, I);
    for (int d = 1; d <= g->D; ++d) {
        if (d == dom) {
            T s = tsign(VC(nh, d)[I]);
            VC(nh, d)[I] = s == 0 ? (T)1 : s;
            continue;
        }
        T hl = h3(g, f, I - ST(d), dom), hc = h3(g, f, I, dom), hr = h3(g, f, I + ST(d), dom);
        T n = (hl - hr) / 2;
        if (fabs(n) > (T)0.5) n = ((double)n * ((double)hc - 1.5) >= 0.0) ? hc - hr : hl - hc;
        VC(nh, d)[I] = n;
    }
}
/* normalEstimation.jl:248-257 YoungSum */
static T young_sum(const grid_t *g, const T *f, size_t I, int d) {
    /* II ∈ I-δxy:I, III ∈ II:II+δxy  with δxy = oneunit - δ(d), column-major order */
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) { lo[a] = 0; hi[a] = 0; }
    for (int a = 0; a < g->D; ++a) if (a != d - 1) { lo[a] = -1; hi[a] = 0; }
    T s = 0;
    for (int c2 = lo[2]; c2 <= hi[2]; ++c2)
        for (int c1 = lo[1]; c1 <= hi[1]; ++c1)
            for (int c0 = lo[0]; c0 <= hi[0]; ++c0) {
                long II = (long)I + c0 * g->st[0] + c1 * g->st[1] + c2 * g->st[2];
                int l2[3] = {0, 0, 0}, h2[3] = {0, 0, 0};
                for (int a = 0; a < g->D; ++a) if (a != d - 1) h2[a] = 1;
                for (int e2 = l2[2]; e2 <= h2[2]; ++e2)
                    for (int e1 = l2[1]; e1 <= h2[1]; ++e1)
                        for (int e0 = l2[0]; e0 <= h2[0]; ++e0)
                            s += f[II + e0 * g->st[0] + e1 * g->st[1] + e2 * g->st[2]];
            }
    return s;
}
/* normalEstimation.jl:232-247 (Youngs) */
static void normal_youngs(const grid_t *g, const T *f, T *nh, size_t I) {
    T a = 0;
    for (int d = 1; d <= g->D; ++d) {
        VC(nh, d)[I] = (T)((double)(young_sum(g, f, I - ST(d), d) - young_sum(g, f, I + ST(d), d)) * 0.5);
        a += fabs(VC(nh, d)[I]);
    }
    for (int d = 1; d <= g->D; ++d) VC(nh, d)[I] = a == 0 ? (T)(1.0 / g->D) : VC(nh, d)[I] / a;
}
/* normalEstimation.jl:210-224 (CCi) */
static void normal_cci(const grid_t *g, const T *f, const T *nh, size_t I, int dc, T *out) {
    T s = 0;
    for (int d = 1; d <= g->D; ++d) {
        if (d == dc) {
            T sg = tsign(VC(nh, d)[I]);
            out[d - 1] = sg == 0 ? (T)1 : sg;
        } else {
            T hu = h3(g, f, I + ST(d), dc), hd = h3(g, f, I - ST(d), dc);
            out[d - 1] = -(hu - hd) / 2;
        }
    }
    for (int d = 0; d < g->D; ++d) s += fabs(out[d]);
    for (int d = 0; d < g->D; ++d) out[d] /= s;
}
/* normalEstimation.jl:185-202 (MYC) */
static void normal_myc(const grid_t *g, const T *f, T *nh, size_t I) {
    normal_youngs(g, f, nh, I);
    T mx = 0;
    for (int i = 1; i <= g->D; ++i) mx = fabs(VC(nh, i)[I]) > mx ? fabs(VC(nh, i)[I]) : mx;
    T curm0 = 0, cc[3];
    int ccz = 1;
    for (int iz = 1; iz <= g->D; ++iz) {
        normal_cci(g, f, nh, I, iz, cc);
        if (fabs(cc[iz - 1]) > curm0) ccz = iz;
        curm0 = fabs(cc[iz - 1]);
    }
    normal_cci(g, f, nh, I, ccz, cc);
    if (fabs(cc[ccz - 1]) < mx)
        for (int i = 1; i <= g->D; ++i) VC(nh, i)[I] = cc[i - 1];
}
/* normalEstimation.jl:172-176 (SLIC): the dominant PCD axis only */
static void normal_slic(const grid_t *g, const T *f, T *nh, size_t I) {
    normal_pcd(g, f, nh, I);
    int d = argabsmax(g, nh, I);
    for (int i = 1; i <= g->D; ++i) VC(nh, i)[I] = i == d ? tsign(VC(nh, i)[I]) : (T)0;
}
static void normal_scheme(const grid_t *g, int scheme, const T *f, T *nh, size_t I) {
    switch (scheme) {
    case 1: normal_column(g, f, nh, I); break;
    case 2: normal_wy(g, f, nh, I); break;
    case 3: normal_myc(g, f, nh, I); break;
    case 4: normal_pcd(g, f, nh, I); break;    /* normalEstimation.jl:161-165 */
    case 5: normal_youngs(g, f, nh, I); break; /* normalEstimation.jl:232-247 */
    case 6: normal_slic(g, f, nh, I); break;
    default: normal_wh(g, f, nh, I); break;
    }
}

/* ------------------------------------------------------------------ */
/* boundary conditions (VOFutil.jl:44-117, WaterLily BC!, perBC!)      */
/* pinned by maintests.jl:79-129                                       */
/* ------------------------------------------------------------------ */
/* VOFutil.jl:64-74 BCf!(f) (d == 0) and :76-89 BCf!(d,f) / :106-117 BCv1D! */
static void bc_f(const grid_t *g, T *f, int dd) {
    for (int jj = 1; jj <= g->D; ++jj) {
        if (PER(jj)) {
            rng_t R1 = r_slice(g, 1, jj, 1), R2 = r_slice(g, NG(jj), jj, 1);
            FOR_R(R1) { size_t I = IX(i, j, k); f[I] = f[I + (NG(jj) - 2) * ST(jj)]; }
            FOR_R(R2) { size_t I = IX(i, j, k); f[I] = f[I + (2 - NG(jj)) * ST(jj)]; }
        } else if (jj == dd) {
            rng_t R1 = r_slice(g, 1, jj, 1);
            FOR_R(R1) { size_t I = IX(i, j, k); f[I] = f[I + 2 * ST(jj)]; }
        } else {
            rng_t R1 = r_slice(g, 1, jj, 1), R2 = r_slice(g, NG(jj), jj, 1);
            FOR_R(R1) { size_t I = IX(i, j, k); f[I] = f[I + ST(jj)]; }
            FOR_R(R2) { size_t I = IX(i, j, k); f[I] = f[I - ST(jj)]; }
        }
    }
}
/* VOFutil.jl:91-104 BCv! */
static void bc_v(const grid_t *g, T *v) {
    for (int d = 1; d <= g->D; ++d) bc_f(g, VC(v, d), d);
}
/* VOFutil.jl:44-62 BCVOF! */
static void bc_vof(const grid_t *g, T *f, T *al, T *nh) {
    for (int jj = 1; jj <= g->D; ++jj) {
        rng_t R1 = r_slice(g, 1, jj, 1), R2 = r_slice(g, NG(jj), jj, 1);
        if (PER(jj)) {
            FOR_R(R1) {
                size_t I = IX(i, j, k), S = I + (NG(jj) - 2) * ST(jj);
                f[I] = f[S];
                for (int c = 1; c <= g->D; ++c) VC(nh, c)[I] = VC(nh, c)[S];
                al[I] = al[S];
            }
            FOR_R(R2) {
                size_t I = IX(i, j, k), S = I + (2 - NG(jj)) * ST(jj);
                f[I] = f[S];
                for (int c = 1; c <= g->D; ++c) VC(nh, c)[I] = VC(nh, c)[S];
                al[I] = al[S];
            }
        } else {
            FOR_R(R1) { size_t I = IX(i, j, k); f[I] = f[I + ST(jj)]; }
            FOR_R(R2) { size_t I = IX(i, j, k); f[I] = f[I - ST(jj)]; }
        }
    }
}
/* WaterLily BC!(a,A,false,perdir) for a staggered vector field */
/* WaterLily BC!(a, A, saveexit, perdir): periodic copies; normal component
 * Dirichlet A on slices 1, 2 and N (the exit face N of u_x skipped when
 * saveexit); tangential components Neumann */
static void bc_vec_e(const grid_t *g, T *a, const T *A, int saveexit) {
    for (int ii = 1; ii <= g->D; ++ii) {
        T *ai = VC(a, ii);
        for (int jj = 1; jj <= g->D; ++jj) {
            rng_t R1 = r_slice(g, 1, jj, 1), R2 = r_slice(g, 2, jj, 1), R3 = r_slice(g, NG(jj), jj, 1);
            if (PER(jj)) {
                FOR_R(R1) { size_t I = IX(i, j, k); ai[I] = ai[I + (NG(jj) - 2) * ST(jj)]; }
                FOR_R(R3) { size_t I = IX(i, j, k); ai[I] = ai[I + (2 - NG(jj)) * ST(jj)]; }
            } else if (ii == jj) {
                FOR_R(R1) ai[IX(i, j, k)] = A[ii - 1];
                FOR_R(R2) ai[IX(i, j, k)] = A[ii - 1];
                if (!saveexit || ii > 1) FOR_R(R3) ai[IX(i, j, k)] = A[ii - 1];
            } else {
                FOR_R(R1) { size_t I = IX(i, j, k); ai[I] = ai[I + ST(jj)]; }
                FOR_R(R3) { size_t I = IX(i, j, k); ai[I] = ai[I - ST(jj)]; }
            }
        }
    }
}
/* BC!(a, A, exitBC, perdir) as the step calls it (flow.jl:69, :79, :82, :197, :207) */
static void bc_vec(const grid_t *g, T *a, const T *A) { bc_vec_e(g, a, A, g->exitbc); }
/* WaterLily perBC! */
static void per_bc(const grid_t *g, T *a) {
    for (int jj = 1; jj <= g->D; ++jj) {
        if (!PER(jj)) continue;
        rng_t R1 = r_slice(g, 1, jj, 1), R3 = r_slice(g, NG(jj), jj, 1);
        FOR_R(R1) { size_t I = IX(i, j, k); a[I] = a[I + (NG(jj) - 2) * ST(jj)]; }
        FOR_R(R3) { size_t I = IX(i, j, k); a[I] = a[I + (2 - NG(jj)) * ST(jj)]; }
    }
}
/* VOFutil.jl:127-137 cleanWisp! */
static void clean_wisp(const grid_t *g, T *f, T tol) {
    rng_t R = r_in(g);
    PFOR_R(R) {
        size_t I = IX(i, j, k);
        f[I] = f[I] < tol ? (T)0 : (f[I] > 1 - tol ? (T)1 : f[I]);
    }
}
/* VOFutil.jl:229-234 f2face! */
static void f2face(const grid_t *g, T *ff, const T *fc) {
    rng_t R = r_in(g);
    for (int d = 1; d <= g->D; ++d) {
        T *fd = VC(ff, d);
        PFOR_R(R) { size_t I = IX(i, j, k); fd[I] = (fc[I] + fc[I - ST(d)]) / 2; }
    }
    bc_v(g, ff);
}
/* VOFutil.jl:186-191 getμ */
static inline T getmu(const grid_t *g, int i, int jj, size_t I, const T *ff, T lm, T mu, T lr) {
    T f1 = VC(ff, i)[I - ST(jj)], f2 = VC(ff, i)[I], f3 = VC(ff, jj)[I - ST(i)], f4 = VC(ff, jj)[I];
    T s = (((f1 + f2) + f3) + f4) / 4;
    T fr = lr < 1 ? tmin(tmin(tmin(f1, f2), f3), f4) : tmax(tmax(tmax(f1, f2), f3), f4);
    return mu * tmin(lin(s, lm), (s > (T)0.5 ? (T)1 : lm / lr) * lin(fr, lr));
}
/* VOFutil.jl:198-201 ρu2u! */
static void rhou2u(const grid_t *g, T *u, const T *ru, const T *f, T lr) {
    rng_t R = r_in(g);
    PFOR_R(R) {
        size_t I = IX(i, j, k);
        for (int d = 1; d <= g->D; ++d) VC(u, d)[I] = VC(ru, d)[I] / lin((f[I] + f[I - ST(d)]) / 2, lr);
    }
}
/* VOFutil.jl:208-211 u2ρu! */
static void u2rhou(const grid_t *g, T *ru, const T *u, const T *f, T lr) {
    rng_t R = r_in(g);
    PFOR_R(R) {
        size_t I = IX(i, j, k);
        for (int d = 1; d <= g->D; ++d) VC(ru, d)[I] = VC(u, d)[I] * lin((f[I] + f[I - ST(d)]) / 2, lr);
    }
}

/* ------------------------------------------------------------------ */
/* context                                                              */
/* ------------------------------------------------------------------ */
typedef struct {
    grid_t g;
    T *L, *Dg, *iD, *x, *eps, *r, *z;
    int own;
} lev_t;

struct oc_ctx {
    grid_t g;
    int D;
    T uBC[3], gv[3], zero3[3];
    int has_g, has_mu;
    T nu, mu, lr, lm, eta;
    int has_eta;
    int scheme, limiter;
    T *u, *u0, *fv, *p, *sig, *mu0;
    T *f, *f0, *al, *nh, *ff, *ru, *ruf, *dr;
    int8_t *cb;
    lev_t lev[12];
    int nlev;
    double mg_tol;
    int mg_itmx;
    int coarse_it; /* pcg! iterations on coarse levels inside Vcycle! (WaterLily: 6) */
    int fcg_cit[16]; /* the same per level in the flexible CG's preconditioner (by level size) */
    int fcg_cg1;     /* coarse levels above the tail use mg_pcg1 in the flexible CG's preconditioner */
    int tail_level;  /* first level of at most 4096 cells (the HIP library's one-workgroup tail) */
    T *cg1_u, *cg1_w, *cg1_s; /* mg_pcg1 scratch, level-1 sized */
    int poisson_mode, smooth_it;
    int fine_cheb;   /* the flexible CG's fine smoother: mg_cheb2 (else smooth_it pcg! iterations) */
    T *cheb_d0;      /* its d₀ (level-0 sized) */
    double poisson_rtol;
    int last_iters[2];
    char msg[512];
};

static T *alloc_fill(size_t n, T v) {
    T *a = (T *)malloc(n * sizeof(T));
    for (size_t q = 0; q < n; ++q) a[q] = v;
    return a;
}

/* ------------------------------------------------------------------ */
/* WaterLily Poisson / MultiLevelPoisson (semantics restated; the       */
/* iterates are parity unpinned -- the converged solution is pinned by  */
/* a direct sparse solve in tests/test_oracle_poisson.py)               */
/* ------------------------------------------------------------------ */
static inline T mg_mult(const grid_t *g, size_t I, const T *L, const T *Dg, const T *x) {
    T sl = 0, su = 0;
    for (int i = 1; i <= g->D; ++i) sl += x[I - ST(i)] * VC(L, i)[I];
    for (int i = 1; i <= g->D; ++i) su += x[I + ST(i)] * VC(L, i)[I + ST(i)];
    return (x[I] * Dg[I] + sl) + su;
}
static void mg_set_diag(lev_t *lv) {
    grid_t *g = &lv->g;
    rng_t R = r_in(g);
    PFOR_R(R) {
        size_t I = IX(i, j, k);
        T s = 0;
        for (int d = 1; d <= g->D; ++d) s -= (VC(lv->L, d)[I] + VC(lv->L, d)[I + ST(d)]);
        lv->Dg[I] = s;
    }
    PFOR_R(R) {
        size_t I = IX(i, j, k);
        lv->iD[I] = (lv->Dg[I] * lv->Dg[I]) < 2 * EPS_T ? (T)0 : (T)1 / lv->Dg[I];
    }
}
static void mg_residual(lev_t *lv) {
    grid_t *g = &lv->g;
    rng_t R = r_in(g);
    per_bc(g, lv->x);
    PFOR_R(R) {
        size_t I = IX(i, j, k);
        lv->r[I] = lv->iD[I] == 0 ? (T)0 : lv->z[I] - mg_mult(g, I, lv->L, lv->Dg, lv->x);
    }
    T s = (T)r_dot(g, r_all(g), lv->r, NULL) / (T)rng_count(R);
    if (fabs(s) <= 2 * EPS_T) return;
    PFOR_R(R) { size_t I = IX(i, j, k); lv->r[I] = lv->r[I] - s; }
}
static void mg_increment(lev_t *lv) {
    grid_t *g = &lv->g;
    rng_t R = r_in(g);
    per_bc(g, lv->eps);
    PFOR_R(R) {
        size_t I = IX(i, j, k);
        lv->r[I] = lv->r[I] - mg_mult(g, I, lv->L, lv->Dg, lv->eps);
        lv->x[I] = lv->x[I] + lv->eps[I];
    }
}
static void mg_jacobi(lev_t *lv) {
    grid_t *g = &lv->g;
    rng_t R = r_in(g);
    PFOR_R(R) { size_t I = IX(i, j, k); lv->eps[I] = lv->r[I] * lv->iD[I]; }
    mg_increment(lv);
}
static void mg_pcg(lev_t *lv, int it) {
    grid_t *g = &lv->g;
    rng_t R = r_in(g), A = r_all(g);
    T *x = lv->x, *r = lv->r, *e = lv->eps, *z = lv->z;
    PFOR_R(R) { size_t I = IX(i, j, k); z[I] = e[I] = r[I] * lv->iD[I]; }
    T rho = (T)r_dot(g, A, r, z);
    if (fabs(rho) < 10 * EPS_T) return;
    for (int q = 1; q <= it; ++q) {
        per_bc(g, e);
        PFOR_R(R) { size_t I = IX(i, j, k); z[I] = mg_mult(g, I, lv->L, lv->Dg, e); }
        T alpha = rho / (T)r_dot(g, R, z, e);
        PFOR_R(R) {
            size_t I = IX(i, j, k);
            x[I] += alpha * e[I];
            r[I] -= alpha * z[I];
        }
        if (q == it || fabs((double)alpha) < 1e-2) return;
        PFOR_R(R) { size_t I = IX(i, j, k); z[I] = r[I] * lv->iD[I]; }
        T rho2 = (T)r_dot(g, A, r, z);
        if (fabs(rho2) < 10 * EPS_T) return;
        T beta = rho2 / rho;
        PFOR_R(R) { size_t I = IX(i, j, k); e[I] = beta * e[I] + z[I]; }
        rho = rho2;
    }
}
/* Jacobi-preconditioned CG in the single-reduction form of Chronopoulos & Gear
 * (1989), the flexible CG preconditioner's smoother on coarse levels (no
 * reference counterpart; WaterLily's pcg! needs two reductions per
 * iteration).  With u = r·iD and w = A u, each step takes
 * β = γ/γ₋, α = γ/(δ − βγ/α₋)  (α = γ/δ at the start), p = u + βp,
 * s = w + βs, x += αp, r −= αs, then γ = r·u, δ = u·w of the new r.
 * Stops after `it` updates of x, after an update with |α| < 1e-2, or when
 * |γ| < 10eps (as pcg!).  `w`, `s` are scratch of the level's size. */
static void mg_pcg1(lev_t *lv, int it, T *u, T *w, T *s) {
    grid_t *g = &lv->g;
    rng_t R = r_in(g);
    T *x = lv->x, *r = lv->r, *p = lv->eps;
    PFOR_R(R) { size_t I = IX(i, j, k); u[I] = r[I] * lv->iD[I]; }
    per_bc(g, u);
    PFOR_R(R) { size_t I = IX(i, j, k); w[I] = mg_mult(g, I, lv->L, lv->Dg, u); }
    T gam = (T)r_dot(g, R, r, u), del = (T)r_dot(g, R, u, w);
    T gprev = 0, aprev = 0;
    for (int q = 1; q <= it; ++q) {
        if (fabs(gam) < 10 * EPS_T) return;
        T alpha, beta = 0;
        if (q == 1) {
            alpha = gam / del;
        } else {
            beta = gam / gprev;
            alpha = gam / (del - (beta * gam) / aprev);
        }
        PFOR_R(R) {
            size_t I = IX(i, j, k);
            T pp = q == 1 ? u[I] : u[I] + beta * p[I];
            T ss = q == 1 ? w[I] : w[I] + beta * s[I];
            p[I] = pp;
            s[I] = ss;
            x[I] = x[I] + alpha * pp;
            r[I] = r[I] - alpha * ss;
        }
        if (q == it || fabs((double)alpha) < 1e-2) return;
        PFOR_R(R) { size_t I = IX(i, j, k); u[I] = r[I] * lv->iD[I]; }
        per_bc(g, u);
        PFOR_R(R) { size_t I = IX(i, j, k); w[I] = mg_mult(g, I, lv->L, lv->Dg, u); }
        gprev = gam;
        aprev = alpha;
        gam = (T)r_dot(g, R, r, u);
        del = (T)r_dot(g, R, u, w);
    }
}
/* restrict!(a,b): a[I] = Σ_{J∈up(I)} b[J],  up(I) = 2I-2 : 2I-1 */
static void mg_restrict(lev_t *c, const lev_t *fl) {
    const grid_t *g = &c->g, *gf = &fl->g;
    rng_t R = r_in(g);
    PFOR_R(R) {
        PCOORD(P);
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = a < g->D ? 2 * P[a] - 2 : 1;
            hi[a] = a < g->D ? 2 * P[a] - 1 : 1;
        }
        T s = 0;
        for (int c2 = lo[2]; c2 <= hi[2]; ++c2)
            for (int c1 = lo[1]; c1 <= hi[1]; ++c1)
                for (int c0 = lo[0]; c0 <= hi[0]; ++c0) s += fl->r[IXg(gf, c0, c1, c2)];
        c->r[IX(i, j, k)] = s;
    }
}
/* restrictL!(a,b): a[I,i] = 0.5 Σ_{J∈up(I,i)} b[J,i]; then BC!(a,0) */
static void mg_restrictL(lev_t *c, const lev_t *fl) {
    const grid_t *g = &c->g, *gf = &fl->g;
    rng_t R = r_in(g);
    for (int ii = 1; ii <= g->D; ++ii) {
        PFOR_R(R) {
            PCOORD(P);
            int lo[3], hi[3];
            for (int a = 0; a < 3; ++a) {
                lo[a] = a < g->D ? 2 * P[a] - 2 : 1;
                hi[a] = a < g->D ? (a == ii - 1 ? 2 * P[a] - 2 : 2 * P[a] - 1) : 1;
            }
            T s = 0;
            for (int c2 = lo[2]; c2 <= hi[2]; ++c2)
                for (int c1 = lo[1]; c1 <= hi[1]; ++c1)
                    for (int c0 = lo[0]; c0 <= hi[0]; ++c0) s += fl->L[IXg(gf, c0, c1, c2) + gf->ns * (ii - 1)];
            VC(c->L, ii)[IX(i, j, k)] = (T)(0.5 * (double)s);
        }
    }
    T z3[3] = {0, 0, 0};
    bc_vec_e(g, c->L, z3, 0);
}
/* prolongate!(a,b): a[I] = b[down(I)],  down(I) = (I+2)÷2 */
static void mg_prolongate(lev_t *fl, const lev_t *c) {
    const grid_t *g = &fl->g, *gc = &c->g;
    rng_t R = r_in(g);
    PFOR_R(R) {
        int ci = g->D > 0 ? (i + 2) / 2 : 1, cj = g->D > 1 ? (j + 2) / 2 : 1, ck = g->D > 2 ? (k + 2) / 2 : 1;
        fl->eps[IX(i, j, k)] = c->x[IXg(gc, ci, cj, ck)];
    }
}
static void mg_vcycle_it(oc_ctx *c, int l, const int *cit) {
    lev_t *fl = &c->lev[l], *co = &c->lev[l + 1];
    mg_jacobi(fl);
    mg_restrict(co, fl);
    for (size_t q = 0; q < co->g.ns; ++q) co->x[q] = 0;
    if (l + 2 < c->nlev) mg_vcycle_it(c, l + 1, cit);
    if (cit && c->fcg_cg1 && l + 1 < c->tail_level)
        mg_pcg1(co, cit[l + 1], c->cg1_u, c->cg1_w, c->cg1_s);
    else
        mg_pcg(co, cit ? cit[l + 1] : c->coarse_it);
    mg_prolongate(fl, co);
    mg_increment(fl);
}
static void mg_vcycle(oc_ctx *c, int l) { mg_vcycle_it(c, l, NULL); }
static void mg_update(oc_ctx *c) {
    mg_set_diag(&c->lev[0]);
    for (int l = 1; l < c->nlev; ++l) {
        mg_restrictL(&c->lev[l], &c->lev[l - 1]);
        mg_set_diag(&c->lev[l]);
    }
}
/* WaterLily solver!(ml; tol, itmx) */
static int mg_solver(oc_ctx *c, double tol, int itmx, double *r2out) {
    lev_t *p = &c->lev[0];
    mg_residual(p);
    T r2 = (T)r_dot(&p->g, r_all(&p->g), p->r, p->r);
    int n = 0;
    while ((double)r2 > tol && n < itmx) {
        mg_vcycle(c, 0);
        mg_pcg(p, 6);
        r2 = (T)r_dot(&p->g, r_all(&p->g), p->r, p->r);
        n++;
    }
    per_bc(&p->g, p->x);
    if (r2out) *r2out = (double)r2;
    return n;
}

static int mg_build(oc_ctx *c) {
    grid_t *g = &c->g;
    c->nlev = 0;
    lev_t *l0 = &c->lev[0];
    l0->g = *g;
    l0->L = c->mu0;
    l0->x = c->p;
    l0->z = c->sig;
    l0->Dg = alloc_fill(g->ns, 0);
    l0->iD = alloc_fill(g->ns, 0);
    l0->eps = alloc_fill(g->ns, 0);
    l0->r = alloc_fill(g->ns, 0);
    l0->own = 0;
    mg_set_diag(l0);
    c->nlev = 1;
    for (;;) {
        grid_t *gp = &c->lev[c->nlev - 1].g;
        int ok = 1;
        for (int d = 0; d < g->D; ++d) ok &= (gp->Ng[d] % 2 == 0 && gp->Ng[d] > 4);
        if (!(ok && c->nlev <= 10)) break;
        int Na[3];
        for (int d = 0; d < 3; ++d) Na[d] = d < g->D ? 1 + gp->Ng[d] / 2 : 1;
        int perdir = g->per[0] | (g->per[1] << 1) | (g->per[2] << 2);
        lev_t *lv = &c->lev[c->nlev];
        grid_init(&lv->g, g->D, Na, perdir);
        size_t n = lv->g.ns;
        lv->L = alloc_fill(n * g->D, 0);
        lv->x = alloc_fill(n, 0);
        lv->z = alloc_fill(n, 0);
        lv->Dg = alloc_fill(n, 0);
        lv->iD = alloc_fill(n, 0);
        lv->eps = alloc_fill(n, 0);
        lv->r = alloc_fill(n, 0);
        lv->own = 1;
        mg_restrictL(lv, &c->lev[c->nlev - 1]);
        mg_set_diag(lv);
        c->nlev++;
    }
    if (c->nlev <= 2) {
        snprintf(c->msg, sizeof c->msg, "MultiLevelPoisson requires size=a2^n, where n>2");
        return OC_BADARG;
    }
    /* the first level of at most 4096 interior cells: where the HIP library's
     * one-workgroup tail (which runs pcg!) begins */
    c->tail_level = c->nlev;
    for (int l = 1; l < c->nlev; ++l) {
        long n = rng_count(r_in(&c->lev[l].g));
        if (n <= 4096) { c->tail_level = l; break; }
    }
    size_t n1 = c->lev[1].g.ns;
    c->cg1_u = alloc_fill(n1, 0);
    c->cheb_d0 = alloc_fill(c->lev[0].g.ns, 0);
    c->cg1_w = alloc_fill(n1, 0);
    c->cg1_s = alloc_fill(n1, 0);
    return OC_OK;
}

/* ------------------------------------------------------------------ */
/* cVOF advection (advection.jl)                                        */
/* ------------------------------------------------------------------ */
/* normalEstimation.jl:10-28 reconstructInterface! */
static void reconstruct(oc_ctx *c, T *f, T *al, T *nh) {
    grid_t *g = &c->g;
    /* deviation: zero ghost n̂/α on non-periodic sides (see file header) */
    for (int jj = 1; jj <= g->D; ++jj) {
        if (PER(jj)) continue;
        rng_t R1 = r_slice(g, 1, jj, 1), R2 = r_slice(g, NG(jj), jj, 1);
        FOR_R(R1) { size_t I = IX(i, j, k); al[I] = 0; for (int d = 1; d <= g->D; ++d) VC(nh, d)[I] = 0; }
        FOR_R(R2) { size_t I = IX(i, j, k); al[I] = 0; for (int d = 1; d <= g->D; ++d) VC(nh, d)[I] = 0; }
    }
    rng_t R = r_in(g);
    PFOR_R(R) {
        size_t I = IX(i, j, k);
        if (fullorempty(f[I])) {
            for (int d = 1; d <= g->D; ++d) VC(nh, d)[I] = 0;
            al[I] = 0;
        } else {
            normal_scheme(g, c->scheme, f, nh, I);
            T n[3] = {0, 0, 0};
            for (int d = 1; d <= g->D; ++d) n[d - 1] = VC(nh, d)[I];
            al[I] = getInterceptN(g->D, n, f[I]);
        }
    }
    bc_vof(g, f, al, nh);
}
/* advection.jl:113-137 getVOFFlux! (one face) */
static void vof_flux_face(const grid_t *g, T *ff, const T *f, const T *al, const T *nh, T dl, int d, size_t IF,
                          T *ruf, T lr) {
    if (dl == 0) return;
    size_t IC = dl > 0 ? IF - ST(d) : IF;
    T s = 0;
    for (int ii = 1; ii <= g->D; ++ii) s += fabs(VC(nh, ii)[IC]);
    if (s == 0 || fullorempty(f[IC])) {
        ff[IF] = f[IC] * dl;
        VC(ruf, d)[IF] += dl * lr + (1 - lr) * ff[IF];
        return;
    }
    T a = dl > 0 ? al[IC] - VC(nh, d)[IC] * (1 - dl) : al[IC];
    T n[3];
    for (int ii = 1; ii <= g->D; ++ii) n[ii - 1] = VC(nh, ii)[IC] * (ii == d ? fabs(dl) : (T)1);
    ff[IF] = getVolumeFractionN(g->D, n, a) * dl;
    VC(ruf, d)[IF] += dl * lr + (1 - lr) * ff[IF];
}
/* advection.jl:108-112 getVOFFlux! (whole field) */
static void vof_flux(oc_ctx *c, T *ff, const T *f, const T *al, const T *nh, const T *u, const T *u0, T dt, int d,
                     T *ruf) {
    grid_t *g = &c->g;
    memset(ff, 0, g->ns * sizeof(T));
    rng_t R = r_in_uWB(g, d);
    const T *ud = VC(u, d), *u0d = VC(u0, d);
    PFOR_R(R) {
        size_t I = IX(i, j, k);
        vof_flux_face(g, ff, f, al, nh, (dt / 2) * (ud[I] + u0d[I]), d, I, ruf, c->lr);
    }
}
static inline T divu(const grid_t *g, const T *u, size_t I) {
    T s = 0;
    for (int i = 1; i <= g->D; ++i) s += VC(u, i)[I + ST(i)] - VC(u, i)[I];
    return s;
}
/* advection.jl:145-189 reportFillError: NaN -> OC_NAN, |div|>10 -> OC_EXPLODE,
 * fill error -> OC_FILLWARN (printed, not thrown, by the reference) */
static int report_fill(oc_ctx *c, const T *f, const T *u, const T *u0, T tol) {
    grid_t *g = &c->g;
    size_t imax = 0, imin = 0;
    T mx = f[0], mn = f[0];
    int nanseen = (f[0] != f[0]);
    for (size_t q = 1; q < g->ns; ++q) {
        T v = f[q];
        if (v != v) { nanseen = 1; continue; }
        if (v > mx || mx != mx) { mx = v; imax = q; }
        if (v < mn || mn != mn) { mn = v; imin = q; }
    }
    if (nanseen) { snprintf(c->msg, sizeof c->msg, "NaN!"); return OC_NAN; }
    int st = OC_OK;
    if (mx - 1 > tol) {
        T du0 = fabs(divu(g, u0, imax)), du = fabs(divu(g, u, imax));
        if (du0 + du > 10 || (du0 + du) != (du0 + du)) {
            snprintf(c->msg, sizeof c->msg, "divergence, %g, is exploding!", (double)(du0 + du));
            return OC_EXPLODE;
        }
        snprintf(c->msg, sizeof c->msg, "max VOF @ %zu not in [0,1], df = %g", imax, (double)(mx - 1));
        st = OC_FILLWARN;
    }
    if (mn < -tol) {
        T du0 = fabs(divu(g, u0, imin)), du = fabs(divu(g, u, imin));
        if (du0 + du > 10 || (du0 + du) != (du0 + du)) {
            snprintf(c->msg, sizeof c->msg, "divergence, %g, is exploding!", (double)(du0 + du));
            return OC_EXPLODE;
        }
        snprintf(c->msg, sizeof c->msg, "min VOF @ %zu not in [0,1], df = %g", imin, (double)(-mn));
        st = OC_FILLWARN;
    }
    return st;
}
/* advection.jl:80-89 advectVOF1d! (report with rtol, clean with ctol) */
static int advect_vof_1d(oc_ctx *c, T *f, const T *u, const T *u0, T dt, int d, T rtol, T ctol) {
    grid_t *g = &c->g;
    reconstruct(c, f, c->al, c->nh);
    vof_flux(c, c->ff, f, c->al, c->nh, u, u0, dt, d, c->ruf);
    rng_t R = r_in(g);
    const T *ud = VC(u, d), *u0d = VC(u0, d);
    const T *ff = c->ff;
    const int8_t *cb = c->cb;
    PFOR_R(R) {
        size_t I = IX(i, j, k);
        T dud = ud[I + ST(d)] - ud[I], du0d = u0d[I + ST(d)] - u0d[I];
        f[I] += (ff[I] - ff[I + ST(d)]) + (((T)cb[I] * (dud + du0d)) * dt) / 2;
    }
    int st = report_fill(c, f, u, u0, rtol);
    if (st == OC_NAN || st == OC_EXPLODE) return st;
    clean_wisp(g, f, ctol);
    bc_f(g, f, 0);
    return st;
}
static void set_cbar(oc_ctx *c, const T *f) {
    for (size_t q = 0; q < c->g.ns; ++q) c->cb[q] = f[q] < (T)0.5 ? 0 : 1;
}
static void dir_order(int D, int step_index, int *dirO) {
    /* flow.jl:163, advection.jl:22: dirO = ntuple(i->mod(length(Δt)+i,D)+1, D) */
    for (int i = 1; i <= D; ++i) dirO[i - 1] = ((step_index + i) % D) + 1;
}
static int worst(int a, int b) {
    if (a == OC_NAN || b == OC_NAN) return OC_NAN;
    if (a == OC_EXPLODE || b == OC_EXPLODE) return OC_EXPLODE;
    return a > b ? a : b;
}

/* ------------------------------------------------------------------ */
/* CMOM + SynDRoM (flow.jl:5-57, :157-241)                               */
/* ------------------------------------------------------------------ */
static T limiter(int lim, T u, T c, T d) {
    switch (lim) {
    case 1: return median3((3 * c - u) / 2, c, (c + d) / 2);                   /* minmod */
    case 2: return c;                                                        /* upwind */
    case 3: {                                                                /* vanAlbada1 */
        T al = c - u, be = d - c;
        T w = (al == be && al == 0) ? (T)0 : (al + be) / (al * al + be * be);
        return c + (tmax(al * be, 0) * w) / 2;
    }
    case 4: {                                                                /* superbee = Sweby(γ=2) */
        T s = tsign(d - u), gm = 2;
        if (c <= tmin(u, d) || c >= tmax(u, d)) return c;
        return c + (s * tmax(tmax(0, tmin((s * gm) * (c - u), s * (d - c))), tmin(s * (c - u), (s * gm) * (d - c)))) / 2;
    }
    case 5: {                                                                /* TVDcen */
        T s = tsign(d - u);
        if (c <= tmin(u, d) || c >= tmax(u, d)) return c;
        return c + s * tmin(s * (c - u), (s * (d - c)) / 2);
    }
    case 6: {                                                                /* TVDdown */
        T s = tsign(d - u);
        if (c <= tmin(u, d) || c >= tmax(u, d)) return c;
        return c + s * tmin(s * (c - u), s * (d - c));
    }
    case 7: {                                                                /* Sweby, γ = 1.5 */
        T s = tsign(d - u), gm = (T)1.5;
        if (c <= tmin(u, d) || c >= tmax(u, d)) return c;
        return c + (s * tmax(tmax(0, tmin((s * gm) * (c - u), s * (d - c))), tmin(s * (c - u), (s * gm) * (d - c)))) / 2;
    }
    default: return median3(((7 * c + d) - 2 * u) / 6, c, median3(2 * c - u, c, d)); /* Koren */
    }
}
typedef struct {
    const grid_t *g;
    const T *ruf, *us, *drho;
    T dt, lr;
    int lim;
} mom_args;
/* flow.jl:43-56: the SynDRoM face value of ϕq once Ψ and the donor's old
 * face fraction fd = fOld[IiCell] are known */
static T phiq_value(T Psi, T fd, T dt, T lr, int lim, T uu, T cc, T dd) {
    T vI = cc, vd = limiter(lim, uu, cc, dd);
    T va = 2 * vI - vd;
    T mOut = fabs(Psi) * dt;
    T mOld = lin(fd, lr);
    if (mOut > mOld) return Psi * vI;
    T l2 = fabs(mOut) / mOld, l1 = 1 - l2;
    T vb = l2 * va + l1 * vd;
    return (Psi * (vb + vd)) / 2;
}
/* flow.jl:37-57 ϕq;  II = linear vector index (cell I, comp i), I = cell index */
static T phiq(const mom_args *m, int jd, int i, size_t Icell, size_t II, T uu, T cc, T dd) {
    const grid_t *g = m->g;
    T Psi = (VC(m->ruf, jd)[Icell] + VC(m->ruf, jd)[Icell - ST(i)]) / 2;
    size_t IiC = Psi > 0 ? II - ST(jd) : II;
    return phiq_value(Psi, m->drho[IiC], m->dt, m->lr, m->lim, uu, cc, dd);
}
static inline T psi_of(const mom_args *m, int jd, int i, size_t I) {
    const grid_t *g = m->g;
    return (VC(m->ruf, jd)[I] + VC(m->ruf, jd)[I - ST(i)]) / 2;
}
/* flow.jl:20-23 ϕu */
static T phiu(const mom_args *m, int jd, int i, size_t I) {
    const grid_t *g = m->g;
    size_t II = I + g->ns * (i - 1);
    const T *f = m->us;
    long s = ST(jd);
    T Psi = psi_of(m, jd, i, I);
    return Psi > 0 ? phiq(m, jd, i, I, II, f[II - 2 * s], f[II - s], f[II])
                   : phiq(m, jd, i, I, II, f[II + s], f[II], f[II - s]);
}
/* flow.jl:24-27 ϕuP (Ip: the periodic image of I-2δ) */
static T phiuP(const mom_args *m, int jd, int i, size_t IpII, size_t I) {
    const grid_t *g = m->g;
    size_t II = I + g->ns * (i - 1);
    const T *f = m->us;
    long s = ST(jd);
    T Psi = psi_of(m, jd, i, I);
    return Psi > 0 ? phiq(m, jd, i, I, II, f[IpII], f[II - s], f[II])
                   : phiq(m, jd, i, I, II, f[II + s], f[II], f[II - s]);
}
/* flow.jl:28-31 ϕuL */
static T phiuL(const mom_args *m, int jd, int i, size_t I) {
    const grid_t *g = m->g;
    size_t II = I + g->ns * (i - 1);
    const T *f = m->us;
    long s = ST(jd);
    T Psi = psi_of(m, jd, i, I);
    return Psi > 0 ? phiq(m, jd, i, I, II, 2 * f[II - s] - f[II], f[II - s], f[II])
                   : phiq(m, jd, i, I, II, f[II + s], f[II], f[II - s]);
}
/* flow.jl:32-35 ϕuR */
static T phiuR(const mom_args *m, int jd, int i, size_t I) {
    const grid_t *g = m->g;
    size_t II = I + g->ns * (i - 1);
    const T *f = m->us;
    long s = ST(jd);
    T Psi = psi_of(m, jd, i, I);
    return Psi < 0 ? phiq(m, jd, i, I, II, 2 * f[II] - f[II - s], f[II], f[II - s])
                   : phiq(m, jd, i, I, II, f[II - 2 * s], f[II - s], f[II]);
}
/* flow.jl:212-241 advectρuu1D! */
static void advect_rhouu_1d(oc_ctx *c, T *ru, T *r, T *Phi, const T *ruf, const T *us, const T *uOld, const T *drho,
                            T *dila, const T *u, const T *u0, int d, T dt) {
    grid_t *g = &c->g;
    memset(r, 0, g->ns * g->D * sizeof(T));
    int jd = d;
    rng_t R = r_in(g);
    const T *ud = VC(u, d), *u0d = VC(u0, d);
    PFOR_R(R) {
        size_t I = IX(i, j, k);
        dila[I] = (lin((T)c->cb[I], c->lr) * ((ud[I + ST(d)] - ud[I]) + (u0d[I + ST(d)] - u0d[I]))) / 2;
    }
    bc_f(g, dila, 0);
    mom_args m = {g, ruf, us, drho, dt, c->lr, c->limiter};
    for (int ii = 1; ii <= g->D; ++ii) {
        T *ri = VC(r, ii);
        int tagper = PER(jd);
        rng_t RL = r_slice(g, 2, jd, 2), RU = r_slice(g, NG(jd), jd, 2), RI = r_in_u(g, jd);
        if (!tagper) {
            FOR_R(RL) { size_t I = IX(i, j, k); ri[I] += phiuL(&m, jd, ii, I); }
        } else {
            FOR_R(RL) {
                size_t I = IX(i, j, k);
                size_t Ip = I + (NG(jd) - 2 - 2) * ST(jd) + g->ns * (ii - 1);
                Phi[I] = phiuP(&m, jd, ii, Ip, I);
                ri[I] += Phi[I];
            }
        }
        PFOR_R(RI) {
            size_t I = IX(i, j, k);
            Phi[I] = phiu(&m, jd, ii, I);
            ri[I] += Phi[I];
        }
        PFOR_R(RI) { size_t I = IX(i, j, k); ri[I - ST(jd)] -= Phi[I]; }
        if (!tagper) {
            FOR_R(RU) { size_t I = IX(i, j, k); ri[I - ST(jd)] += -phiuR(&m, jd, ii, I); }
        } else {
            FOR_R(RU) { size_t I = IX(i, j, k); ri[I - ST(jd)] -= Phi[I + (2 - NG(jd)) * ST(jd)]; }
        }
        const T *uo = VC(uOld, ii);
        PFOR_R(R) { size_t I = IX(i, j, k); ri[I] += uo[I] * ((dila[I] + dila[I - ST(ii)]) / 2); }
    }
    size_t n = g->ns * g->D;
    for (size_t q = 0; q < n; ++q) ru[q] = ru[q] + dt * r[q];
}
/* flow.jl:157-210 advectfq! / advectVOFρuu!  (f advected in place) */
static int advect_fq(oc_ctx *c, T *f, const T *u1, const T *u2, const T *uOld, T Dt, int step_index) {
    grid_t *g = &c->g;
    T tol = 10 * EPS_T;
    set_cbar(c, f);
    int dirO[3];
    dir_order(g->D, step_index, dirO);
    T *r = c->fv, *Phi = c->sig, *us = c->nh, *dila = c->al;
    int st = OC_OK;
    for (int op = 0; op < g->D; ++op) {
        int d = dirO[op];
        T dt = (T)1 * Dt;
        rhou2u(g, r, c->ru, f, c->lr);
        bc_vec(g, r, c->uBC);
        memcpy(Phi, f, g->ns * sizeof(T));
        memset(c->ruf, 0, g->ns * g->D * sizeof(T));
        int s = advect_vof_1d(c, f, u1, u2, dt, d, 10 * tol, tol);
        st = worst(st, s);
        if (s == OC_NAN || s == OC_EXPLODE) return st;
        f2face(g, c->dr, Phi);
        memcpy(us, r, g->ns * g->D * sizeof(T));
        T idt = (T)1 / dt;
        for (size_t q = 0; q < g->ns * g->D; ++q) c->ruf[q] = c->ruf[q] * idt;
        bc_vec(g, c->ruf, c->uBC);
        advect_rhouu_1d(c, c->ru, r, Phi, c->ruf, us, uOld, c->dr, dila, u1, u2, d, dt);
    }
    return st;
}

/* ------------------------------------------------------------------ */
/* viscosity, RK update, coefficients, CFL, projection (flow.jl)        */
/* ------------------------------------------------------------------ */
/* flow.jl:144 viscF */
static inline T viscF(const oc_ctx *c, int i, int jj, size_t I, const T *u, const T *ff) {
    const grid_t *g = &c->g;
    T mu = getmu(g, i, jj, I, ff, c->lm, c->mu, c->lr);
    return mu * ((VC(u, i)[I] - VC(u, i)[I - ST(jj)]) + (VC(u, jj)[I] - VC(u, jj)[I - ST(i)]));
}
/* ------------------------------------------------------------------ */
/* metrics.jl:15-50 (per-cell diagnostics; the Float64 literals make    */
/* them Float64 for either element type) and their domain integrals     */
/* ------------------------------------------------------------------ */
/* ρkeI(I,u,f,λρ,U) = 0.25 Σ_i ((u[I,i]-U_i)² + (u[I+δi,i]-U_i)²) ρ(I) */
static double rho_ke(const grid_t *g, size_t I, const T *u, const T *f, T lr, const double *U) {
    double s = 0, rho = (double)lin(f[I], lr);
    for (int i = 1; i <= g->D; ++i) {
        double a = (double)VC(u, i)[I] - U[i - 1], b = (double)VC(u, i)[I + ST(i)] - U[i - 1];
        s += (a * a + b * b) * rho;
    }
    return 0.25 * s;
}
/* ρuI(i,I,u,f,λρ,U) = 0.5 (u[I,i] + u[I+δi,i] - 2U_i) ρ(I) */
static double rho_u(const grid_t *g, int i, size_t I, const T *u, const T *f, T lr, const double *U) {
    return 0.5 * (((double)VC(u, i)[I] + (double)VC(u, i)[I + ST(i)]) - 2 * U[i - 1]) * (double)lin(f[I], lr);
}
/* ρgh(I,g,f,λρ,StatWL) = -ρ(I) Σ_i g_i (loc(0,I)_i - StatWL_i), loc(0,I) = I - 1.5 (1-based) */
static double rho_gh(const grid_t *g, const int q[3], size_t I, const double *gv, const T *f, T lr, const double *wl) {
    double s = 0;
    for (int i = 0; i < g->D; ++i) s += gv[i] * ((q[i] - 1.5) - wl[i]);
    return -(double)lin(f[I], lr) * s;
}
/* EnsI(I,ω) = ½ · ¼ · Σ over the 4 edges of cell I of ω² (per component in 3D,
 * the edges parallel to axis i, ix = shiftDir(i,3,1), iy = shiftDir(i,3,2)) */
static double ens(const grid_t *g, size_t I, const T *w) {
    if (g->D == 2)
        return 0.5 * 0.25 * (((double)w[I] * w[I] + (double)w[I + ST(1)] * w[I + ST(1)]) +
                             (double)w[I + ST(2)] * w[I + ST(2)] + (double)w[I + ST(1) + ST(2)] * w[I + ST(1) + ST(2)]);
    double s = 0;
    for (int i = 1; i <= 3; ++i) {
        int ix = i % 3 + 1, iy = (i + 1) % 3 + 1;
        const T *wi = VC(w, i);
        double a = wi[I], b = wi[I + ST(ix)], cc = wi[I + ST(iy)], d = wi[I + ST(ix) + ST(iy)];
        s += ((a * a + b * b) + cc * cc) + d * d;
    }
    return 0.5 * 0.25 * s;
}
/* the vorticity the integral uses: ω_i = ∂_j u_k - ∂_k u_j with (i, j, k) cyclic
 * (2D: ω = ∂_x v - ∂_y u) by backward differences, i.e. at the edge of cell I
 * on its lower j and k sides -- the edges EnsI averages */
static void vorticity(const grid_t *g, const T *u, T *w) {
    rng_t R = r_all(g);
    FOR_R(R) {
        size_t I = IX(i, j, k);
        int q[3] = {i, j, k};
        if (g->D == 2) {
            if (q[0] < 2 || q[1] < 2) { w[I] = 0; continue; }
            w[I] = (VC(u, 2)[I] - VC(u, 2)[I - ST(1)]) - (VC(u, 1)[I] - VC(u, 1)[I - ST(2)]);
            continue;
        }
        for (int a = 1; a <= 3; ++a) {
            int jj = a % 3 + 1, kk = (a + 1) % 3 + 1;
            if (q[jj - 1] < 2 || q[kk - 1] < 2) { VC(w, a)[I] = 0; continue; }
            VC(w, a)[I] = (VC(u, kk)[I] - VC(u, kk)[I - ST(jj)]) - (VC(u, jj)[I] - VC(u, jj)[I - ST(kk)]);
        }
    }
}

/* ------------------------------------------------------------------ */
/* surfaceTension.jl                                                    */
/* ------------------------------------------------------------------ */
static inline int contain_interface(T v) { return v > 0 && v < 1; } /* VOFutil.jl:144 */
/* surfaceTension.jl:96-124 getPopinetHeightAdaptive(I, f, i, monotonic=true):
 * column height about cell q (1-based) along the signed axis i (1-based),
 * walking to the air side, then to the water side; validCI (util.jl:82) stops
 * a walk at the array's end */
static T popinet_height(const grid_t *g, const T *f, const int q[3], int i) {
    int ax = abs(i), s = i > 0 ? 1 : -1;
    long st = (long)ST(ax) * s;
    size_t I0 = IX(q[0], q[1], q[2]);
    T fnow = f[I0], H = fnow - (T)0.5;
    int fin = fnow < 1, c = q[ax - 1];
    long In = (long)I0;
    while (!fin || contain_interface(fnow)) {
        c += s;
        In += st;
        if (c < 1 || c > NG(ax)) break;
        T v = f[In];
        fnow = v > fnow ? (T)0 : v;
        H += fnow;
        if (contain_interface(fnow)) fin = 1;
    }
    fnow = f[I0];
    fin = fnow > 0;
    c = q[ax - 1];
    In = (long)I0;
    while (!fin || contain_interface(fnow)) {
        c -= s;
        In -= st;
        if (c < 1 || c > NG(ax)) break;
        T v = f[In];
        fnow = v < fnow ? (T)1 : v;
        H += fnow - 1;
        if (contain_interface(fnow)) fin = 1;
    }
    return H;
}
/* surfaceTension.jl:89 root1p5 */
static inline T root1p5(T a) { return sqrt((a * a) * a); }
/* surfaceTension.jl:35-80 getCurvature (util.jl:62-63 getXdir / getXYdir) */
static T curvature(const grid_t *g, const T *f, const int q[3], int i) {
    int s = i > 0 ? 1 : -1, ai = abs(i);
    if (g->D == 2) {
        int ix = ai == 1 ? -2 * s : s, ax = abs(ix), sx = ix > 0 ? 1 : -1;
        T H[3];
        for (int u = -1; u <= 1; ++u) {
            int p[3] = {q[0], q[1], q[2]};
            p[ax - 1] += u * sx;
            H[u + 1] = popinet_height(g, f, p, i);
        }
        T Hx = (H[2] - H[0]) / 2;
        T Hxx = (H[2] + H[0]) - 2 * H[1];
        return Hxx / root1p5(1 + Hx * Hx);
    }
    int ix = s * (ai % 3 + 1), iy = (ai + 1) % 3 + 1;
    int axx = abs(ix), sx = ix > 0 ? 1 : -1;
    T H[3][3], flt = (T)0.2;
    for (int u = -1; u <= 1; ++u)
        for (int v = -1; v <= 1; ++v) {
            int p[3] = {q[0], q[1], q[2]};
            p[axx - 1] += u * sx;
            p[iy - 1] += v;
            H[u + 1][v + 1] = popinet_height(g, f, p, i);
        }
    T Hx = (H[2][1] - H[0][1]) / 2, Hy = (H[1][2] - H[1][0]) / 2;
    T Hxx = (((H[2][1] + H[0][1]) - 2 * H[1][1]) + ((H[2][0] + H[0][0]) - 2 * H[1][0]) * flt +
             ((H[2][2] + H[0][2]) - 2 * H[1][2]) * flt) / (1 + 2 * flt);
    T Hyy = (((H[1][2] + H[1][0]) - 2 * H[1][1]) + ((H[0][2] + H[0][0]) - 2 * H[0][1]) * flt +
             ((H[2][2] + H[2][0]) - 2 * H[2][1]) * flt) / (1 + 2 * flt);
    T Hxy = (((H[2][2] + H[0][0]) - H[2][0]) - H[0][2]) / 4;
    return ((Hxx * (1 + Hy * Hy) + Hyy * (1 + Hx * Hx)) - ((2 * Hxy) * Hx) * Hy) / root1p5((1 + Hx * Hx) + Hy * Hy);
}
/* surfaceTension.jl:7-19 surfTen!: per direction d, the face-averaged f in the
 * fᶠ buffer with BCf!(d), the WY normal of it in n̂ at interface cells, and
 * forcing[I,d] += η κ (-∂_d f).  BCf!(d) leaves the upper ghost plane of axis
 * d unwritten (it holds whatever fᶠ held); here it is set to 0, which is what
 * the advection sweeps leave there when the wall-normal velocity is 0. */
static void surften(oc_ctx *c, T *r, const T *f) {
    grid_t *g = &c->g;
    rng_t R = r_in(g);
    for (int d = 1; d <= g->D; ++d) {
        T *fb = c->ff, *rd = VC(r, d);
        FOR_R(R) {
            size_t I = IX(i, j, k);
            fb[I] = (f[I] + f[I - ST(d)]) / 2;
        }
        bc_f(g, fb, d);
        if (!PER(d)) {
            rng_t R2 = r_slice(g, NG(d), d, 1);
            FOR_R(R2) { fb[IX(i, j, k)] = 0; }
        }
        FOR_R(R) {
            size_t I = IX(i, j, k);
            if (!contain_interface(fb[I])) continue;
            normal_wy(g, fb, c->nh, I);
            int im = argabsmax(g, c->nh, I);
            int sgn = signbit(VC(c->nh, im)[I]) ? -1 : 1;
            int q[3] = {i, j, k};
            rd[I] += c->eta * curvature(g, fb, q, sgn * im) * -(f[I] - f[I - ST(d)]);
        }
    }
}
/* flow.jl:113-139 viscSurfTenρu! / visc!  (+ surfTen! when η is set) */
static void visc_only(oc_ctx *c, T *r, const T *u, T *ffc, T *Phi, const T *f);
static void visc(oc_ctx *c, T *r, const T *u, T *ffc, T *Phi, const T *f) {
    visc_only(c, r, u, ffc, Phi, f);
    if (c->has_eta) surften(c, r, f);
}
static void visc_only(oc_ctx *c, T *r, const T *u, T *ffc, T *Phi, const T *f) {
    grid_t *g = &c->g;
    memset(r, 0, g->ns * g->D * sizeof(T));
    if (!c->has_mu) return;
    f2face(g, ffc, f);
    for (int ii = 1; ii <= g->D; ++ii)
        for (int jj = 1; jj <= g->D; ++jj) {
            T *ri = VC(r, ii);
            int tagper = PER(jj);
            rng_t RL = r_slice(g, 2, jj, 2), RU = r_slice(g, NG(jj), jj, 2), RI = r_in_u(g, jj);
            if (!tagper) {
                FOR_R(RL) { size_t I = IX(i, j, k); ri[I] += -viscF(c, ii, jj, I, u, ffc); }
            } else {
                FOR_R(RL) {
                    size_t I = IX(i, j, k);
                    Phi[I] = -viscF(c, ii, jj, I, u, ffc);
                    ri[I] += Phi[I];
                }
            }
            PFOR_R(RI) {
                size_t I = IX(i, j, k);
                Phi[I] = -viscF(c, ii, jj, I, u, ffc);
                ri[I] += Phi[I];
            }
            PFOR_R(RI) { size_t I = IX(i, j, k); ri[I - ST(jj)] -= Phi[I]; }
            if (!tagper) {
                FOR_R(RU) { size_t I = IX(i, j, k); ri[I - ST(jj)] += viscF(c, ii, jj, I, u, ffc); }
            } else {
                FOR_R(RU) { size_t I = IX(i, j, k); ri[I - ST(jj)] -= Phi[I + (2 - NG(jj)) * ST(jj)]; }
            }
        }
}
/* flow.jl:244-252 updateU! */
static void update_u(oc_ctx *c, T *u, T *ru, const T *ru0, T *F, T dt, const T *f, T w) {
    grid_t *g = &c->g;
    T wT = w, a = (T)1 / wT - 1;
    size_t n = g->ns * g->D;
    for (size_t q = 0; q < n; ++q) ru[q] = ((a * ru0[q] + ru[q]) + F[q] * dt) * wT;
    rhou2u(g, u, ru, f, c->lr);
    memset(F, 0, n * sizeof(T));
    if (c->has_g)
        for (int d = 1; d <= g->D; ++d) {
            T *Fd = VC(F, d);
            for (size_t q = 0; q < g->ns; ++q) Fd[q] += c->gv[d - 1];
        }
    T al = dt * wT;
    for (size_t q = 0; q < n; ++q) u[q] = u[q] + al * F[q];
}
/* flow.jl:254-259 updateL! (after fill!(μ₀,1)) */
static void update_L(oc_ctx *c, const T *f) {
    grid_t *g = &c->g;
    size_t n = g->ns * g->D;
    for (size_t q = 0; q < n; ++q) c->mu0[q] = 1;
    rng_t R = r_in(g);
    for (int d = 1; d <= g->D; ++d) {
        T *Ld = VC(c->mu0, d);
        PFOR_R(R) { size_t I = IX(i, j, k); Ld[I] /= lin((f[I] + f[I - ST(d)]) / 2, c->lr); }
    }
    bc_vec_e(g, c->mu0, c->zero3, 0); /* BC!(μ₀, zeros, false, perdir), flow.jl:258 */
}
/* flow.jl:262-281 MPCFL (+ WaterLily flux_out, whose `0.` literal is Float64) */
static T mpcfl(oc_ctx *c) {
    grid_t *g = &c->g;
    rng_t R = r_in(g);
    T *s = c->sig, *u = c->u;
    memset(s, 0, g->ns * sizeof(T));
    PFOR_R(R) {
        size_t I = IX(i, j, k);
        double acc = 0;
        for (int d = 1; d <= g->D; ++d) {
            double up = (double)VC(u, d)[I + ST(d)], lo = -(double)VC(u, d)[I];
            acc += (up > 0.0 ? up : 0.0) + (lo > 0.0 ? lo : 0.0);
        }
        s[I] = (T)acc;
    }
    T dtAdv = (T)1 / (r_max(g, r_all(g), s) + 5 * c->nu);
    PFOR_R(R) {
        size_t I = IX(i, j, k);
        T acc = 0;
        for (int d = 1; d <= g->D; ++d) acc += tmax(fabs(VC(u, d)[I]), fabs(VC(u, d)[I + ST(d)]));
        s[I] = acc;
    }
    T dtV = (T)1 / (2 * r_max(g, r_all(g), s));
    T dtG = 1;
    if (c->has_g) {
        T g2 = 0;
        for (int d = 0; d < g->D; ++d) g2 += c->gv[d] * c->gv[d];
        dtG = (T)1 / (2 * sqrt(g2));
    }
    T dtVisc = c->has_mu ? (T)3 / ((14 * c->mu) * tmax(1, c->lm / c->lr)) : (T)1;
    T dtST = c->has_eta ? (T)sqrt((1 + c->lr) / ((T)(8 * M_PI) * c->eta)) : (T)1;
    T m = tmin(tmin(tmin(tmin(dtV, dtAdv), dtG), dtVisc), dtST);
    return (T)0.8 * m;
}
/* flow.jl:343-347 inproject!  (+ the FP64 refinement variant below) */
static void proj_apply(oc_ctx *c, T dt) {
    grid_t *g = &c->g;
    rng_t R = r_in(g);
    for (int i_ = 1; i_ <= g->D; ++i_) {
        T *ui = VC(c->u, i_), *Li = VC(c->mu0, i_);
        PFOR_R(R) { size_t I = IX(i, j, k); ui[I] -= Li[I] * (c->p[I] - c->p[I - ST(i_)]); }
    }
    T idt = (T)1 / dt;
    for (size_t q = 0; q < g->ns; ++q) c->p[q] = c->p[q] * idt;
}
int oc_project_fcg(oc_ctx *c, double dt_w, double rtol, int max_it, int smooth_it, int *iters, double *rel);
/* flow.jl:300-326 psolver!(p; tol, itmx) on the single-level Poisson (the
 * finest MultiLevelPoisson level: same L, D, iD, x, ϵ, r, z): Jacobi-
 * preconditioned CG without early exits, stopping on the absolute Σr² */
static int psolver(oc_ctx *c, T tol, int itmx, double *r2o) {
    lev_t *lv = &c->lev[0];
    grid_t *g = &lv->g;
    rng_t R = r_in(g), A = r_all(g);
    T *x = lv->x, *r = lv->r, *e = lv->eps, *z = lv->z;
    per_bc(g, x);
    mg_residual(lv);
    T r2 = (T)r_dot(g, R, r, r);
    int n = 0;
    PFOR_R(R) { size_t I = IX(i, j, k); z[I] = e[I] = r[I] * lv->iD[I]; }
    T rho = (T)r_dot(g, A, r, z);
    while ((r2 > tol || (r2 > tol / 4 && n == 0)) && n < itmx) {
        per_bc(g, e);
        PFOR_R(R) { size_t I = IX(i, j, k); z[I] = mg_mult(g, I, lv->L, lv->Dg, e); }
        T alpha = rho / (T)r_dot(g, R, z, e);
        PFOR_R(R) {
            size_t I = IX(i, j, k);
            x[I] += alpha * e[I];
            r[I] -= alpha * z[I];
        }
        PFOR_R(R) { size_t I = IX(i, j, k); z[I] = r[I] * lv->iD[I]; }
        T rho2 = (T)r_dot(g, A, r, z);
        T beta = rho2 / rho;
        PFOR_R(R) { size_t I = IX(i, j, k); e[I] = beta * e[I] + z[I]; }
        rho = rho2;
        r2 = (T)r_dot(g, R, r, r);
        n++;
    }
    per_bc(g, x);
    if (r2o) *r2o = (double)r2;
    return n;
}
/* flow.jl:328-335 myproject! */
static int myproject(oc_ctx *c, T dt, int *iters, double *r2) {
    if (c->poisson_mode == 2) {
        /* inproject!(a, b::Poisson, dt) (flow.jl:337-341) */
        grid_t *g = &c->g;
        rng_t R = r_in(g);
        lev_t *l0 = &c->lev[0];
        memset(c->sig, 0, g->ns * sizeof(T));
        memset(l0->eps, 0, g->ns * sizeof(T));
        memset(l0->r, 0, g->ns * sizeof(T));
        PFOR_R(R) { size_t I = IX(i, j, k); c->sig[I] = divu(g, c->u, I); }
        for (size_t q = 0; q < g->ns; ++q) c->p[q] = c->p[q] * dt;
        int n = psolver(c, (T)(50 * EPS_T), 2000, r2);
        if (iters) *iters = n;
        proj_apply(c, dt);
        return OC_OK;
    }
    if (c->poisson_mode == 1) {
        double rel;
        int e = oc_project_fcg(c, (double)dt, c->poisson_rtol, 200, c->smooth_it, iters, &rel);
        if (r2) *r2 = rel;
        return e;
    }
    grid_t *g = &c->g;
    rng_t R = r_in(g);
    PFOR_R(R) { size_t I = IX(i, j, k); c->sig[I] = divu(g, c->u, I); }
    for (size_t q = 0; q < g->ns; ++q) c->p[q] = c->p[q] * dt;
    int n = mg_solver(c, c->mg_tol, c->mg_itmx, r2);
    if (iters) *iters = n;
    proj_apply(c, dt);
    return OC_OK;
}
int oc_last_iters(oc_ctx *c, int q) { return c->last_iters[q & 1]; }

/* ------------------------------------------------------------------ */
/* MPFMomStep! (flow.jl:60-109) and advect! (advection.jl:17-78)         */
/* ------------------------------------------------------------------ */
int oc_mom_step(oc_ctx *c, double dtd, int step_index, double *dt_next) {
    grid_t *g = &c->g;
    size_t nv = g->ns * g->D, n = g->ns;
    T dt = (T)dtd;
    memcpy(c->u0, c->u, nv * sizeof(T));
    memcpy(c->f0, c->f, n * sizeof(T));
    /* predictor */
    T dtc = (T)0.5;
    u2rhou(g, c->ru, c->u0, c->f0, c->lr);
    bc_vec(g, c->ru, c->uBC);
    int st = advect_fq(c, c->f0, c->u0, c->u, c->u, dt, step_index);
    if (st == OC_NAN || st == OC_EXPLODE) return st;
    for (size_t q = 0; q < n; ++q) c->f0[q] = (c->f0[q] + c->f[q]) / 2;
    visc(c, c->fv, c->u, c->nh, c->sig, c->f0);
    u2rhou(g, c->nh, c->u0, c->f, c->lr);
    update_u(c, c->u, c->ru, c->nh, c->fv, dt, c->f0, dtc);
    bc_vec(g, c->u, c->uBC);
    update_L(c, c->f0);
    mg_update(c);
    myproject(c, dtc * dt, &c->last_iters[0], NULL);
    bc_vec(g, c->u, c->uBC);
    /* corrector */
    memcpy(c->f0, c->f, n * sizeof(T));
    u2rhou(g, c->ru, c->u0, c->f, c->lr);
    bc_vec(g, c->ru, c->uBC);
    int s2 = advect_fq(c, c->f, c->u, c->u, c->u0, dt, step_index);
    st = worst(st, s2);
    if (st == OC_NAN || st == OC_EXPLODE) return st;
    visc(c, c->fv, c->u, c->nh, c->sig, c->f);
    u2rhou(g, c->nh, c->u0, c->f, c->lr);
    memcpy(c->u0, c->u, nv * sizeof(T));
    update_u(c, c->u, c->ru, c->nh, c->fv, dt, c->f, (T)1);
    bc_vec(g, c->u, c->uBC);
    update_L(c, c->f);
    mg_update(c);
    myproject(c, dt, &c->last_iters[1], NULL);
    bc_vec(g, c->u, c->uBC);
    T cfl = mpcfl(c);
    double nx = dmin((double)cfl, 1.2 * (double)dt);
    if (dt_next) *dt_next = (double)(T)nx;
    return st;
}

int oc_advect(oc_ctx *c, double dtd, int step_index) {
    grid_t *g = &c->g;
    T tol = 10 * EPS_T, Dt = (T)dtd;
    memset(c->ruf, 0, g->ns * g->D * sizeof(T));
    set_cbar(c, c->f);
    int dirO[3];
    dir_order(g->D, step_index, dirO);
    int st = OC_OK;
    for (int op = 0; op < g->D; ++op) {
        int d = dirO[op];
        int s = advect_vof_1d(c, c->f, c->u0, c->u, (T)1 * Dt, d, tol, tol);
        st = worst(st, s);
        if (s == OC_NAN || s == OC_EXPLODE) return st;
    }
    return st;
}

/* ------------------------------------------------------------------ */
/* construction                                                         */
/* ------------------------------------------------------------------ */
oc_ctx *oc_create(const oc_params *p) {
    if (!(p->D == 2 || p->D == 3)) return NULL;
    oc_ctx *c = (oc_ctx *)calloc(1, sizeof(oc_ctx));
    int Ng[3];
    for (int d = 0; d < 3; ++d) Ng[d] = d < p->D ? p->N[d] + 2 : 1;
    grid_init(&c->g, p->D, Ng, p->perdir);
    c->g.exitbc = p->exit_bc && !(p->perdir & 1);
    grid_t *g = &c->g;
    c->D = p->D;
    for (int d = 0; d < 3; ++d) {
        c->uBC[d] = (T)p->u_bc[d];
        c->gv[d] = (T)p->g[d];
        c->zero3[d] = 0;
    }
    c->has_g = p->has_g;
    c->nu = (T)p->nu;
    c->has_mu = p->nu != 0;
    c->mu = (T)p->nu;
    c->lr = (T)p->lambda_rho;
    c->lm = (T)p->lambda_mu;
    c->has_eta = p->eta != 0;
    c->eta = (T)p->eta;
    c->scheme = p->normal_scheme;
    c->limiter = p->limiter;
    c->mg_tol = p->mg_tol > 0 ? p->mg_tol : 1e-4;
    c->mg_itmx = p->mg_itmx > 0 ? p->mg_itmx : 200;
    c->coarse_it = 6;
    /* the flexible CG preconditioner's solver iterations on coarse level l of
     * n interior cells (the HIP library's fcg_coarse_its) */
    for (int l = 0; l < 16; ++l) {
        long n = 1;
        for (int a = 0; a < p->D; ++a) n *= (p->N[a] >> l) > 1 ? (p->N[a] >> l) : 1;
        c->fcg_cit[l] = (l == 1 && n > 1000000) ? 6 : n > 100000 ? 56 : n > 512 ? 64 : 32;
    }
    c->fcg_cg1 = 1;
    c->fine_cheb = 1;
    c->poisson_mode = p->poisson_mode;
    c->poisson_rtol = p->poisson_rtol > 0 ? p->poisson_rtol : 1e-6;
    c->smooth_it = p->smooth_it;
    size_t n = g->ns, nv = n * g->D;
    /* WaterLily Flow: u = uλ(i,x) = u_BC[i]; BC!; u⁰ = copy(u); μ₀ = 1; BC!(μ₀,0) */
    c->u = alloc_fill(nv, 0);
    for (int d = 1; d <= g->D; ++d)
        for (size_t q = 0; q < n; ++q) VC(c->u, d)[q] = c->uBC[d - 1];
    bc_vec(g, c->u, c->uBC);
    c->u0 = alloc_fill(nv, 0);
    memcpy(c->u0, c->u, nv * sizeof(T));
    c->fv = alloc_fill(nv, 0);
    c->p = alloc_fill(n, 0);
    c->sig = alloc_fill(n, 0);
    c->mu0 = alloc_fill(nv, 1);
    bc_vec_e(g, c->mu0, c->zero3, 0); /* BC!(μ₀, zeros, false, perdir), flow.jl:258 */
    /* cVOF.jl:45-89 */
    c->f = alloc_fill(n, 1);
    c->al = alloc_fill(n, 0);
    c->nh = alloc_fill(nv, 0);
    c->cb = (int8_t *)calloc(n, 1);
    c->f0 = alloc_fill(n, 1);
    c->ff = alloc_fill(n, 0);
    c->ru = alloc_fill(nv, 0);
    c->ruf = alloc_fill(nv, 0);
    c->dr = alloc_fill(nv, 1);
    if (mg_build(c) != OC_OK) { oc_destroy(c); return NULL; }
    return c;
}
void oc_destroy(oc_ctx *c) {
    if (!c) return;
    T *bufs[] = {c->u, c->u0, c->fv, c->p, c->sig, c->mu0, c->f, c->f0, c->al, c->nh, c->ff, c->ru, c->ruf, c->dr};
    for (unsigned q = 0; q < sizeof bufs / sizeof bufs[0]; ++q) free(bufs[q]);
    free(c->cb);
    free(c->cg1_u); free(c->cg1_w); free(c->cg1_s); free(c->cheb_d0);
    for (int l = 0; l < c->nlev; ++l) {
        lev_t *lv = &c->lev[l];
        free(lv->Dg); free(lv->iD); free(lv->eps); free(lv->r);
        if (lv->own) { free(lv->L); free(lv->x); free(lv->z); }
    }
    free(c);
}
/* VOFutil.jl:8-37 applyVOF! from SDF samples (per interior cell, x fastest:
 * [sdf(centre), sdf(centre+Δx e1), sdf(centre-Δx e1), ...]), then
 * cleanWisp!, BCf! and f⁰ = copy(f) (cVOF.jl:64-65) */
int oc_apply_vof(oc_ctx *c, const T *sdf) {
    grid_t *g = &c->g;
    rng_t R = r_in(g);
    int ns = 1 + 2 * g->D;
    long q = 0;
    FOR_R(R) {
        size_t I = IX(i, j, k);
        const T *s = sdf + (size_t)q * ns;
        T sumN = 0, sumN2 = 0, n[3];
        for (int d = 1; d <= g->D; ++d) {
            T dd = s[2 * d - 1] - s[2 * d];
            VC(c->nh, d)[I] = dd;
            n[d - 1] = dd;
            sumN += dd;
            sumN2 += dd * dd;
        }
        c->al[I] = sumN / 2 - sqrt(sumN2) * s[0];
        c->f[I] = getVolumeFractionN(g->D, n, c->al[I]);
        q++;
    }
    clean_wisp(g, c->f, 10 * EPS_T);
    bc_f(g, c->f, 0);
    memcpy(c->f0, c->f, g->ns * sizeof(T));
    return OC_OK;
}
int oc_set_initial_velocity(oc_ctx *c, const T *u) {
    grid_t *g = &c->g;
    size_t nv = g->ns * g->D;
    memcpy(c->u, u, nv * sizeof(T));
    bc_vec(g, c->u, c->uBC);
    memcpy(c->u0, c->u, nv * sizeof(T));
    return OC_OK;
}
double oc_cfl(oc_ctx *c) { return (double)mpcfl(c); }

static T *field_ptr(oc_ctx *c, int id, size_t *len) {
    grid_t *g = &c->g;
    size_t n = g->ns, nv = n * g->D;
    switch (id) {
    case OC_F: *len = n; return c->f;
    case OC_F0: *len = n; return c->f0;
    case OC_ALPHA: *len = n; return c->al;
    case OC_NHAT: *len = nv; return c->nh;
    case OC_FFLUX: *len = n; return c->ff;
    case OC_RHOU: *len = nv; return c->ru;
    case OC_RHOUF: *len = nv; return c->ruf;
    case OC_DRHO: *len = nv; return c->dr;
    case OC_U: *len = nv; return c->u;
    case OC_U0: *len = nv; return c->u0;
    case OC_FORCE: *len = nv; return c->fv;
    case OC_P: *len = n; return c->p;
    case OC_SIGMA: *len = n; return c->sig;
    case OC_MU0: *len = nv; return c->mu0;
    case OC_MG_D: *len = n; return c->lev[0].Dg;
    case OC_MG_ID: *len = n; return c->lev[0].iD;
    case OC_MG_R: *len = n; return c->lev[0].r;
    case OC_MG_EPS: *len = n; return c->lev[0].eps;
    default: *len = 0; return NULL;
    }
}
int oc_field_len(oc_ctx *c, int id) {
    size_t len;
    if (id == OC_CBAR) return (int)c->g.ns;
    field_ptr(c, id, &len);
    return (int)len;
}
int oc_set_field(oc_ctx *c, int id, const void *src) {
    size_t len;
    if (id == OC_CBAR) { memcpy(c->cb, src, c->g.ns); return OC_OK; }
    T *a = field_ptr(c, id, &len);
    if (!a) return OC_BADARG;
    memcpy(a, src, len * sizeof(T));
    return OC_OK;
}
int oc_get_field(oc_ctx *c, int id, void *dst) {
    size_t len;
    if (id == OC_CBAR) { memcpy(dst, c->cb, c->g.ns); return OC_OK; }
    T *a = field_ptr(c, id, &len);
    if (!a) return OC_BADARG;
    memcpy(dst, a, len * sizeof(T));
    return OC_OK;
}
int oc_mg_nlevels(oc_ctx *c) { return c->nlev; }
int oc_mg_level_dims(oc_ctx *c, int l, int *Ng) {
    if (l < 0 || l >= c->nlev) return OC_BADARG;
    for (int d = 0; d < 3; ++d) Ng[d] = c->lev[l].g.Ng[d];
    return OC_OK;
}
/* a copy of level l's array (ghost-inclusive): which 0 = L (D components),
 * 1 = D (diagonal), 2 = iD, 3 = r, 4 = x, 5 = ϵ.  For the tests of the
 * multigrid's transfer operators and coarse coefficients. */
int oc_mg_level_get(oc_ctx *c, int l, int which, real_t *out) {
    if (l < 0 || l >= c->nlev || which < 0 || which > 5) return OC_BADARG;
    lev_t *lv = &c->lev[l];
    size_t n = lv->g.ns * (which == 0 ? (size_t)lv->g.D : 1);
    const T *a = which == 0 ? lv->L : which == 1 ? lv->Dg : which == 2 ? lv->iD : which == 3 ? lv->r
                                                                    : which == 4 ? lv->x : lv->eps;
    for (size_t q = 0; q < n; ++q) out[q] = a[q];
    return OC_OK;
}
/* level l's r set from `in`, then restrict! into level l+1 (whose r is
 * returned by oc_mg_level_get), or prolongate! level l+1's x (set from `in`)
 * into level l's ϵ: dir 0 restricts, 1 prolongates. */
int oc_mg_transfer(oc_ctx *c, int l, int dir, const real_t *in) {
    if (l < 0 || l + 1 >= c->nlev) return OC_BADARG;
    lev_t *fl = &c->lev[l], *co = &c->lev[l + 1];
    if (dir == 0) {
        for (size_t q = 0; q < fl->g.ns; ++q) fl->r[q] = in[q];
        mg_restrict(co, fl);
    } else {
        for (size_t q = 0; q < co->g.ns; ++q) co->x[q] = in[q];
        mg_prolongate(fl, co);
    }
    return OC_OK;
}
/* level l's array `which` (as oc_mg_level_get; 3 = r, 4 = x only) set from `in` */
int oc_mg_level_set(oc_ctx *c, int l, int which, const real_t *in) {
    if (l < 0 || l >= c->nlev || (which != 3 && which != 4)) return OC_BADARG;
    lev_t *lv = &c->lev[l];
    T *a = which == 3 ? lv->r : lv->x;
    for (size_t q = 0; q < lv->g.ns; ++q) a[q] = in[q];
    return OC_OK;
}
/* pcg!(level l; it) on the level's r and x: WaterLily's two-reduction form
 * (single = 0) or the single-reduction form of the flexible CG's coarse
 * levels (single = 1; levels >= 1, whose size the scratch covers). */
int oc_mg_pcg(oc_ctx *c, int l, int it, int single) {
    if (l < 0 || l >= c->nlev || it < 1 || (single && l < 1)) return OC_BADARG;
    if (single) mg_pcg1(&c->lev[l], it, c->cg1_u, c->cg1_w, c->cg1_s);
    else mg_pcg(&c->lev[l], it);
    return OC_OK;
}
const char *oc_last_error(oc_ctx *c) { return c->msg; }
void oc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* stage entry points                                                   */
/* ------------------------------------------------------------------ */
int oc_reconstruct(oc_ctx *c) { reconstruct(c, c->f, c->al, c->nh); return OC_OK; }
int oc_vof_sweep(oc_ctx *c, int d, double dt) {
    set_cbar(c, c->f);
    memset(c->ruf, 0, c->g.ns * c->g.D * sizeof(T));
    T tol = 10 * EPS_T;
    return advect_vof_1d(c, c->f, c->u0, c->u, (T)dt, d, 10 * tol, tol);
}
int oc_visc(oc_ctx *c, int use_f0) {
    visc(c, c->fv, c->u, c->nh, c->sig, use_f0 ? c->f0 : c->f);
    return OC_OK;
}
int oc_update_L(oc_ctx *c, int use_f0) { update_L(c, use_f0 ? c->f0 : c->f); return OC_OK; }
int oc_mg_update(oc_ctx *c) { mg_update(c); return OC_OK; }
int oc_project(oc_ctx *c, double dt_w, int *iters, double *r2) {
    int st = myproject(c, (T)dt_w, iters, r2);
    bc_vec(&c->g, c->u, c->uBC);
    return st;
}
int oc_mg_solve(oc_ctx *c, double tol, int itmx, int *iters, double *r2) {
    int n = mg_solver(c, tol, itmx, r2);
    if (iters) *iters = n;
    return OC_OK;
}

/* FP64 iterative refinement around the MultiLevelPoisson V-cycle (this
 * build's mixed-precision solve; no reference counterpart).  The operator is
 * the same coefficient field L (values of type T, evaluated in double). */
static double mult64(const grid_t *g, size_t I, const T *L, const double *x) {
    double dg = 0, sl = 0, su = 0;
    for (int i = 1; i <= g->D; ++i) dg -= ((double)VC(L, i)[I] + (double)VC(L, i)[I + ST(i)]);
    for (int i = 1; i <= g->D; ++i) sl += x[I - ST(i)] * (double)VC(L, i)[I];
    for (int i = 1; i <= g->D; ++i) su += x[I + ST(i)] * (double)VC(L, i)[I + ST(i)];
    return (x[I] * dg + sl) + su;
}
static void per_bc64(const grid_t *g, double *a) {
    for (int jj = 1; jj <= g->D; ++jj) {
        if (!PER(jj)) continue;
        rng_t R1 = r_slice(g, 1, jj, 1), R3 = r_slice(g, NG(jj), jj, 1);
        FOR_R(R1) { size_t I = IX(i, j, k); a[I] = a[I + (NG(jj) - 2) * ST(jj)]; }
        FOR_R(R3) { size_t I = IX(i, j, k); a[I] = a[I + (2 - NG(jj)) * ST(jj)]; }
    }
}
static double resid64(oc_ctx *c, const double *z, const double *x, double *r) {
    grid_t *g = &c->g;
    rng_t R = r_in(g);
    lev_t *l0 = &c->lev[0];
    double s = 0;
    long cnt = rng_count(R);
    PFOR_R(R) {
        size_t I = IX(i, j, k);
        r[I] = l0->iD[I] == 0 ? 0.0 : z[I] - mult64(g, I, l0->L, x);
    }
    FOR_R(R) s += r[IX(i, j, k)];
    s /= (double)cnt;
    double nn = 0;
    FOR_R(R) {
        size_t I = IX(i, j, k);
        r[I] -= s;
        nn += r[I] * r[I];
    }
    return sqrt(nn);
}
int oc_project_refined(oc_ctx *c, double dt_w, double rtol, int max_outer, int *outer, double *rel) {
    grid_t *g = &c->g;
    rng_t R = r_in(g);
    lev_t *l0 = &c->lev[0];
    size_t n = g->ns;
    double *z = (double *)calloc(n, sizeof(double)), *x = (double *)calloc(n, sizeof(double)),
           *r = (double *)calloc(n, sizeof(double));
    double zs = 0;
    long cnt = rng_count(R);
    FOR_R(R) {
        size_t I = IX(i, j, k);
        double s = 0;
        for (int d = 1; d <= g->D; ++d) s += (double)VC(c->u, d)[I + ST(d)] - (double)VC(c->u, d)[I];
        z[I] = s;
        zs += s;
    }
    zs /= (double)cnt;
    double zn = 0;
    FOR_R(R) {
        size_t I = IX(i, j, k);
        z[I] -= zs;
        zn += z[I] * z[I];
    }
    zn = sqrt(zn);
    for (size_t q = 0; q < n; ++q) x[q] = (double)c->p[q] * dt_w;
    per_bc64(g, x);
    double rn = resid64(c, z, x, r);
    double rl = zn > 0 ? rn / zn : 0.0;
    int it = 0;
    while (rl > rtol && it < max_outer) {
        /* the inner FP32 solve sees the residual scaled to unit RMS: WaterLily's
         * pcg! stops on absolute thresholds (|ρ| < 10eps), which a small
         * residual would trip before any smoothing is done */
        double sc = rn / sqrt((double)cnt);
        for (size_t q = 0; q < n; ++q) { l0->r[q] = (T)(r[q] / sc); l0->x[q] = 0; }
        mg_vcycle(c, 0);
        mg_pcg(l0, 6);
        per_bc(g, l0->x);
        for (size_t q = 0; q < n; ++q) x[q] += sc * (double)l0->x[q];
        per_bc64(g, x);
        rn = resid64(c, z, x, r);
        rl = rn / zn;
        it++;
    }
    for (int i_ = 1; i_ <= g->D; ++i_) {
        T *ui = VC(c->u, i_), *Li = VC(c->mu0, i_);
        PFOR_R(R) {
            size_t I = IX(i, j, k);
            ui[I] = (T)((double)ui[I] - (double)Li[I] * (x[I] - x[I - ST(i_)]));
        }
    }
    for (size_t q = 0; q < n; ++q) c->p[q] = (T)(x[q] / dt_w);
    bc_vec(g, c->u, c->uBC);
    if (outer) *outer = it;
    if (rel) *rel = rl;
    free(z); free(x); free(r);
    return OC_OK;
}



/* symmetric V(npre,npost) cycle with damped Jacobi smoothing on WaterLily's
 * levels (same transfer and coarse operators); the coarsest level gets
 * `ncoarse` smoothing sweeps.  Preconditioner of the flexible CG. */
static void mg_djacobi(lev_t *lv, T omega, int sweeps) {
    grid_t *g = &lv->g;
    rng_t R = r_in(g);
    for (int s = 0; s < sweeps; ++s) {
        PFOR_R(R) { size_t I = IX(i, j, k); lv->eps[I] = omega * lv->r[I] * lv->iD[I]; }
        mg_increment(lv);
    }
}
static void mg_vcycle_sym(oc_ctx *c, int l, int npre, int npost, T omega, int ncoarse) {
    lev_t *fl = &c->lev[l];
    if (l == c->nlev - 1) { mg_djacobi(fl, omega, ncoarse); return; }
    lev_t *co = &c->lev[l + 1];
    mg_djacobi(fl, omega, npre);
    mg_restrict(co, fl);
    for (size_t q = 0; q < co->g.ns; ++q) co->x[q] = 0;
    mg_vcycle_sym(c, l + 1, npre, npost, omega, ncoarse);
    mg_prolongate(fl, co);
    mg_increment(fl);
    mg_djacobi(fl, omega, npost);
}
int g_cycle_type = 0, g_npre = 2, g_npost = 2, g_ncoarse = 16;
double g_omega = 0.8;
void oc_set_coarse_it(oc_ctx *c, int it) { c->coarse_it = it; }
void oc_set_fcg_cg1(oc_ctx *c, int on) { c->fcg_cg1 = on; }
void oc_set_fine_cheb(oc_ctx *c, int on) { c->fine_cheb = on; }
void oc_set_fcg_cit(oc_ctx *c, const int *cit, int n) {
    for (int l = 0; l < 16; ++l) c->fcg_cit[l] = l < n ? cit[l] : cit[n - 1];
}
void oc_set_cycle(int type, int npre, int npost, double omega, int ncoarse) {
    g_cycle_type = type; g_npre = npre; g_npost = npost; g_omega = omega; g_ncoarse = ncoarse;
}
/* Flexible preconditioned CG in FP64 (Notay's FCG(1)) with the FP32
 * MultiLevelPoisson V-cycle (+ `smooth_it` pcg! iterations on the finest
 * level) as the preconditioner.  This build's solver for the 1e-6 relative
 * residual target; no reference counterpart.  Returns the iteration count. */
/* The fine-level smoother after the V-cycle (the HIP library's k_cheb2_dots):
 * two Chebyshev steps on D⁻¹A for the interval [0.5, 2]: d₀ = c0·(r·iD),
 * r₁ = r − A d₀, d₁ = ca·d₀ + cb·(r₁·iD), x = (x + d₀) + d₁, with
 * θ = 1.25, δ = 0.75, σ = θ/δ, ρ₀ = 1/σ, ρ₁ = 1/(2σ − ρ₀), c0 = 1/θ,
 * ca = ρ₁ρ₀, cb = 2ρ₁/δ (Saad's three-term Chebyshev recurrence). */
static void mg_cheb2(lev_t *lv, T *d0) {
    const double lmin = 0.5, lmax = 2.0, th = (lmax + lmin) / 2, de = (lmax - lmin) / 2, sg = th / de;
    const double ro0 = 1 / sg, ro1 = 1 / (2 * sg - ro0);
    const T c0 = (T)(1 / th), ca = (T)(ro1 * ro0), cb = (T)(2 * ro1 / de);
    grid_t *g = &lv->g;
    rng_t R = r_in(g);
    PFOR_R(R) { size_t I = IX(i, j, k); d0[I] = (lv->r[I] * lv->iD[I]) * c0; }
    per_bc(g, d0);
    PFOR_R(R) {
        size_t I = IX(i, j, k);
        T r1 = lv->r[I] - mg_mult(g, I, lv->L, lv->Dg, d0);
        T d1 = ca * d0[I] + cb * (r1 * lv->iD[I]);
        lv->x[I] = (lv->x[I] + d0[I]) + d1;
    }
}
static void precond(oc_ctx *c, const double *r, double *w, double sc, int smooth_it) {
    lev_t *l0 = &c->lev[0];
    size_t n = c->g.ns;
    for (size_t q = 0; q < n; ++q) { l0->r[q] = (T)(r[q] / sc); l0->x[q] = 0; }
    if (g_cycle_type == 1) {
        mg_vcycle_sym(c, 0, g_npre, g_npost, (T)g_omega, g_ncoarse);
    } else {
        mg_vcycle_it(c, 0, c->fcg_cit);
        if (c->fine_cheb) mg_cheb2(l0, c->cheb_d0);
        else if (smooth_it > 0) mg_pcg(l0, smooth_it);
    }
    per_bc(&c->g, l0->x);
    for (size_t q = 0; q < n; ++q) w[q] = sc * (double)l0->x[q];
}
static double dot64(const grid_t *g, const double *a, const double *b) {
    rng_t R = r_in(g);
    double s = 0;
    FOR_R(R) { size_t I = IX(i, j, k); s += a[I] * b[I]; }
    return s;
}
int oc_project_fcg(oc_ctx *c, double dt_w, double rtol, int max_it, int smooth_it, int *iters, double *rel) {
    grid_t *g = &c->g;
    rng_t R = r_in(g);
    lev_t *l0 = &c->lev[0];
    size_t n = g->ns;
    long cnt = rng_count(R);
    double *z = calloc(n, 8), *x = calloc(n, 8), *r = calloc(n, 8), *w = calloc(n, 8), *p = calloc(n, 8),
           *q = calloc(n, 8);
    double zs = 0;
    FOR_R(R) {
        size_t I = IX(i, j, k);
        double s = 0;
        for (int d = 1; d <= g->D; ++d) s += (double)VC(c->u, d)[I + ST(d)] - (double)VC(c->u, d)[I];
        z[I] = s;
        zs += s;
    }
    zs /= (double)cnt;
    double zn = 0;
    FOR_R(R) { size_t I = IX(i, j, k); z[I] -= zs; zn += z[I] * z[I]; }
    zn = sqrt(zn);
    for (size_t qq = 0; qq < n; ++qq) x[qq] = (double)c->p[qq] * dt_w;
    per_bc64(g, x);
    double rn = resid64(c, z, x, r);
    double rl = zn > 0 ? rn / zn : 0.0;
    int it = 0, kk = 0;
    double pq = 0;
    while (rl > rtol && it < max_it) {
        precond(c, r, w, rn / sqrt((double)cnt), smooth_it);
        double wr = dot64(g, w, r);
        double beta = 0;
        if (kk > 0) beta = -dot64(g, w, q) / pq;
        FOR_R(R) { size_t I = IX(i, j, k); p[I] = w[I] + beta * p[I]; }
        per_bc64(g, p);
        PFOR_R(R) { size_t I = IX(i, j, k); PCOORD(P); (void)P; q[I] = mult64(g, I, l0->L, p); }
        pq = dot64(g, p, q);
        double alpha = wr / pq;
        double nn = 0;
        FOR_R(R) {
            size_t I = IX(i, j, k);
            x[I] += alpha * p[I];
            r[I] -= alpha * q[I];
            nn += r[I] * r[I];
        }
        rn = sqrt(nn);
        rl = rn / zn;
        it++;
        kk++;
        if (rl <= rtol) {  /* confirm with the true residual; restart the recurrence if it drifted */
            per_bc64(g, x);
            rn = resid64(c, z, x, r);
            rl = rn / zn;
            kk = 0;
        }
    }
    per_bc64(g, x);
    for (int i_ = 1; i_ <= g->D; ++i_) {
        T *ui = VC(c->u, i_), *Li = VC(c->mu0, i_);
        PFOR_R(R) {
            size_t I = IX(i, j, k);
            ui[I] = (T)((double)ui[I] - (double)Li[I] * (x[I] - x[I - ST(i_)]));
        }
    }
    for (size_t qq = 0; qq < n; ++qq) c->p[qq] = (T)(x[qq] / dt_w);
    bc_vec(g, c->u, c->uBC);
    if (iters) *iters = it;
    if (rel) *rel = rl;
    free(z); free(x); free(r); free(w); free(p); free(q);
    return OC_OK;
}

/* ------------------------------------------------------------------ */
/* point functions for known-answer tests                               */
/* ------------------------------------------------------------------ */
real_t oc_getIntercept2(real_t n1, real_t n2, real_t g) { return getIntercept2(n1, n2, g); }
real_t oc_getIntercept3(real_t n1, real_t n2, real_t n3, real_t g) { return getIntercept3(n1, n2, n3, g); }
real_t oc_getVolumeFraction2(real_t n1, real_t n2, real_t b) { return getVolumeFraction2(n1, n2, b); }
real_t oc_getVolumeFraction3(real_t n1, real_t n2, real_t n3, real_t b) { return getVolumeFraction3(n1, n2, n3, b); }
real_t oc_f2alpha3(real_t m1, real_t m2, real_t m3, real_t v) { return f2a3(m1, m2, m3, v); }
real_t oc_alpha2f3(real_t m1, real_t m2, real_t m3, real_t a) { return a2f3(m1, m2, m3, a); }
real_t oc_koren(real_t u, real_t c, real_t d) { return limiter(0, u, c, d); }
real_t oc_limiter(int lim, real_t u, real_t c, real_t d) { return limiter(lim, u, c, d); }
real_t oc_phiq_value(real_t Psi, real_t fd, real_t dt, real_t lr, int lim, real_t u, real_t c, real_t d) {
    return phiq_value(Psi, fd, dt, lr, lim, u, c, d);
}
double oc_rhoke(int D, const int *Ng, const real_t *u, const real_t *f, real_t lr, const int *I, const double *U) {
    grid_t g;
    grid_init(&g, D, Ng, 0);
    const grid_t *gp = &g;
    return rho_ke(gp, IXg(gp, I[0], I[1], D > 2 ? I[2] : 1), u, f, lr, U);
}
double oc_rhou(int D, const int *Ng, int i, const real_t *u, const real_t *f, real_t lr, const int *I, const double *U) {
    grid_t g;
    grid_init(&g, D, Ng, 0);
    const grid_t *gp = &g;
    return rho_u(gp, i, IXg(gp, I[0], I[1], D > 2 ? I[2] : 1), u, f, lr, U);
}
double oc_rhogh(int D, const int *Ng, const double *gv, const real_t *f, real_t lr, const int *I, const double *wl) {
    grid_t g;
    grid_init(&g, D, Ng, 0);
    const grid_t *gp = &g;
    int q[3] = {I[0], I[1], D > 2 ? I[2] : 1};
    return rho_gh(gp, q, IXg(gp, q[0], q[1], q[2]), gv, f, lr, wl);
}
double oc_ens(int D, const int *Ng, const real_t *w, const int *I) {
    grid_t g;
    grid_init(&g, D, Ng, 0);
    const grid_t *gp = &g;
    return ens(gp, IXg(gp, I[0], I[1], D > 2 ? I[2] : 1), w);
}
/* domain integrals over the interior: out = {Σρke, Σρu_1..3, Σρgh, Σens, Σf} */
int oc_metrics(oc_ctx *c, const double *U, const double *wl, double *out) {
    grid_t *g = &c->g;
    T *w = (T *)calloc(g->ns * (g->D == 3 ? 3 : 1), sizeof(T));
    vorticity(g, c->u, w);
    double gv[3] = {0, 0, 0};
    if (c->has_g)
        for (int d = 0; d < g->D; ++d) gv[d] = (double)c->gv[d];
    for (int q = 0; q < 7; ++q) out[q] = 0;
    rng_t R = r_in(g);
    FOR_R(R) {
        size_t I = IX(i, j, k);
        int q[3] = {i, j, k};
        out[0] += rho_ke(g, I, c->u, c->f, c->lr, U);
        for (int a = 1; a <= g->D; ++a) out[a] += rho_u(g, a, I, c->u, c->f, c->lr, U);
        out[4] += rho_gh(g, q, I, gv, c->f, c->lr, wl);
        out[5] += ens(g, I, w);
        out[6] += (double)c->f[I];
    }
    free(w);
    return OC_OK;
}
real_t oc_popinet_height(int D, const int *Ng, const real_t *f, const int *I, int i) {
    grid_t g;
    grid_init(&g, D, Ng, 0);
    int q[3] = {I[0], D > 1 ? I[1] : 1, D > 2 ? I[2] : 1};
    return popinet_height(&g, f, q, i);
}
real_t oc_curvature(int D, const int *Ng, const real_t *f, const int *I, int i) {
    grid_t g;
    grid_init(&g, D, Ng, 0);
    int q[3] = {I[0], D > 1 ? I[1] : 1, D > 2 ? I[2] : 1};
    return curvature(&g, f, q, i);
}
int oc_normal_block(int D, const int *Ng, const real_t *f, int scheme, const int *I, real_t *n_out) {
    grid_t gg, *g = &gg;
    grid_init(g, D, Ng, 0);
    T *nh = alloc_fill(g->ns * D, 0);
    size_t Ii = IX(I[0], I[1], D > 2 ? I[2] : 1);
    normal_scheme(g, scheme, f, nh, Ii);
    for (int d = 1; d <= D; ++d) n_out[d - 1] = VC(nh, d)[Ii];
    free(nh);
    return OC_OK;
}
real_t oc_getmu(int D, const int *Ng, const real_t *fface, int i, int jj, const int *I, real_t lm, real_t mu,
                real_t lr) {
    grid_t gg, *g = &gg;
    grid_init(g, D, Ng, 0);
    return getmu(g, i, jj, IX(I[0], I[1], D > 2 ? I[2] : 1), fface, lm, mu, lr);
}
int oc_bc_scalar(int D, const int *Ng, real_t *f, int perdir, int d) {
    grid_t g;
    grid_init(&g, D, Ng, perdir);
    bc_f(&g, f, d);
    return OC_OK;
}
int oc_bc_vel(int D, const int *Ng, real_t *a, const real_t *A, int perdir, int saveexit) {
    grid_t G, *g = &G;
    grid_init(g, D, Ng, perdir);
    bc_vec_e(g, a, A, saveexit);
    return OC_OK;
}
int oc_bc_facevec(int D, const int *Ng, real_t *v, int perdir) {
    grid_t g;
    grid_init(&g, D, Ng, perdir);
    bc_v(&g, v);
    return OC_OK;
}
int oc_bc_vof(int D, const int *Ng, real_t *f, real_t *alpha, real_t *nhat, int perdir) {
    grid_t g;
    grid_init(&g, D, Ng, perdir);
    bc_vof(&g, f, alpha, nhat);
    return OC_OK;
}
int oc_f2face(int D, const int *Ng, const real_t *fc, real_t *ff, int perdir) {
    grid_t g;
    grid_init(&g, D, Ng, perdir);
    f2face(&g, ff, fc);
    return OC_OK;
}
int oc_vof_flux_face(int D, const int *Ng, real_t *ff, const real_t *f, const real_t *alpha, const real_t *nhat,
                     real_t dl, int d, const int *IFace, real_t *rhouf, real_t lr) {
    grid_t gg, *g = &gg;
    grid_init(g, D, Ng, 0);
    vof_flux_face(g, ff, f, alpha, nhat, dl, d, IX(IFace[0], IFace[1], D > 2 ? IFace[2] : 1), rhouf, lr);
    return OC_OK;
}

/* ------------------------------------------------------------------ */
/* redistaning.jl: level-set reinitialisation (LevelSet, redistaning!) */
/* pinned by maintests.jl:264-319 (ramp, diagonal plane, planar front)  */
/* plus the closed forms of tests/test_oracle_pins.py                   */
/* ------------------------------------------------------------------ */
/* redistaning.jl:110 minmod */
static inline T ls_minmod(T a, T b) { return fabs(a) <= fabs(b) ? a : b; }
/* redistaning.jl:122-143: second-order ENO upwind contribution to |∇ϕ|² along one axis */
static T ls_grad2(T a, T b, T c, T d, T e, T s) {
    T dp = d - c, dm = c - b;
    T d2p = (e + c) - 2 * d, d2o = (d + b) - 2 * c, d2m = (c + a) - 2 * b;
    T dR = dp - ls_minmod(d2p, d2o) / 2;
    T dL = dm + ls_minmod(d2o, d2m) / 2;
    T wR = dR * s, wL = dL * s;
    if (wR < 0 && (wR + wL) < 0) return dR * dR;
    if (wL > 0 && (wR + wL) > 0) return dL * dL;
    return 0;
}
/* redistaning.jl:67-108 computeL!: L = ϕini (1 - |∇ϕ|) on the interior; per axis the
 * first and last interior cells take the Neumann (mirrored) or periodic (wrapped)
 * building block, the others the centred 5-point stencil */
static void ls_computeL(const grid_t *g, T *L, const T *phi, const T *ini) {
    for (size_t q = 0; q < g->ns; ++q) L[q] = 0;
    for (int ii = 1; ii <= g->D; ++ii) {
        long s = ST(ii);
        int N = NG(ii);
        rng_t lo = r_slice(g, 2, ii, 2), hi = r_slice(g, N - 1, ii, 2), in = r_in(g);
        in.lo[ii - 1] = 3;
        in.hi[ii - 1] = N - 2;
        FOR_R(lo) {
            size_t I = IX(i, j, k);
            T a = PER(ii) ? phi[I + (long)(N - 4) * s] : phi[I - s];
            L[I] += ls_grad2(a, phi[I - s], phi[I], phi[I + s], phi[I + 2 * s], tsign(ini[I]));
        }
        PFOR_R(in) {
            size_t I = IX(i, j, k);
            L[I] += ls_grad2(phi[I - 2 * s], phi[I - s], phi[I], phi[I + s], phi[I + 2 * s], tsign(ini[I]));
        }
        FOR_R(hi) {
            size_t I = IX(i, j, k);
            T e = PER(ii) ? phi[I - (long)(N - 4) * s] : phi[I + s];
            L[I] += ls_grad2(phi[I - 2 * s], phi[I - s], phi[I], phi[I + s], e, tsign(ini[I]));
        }
    }
    rng_t in = r_in(g);
    PFOR_R(in) {
        size_t I = IX(i, j, k);
        L[I] = ini[I] * (1 - sqrt(L[I]));
    }
}
/* redistaning.jl:28-31 _redistaningStage!: ϕ = α ϕ⁰ + (1-α)(ϕ + dτ L(ϕ)) on the interior */
static void ls_stage(const grid_t *g, T *phi, const T *phi0, const T *ini, T *L, T dtau, T alpha) {
    ls_computeL(g, L, phi, ini);
    rng_t in = r_in(g);
    PFOR_R(in) {
        size_t I = IX(i, j, k);
        phi[I] = alpha * phi0[I] + (1 - alpha) * (phi[I] + dtau * L[I]);
    }
}
/* Julia round(x): nearest integer, ties to even */
static double round_even(double x) {
    double r = floor(x), f = x - r;
    if (f > 0.5 || (f == 0.5 && fmod(r, 2.0) != 0)) r += 1;
    return r;
}
/* redistaning.jl:44-56 redistaning!: round(d/dτ) SSP-RK3 steps, BCf! after each stage */
static void ls_redistance(const grid_t *g, T *phi, T *phi0, const T *ini, T *L, double d, double dtau) {
    long itmx = (long)round_even(d / dtau);
    T dt = (T)dtau;
    for (long it = 0; it < itmx; ++it) {
        memcpy(phi0, phi, g->ns * sizeof(T));
        ls_stage(g, phi, phi0, ini, L, dt, (T)0);
        bc_f(g, phi, 0);
        ls_stage(g, phi, phi0, ini, L, dt, (T)(3.0 / 4));
        bc_f(g, phi, 0);
        ls_stage(g, phi, phi0, ini, L, dt, (T)(1.0 / 3));
        bc_f(g, phi, 0);
    }
}
real_t oc_ls_grad2(real_t a, real_t b, real_t c, real_t d, real_t e, real_t s) { return ls_grad2(a, b, c, d, e, s); }
int oc_ls_computeL(int D, const int *Ng, real_t *L, const real_t *phi, const real_t *ini, int perdir) {
    grid_t g;
    grid_init(&g, D, Ng, perdir);
    for (int d = 0; d < D; ++d)
        if (Ng[d] < 4) return OC_BADARG;
    ls_computeL(&g, L, phi, ini);
    return OC_OK;
}
int oc_ls_stage(int D, const int *Ng, real_t *phi, const real_t *phi0, const real_t *ini, real_t *L, real_t dtau,
                real_t alpha, int perdir) {
    grid_t g;
    grid_init(&g, D, Ng, perdir);
    for (int d = 0; d < D; ++d)
        if (Ng[d] < 4) return OC_BADARG;
    ls_stage(&g, phi, phi0, ini, L, dtau, alpha);
    return OC_OK;
}
int oc_redistance(int D, const int *Ng, real_t *phi, real_t *phi0, const real_t *ini, real_t *L, double d, double dtau,
                  int perdir) {
    grid_t g;
    grid_init(&g, D, Ng, perdir);
    for (int q = 0; q < D; ++q)
        if (Ng[q] < 4) return OC_BADARG;
    ls_redistance(&g, phi, phi0, ini, L, d, dtau);
    return OC_OK;
}
