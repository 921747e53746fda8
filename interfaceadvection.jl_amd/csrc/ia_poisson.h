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
#include <type_traits>

namespace ia {

// column tiling: a block covers nx (x) by NT/nx (y) columns of zb cells.
// nx = TX (64) everywhere, except coarse levels of at most 32 cells per row
// with one cell per thread, where nx = 32 fills every lane (32 by 8).  The
// LDS-tiled traversals assume nx = TX; they run only where zb > 1.
struct CTiling {
    int nt[3];
    int zb;
    long nblk;
    int nx;
};
__device__ inline bool thread_col(const Grid &g, const CTiling &t, int p[3], int &zlo, int &zhi) {
    long b = swizzle_block((long)blockIdx.x, t.nblk);
    long bx = b % t.nt[0];
    long r = b / t.nt[0];
    long by = r % t.nt[1];
    long bz = r / t.nt[1];
    int tx = threadIdx.x % t.nx, ty = threadIdx.x / t.nx;
    p[0] = 1 + (int)bx * t.nx + tx;
    p[1] = g.D > 1 ? 1 + (int)by * (NT / t.nx) + ty : 0;
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
    double rr;      // ‖r‖² after the last accepted update
    int halt, pad;  // set when the recursive residual met the tolerance:
                    // every solver kernel queued after that returns at once
};

// every block of a consumer sums the producer's per-block partials itself
// (fixed order: strided per-thread sums, then the block tree), so a
// reduction costs neither an atomic nor an extra launch
// every thread sums the wavefronts' sums in wavefront order (block_sum's
// order, so the same value) instead of waiting for thread 0 to broadcast:
// two barriers instead of four
__device__ inline double block_total(const double *part, long n, int K = 1, int k = 0) {
    __shared__ double sht[NT / 64];
    double v = 0;
    for (long q = threadIdx.x; q < n; q += NT) v += part[K * q + k];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) sht[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0;
    for (int q = 0; q < NT / 64; ++q) s += sht[q];
    __syncthreads();  // before another block_total of this kernel rewrites sht
    return s;
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
    const int *halt;  // nullptr, or the flexible CG's stop flag
};
// a kernel of a queued-ahead solver iteration returns at once when the
// solve has already converged (uniform for the whole grid: before any barrier)
#define HALT_IF(l)                      \
    if ((l).halt && *(l).halt) return;

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
// mult(I,L,D,x)  (WaterLily: x[I]D[I] + Σ_i x[I-δi]L[I,i] + Σ_i x[I+δi]L[I+δi,i]).
// D[I] = -Σ_i (L[I,i] + L[I+δi,i]) is formed from the L values the stencil
// loads anyway, in the same order as update! stores it (bit-identical, one
// array less to read).  `xv(K)` gives x at flat index K.
template <int D, class T, class X>
__device__ inline T mgmult_fd(const Grid &g, const int p[3], long I, const T *__restrict__ L, X xv) {
    Nb n = neighbours_d<D>(g, p);
    T xm[D], xp[D], lm[D], lp[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        xm[i] = xv(I + n.m[i]);
        xp[i] = xv(I + n.p[i]);
        lm[i] = L[g.ns * i + I];
        lp[i] = L[g.ns * i + I + g.st[i]];
    }
    T dg = 0, sl = 0, su = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) dg -= (lm[i] + lp[i]);
#pragma unroll
    for (int i = 0; i < D; ++i) sl += xm[i] * lm[i];
#pragma unroll
    for (int i = 0; i < D; ++i) su += xp[i] * lp[i];
    return (xv(I) * dg + sl) + su;
}
// the same with x given per cell coordinate (for values formed on the fly
// from another grid, e.g. the prolongated coarse solution)
template <int D, class T, class X>
__device__ inline T mgmult_fc(const Grid &g, const int p[3], long I, const T *__restrict__ L, X xv) {
    T xm[D], xp[D], lm[D], lp[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        int q[3] = {p[0], p[1], p[2]};
        q[i] = (g.per[i] && p[i] == 1) ? g.Ng[i] - 2 : p[i] - 1;
        xm[i] = xv(q);
        q[i] = (g.per[i] && p[i] == g.Ng[i] - 2) ? 1 : p[i] + 1;
        xp[i] = xv(q);
        lm[i] = L[g.ns * i + I];
        lp[i] = L[g.ns * i + I + g.st[i]];
    }
    T dg = 0, sl = 0, su = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) dg -= (lm[i] + lp[i]);
#pragma unroll
    for (int i = 0; i < D; ++i) sl += xm[i] * lm[i];
#pragma unroll
    for (int i = 0; i < D; ++i) su += xp[i] * lp[i];
    return (xv(p) * dg + sl) + su;
}
template <int D, class T>
__device__ inline T mgmult_d(const Grid &g, const int p[3], long I, const T *__restrict__ L, const T *__restrict__ x) {
    return mgmult_fd<D>(g, p, I, L, [&](long K) { return x[K]; });
}
template <class T> __device__ inline T mgmult(const Grid &g, const int p[3], long I, const T *L, const T *, const T *x) {
    return g.D == 3 ? mgmult_d<3>(g, p, I, L, x) : mgmult_d<2>(g, p, I, L, x);
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

// Pipelined column traversal: every load of plane z+1 is issued before the
// stores of plane z, so a wave keeps two planes of loads in flight instead
// of waiting a full memory round trip per plane.  load(p, I) returns the
// values a cell needs; store(p, I, values) computes and writes.
template <class LF, class SF>
__device__ inline void col_pipe(const Grid &g, const CTiling &t, LF load, SF store) {
    int p[3], zlo, zhi;
    if (!thread_col(g, t, p, zlo, zhi)) return;
    p[2] = zlo;
    long I = (long)p[0] + g.st[1] * p[1] + g.st[2] * zlo;
    auto cur = load(p, I);
    for (int z = zlo; z <= zhi; ++z) {
        int pn[3] = {p[0], p[1], z + 1};
        long In = I + g.st[2];
        auto nxt = cur;
        if (z < zhi) nxt = load(pn, In);
        p[2] = z;
        store(p, I, cur);
        cur = nxt;
        I = In;
    }
}
// the operand of mult() at a cell and its neighbours, with the L values
template <int D, class T> struct S7 {
    T c, m[D], p[D], lm[D], lp[D];
};
template <int D, class T, class X>
__device__ inline S7<D, T> ld7(const Grid &g, const int p[3], long I, const T *__restrict__ L, X xv) {
    Nb n = neighbours_d<D>(g, p);
    S7<D, T> s;
    s.c = xv(I);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        s.m[i] = xv(I + n.m[i]);
        s.p[i] = xv(I + n.p[i]);
        s.lm[i] = L[g.ns * i + I];
        s.lp[i] = L[g.ns * i + I + g.st[i]];
    }
    return s;
}
// mult() on loaded values, in mgmult_fd's operation order
template <int D, class T> __device__ inline T mul7(const S7<D, T> &s) {
    T dg = 0, sl = 0, su = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) dg -= (s.lm[i] + s.lp[i]);
#pragma unroll
    for (int i = 0; i < D; ++i) sl += s.m[i] * s.lm[i];
#pragma unroll
    for (int i = 0; i < D; ++i) su += s.p[i] * s.lp[i];
    return (s.c * dg + sl) + su;
}
// run body<D>() with D the grid's dimension (2 or 3) as a compile-time constant
#define DISPATCH_D(g, ...)                                  \
    if ((g).D == 3) {                                       \
        constexpr int D = 3;                                \
        __VA_ARGS__                                         \
    } else {                                                \
        constexpr int D = 2;                                \
        __VA_ARGS__                                         \
    }

// ---------------------------------------------------------------------------
// 2.5D LDS-tiled 7-point traversal (3D only) for an operand that is formed on
// the fly from other arrays (E(K) at flat index K).  The block's 64x4 column
// tile walks its z range; each plane of the operand, with a one-cell halo in
// x and y, is evaluated once into LDS (double-buffered: one barrier per
// plane); the z-neighbours stay in registers.  Periodic axes are resolved by
// evaluating ghost positions at their periodic image, so the tile holds
// exactly the values the neighbour-offset stencil would read.
//   body(p, I, c, xm, xp, ym, yp, zm, zp) runs for every interior cell.
// ---------------------------------------------------------------------------
__device__ inline int per_img(const Grid &g, int a, int q) {
    if (g.per[a]) {
        if (q == 0) return g.Ng[a] - 2;
        if (q == g.Ng[a] - 1) return 1;
    }
    return q;
}
template <class V, class R, class LF, class CF, class PF, class BF>
__device__ inline void tiled7(const Grid &g, const CTiling &t, LF load, CF comb, PF pre, BF body) {
    __shared__ V tile[2][TY + 2][TX + 2];
    long b = swizzle_block((long)blockIdx.x, t.nblk);
    long bx = b % t.nt[0], rr = b / t.nt[0];
    long by = rr % t.nt[1], bz = rr / t.nt[1];
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    int x = 1 + (int)bx * TX + tx, y = 1 + (int)by * TY + ty;
    int zlo = 1 + (int)bz * t.zb, zhi = min(zlo + t.zb - 1, g.Ng[2] - 2);
    bool active = x <= g.Ng[0] - 2 && y <= g.Ng[1] - 2;
    bool inarr = x <= g.Ng[0] - 1 && y <= g.Ng[1] - 1;
    int xi = per_img(g, 0, x), yi = per_img(g, 1, y);
    // halo roles: an x-halo value (edge lanes) and a y-halo value (edge rows)
    bool hx = tx == 0 || tx == TX - 1, hy = ty == 0 || ty == TY - 1;
    int hxq = tx == 0 ? x - 1 : x + 1, hyq = ty == 0 ? y - 1 : y + 1;
    bool hx_ok = hx && hxq <= g.Ng[0] - 1 && y <= g.Ng[1] - 1;
    bool hy_ok = hy && hyq <= g.Ng[1] - 1 && x <= g.Ng[0] - 1;
    int hxc = tx == 0 ? 0 : TX + 1, hyr = ty == 0 ? 0 : TY + 1;
    hxq = per_img(g, 0, hxq);
    hyq = per_img(g, 1, hyq);
    // load(K) takes the linear index; a load(q0, q1, q2) takes the cell's coordinates
    auto LD = [&](int q0, int q1, int q2) {
        if constexpr (std::is_invocable_v<LF, int, int, int>)
            return load(q0, q1, q2);
        else
            return load((long)q0 + g.st[1] * q1 + g.st[2] * q2);
    };
    R rm{}, rc{}, rp{}, rhx{}, rhy{};
    if (inarr) {
        rm = LD(xi, yi, per_img(g, 2, zlo - 1));
        rc = LD(xi, yi, zlo);
        rp = LD(xi, yi, per_img(g, 2, zlo + 1));
    }
    if (hx_ok) rhx = LD(hxq, yi, zlo);
    if (hy_ok) rhy = LD(xi, hyq, zlo);
    V em = comb(rm), ec = comb(rc);
    long I0 = (long)x + g.st[1] * y + g.st[2] * zlo;
    auto pv = pre(active ? I0 : 0);
    for (int z = zlo; z <= zhi; ++z) {
        int buf = (z - zlo) & 1;
        long I = (long)x + g.st[1] * y + g.st[2] * z;
        // prefetch: own column two planes ahead, halo and pre-loads one plane ahead
        R rn{}, nhx{}, nhy{};
        auto pvn = pv;
        if (z < zhi) {
            if (inarr) rn = LD(xi, yi, per_img(g, 2, z + 2));
            if (hx_ok) nhx = LD(hxq, yi, z + 1);
            if (hy_ok) nhy = LD(xi, hyq, z + 1);
            pvn = pre(active ? I + g.st[2] : 0);
        }
        V ep = comb(rp);
        tile[buf][ty + 1][tx + 1] = ec;
        if (hx) tile[buf][ty + 1][hxc] = hx_ok ? comb(rhx) : V(0);
        if (hy) tile[buf][hyr][tx + 1] = hy_ok ? comb(rhy) : V(0);
        __syncthreads();
        if (active) {
            int p[3] = {x, y, z};
            body(p, I, pv, ec, tile[buf][ty + 1][tx], tile[buf][ty + 1][tx + 2], tile[buf][ty][tx + 1],
                 tile[buf][ty + 2][tx + 1], em, ep);
        }
        em = ec;
        ec = ep;
        rp = rn;
        rhx = nhx;
        rhy = nhy;
        pv = pvn;
    }
}
// A position up to two cells outside the grid, as a level-set stage reads it:
// -1 and Ng are the points two cells beyond the first and last interior cells,
// the wrapped interior point on a periodic axis and the ghost itself
// (mirrored) otherwise; 0 and Ng-1 are the ghosts, every other q itself.
__device__ inline int ls_img2(const Grid &g, int a, int q) {
    if (q == -1) return g.per[a] ? g.Ng[a] - 3 : 0;
    if (q == g.Ng[a]) return g.per[a] ? 2 : g.Ng[a] - 1;
    return q;
}
// One stage of redistaning! in 3D on a column tiling: the block's 64x4
// columns walk their z range; each plane of ϕ with a two-cell x/y halo goes
// through LDS (double-buffered, one barrier per plane) and the four z
// neighbours stay in registers.  Every cell evaluates the same values in the
// same order as k_ls_stage.
template <class T>
__global__ void __launch_bounds__(NT) k_ls_stage_tiled(Grid g, CTiling t, const T *phi, const T *phi0, const T *ini,
                                                       T *out, T dtau, T alpha) {
    constexpr int WX = TX + 4, WY = TY + 4, NW = WX * WY;
    __shared__ T tile[2][WY][WX];
    long b = swizzle_block((long)blockIdx.x, t.nblk);
    long bx = b % t.nt[0], rr = b / t.nt[0];
    long by = rr % t.nt[1], bz = rr / t.nt[1];
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    int x0 = 1 + (int)bx * TX, y0 = 1 + (int)by * TY;
    int x = x0 + tx, y = y0 + ty;
    int zlo = 1 + (int)bz * t.zb, zhi = min(zlo + t.zb - 1, g.Ng[2] - 2);
    bool active = x <= g.Ng[0] - 2 && y <= g.Ng[1] - 2;
    // tile entries k = threadIdx.x + NT*j: positions (x0-2+i, y0-2+j') mapped by
    // ls_img2; entries beyond Ng in x or y are read by no interior cell
    long off[3];
    bool ok[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        int k = (int)threadIdx.x + NT * j;
        int xx = x0 - 2 + k % WX, yy = y0 - 2 + k / WX;
        ok[j] = k < NW && xx <= g.Ng[0] && yy <= g.Ng[1];
        off[j] = ok[j] ? (long)ls_img2(g, 0, xx) + g.st[1] * ls_img2(g, 1, yy) : 0;
    }
    long col = (long)x + g.st[1] * y;
    auto LZ = [&](int z) { return active ? phi[col + g.st[2] * ls_img2(g, 2, z)] : T(0); };
    T zm2 = LZ(zlo - 2), zm1 = LZ(zlo - 1), zc = LZ(zlo), zp1 = LZ(zlo + 1), zp2 = LZ(zlo + 2);
    for (int z = zlo; z <= zhi; ++z) {
        int buf = (z - zlo) & 1;
        long zo = g.st[2] * z;
        T *tb = &tile[buf][0][0];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            int k = (int)threadIdx.x + NT * j;
            if (k < NW) tb[k] = ok[j] ? phi[off[j] + zo] : T(0);
        }
        T zn = z < zhi ? LZ(z + 3) : T(0);
        __syncthreads();
        if (active) {
            long I = col + zo;
            T si = ini[I];
            T s = si > 0 ? T(1) : (si < 0 ? T(-1) : si);
            const T(*w)[WX] = tile[buf];
            int cx = tx + 2, cy = ty + 2;
            T c = w[cy][cx];
            T acc = 0;
            acc = acc + ls_grad2(w[cy][cx - 2], w[cy][cx - 1], c, w[cy][cx + 1], w[cy][cx + 2], s);
            acc = acc + ls_grad2(w[cy - 2][cx], w[cy - 1][cx], c, w[cy + 1][cx], w[cy + 2][cx], s);
            acc = acc + ls_grad2(zm2, zm1, c, zp1, zp2, s);
            T L = si * (T(1) - sqrt(acc));
            T v = alpha * phi0[I] + (T(1) - alpha) * (c + dtau * L);
            int p[3] = {x, y, z};
            write_images<T>(g, out, p, v, 0, 0, T(0));
        }
        zm2 = zm1;
        zm1 = zc;
        zc = zp1;
        zp1 = zp2;
        zp2 = zn;
    }
}
// the L values of a 3D cell (faces I and I+δ per axis)
template <class V> struct L6 {
    V m0, m1, m2, p0, p1, p2;
};
template <class V, class T> __device__ inline L6<V> load_L6(const Grid &g, long I, const T *__restrict__ L) {
    L6<V> l;
    l.m0 = (V)L[I];
    l.m1 = (V)L[g.ns + I];
    l.m2 = (V)L[2 * g.ns + I];
    l.p0 = (V)L[I + 1];
    l.p1 = (V)L[g.ns + I + g.st[1]];
    l.p2 = (V)L[2 * g.ns + I + g.st[2]];
    return l;
}
// mult() on tiled operand values, in mgmult_fd's operation order
template <class V>
__device__ inline V mult7(const L6<V> &l, V c, V xm, V xp, V ym, V yp, V zm, V zp) {
    V dg = 0, sl = 0, su = 0;
    dg -= (l.m0 + l.p0);
    dg -= (l.m1 + l.p1);
    dg -= (l.m2 + l.p2);
    sl += xm * l.m0;
    sl += ym * l.m1;
    sl += zm * l.m2;
    su += xp * l.p0;
    su += yp * l.p1;
    su += zp * l.p2;
    return (c * dg + sl) + su;
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
    HALT_IF(l)
    const T *__restrict__ r = l.r;
    const T *__restrict__ iD = l.iD;
    T *__restrict__ e = l.e;
    struct V2 {
        T r, d;
    };
    col_pipe(l.g, l.t, [&](const int *, long I) { return V2{r[I], iD[I]}; },
             [&](const int *, long I, const V2 &v) { e[I] = v.r * v.d; });
}
// increment!: r -= Aϵ; x += ϵ
template <int D, class T> __device__ inline void increment_d(const LevV<T> &l) {
    const T *__restrict__ e = l.e;
    T *__restrict__ r = l.r;
    T *__restrict__ x = l.x;
    struct V {
        S7<D, T> s;
        T r, x;
    };
    col_pipe(l.g, l.t,
             [&](const int *p, long I) { return V{ld7<D>(l.g, p, I, l.L, [&](long K) { return e[K]; }), r[I], x[I]}; },
             [&](const int *, long I, const V &v) {
                 r[I] = v.r - mul7<D>(v.s);
                 x[I] = v.x + v.s.c;
             });
}
template <class T>
__global__ void __launch_bounds__(NT) k_mg_increment(LevV<T> l) {
    HALT_IF(l)
    DISPATCH_D(l.g, increment_d<D>(l);)
}
// restrict!: coarse r = Σ children; coarse x = 0
template <class T>
__global__ void __launch_bounds__(NT) k_mg_restrict(LevV<T> c, Grid gf, const T *rf, int with_e) {
    HALT_IF(c)
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
        if (with_e) c.e[I] = s * c.iD[I];  // the Jacobi! step that follows on this level
    })
}
// prolongate!: fine ϵ = coarse x[down(I)]
template <class T>
__global__ void __launch_bounds__(NT) k_mg_prolong(LevV<T> f, Grid gc, const T *xc) {
    HALT_IF(f)
    COL_LOOP(f, {
        int q[3];
        for (int a = 0; a < 3; ++a) q[a] = a < f.g.D ? (p[a] + 3) / 2 - 1 : 0;
        f.e[I] = xc[(long)q[0] + gc.st[1] * q[1] + gc.st[2] * q[2]];
    })
}
// prolongate! + increment! + the first statement of the pcg! that follows:
// ϵ = coarse x[down(I)] formed on the fly at the cell and its neighbours;
// r -= Aϵ; x += ϵ; then ϵ = r·iD and partial r·ϵ (-> ρ₀).  pcg!'s z = r·iD
// is not stored: its first use overwrites it with Aϵ.
template <int D, class T>
__device__ inline void prolong_inc_init_d(const LevV<T> &l, const Grid &gc, const T *__restrict__ xc, double &v) {
    T *__restrict__ r = l.r;
    T *__restrict__ x = l.x;
    T *__restrict__ e = l.e;
    const T *__restrict__ iD = l.iD;
    const Grid &g = l.g;
    if (D == 3 && l.t.zb > 1) {
        // column tilings: the prolonged coarse x through an LDS plane tile
        struct Pre {
            L6<T> l;
            T r, x, d;
        };
        tiled7<T, T>(
            g, l.t,
            [&](int q0, int q1, int q2) {
                return xc[(long)((q0 + 3) / 2 - 1) + gc.st[1] * ((q1 + 3) / 2 - 1) + gc.st[2] * ((q2 + 3) / 2 - 1)];
            },
            [&](T q) { return q; },
            [&](long I) {
                Pre w;
                w.l = load_L6<T>(g, I, l.L);
                w.r = r[I];
                w.x = x[I];
                w.d = iD[I];
                return w;
            },
            [&](const int *, long I, const Pre &pv, T c, T xm, T xp, T ym, T yp, T zm, T zp) {
                T rr = pv.r - mult7<T>(pv.l, c, xm, xp, ym, yp, zm, zp);
                r[I] = rr;
                x[I] = pv.x + c;
                T zz = rr * pv.d;
                e[I] = zz;
                v += (double)rr * (double)zz;
            });
        return;
    }
    auto ec = [&](const int q[3]) {
        long Kc = (long)((q[0] + 3) / 2 - 1);
        if (D > 1) Kc += gc.st[1] * ((q[1] + 3) / 2 - 1);
        if (D > 2) Kc += gc.st[2] * ((q[2] + 3) / 2 - 1);
        return xc[Kc];
    };
    struct V {
        S7<D, T> s;
        T r, x, d;
    };
    col_pipe(
        g, l.t,
        [&](const int *p, long I) {
            V w;
            w.s.c = ec(p);
#pragma unroll
            for (int i = 0; i < D; ++i) {
                int q[3] = {p[0], p[1], p[2]};
                q[i] = (g.per[i] && p[i] == 1) ? g.Ng[i] - 2 : p[i] - 1;
                w.s.m[i] = ec(q);
                q[i] = (g.per[i] && p[i] == g.Ng[i] - 2) ? 1 : p[i] + 1;
                w.s.p[i] = ec(q);
                w.s.lm[i] = l.L[g.ns * i + I];
                w.s.lp[i] = l.L[g.ns * i + I + g.st[i]];
            }
            w.r = r[I];
            w.x = x[I];
            w.d = iD[I];
            return w;
        },
        [&](const int *, long I, const V &w) {
            T rr = w.r - mul7<D>(w.s);
            r[I] = rr;
            x[I] = w.x + w.s.c;
            T zz = rr * w.d;
            e[I] = zz;
            v += (double)rr * (double)zz;
        });
}
template <class T>
__global__ void __launch_bounds__(NT) k_mg_prolong_inc_init(LevV<T> l, Grid gc, const T *xc, double *partA) {
    HALT_IF(l)
    double v = 0;
    DISPATCH_D(l.g, prolong_inc_init_d<D>(l, gc, xc, v);)
    block_partial(v, partA);
}
// pcg! (WaterLily):  z = ϵ = r·iD ; partial r·z  (-> ρ₀)
template <class T>
__global__ void __launch_bounds__(NT) k_pcg_init(LevV<T> l, double *partA) {
    HALT_IF(l)
    double v = 0;
    COL_LOOP(l, {
        T zz = l.r[I] * l.iD[I];
        l.z[I] = zz;
        l.e[I] = zz;
        v += (double)l.r[I] * (double)zz;
    })
    block_partial(v, partA);
}
template <int D, class T> __device__ inline void pcg_Ae_d(const LevV<T> &l, double &v) {
    const T *__restrict__ e = l.e;
    T *__restrict__ z = l.z;
    if (D == 3 && l.t.zb > 1) {
        // column tilings: ϵ through an LDS plane tile (95.8 -> 81.5 µs at 256³)
        tiled7<T, T>(
            l.g, l.t, [&](long K) { return e[K]; }, [&](T q) { return q; },
            [&](long I) { return load_L6<T>(l.g, I, l.L); },
            [&](const int *, long I, const L6<T> &lv, T c, T xm, T xp, T ym, T yp, T zm, T zp) {
                T zz = mult7<T>(lv, c, xm, xp, ym, yp, zm, zp);
                z[I] = zz;
                v += (double)zz * (double)c;
            });
        return;
    }
    col_pipe(l.g, l.t, [&](const int *p, long I) { return ld7<D>(l.g, p, I, l.L, [&](long K) { return e[K]; }); },
             [&](const int *, long I, const S7<D, T> &s) {
                 T zz = mul7<D>(s);
                 z[I] = zz;
                 v += (double)zz * (double)s.c;
             });
}
// iteration i: ρ (from ρ₀ when i == 1, else from the previous ϵ-update), z = Aϵ, partial z·ϵ
template <class T>
__global__ void __launch_bounds__(NT) k_pcg_Ae(LevV<T> l, PcgState *S, int sin, int sout, const double *partA,
                                               double *partB) {
    HALT_IF(l)
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
    DISPATCH_D(l.g, pcg_Ae_d<D>(l, v);)
    block_partial(v, partB);
}
// α = ρ/(z·ϵ); x += αϵ; r -= αz; z = r·iD; partial r·z; stop when i == itmax or |α| < 1e-2
template <class T>
__global__ void __launch_bounds__(NT) k_pcg_update(LevV<T> l, PcgState *S, int sin, int sout, const double *partB,
                                                   double *partC, int i, int itmax) {
    HALT_IF(l)
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
    {
        T *__restrict__ x = l.x;
        T *__restrict__ r = l.r;
        T *__restrict__ z = l.z;
        const T *__restrict__ e = l.e;
        const T *__restrict__ iD = l.iD;
        struct V {
            T x, e, r, z, d;
        };
        col_pipe(l.g, l.t, [&](const int *, long I) { return V{x[I], e[I], r[I], z[I], iD[I]}; },
                 [&](const int *, long I, const V &w) {
                     x[I] = w.x + alpha * w.e;
                     T rr = w.r - alpha * w.z;
                     r[I] = rr;
                     T zz = rr * w.d;
                     z[I] = zz;
                     v += (double)rr * (double)zz;
                 });
    }
    block_partial(v, partC);
}
// k_pcg_eps fused with the next k_pcg_Ae: β from ρ₂ = r·z, then ϵ' = βϵ + z
// formed on the fly at the cell and its neighbours, z' = Aϵ'; ϵ' and z' go
// to the other buffer pair (e2, z2) so no thread overwrites what another
// still reads; partial z'·ϵ'
template <int D, class T>
__device__ inline void eps_ae_d(const LevV<T> &l, T beta, T *__restrict__ e2, T *__restrict__ z2, double &v) {
    const T *__restrict__ e = l.e;
    const T *__restrict__ z = l.z;
    col_pipe(l.g, l.t,
             [&](const int *p, long I) { return ld7<D>(l.g, p, I, l.L, [&](long K) { return beta * e[K] + z[K]; }); },
             [&](const int *, long I, const S7<D, T> &s) {
                 T zz = mul7<D>(s);
                 e2[I] = s.c;
                 z2[I] = zz;
                 v += (double)zz * (double)s.c;
             });
}
template <class T>
__global__ void __launch_bounds__(NT) k_pcg_epsAe(LevV<T> l, T *e2, T *z2, PcgState *S, int sin, int sout,
                                                  const double *partC, double *partB) {
    HALT_IF(l)
    PcgState st = S[sin];
    T beta = T(0);
    bool go = false;
    if (!st.done) {
        if (st.stop) {
            st.done = 1;
        } else {
            T rho2 = (T)block_total(partC, l.t.nblk);
            if (fabs(rho2) < 10 * eps_of<T>()) {
                st.done = 1;
            } else {
                beta = rho2 / (T)st.rho;
                st.rho = (double)rho2;
                go = true;
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) S[sout] = st;
    if (!go) return;
    double v = 0;
    DISPATCH_D(l.g, eps_ae_d<D>(l, beta, e2, z2, v);)
    block_partial(v, partB);
}
// ρ₂ = r·z; exit when stop or |ρ₂| < 10eps; β = ρ₂/ρ; ϵ = βϵ + z
template <class T>
__global__ void __launch_bounds__(NT) k_pcg_eps(LevV<T> l, PcgState *S, int sin, int sout, const double *partC) {
    HALT_IF(l)
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
// ---------------------------------------------------------------------------
// Jacobi-preconditioned CG in the single-reduction form of Chronopoulos and
// Gear (1989): the flexible CG preconditioner's solver on coarse levels above
// the one-workgroup tail (no reference counterpart; WaterLily's pcg! needs two
// reductions, so two launches, per iteration).  With u = r·iD and w = A u:
//   init:  u₀, w₀ = A u₀, partials γ₀ = r·u₀, δ₀ = u₀·w₀
//   step:  β = γ/γ₋, α = γ/(δ − βγ/α₋)  (first step: α = γ/δ, p = u, s = w)
//          p = u + βp, s = w + βs, x += αp, r' = r − αs, then
//          w' = A u' with u'(K) = r'(K)·iD(K) formed on the fly at the cell and
//          its neighbours, partials γ' = r'·u', δ' = u'·w'
// so one launch per iteration.  r, w and s are read at neighbours, so their
// new values go to the other buffer of a pair; p and x are updated in place.
// Stops (flags in the PcgState slot, as pcg!) after `it` updates of x, after
// an update with |α| < 1e-2, or when |γ| < 10eps.
// ---------------------------------------------------------------------------
template <class T> struct Cg1V {
    const T *r, *w, *s;   // read (previous iterate), also at neighbours
    T *ro, *wo, *so;      // written
};
template <class T>
__global__ void __launch_bounds__(NT) k_cg1_init(LevV<T> l, T *w, PcgState *S, double *part) {
    HALT_IF(l)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        PcgState st;
        st.rho = 0;
        st.alpha = 0;
        st.done = 0;
        st.stop = 0;
        S[0] = st;
    }
    const T *__restrict__ r = l.r;
    const T *__restrict__ iD = l.iD;
    T *__restrict__ pd = l.e;  // search direction
    T *__restrict__ s = l.z;
    double g = 0, d = 0;
    DISPATCH_D(l.g, {
        col_pipe(l.g, l.t,
                 [&](const int *q, long I) { return ld7<D>(l.g, q, I, l.L, [&](long K) { return r[K] * iD[K]; }); },
                 [&](const int *, long I, const S7<D, T> &st) {
                     T ww = mul7<D>(st);
                     pd[I] = st.c;
                     s[I] = ww;
                     w[I] = ww;
                     g += (double)r[I] * (double)st.c;
                     d += (double)st.c * (double)ww;
                 });
    })
    block_partial(g, part, 2, 0);
    block_partial(d, part, 2, 1);
}
template <class T>
__global__ void __launch_bounds__(NT) k_cg1_step(LevV<T> l, Cg1V<T> b, PcgState *S, int q, int it,
                                                 const double *pin, double *pout) {
    HALT_IF(l)
    PcgState st = S[q - 1];
    T alpha = 0, beta = 0;
    if (!st.done) {
        if (st.stop) {
            st.done = 1;
        } else {
            T gam = (T)block_total(pin, l.t.nblk, 2, 0);
            T del = (T)block_total(pin, l.t.nblk, 2, 1);
            if (fabs(gam) < 10 * eps_of<T>()) {
                st.done = 1;
            } else {
                if (q == 1) {
                    alpha = gam / del;
                } else {
                    beta = gam / (T)st.rho;
                    alpha = gam / (del - (beta * gam) / (T)st.alpha);
                }
                st.rho = (double)gam;
                st.alpha = (double)alpha;
                st.stop = (q == it || fabs((double)alpha) < 1e-2) ? 1 : 0;
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) S[q] = st;
    if (st.done) return;
    const T *__restrict__ r = b.r;
    const T *__restrict__ w = b.w;
    const T *__restrict__ s = b.s;
    const T *__restrict__ iD = l.iD;
    T *__restrict__ pd = l.e;  // search direction (COL_LOOP declares p[3], the cell)
    T *__restrict__ x = l.x;
    T *__restrict__ ro = b.ro;
    T *__restrict__ so = b.so;
    const bool first = q == 1;
    // s and r' at any cell, from the previous iterate
    auto sK = [&](long K) { return first ? w[K] : w[K] + beta * s[K]; };
    auto rK = [&](long K) { return r[K] - alpha * sK(K); };
    if (st.stop) {
        // last update: x and r only
        COL_LOOP(l, {
            T u = r[I] * iD[I];
            T pp = first ? u : u + beta * pd[I];
            T ss = sK(I);
            pd[I] = pp;
            so[I] = ss;
            x[I] = x[I] + alpha * pp;
            ro[I] = r[I] - alpha * ss;
        })
        return;
    }
    T *__restrict__ wo = b.wo;
    double g = 0, d = 0;
    if (l.g.D == 3) {
        // u' = r'·iD evaluated once per cell into the plane tile (z-neighbours
        // in registers) instead of at each of its seven stencil uses
        struct Raw {
            T r, w, s, d;
        };
        struct Pre {
            L6<T> l;
            T r, w, s, d, p, x;
        };
        tiled7<T, Raw>(
            l.g, l.t, [&](long K) { return Raw{r[K], w[K], first ? T(0) : s[K], iD[K]}; },
            [&](const Raw &q) { return (q.r - alpha * (first ? q.w : q.w + beta * q.s)) * q.d; },
            [&](long I) { return Pre{load_L6<T>(l.g, I, l.L), r[I], w[I], first ? T(0) : s[I], iD[I], pd[I], x[I]}; },
            [&](const int *, long I, const Pre &pv, T c, T xm, T xp, T ym, T yp, T zm, T zp) {
                T u = pv.r * pv.d;
                T pp = first ? u : u + beta * pv.p;
                T ss = first ? pv.w : pv.w + beta * pv.s;
                T rr = pv.r - alpha * ss;
                T ww = mult7<T>(pv.l, c, xm, xp, ym, yp, zm, zp);
                pd[I] = pp;
                so[I] = ss;
                x[I] = pv.x + alpha * pp;
                ro[I] = rr;
                wo[I] = ww;
                g += (double)rr * (double)c;
                d += (double)c * (double)ww;
            });
        block_partial(g, pout, 2, 0);
        block_partial(d, pout, 2, 1);
        return;
    }
    DISPATCH_D(l.g, {
        col_pipe(l.g, l.t,
                 [&](const int *qq, long I) { return ld7<D>(l.g, qq, I, l.L, [&](long K) { return rK(K) * iD[K]; }); },
                 [&](const int *, long I, const S7<D, T> &sv) {
                     T u = r[I] * iD[I];
                     T pp = first ? u : u + beta * pd[I];
                     T ss = sK(I);
                     T rr = r[I] - alpha * ss;
                     T ww = mul7<D>(sv);
                     pd[I] = pp;
                     so[I] = ss;
                     x[I] = x[I] + alpha * pp;
                     ro[I] = rr;
                     wo[I] = ww;
                     g += (double)rr * (double)sv.c;
                     d += (double)sv.c * (double)ww;
                 });
    })
    block_partial(g, pout, 2, 0);
    block_partial(d, pout, 2, 1);
}

// two block sums in one pass (same per-thread order and tree as two calls of
// block_sum / block_total, so the same values)
// lead: the barrier that protects sh2 from an earlier block_sum2 of the
// same kernel (not needed on its first use)
template <int NB = NT>
__device__ inline void block_sum2(double &a, double &b, bool lead = true) {  // valid in thread 0
    __shared__ double sh2[2][NB / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lead) __syncthreads();
    if (lane == 0) {
        sh2[0][w] = a;
        sh2[1][w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0, sb = 0;
        for (int q = 0; q < NB / 64; ++q) {
            sa += sh2[0][q];
            sb += sh2[1][q];
        }
        a = sa;
        b = sb;
    }
}
__device__ inline void block_total2(const double *part, long n, double &a, double &b) {
    __shared__ double tot2[2];
    double va = 0, vb = 0;
    for (long q = threadIdx.x; q < n; q += NT) {
        va += part[2 * q];
        vb += part[2 * q + 1];
    }
    block_sum2(va, vb);
    if (threadIdx.x == 0) {
        tot2[0] = va;
        tot2[1] = vb;
    }
    __syncthreads();
    a = tot2[0];
    b = tot2[1];
    __syncthreads();
}
// k_cg1_step for tilings of one cell per thread (zb == 1, or 2D): every load
// the thread needs -- the previous partials, and r, w, s, iD at the cell and
// its neighbours -- is issued before the partial sums are reduced, so the
// launch waits for one memory round trip instead of three.  Same arithmetic.
// NB threads per workgroup (l.t tiles rows of nx cells, NB/nx rows per
// workgroup); nin partial pairs come in (k_cg1_init's, then this kernel's).
template <int NB>
__device__ inline bool thread_row(const Grid &g, const CTiling &t, int p[3]) {
    long b = swizzle_block((long)blockIdx.x, t.nblk);
    long bx = b % t.nt[0];
    long r = b / t.nt[0];
    long by = r % t.nt[1];
    long bz = r / t.nt[1];
    int tx = threadIdx.x % t.nx, ty = threadIdx.x / t.nx;
    p[0] = 1 + (int)bx * t.nx + tx;
    p[1] = g.D > 1 ? 1 + (int)by * (NB / t.nx) + ty : 0;
    p[2] = g.D > 2 ? 1 + (int)bz : 0;
    bool ok = p[0] <= g.Ng[0] - 2;
    if (g.D > 1) ok = ok && p[1] <= g.Ng[1] - 2;
    return ok;
}
template <int D, class T, int NB = NT>
__global__ void __launch_bounds__(NB) k_cg1_step1(LevV<T> l, Cg1V<T> b, PcgState *S, int q, int it,
                                                  const double *pin, long nin, double *pout) {
    HALT_IF(l)
    const Grid &g = l.g;
    int p[3];
    bool act = thread_row<NB>(g, l.t, p);
    long I = (long)p[0] + g.st[1] * p[1] + g.st[2] * p[2];
    PcgState st = S[q - 1];
    // the previous launch's partials (strided per thread, as block_total)
    double va = 0, vb = 0;
    for (long k = threadIdx.x; k < nin; k += NB) {
        va += pin[2 * k];
        vb += pin[2 * k + 1];
    }
    // raw values at the cell and its neighbours
    T rc = 0, wc = 0, sc = 0, dc = 0, pc = 0, xc = 0;
    T rm[D], rp[D], wm[D], wp[D], sm[D], sp[D], dm[D], dp[D], lm[D], lp[D];
    if (act) {
        Nb n = neighbours_d<D>(g, p);
        rc = b.r[I];
        wc = b.w[I];
        sc = b.s[I];
        dc = l.iD[I];
        pc = l.e[I];
        xc = l.x[I];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            long Km = I + n.m[i], Kp = I + n.p[i];
            rm[i] = b.r[Km];
            rp[i] = b.r[Kp];
            wm[i] = b.w[Km];
            wp[i] = b.w[Kp];
            sm[i] = b.s[Km];
            sp[i] = b.s[Kp];
            dm[i] = l.iD[Km];
            dp[i] = l.iD[Kp];
            lm[i] = l.L[g.ns * i + I];
            lp[i] = l.L[g.ns * i + I + g.st[i]];
        }
    }
    T alpha = 0, beta = 0;
    bool reduce = !st.done && !st.stop;
    if (reduce) {
        // block_total2, with the per-thread sums already loaded; every
        // thread sums the wavefronts' sums in order (one barrier)
        __shared__ double shs[2][NB / 64];
        va = wave_sum(va);
        vb = wave_sum(vb);
        if ((threadIdx.x & 63) == 0) {
            shs[0][threadIdx.x >> 6] = va;
            shs[1][threadIdx.x >> 6] = vb;
        }
        __syncthreads();
        double sa = 0, sb = 0;
        for (int w = 0; w < NB / 64; ++w) sa += shs[0][w];
        for (int w = 0; w < NB / 64; ++w) sb += shs[1][w];
        T gam = (T)sa;
        T del = (T)sb;
        if (fabs(gam) < 10 * eps_of<T>()) {
            st.done = 1;
        } else {
            if (q == 1) {
                alpha = gam / del;
            } else {
                beta = gam / (T)st.rho;
                alpha = gam / (del - (beta * gam) / (T)st.alpha);
            }
            st.rho = (double)gam;
            st.alpha = (double)alpha;
            st.stop = (q == it || fabs((double)alpha) < 1e-2) ? 1 : 0;
        }
    } else if (!st.done) {
        st.done = 1;  // the previous launch made the last update
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) S[q] = st;
    if (st.done) return;
    const bool first = q == 1;
    double gd = 0, dd = 0;
    if (act) {
        T u = rc * dc;
        T pp = first ? u : u + beta * pc;
        T ss = first ? wc : wc + beta * sc;
        T rr = rc - alpha * ss;
        l.e[I] = pp;
        b.so[I] = ss;
        l.x[I] = xc + alpha * pp;
        b.ro[I] = rr;
        if (!st.stop) {
            // u' = r'·iD at the cell and its neighbours, w' = A u' (mul7's order)
            S7<D, T> sv;
            sv.c = rr * dc;
#pragma unroll
            for (int i = 0; i < D; ++i) {
                T smv = first ? wm[i] : wm[i] + beta * sm[i];
                T spv = first ? wp[i] : wp[i] + beta * sp[i];
                sv.m[i] = (rm[i] - alpha * smv) * dm[i];
                sv.p[i] = (rp[i] - alpha * spv) * dp[i];
                sv.lm[i] = lm[i];
                sv.lp[i] = lp[i];
            }
            T ww = mul7<D>(sv);
            b.wo[I] = ww;
            gd = (double)rr * (double)sv.c;
            dd = (double)sv.c * (double)ww;
        }
    }
    if (st.stop) return;
    block_sum2<NB>(gd, dd, false);  // sh2's first use in this kernel
    if (threadIdx.x == 0) {
        pout[2 * blockIdx.x] = gd;
        pout[2 * blockIdx.x + 1] = dd;
    }
}

// ---------------------------------------------------------------------------
// psolver! (flow.jl:300-326): Jacobi-preconditioned CG on the single-level
// Poisson (the finest level's L, D, iD, x, ϵ, r, z), without pcg!'s early
// exits; the host tests the absolute Σr² after every iteration.  ρ lives in
// one of two device slots (read one, write the other) so that no block reads
// a value another block of the same launch is writing.
// ---------------------------------------------------------------------------
// ρ₀ = (T) Σ r·z from k_pcg_init's partials (one block)
template <class T>
__global__ void __launch_bounds__(NT) k_ps_rho(const double *part, long nblk, double *rho) {
    T v = (T)block_total(part, nblk);
    if (threadIdx.x == 0) *rho = (double)v;
}
// z = Aϵ, partial z·ϵ
template <class T>
__global__ void __launch_bounds__(NT) k_ps_Ae(LevV<T> l, double *partB) {
    double v = 0;
    DISPATCH_D(l.g, pcg_Ae_d<D>(l, v);)
    block_partial(v, partB);
}
// α = ρ/(z·ϵ); x += αϵ; r -= αz; z = r·iD; partials r·z and r·r
template <class T>
__global__ void __launch_bounds__(NT) k_ps_update(LevV<T> l, const double *rho, const double *partB,
                                                  double *partC) {
    T alpha = (T)*rho / (T)block_total(partB, l.t.nblk);
    double vz = 0, vr = 0;
    COL_LOOP(l, {
        l.x[I] = l.x[I] + alpha * l.e[I];
        T rr = l.r[I] - alpha * l.z[I];
        l.r[I] = rr;
        T zz = rr * l.iD[I];
        l.z[I] = zz;
        vz += (double)rr * (double)zz;
        vr += (double)rr * (double)rr;
    })
    block_partial(vz, partC, 2, 0);
    block_partial(vr, partC, 2, 1);
}
// ρ₂ = r·z, β = ρ₂/ρ, ϵ = βϵ + z, ρ ← ρ₂ (into the other slot)
template <class T>
__global__ void __launch_bounds__(NT) k_ps_eps(LevV<T> l, const double *rho_in, double *rho_out,
                                               const double *partC) {
    T rho2 = (T)block_total(partC, l.t.nblk, 2, 0);
    T beta = rho2 / (T)*rho_in;
    if (blockIdx.x == 0 && threadIdx.x == 0) *rho_out = (double)rho2;
    COL_LOOP(l, { l.e[I] = beta * l.e[I] + l.z[I]; })
}

// MPCFL (flow.jl:262-281): block maxima of flux_out and maxTotalFlux; and,
// for the step's statistics, the block's FP64 Σf and its min / max f over
// the interior (5 partials per block).  Each thread walks a z-column of
// cells, so the five block reductions run once per column tile.
template <class T>
__global__ void __launch_bounds__(NT) k_cfl(Grid g, CTiling t, const T *u, const T *f, double *part) {
    double a = 0, b = 0, fs = 0, fmn = 1e300, fmx = -1e300;
    int p[3], zlo, zhi;
    if (thread_col(g, t, p, zlo, zhi))
        for (int z = zlo; z <= zhi; ++z) {
            p[2] = z;
            long I = (long)p[0] + g.st[1] * p[1] + g.st[2] * p[2];
            double acc = 0;
            T s2 = 0;
            for (int d = 0; d < g.D; ++d) {
                double up = (double)u[g.ns * d + I + g.st[d]], lo = -(double)u[g.ns * d + I];
                acc += (up > 0.0 ? up : 0.0) + (lo > 0.0 ? lo : 0.0);
                s2 += tmax(fabs(u[g.ns * d + I]), fabs(u[g.ns * d + I + g.st[d]]));
            }
            double av = (double)(T)acc, bv = (double)s2, fv = (double)f[I];
            a = av > a ? av : a;
            b = bv > b ? bv : b;
            fs += fv;
            fmn = fv < fmn ? fv : fmn;
            fmx = fv > fmx ? fv : fmx;
        }
    double ma = block_max(a);
    __syncthreads();
    double mb = block_max(b);
    __syncthreads();
    double sf = block_sum(fs);
    __syncthreads();
    double mx = block_max(fmx);
    __syncthreads();
    double mn = -block_max(-fmn);
    if (threadIdx.x == 0) {
        double *o = part + 5 * blockIdx.x;
        o[0] = ma;
        o[1] = mb;
        o[2] = sf;
        o[3] = mn;
        o[4] = mx;
    }
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
    if (f.g.D == 3 && f.t.zb > 1) {
        // column tilings: x through an LDS plane tile (mult64's operations)
        const double *__restrict__ x = f.x;
        const double *__restrict__ z = f.z;
        const T *__restrict__ iD = f.iD;
        double *__restrict__ r = f.r;
        struct Pre {
            L6<double> l;
            double z;
            T d;
        };
        tiled7<double, double>(
            f.g, f.t, [&](long K) { return x[K]; }, [&](double q) { return q; },
            [&](long I) { return Pre{load_L6<double>(f.g, I, f.L), z[I], iD[I]}; },
            [&](const int *, long I, const Pre &pv, double c, double xm, double xp, double ym, double yp, double zm,
                double zp) {
                double rr = pv.d == T(0) ? 0.0 : pv.z - mult7<double>(pv.l, c, xm, xp, ym, yp, zm, zp);
                r[I] = rr;
                v += rr;
            });
        block_partial(v, part);
        return;
    }
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
// k_fcg_load fused with the V-cycle's first Jacobi! + increment! on level 0:
// ϵ = (r/s)·iD is formed on the fly at the cell and its neighbours, so
// r_T = r/s - Aϵ and x_T = ϵ are written without storing r/s or ϵ.
template <int D, class T>
__device__ inline void load_inc_cell(const FcgV<T> &f, const LevV<T> &l, double s, const int p[3], long I) {
    const double *__restrict__ r64 = f.r;
    const T *__restrict__ iD = l.iD;
    auto ev = [&](long K) { return (T)(r64[K] / s) * iD[K]; };
    T wr = (T)(r64[I] / s);
    l.r[I] = wr - mgmult_fd<D>(l.g, p, I, l.L, ev);
    l.x[I] = T(0) + ev(I);
}
template <class T>
__global__ void __launch_bounds__(NT) k_fcg_load_inc(FcgV<T> f, LevV<T> l, const double *part_rr) {
    HALT_IF(l)
    double s = sqrt(block_total(part_rr, f.t.nblk) / (double)f.ncell);
    if (blockIdx.x == 0 && threadIdx.x == 0) f.S->sload = s;
    if (l.g.D == 3) {
        const double *__restrict__ r64 = f.r;
        const T *__restrict__ iD = l.iD;
        T *__restrict__ r = l.r;
        T *__restrict__ x = l.x;
        struct Raw {
            double r;
            T d;
        };
        struct Pre {
            L6<T> l;
            double rc;
        };
        tiled7<T, Raw>(
            l.g, l.t, [&](long K) { return Raw{r64[K], iD[K]}; }, [&](const Raw &q) { return (T)(q.r / s) * q.d; },
            [&](long I) { return Pre{load_L6<T>(l.g, I, l.L), r64[I]}; },
            [&](const int *, long I, const Pre &pv, T c, T xm, T xp, T ym, T yp, T zm, T zp) {
                T wr = (T)(pv.rc / s);
                r[I] = wr - mult7<T>(pv.l, c, xm, xp, ym, yp, zm, zp);
                x[I] = T(0) + c;
            });
    } else {
        COL_LOOP(f, { load_inc_cell<2>(f, l, s, p, I); })
    }
}
// the last pcg! iteration on level 0 inside the preconditioner, fused with
// the FCG dots: α = ρ/(z·ϵ); x += αϵ (r and z are dead afterwards); then
// partials of w·r and w·q with w = s·x_T
template <class T>
__global__ void __launch_bounds__(NT) k_pcg_last_dots(LevV<T> l, FcgV<T> f, PcgState *S, int sin, int sout,
                                                      const double *partB, double *part, int first) {
    HALT_IF(l)
    PcgState st = S[sin];
    T alpha = T(0);
    if (!st.done) {
        alpha = (T)st.rho / (T)block_total(partB, l.t.nblk);
        st.alpha = (double)alpha;
        st.stop = 1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) S[sout] = st;
    double s = f.S->sload;
    double a = 0, b = 0;
    bool upd = !st.done;
    {
        T *__restrict__ x = l.x;
        const T *__restrict__ e = l.e;
        const double *__restrict__ r64 = f.r;
        const double *__restrict__ q64 = f.q;
        struct V {
            double r, q;
            T x, e;
        };
        col_pipe(f.g, f.t,
                 [&](const int *, long I) { return V{r64[I], first ? 0.0 : q64[I], x[I], upd ? e[I] : T(0)}; },
                 [&](const int *, long I, const V &w) {
                     T xn = w.x;
                     if (upd) {
                         xn = xn + alpha * w.e;
                         x[I] = xn;
                     }
                     double ww = s * (double)xn;
                     a += ww * w.r;
                     if (!first) b += ww * w.q;
                 });
    }
    block_partial(a, part, 2, 0);
    block_partial(b, part, 2, 1);
}
// coefficients of the fine smoother below: Chebyshev on [λmin, λmax] =
// [0.5, 2] (D⁻¹A has its spectrum in (0, 2]; the smoother damps the upper
// three quarters), θ = (λmax+λmin)/2, δ = (λmax-λmin)/2, σ = θ/δ, ρ₀ = 1/σ,
// ρ₁ = 1/(2σ - ρ₀): c0 = 1/θ, ca = ρ₁ρ₀, cb = 2ρ₁/δ (the oracle's mg_cheb2)
inline void cheb2_coeffs(double &c0, double &ca, double &cb) {
    const double lmin = 0.5, lmax = 2.0, th = (lmax + lmin) / 2, de = (lmax - lmin) / 2, sg = th / de;
    const double ro0 = 1 / sg, ro1 = 1 / (2 * sg - ro0);
    c0 = 1 / th;
    ca = ro1 * ro0;
    cb = 2 * ro1 / de;
}
// the preconditioner's fine-level smoother: two Chebyshev steps on D⁻¹A for
// the eigenvalue interval [0.5, 2] (Saad's three-term recurrence; c0 = 1/θ,
// ca = ρ₁ρ₀, cb = 2ρ₁/δ), fused with the FCG dots.  ϵ = r·iD is on hand from
// the V-cycle's last prolongate!/increment!, so d₀ = c0·ϵ at every stencil
// point; r₁ = r - A d₀; d₁ = ca·d₀ + cb·(r₁·iD); x = (x + d₀) + d₁ (the
// second step's residual update is dead: only x leaves the preconditioner).
// Then the partials of w·r and w·q with w = s·x.
template <class T>
__global__ void __launch_bounds__(NT) k_cheb2_dots(LevV<T> l, FcgV<T> f, T c0, T ca, T cb, double *part, int first) {
    HALT_IF(l)
    double s = f.S->sload;
    double a = 0, b = 0;
    T *__restrict__ x = l.x;
    const T *__restrict__ e = l.e;
    const T *__restrict__ r = l.r;
    const T *__restrict__ iD = l.iD;
    const double *__restrict__ r64 = f.r;
    const double *__restrict__ q64 = f.q;
    if (l.g.D == 3 && l.t.zb > 1) {
        struct Pre {
            L6<T> l;
            T r, x, d;
            double r64, q64;
        };
        tiled7<T, T>(
            l.g, l.t, [&](long K) { return e[K]; }, [&](T q) { return q * c0; },
            [&](long I) { return Pre{load_L6<T>(l.g, I, l.L), r[I], x[I], iD[I], r64[I], first ? 0.0 : q64[I]}; },
            [&](const int *, long I, const Pre &pv, T c, T xm, T xp, T ym, T yp, T zm, T zp) {
                T r1 = pv.r - mult7<T>(pv.l, c, xm, xp, ym, yp, zm, zp);
                T d1 = ca * c + cb * (r1 * pv.d);
                T xn = (pv.x + c) + d1;
                x[I] = xn;
                double ww = s * (double)xn;
                a += ww * pv.r64;
                if (!first) b += ww * pv.q64;
            });
    } else {
        DISPATCH_D(l.g, {
            struct V {
                S7<D, T> s7;
                T r, x, d;
                double r64, q64;
            };
            col_pipe(l.g, l.t,
                     [&](const int *p, long I) {
                         return V{ld7<D>(l.g, p, I, l.L, [&](long K) { return e[K] * c0; }), r[I], x[I], iD[I],
                                  r64[I], first ? 0.0 : q64[I]};
                     },
                     [&](const int *, long I, const V &v) {
                         T r1 = v.r - mul7<D>(v.s7);
                         T d1 = ca * v.s7.c + cb * (r1 * v.d);
                         T xn = (v.x + v.s7.c) + d1;
                         x[I] = xn;
                         double ww = s * (double)xn;
                         a += ww * v.r64;
                         if (!first) b += ww * v.q64;
                     });
        })
    }
    block_partial(a, part, 2, 0);
    block_partial(b, part, 2, 1);
}
// partials of w·r and w·q (w = s·x_T)
template <class T>
__global__ void __launch_bounds__(NT) k_fcg_dots(FcgV<T> f, double *part, int first) {
    if (f.S->halt) return;
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
    if (f.S->halt) return;
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
    if (g.D == 3) {
        const T *__restrict__ w = f.w;
        const double *__restrict__ pa = f.pa;
        double *__restrict__ pb = f.pb;
        double *__restrict__ q = f.q;
        struct Raw {
            double pa;
            T w;
        };
        tiled7<double, Raw>(
            g, f.t, [&](long K) { return Raw{pa[K], w[K]}; },
            [&](const Raw &q) { return s * (double)q.w + beta * q.pa; },
            [&](long I) { return load_L6<double>(g, I, f.L); },
            [&](const int *, long I, const L6<double> &lv, double c, double xm, double xp, double ym, double yp,
                double zm, double zp) {
                double qq = mult7<double>(lv, c, xm, xp, ym, yp, zm, zp);
                pb[I] = c;
                q[I] = qq;
                v += c * qq;
            });
    } else {
        COL_LOOP(f, {
            double pc, qq;
            fcg_apply<2>(f, p, I, s, beta, pc, qq);
            f.pb[I] = pc;
            f.q[I] = qq;
            v += pc * qq;
        })
    }
    block_partial(v, part);
}
// α = ρ/(p·q); x += αp ; r -= αq ; partial ‖r‖²
template <class T>
__global__ void __launch_bounds__(NT) k_fcg_update(FcgV<T> f, const double *part_pq, double *part) {
    if (f.S->halt) return;
    double pq = block_total(part_pq, f.t.nblk);
    double alpha = f.S->wr / pq;
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        f.S->pq = pq;
        f.S->alpha = alpha;
    }
    double v = 0;
    {
        double *__restrict__ x = f.x;
        double *__restrict__ r = f.r;
        const double *__restrict__ pb = f.pb;
        const double *__restrict__ q = f.q;
        struct V {
            double x, r, p, q;
        };
        col_pipe(f.g, f.t, [&](const int *, long I) { return V{x[I], r[I], pb[I], q[I]}; },
                 [&](const int *, long I, const V &w) {
                     x[I] = w.x + alpha * w.p;
                     double rr = w.r - alpha * w.q;
                     r[I] = rr;
                     v += rr * rr;
                 });
    }
    block_partial(v, part);
}
// convergence test of the recursive residual on the device, so the host can
// queue the next iteration before it reads this one's result
template <class T>
__global__ void __launch_bounds__(NT) k_fcg_check(FcgV<T> f, const double *part_rr, double zn, double rtol) {
    if (f.S->halt) return;
    double rr = block_total(part_rr, f.t.nblk);
    if (threadIdx.x == 0) {
        f.S->rr = rr;
        if (sqrt(rr) / zn <= rtol) f.S->halt = 1;
    }
}
__global__ void k_fcg_halt(FcgState *S, int v) {
    if (threadIdx.x == 0) S->halt = v;
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
    if (g.D == 3 && f.t.zb > 1) {
        // column tilings: x through an LDS plane tile (its low neighbours are
        // the periodic images xper reads)
        const double *__restrict__ x = f.x;
        struct Pre {
            T l0, l1, l2, u0, u1, u2;
        };
        tiled7<double, double>(
            g, f.t, [&](long K) { return x[K]; }, [&](double q) { return q; },
            [&](long I) {
                return Pre{f.L[I], f.L[g.ns + I], f.L[2 * g.ns + I], u[I], u[g.ns + I], u[2 * g.ns + I]};
            },
            [&](const int *pc, long I, const Pre &pv, double c, double xm, double, double ym, double, double zm,
                double) {
                int p[3] = {pc[0], pc[1], pc[2]};
                T un0 = (T)((double)pv.u0 - (double)pv.l0 * (c - xm));
                T un1 = (T)((double)pv.u1 - (double)pv.l1 * (c - ym));
                T un2 = (T)((double)pv.u2 - (double)pv.l2 * (c - zm));
                T xv = g.exitx ? u[I + 1] : T(0);  // exitBC: the exit face is left alone
                write_images<T>(g, u, p, un0, 2, 0, ub[0], xv);
                write_images<T>(g, u + g.ns, p, un1, 1, 1, ub[1]);
                write_images<T>(g, u + 2 * g.ns, p, un2, 1, 2, ub[2]);
                pT[I] = (T)(c / dtw);
            });
        return;
    }
    COL_LOOP(f, {
        for (int i = 0; i < g.D; ++i) {
            int q[3] = {p[0], p[1], p[2]};
            q[i] -= 1;
            double xm = xper(g, (const double *)f.x, q[0], q[1], q[2]);
            T un = (T)((double)u[g.ns * i + I] - (double)f.L[g.ns * i + I] * (f.x[I] - xm));
            T xv = i == 0 && g.exitx ? u[I + 1] : T(0);  // exitBC: the exit face is left alone
            write_images<T>(g, u + g.ns * i, p, un, 2, i, ub[i], xv);
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
// pcg! iterations per coarse level: at most CIT_MAX, three state slots each
constexpr int CIT_MAX = 128;
constexpr int PCG_SLOTS = 3 * CIT_MAX + 8;
// the flexible CG preconditioner's solver iterations on coarse level l of n
// interior cells.  Level 1 above 1M cells is bandwidth-bound and costs like
// the fine level, so it keeps WaterLily's 6; deeper levels cost little more
// than their launches per iteration, and solving their problems more exactly
// cuts the outer iterations (256^3 pool: 39 -> 8.5 per solve; 512^3 slab:
// 32 -> 12 with 56 instead of 6 on its second level, 128^3).  Levels of at
// most 512 cells (8^3 and below) keep 32: more there bought nothing.
__host__ __device__ inline int fcg_coarse_its(long n, int l) {
    return (l == 1 && n > 1000000) ? 6 : n > 100000 ? 56 : n > 512 ? 64 : 32;
}
// workgroup size of the single-reduction CG step on a 3D level of n cells
// with one cell per thread: 512 above 100k cells (64³: 512 workgroups, so
// each re-reads half as many partial pairs; 6.3 µs a step against 6.9 with
// 256 threads, 6.5 with 1024), else 256 (32³: 4.4 µs, against 4.7 with 512
// and 4.5 with 128)
inline int cg1_workgroup(long n) { return n > 100000 ? 512 : NT; }
constexpr size_t TAIL_LDS_MAX = 152 * 1024;  // of the 160 KiB of LDS per CU
template <class T> struct TailLev {
    Grid g;
    const T *L, *Dg, *iD;
    T *x, *e, *r, *z;
    int n;  // interior cells
};
template <class T> struct TailArgs {
    TailLev<T> lv[MAXLEV];
    int cit[MAXLEV];   // pcg! iterations per level
    int l0, nlev;
    const int *halt;
};
// workgroup sum (every thread gets it) with one barrier: the per-wave sums
// alternate between two halves of sh, so a wave writing reduction k+1's
// half cannot overtake a wave still reading reduction k's -- it has passed
// reduction k+1's barrier only after every wave finished reading k-1's half.
__device__ inline double bsum1024(double v, double *sh, int &par) {
    v = wave_sum(v);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double *h = sh + par * (NT1 / 64);
    par ^= 1;
    if (lane == 0) h[w] = v;
    __syncthreads();
    double s = 0;
    for (int q = 0; q < NT1 / 64; ++q) s += h[q];
    return s;
}
// ---------------------------------------------------------------------------
// The coarsest levels in one workgroup with every array resident in LDS.
// Each level is stored interior-only (n0 x n1 x n2, x fastest): L (one array
// per axis: the face at the cell's low side), iD, x, ϵ, r, z.  A neighbour
// across a non-periodic wall contributes nothing (its coefficient L is the
// zero wall face) and is read as 0; a periodic axis wraps.  The face above
// the last cell is the wall (0) or, periodic, the image of face 1.  The
// arithmetic is that of the global-memory kernels, term for term.
// ---------------------------------------------------------------------------
template <class T> struct SLev {
    int n[3], D, ncell;
    bool per[3];
    T *L[3], *iD, *x, *e, *r, *z;
};
template <int D, class T>
__device__ inline T smult(const SLev<T> &s, const T *__restrict__ xv, int c, const int q[3]) {
    T xm[D], xp[D], lm[D], lp[D];
    int st = 1;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        {
            int n = s.n[a];
            lm[a] = s.L[a][c];
            if (q[a] > 0) xm[a] = xv[c - st];
            else xm[a] = s.per[a] ? xv[c + (n - 1) * st] : T(0);
            if (q[a] < n - 1) {
                xp[a] = xv[c + st];
                lp[a] = s.L[a][c + st];
            } else if (s.per[a]) {
                xp[a] = xv[c - (n - 1) * st];
                lp[a] = s.L[a][c - (n - 1) * st];
            } else {
                xp[a] = T(0);
                lp[a] = T(0);
            }
            st *= n;
        }
    }
    T dg = 0, sl = 0, su = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) dg -= (lm[a] + lp[a]);
#pragma unroll
    for (int a = 0; a < D; ++a) sl += xm[a] * lm[a];
#pragma unroll
    for (int a = 0; a < D; ++a) su += xp[a] * lp[a];
    return (xv[c] * dg + sl) + su;
}
__device__ inline void scoord(const int n[3], int c, int q[3]) {
    q[0] = c % n[0];
    int r = c / n[0];
    q[1] = r % n[1];
    q[2] = r / n[1];
}
// every cell of a level, each thread the cells threadIdx.x + k*NT1; SLOOP
// then waits for the workgroup, SLOOP_OWN does not (for a loop whose cells'
// values the next phase reads only at the same cells, i.e. by the same thread)
#define SLOOP_OWN(S, ...)                                    \
    for (int c = threadIdx.x; c < (S).ncell; c += NT1) {     \
        int q[3];                                            \
        scoord((S).n, c, q);                                 \
        __VA_ARGS__                                          \
    }
#define SLOOP(S, ...)       \
    SLOOP_OWN(S, __VA_ARGS__) \
    __syncthreads();

template <int D, class T> __device__ void stail_increment(const SLev<T> &l) {
    SLOOP(l, {
        l.r[c] = l.r[c] - smult<D>(l, l.e, c, q);
        l.x[c] = l.x[c] + l.e[c];
    })
}
// pcg!(l; it).  A thread's cells are threadIdx.x + k*NT1 (at most TAIL_KC
// of them), and their x, r, z, ϵ and iD stay in its registers for the whole
// solve: only ϵ, which A·ϵ reads at the neighbours, is also written to LDS.
// Three barriers per iteration: the two sums, and the one after the ϵ update
// (the first sum orders the initial ϵ).  The arithmetic is that of pcg!'s
// global-memory kernels, term for term.  On exit x and r are written back,
// past a barrier, so the caller may read any cell's x.
constexpr int TAIL_KC = 4;  // 4096-cell levels over NT1 threads
template <int D, class T> __device__ void stail_pcg(const SLev<T> &l, int it, double *sh, int &par) {
    int cc[TAIL_KC], q[TAIL_KC][3];
    bool on[TAIL_KC];
    T x[TAIL_KC], r[TAIL_KC], z[TAIL_KC], e[TAIL_KC], iD[TAIL_KC];
    double v = 0;
#pragma unroll
    for (int k = 0; k < TAIL_KC; ++k) {
        cc[k] = threadIdx.x + k * NT1;
        on[k] = cc[k] < l.ncell;
        x[k] = r[k] = z[k] = e[k] = iD[k] = T(0);
        q[k][0] = q[k][1] = q[k][2] = 0;
        if (on[k]) {
            int c = cc[k];
            scoord(l.n, c, q[k]);
            x[k] = l.x[c];
            r[k] = l.r[c];
            iD[k] = l.iD[c];
            T zz = r[k] * iD[k];
            z[k] = zz;
            e[k] = zz;
            l.e[c] = zz;
            v += (double)r[k] * (double)zz;
        }
    }
    auto write_back = [&]() {
#pragma unroll
        for (int k = 0; k < TAIL_KC; ++k)
            if (on[k]) {
                l.x[cc[k]] = x[k];
                l.r[cc[k]] = r[k];
            }
        __syncthreads();
    };
    T rho = (T)bsum1024(v, sh, par);
    if (fabs(rho) < 10 * eps_of<T>()) return;  // x and r unchanged
    for (int i = 1; i <= it; ++i) {
        v = 0;
        T zz[TAIL_KC];
#pragma unroll
        for (int k = 0; k < TAIL_KC; ++k) zz[k] = on[k] ? smult<D>(l, l.e, cc[k], q[k]) : T(0);
#pragma unroll
        for (int k = 0; k < TAIL_KC; ++k)
            if (on[k]) {
                z[k] = zz[k];
                v += (double)zz[k] * (double)e[k];
            }
        T alpha = rho / (T)bsum1024(v, sh, par);
        bool last = i == it || fabs((double)alpha) < 1e-2;
        v = 0;
#pragma unroll
        for (int k = 0; k < TAIL_KC; ++k)
            if (on[k]) {
                x[k] += alpha * e[k];
                T rr = r[k] - alpha * z[k];
                r[k] = rr;
                T z2 = rr * iD[k];
                z[k] = z2;
                v += (double)rr * (double)z2;
            }
        if (last) {
            write_back();
            return;
        }
        T rho2 = (T)bsum1024(v, sh, par);
        if (fabs(rho2) < 10 * eps_of<T>()) {
            write_back();
            return;
        }
        T beta = rho2 / rho;
#pragma unroll
        for (int k = 0; k < TAIL_KC; ++k)
            if (on[k]) {
                e[k] = beta * e[k] + z[k];
                l.e[cc[k]] = e[k];
            }
        __syncthreads();
        rho = rho2;
    }
    write_back();  // it == 0
}
// pcg!(l; it) on a level of at most 64 cells, by wavefront 0 alone, one
// cell per lane.  Its sums are the wavefront's butterfly sums plus 0.0: in
// stail_pcg the other wavefronts' partials would all be 0, so the values are
// the same.  The lanes exchange ϵ through LDS with no workgroup barrier; a
// wavefront's LDS accesses complete in program order, and the fences keep
// the compiler from moving them across one another.
__device__ inline void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <int D, class T> __device__ void stail_pcg_wave(const SLev<T> &l, int it) {
    const int c = threadIdx.x;
    const bool on = c < l.ncell;
    int q[3] = {0, 0, 0};
    T x = 0, r = 0, z = 0, e = 0, iD = 0;
    double v = 0;
    if (on) {
        scoord(l.n, c, q);
        x = l.x[c];
        r = l.r[c];
        iD = l.iD[c];
        T zz = r * iD;
        z = zz;
        e = zz;
        l.e[c] = zz;
        v += (double)r * (double)zz;
    }
    T rho = (T)(0.0 + wave_sum(v));
    if (fabs(rho) < 10 * eps_of<T>()) return;  // x and r unchanged
    wave_lds_order();
    for (int i = 1; i <= it; ++i) {
        v = 0;
        if (on) {
            T zz = smult<D>(l, l.e, c, q);
            z = zz;
            v += (double)zz * (double)e;
        }
        T alpha = rho / (T)(0.0 + wave_sum(v));
        bool last = i == it || fabs((double)alpha) < 1e-2;
        v = 0;
        if (on) {
            x += alpha * e;
            T rr = r - alpha * z;
            r = rr;
            T z2 = rr * iD;
            z = z2;
            v += (double)rr * (double)z2;
        }
        if (last) break;
        T rho2 = (T)(0.0 + wave_sum(v));
        if (fabs(rho2) < 10 * eps_of<T>()) break;
        T beta = rho2 / rho;
        wave_lds_order();
        if (on) {
            e = beta * e + z;
            l.e[c] = e;
        }
        wave_lds_order();
        rho = rho2;
    }
    if (on) {
        l.x[c] = x;
        l.r[c] = r;
    }
}
// pcg! on a tail level; returns past a workgroup barrier
template <int D, class T> __device__ void stail_solve(const SLev<T> &l, int it, double *sh, int &par) {
    if (l.ncell <= 64) {
        if (threadIdx.x < 64) stail_pcg_wave<D>(l, it);
        __syncthreads();
    } else {
        stail_pcg<D>(l, it, sh, par);
    }
}
// LDS bytes the levels l0..nlev-1 need
template <class T> inline size_t tail_lds_bytes(const Grid *gs, int l0, int nlev) {
    size_t b = 0;
    for (int l = l0; l < nlev; ++l) {
        size_t n = 1;
        for (int a = 0; a < gs[l].D; ++a) n *= (size_t)(gs[l].Ng[a] - 2);
        b += n * (gs[l].D + 5) * sizeof(T);
    }
    return b;
}
// the LDS view of tail level l (levels a.l0 .. l-1 are stored before it)
template <class T> __device__ inline SLev<T> slev(const TailArgs<T> &a, int l, T *smem) {
    T *base = smem;
    for (int m = a.l0; m < l; ++m) base += (long)a.lv[m].n * (a.lv[m].g.D + 5);
    const TailLev<T> &G = a.lv[l];
    SLev<T> s;
    s.D = G.g.D;
    s.ncell = G.n;
    for (int d = 0; d < 3; ++d) {
        s.n[d] = d < s.D ? G.g.Ng[d] - 2 : 1;
        s.per[d] = d < s.D && G.g.per[d];
        s.L[d] = base + (long)d * s.ncell;
    }
    base += (long)s.D * s.ncell;
    s.iD = base;
    s.x = base + s.ncell;
    s.e = base + 2L * s.ncell;
    s.r = base + 3L * s.ncell;
    s.z = base + 4L * s.ncell;
    return s;
}
template <int D, class T> __device__ void tail_run(const TailArgs<T> &a, T *smem, double *sh) {
    int par = 0;
    for (int l = a.l0; l < a.nlev; ++l) {
        const TailLev<T> &G = a.lv[l];
        SLev<T> s = slev(a, l, smem);
        // constants of the level, and the top level's restricted residual
        SLOOP(s, {
            long I = (long)(q[0] + 1) + G.g.st[1] * (s.D > 1 ? q[1] + 1 : 0) + G.g.st[2] * (s.D > 2 ? q[2] + 1 : 0);
            s.L[0][c] = G.L[I];
            s.L[1][c] = G.L[G.g.ns + I];
            if (D > 2) s.L[2][c] = G.L[2 * G.g.ns + I];
            s.iD[c] = G.iD[I];
            if (l == a.l0) {
                s.r[c] = G.r[I];
                s.x[c] = T(0);
            }
        })
    }
    int last = a.nlev - 1;
    // descend: Jacobi!(l), restrict!(l -> l+1), coarse x = 0
    for (int l = a.l0; l < last; ++l) {
        SLev<T> f = slev(a, l, smem), cs = slev(a, l + 1, smem);
        SLOOP(f, { f.e[c] = f.r[c] * f.iD[c]; })
        stail_increment<D>(f);
        // each thread then reads, in the next Jacobi! or pcg!, only the
        // coarse cells it restricts here
        SLOOP_OWN(cs, {
            T s = 0;
            for (int c2 = (cs.D > 2 ? 2 * q[2] : 0); c2 <= (cs.D > 2 ? 2 * q[2] + 1 : 0); ++c2)
                for (int c1 = 2 * q[1]; c1 <= 2 * q[1] + 1; ++c1)
                    for (int c0 = 2 * q[0]; c0 <= 2 * q[0] + 1; ++c0) s += f.r[c0 + f.n[0] * (c1 + f.n[1] * c2)];
            cs.r[c] = s;
            cs.x[c] = T(0);
        })
    }
    // ascend: pcg!(l+1), prolongate!, increment!(l)
    for (int l = last - 1; l >= a.l0; --l) {
        SLev<T> f = slev(a, l, smem), cs = slev(a, l + 1, smem);
        stail_solve<D>(cs, a.cit[l + 1], sh, par);
        SLOOP(f, {
            int k0 = q[0] / 2, k1 = f.D > 1 ? q[1] / 2 : 0, k2 = f.D > 2 ? q[2] / 2 : 0;
            f.e[c] = cs.x[k0 + cs.n[0] * (k1 + cs.n[1] * k2)];
        })
        stail_increment<D>(f);
    }
    SLev<T> s = slev(a, a.l0, smem);
    stail_solve<D>(s, a.cit[a.l0], sh, par);
    // the top level's solution back to global memory (interior)
    const TailLev<T> &G = a.lv[a.l0];
    SLOOP(s, {
        long I = (long)(q[0] + 1) + G.g.st[1] * (s.D > 1 ? q[1] + 1 : 0) + G.g.st[2] * (s.D > 2 ? q[2] + 1 : 0);
        G.x[I] = s.x[c];
    })
}
template <class T> __global__ void __launch_bounds__(NT1) k_mg_tail(TailArgs<T> a) {
    HALT_IF(a)
    __shared__ double sh[2 * (NT1 / 64)];
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *smem = (T *)smem_raw;
    if (a.lv[a.l0].g.D == 3) tail_run<3>(a, smem, sh);
    else tail_run<2>(a, smem, sh);
}
}  // namespace ia
