// ia_kernels.h -- kernels of the two-phase step (templated on the field type).
#pragma once
#include "ia_device.h"

namespace ia {

constexpr int NT = TX * TY;  // threads per block (256 = 4 wave64)

// ---------------------------------------------------------------------------
// block reductions (deterministic: fixed shuffle tree, fixed LDS order)
// ---------------------------------------------------------------------------
// a double moved across lanes by a DPP pattern (both halves; every source
// lane is inside the row, so the bound control never applies)
template <int CTRL> __device__ inline double dpp_d(double v) {
    int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, true);
    int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
__device__ inline double lane_d(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}
// sum over the 64 lanes of a fully active wavefront, the same value in every
// lane: pairwise within each 16-lane row (lane pairs l^1, l^2 by quad
// permutes, then the half-row and row mirrors), each pair added in the same
// two operands by both of its lanes; then the four row sums in row order.
// DPP and readlane instead of LDS permutes keep it off the LDS pipeline.
__device__ inline double wave_sum(double v) {
    v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_d<0x141>(v);  // row_half_mirror
    v += dpp_d<0x140>(v);  // row_mirror
    return ((lane_d(v, 0) + lane_d(v, 16)) + lane_d(v, 32)) + lane_d(v, 48);
}
__device__ inline double block_sum(double v) {
    __shared__ double sh[NT / 64];
    v = wave_sum(v);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int q = 0; q < NT / 64; ++q) s += sh[q];
    return s;  // valid in thread 0
}
__device__ inline double block_max(double v) {
    __shared__ double sh[NT / 64];
    for (int o = 32; o > 0; o >>= 1) {
        double w2 = __shfl_xor(v, o, 64);
        v = w2 > v ? w2 : v;
    }
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = sh[0];
    if (threadIdx.x == 0)
        for (int q = 1; q < NT / 64; ++q) s = sh[q] > s ? sh[q] : s;
    return s;
}

// argmax/argmin partials of a VOF sweep (findmax/findmin, advection.jl:146-147)
struct FillPart {
    double vmax, vmin;
    long imax, imin;
    int nan, pad;
};
__device__ inline void fill_combine(FillPart &a, const FillPart &b) {
    if (b.vmax > a.vmax || (b.vmax == a.vmax && b.imax < a.imax)) { a.vmax = b.vmax; a.imax = b.imax; }
    if (b.vmin < a.vmin || (b.vmin == a.vmin && b.imin < a.imin)) { a.vmin = b.vmin; a.imin = b.imin; }
    a.nan |= b.nan;
}
// a wavefront's max / min of a double and the smallest index holding it, by
// the DPP pattern of wave_sum (order-independent, so equal to block_fill's)
template <int CTRL> __device__ inline long dpp_l(long v) {
    int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(v & 0xffffffffL), CTRL, 0xF, 0xF, true);
    int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), CTRL, 0xF, 0xF, true);
    return (long)(((unsigned long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ inline long lane_l(long v, int l) {
    unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v & 0xffffffffL), l);
    unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(v >> 32), l);
    return (long)(((unsigned long)hi << 32) | lo);
}
__device__ inline double wave_max_d(double v) {
    v = fmax(v, dpp_d<0xB1>(v));
    v = fmax(v, dpp_d<0x4E>(v));
    v = fmax(v, dpp_d<0x141>(v));
    v = fmax(v, dpp_d<0x140>(v));
    return fmax(fmax(fmax(lane_d(v, 0), lane_d(v, 16)), lane_d(v, 32)), lane_d(v, 48));
}
__device__ inline double wave_min_d(double v) {
    v = fmin(v, dpp_d<0xB1>(v));
    v = fmin(v, dpp_d<0x4E>(v));
    v = fmin(v, dpp_d<0x141>(v));
    v = fmin(v, dpp_d<0x140>(v));
    return fmin(fmin(fmin(lane_d(v, 0), lane_d(v, 16)), lane_d(v, 32)), lane_d(v, 48));
}
__device__ inline long wave_min_l(long v) {
    long t;
    t = dpp_l<0xB1>(v); v = t < v ? t : v;
    t = dpp_l<0x4E>(v); v = t < v ? t : v;
    t = dpp_l<0x141>(v); v = t < v ? t : v;
    t = dpp_l<0x140>(v); v = t < v ? t : v;
    long a = lane_l(v, 0), b = lane_l(v, 16), c = lane_l(v, 32), d = lane_l(v, 48);
    a = b < a ? b : a;
    a = c < a ? c : a;
    return d < a ? d : a;
}
// block_fill of one value per thread (ok: the thread has a cell and its value
// is not NaN; nan: its value is NaN), equal to building each thread's
// FillPart and calling block_fill: the wavefront reduces by DPP, then thread
// 0 combines the wavefronts' parts.  Valid in thread 0.
__device__ inline FillPart block_fill_cells(double v, long I, bool ok, bool nan) {
    __shared__ FillPart shw[NT / 64];
    double wmx = wave_max_d(ok ? v : -1e300), wmn = wave_min_d(ok ? v : 1e300);
    long imx = wave_min_l(ok && v == wmx ? I : (1L << 60));
    long imn = wave_min_l(ok && v == wmn ? I : (1L << 60));
    unsigned long long anynan = __ballot(nan);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        FillPart p;
        p.vmax = wmx;
        p.vmin = wmn;
        p.imax = imx;
        p.imin = imn;
        p.nan = anynan != 0;
        p.pad = 0;
        shw[w] = p;
    }
    __syncthreads();
    FillPart r = shw[0];
    if (threadIdx.x == 0)
        for (int q = 1; q < NT / 64; ++q) fill_combine(r, shw[q]);
    return r;
}
__device__ inline FillPart block_fill(FillPart v) {
    __shared__ FillPart sh[NT];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) fill_combine(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    return sh[0];
}

// device-resident state of a step (flags) and of a solve (pcg scalars)
struct StepState {
    int status, nwarn;
    double maxf, minf;
    long imax, imin;
};
struct SolveCtl {
    double rho, alpha, beta, s, r2, dot;
    int done, stop, it, pad;
};

// ---------------------------------------------------------------------------
// c̄ = f≥½ (advection.jl:40, flow.jl:172) and ρu = u·ρ(f) (VOFutil.jl:208-211)
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(NT) k_prepare(Grid g, Tiling t, const T *__restrict__ f, const T *__restrict__ usrc,
                                                 T *__restrict__ ru, int8_t *__restrict__ cb, T lr, int do_ru) {
    int p[3];
    if (!thread_cell(g, t, p)) return;
    long I = lin(g, p);
    int8_t c = f[I] < T(0.5) ? 0 : 1;
    write_images<int8_t>(g, cb, p, c, 0, 0, 0);
    if (do_ru)
        for (int d = 0; d < g.D; ++d) {
            long o = g.ns * d;
            ru[o + I] = usrc[o + I] * linp((f[I] + f[I - g.st[d]]) / 2, lr);
        }
}

// ---------------------------------------------------------------------------
// one cVOF sweep along d in two passes: the face pass reconstructs each
// donor cell on the fly and computes every swept-volume face flux once; the
// cell pass makes the conservative update with the c̄ dilation term, the fill
// report, cleanWisp! and BCf! (advection.jl:80-137, normalEstimation.jl:10-28)
// ---------------------------------------------------------------------------
// the coordinates of flat index K of the ghost-inclusive grid; 32-bit
// division whenever the grid has fewer than 2^31 entries (64-bit integer
// division costs several times more)
__device__ inline void ghost_coord(const Grid &g, long K, int q[3]) {
    if (g.ns < (1L << 31)) {
        unsigned k = (unsigned)K, n0 = (unsigned)g.Ng[0], n1 = (unsigned)g.Ng[1];
        unsigned r = k / n0;
        q[0] = (int)(k - r * n0);
        q[1] = (int)(r % n1);
        q[2] = (int)(r / n1);
    } else {
        q[0] = (int)(K % g.Ng[0]);
        long r = K / g.Ng[0];
        q[1] = (int)(r % g.Ng[1]);
        q[2] = (int)(r / g.Ng[1]);
    }
}
template <class T> struct SweepArgs {
    Grid g;
    Tiling t;
    const T *fin;
    T *fout;
    const T *u1, *u2;    // face velocities, component d
    const int8_t *cb;
    T *ruf;              // mass flux, component d
    const T *fn;         // f^n for the mid-point f (predictor) or null
    T *fmid;
    FillPart *part;
    int d, scheme, mode; // mode 0: MPFMomStep! (ρuf·inv(δt) + BC!); 1: advect! (ρuf as computed)
    T dt, lr, idt, rtol, ctol, ubc_d;
};

// fup, fdn: f at the face's upper and lower cells, loaded with the face
// velocities (the donor is known only once δl's sign is)
template <class T>
__device__ inline T vof_face_flux(const SweepArgs<T> &a, const int q[3], long IF, T dl, T fup, T fdn) {
    const Grid &g = a.g;
    int d = a.d;
    int c[3] = {q[0], q[1], q[2]};
    if (dl > 0) c[d] -= 1;
    long IC = IF - (dl > 0 ? g.st[d] : 0);
    T fc = dl > 0 ? fdn : fup;
    bool ghost = c[d] == 0 || c[d] == g.Ng[d] - 1;
    if (fullorempty(fc) || (ghost && !g.per[d])) return fc * dl;
    long ICr = IC;
    if (ghost) {  // BCVOF! copies n̂, α of the periodic image (VOFutil.jl:44-62)
        int cc[3] = {c[0], c[1], c[2]};
        cc[d] = map_scalar(g, d, c[d]);
        ICr = lin(g, cc);
    }
    T n[3];
    if (g.D == 3 && a.scheme == 0) normal_wh_block3(g, a.fin, ICr, n);
    else normal_scheme(a.scheme, g, a.fin, ICr, n);
    T al = getIntercept(g.D, n, fc);
    T aa = dl > 0 ? al - n[d] * (1 - dl) : al;
    T nc[3];
    for (int ii = 0; ii < g.D; ++ii) nc[ii] = n[ii] * (ii == d ? fabs(dl) : T(1));
    return getVolumeFraction(g.D, nc, aa) * dl;
}

// pass 1 of a sweep: the swept-volume flux of every face along d, once (each
// cell's lower face; the last cell also its upper face), into `face`, and the
// mass flux ρuf of that face (advection.jl:108-137, flow.jl:207)
// the flux through the lower d-face of cell p (flat index I)
template <class T> __device__ inline void vof_face_one(const SweepArgs<T> &a, T *__restrict__ face, const int p[3],
                                                       long I) {
    const Grid &g = a.g;
    int d = a.d;
    T fup = a.fin[I], fdn = a.fin[I - g.st[d]];
    T dl = (a.dt / 2) * (a.u1[I] + a.u2[I]);
    T ff = dl == T(0) ? T(0) : vof_face_flux(a, p, I, dl, fup, fdn);
    face[I] = ff;
    T m = dl == T(0) ? T(0) : T(0) + (dl * a.lr + (1 - a.lr) * ff);
    if (a.mode == 0)
        m = !g.per[d] && (p[d] == 1 || (p[d] == g.Ng[d] - 1 && !(g.exitx && d == 0))) ? a.ubc_d : m * a.idt;
    a.ruf[I] = m;
}
template <class T>
__global__ void __launch_bounds__(NT) k_vof_face(SweepArgs<T> a, T *__restrict__ face) {
    const Grid &g = a.g;
    int p[3];
    if (a.d == 0 && (long)blockIdx.x >= a.t.nblk) {
        // the blocks past the tiling: the upper x-face of each x-row
        if (!top_face_x(g, (long)(blockIdx.x - a.t.nblk) * NT + threadIdx.x, p)) return;
        vof_face_one(a, face, p, lin(g, p));
        return;
    }
    if (!thread_cell(g, a.t, p)) return;
    int d = a.d;
    long I = lin(g, p);
    vof_face_one(a, face, p, I);
    if (d != 0 && p[d] == g.Ng[d] - 2) {
        int pu[3] = {p[0], p[1], p[2]};
        pu[d] += 1;
        vof_face_one(a, face, pu, I + g.st[d]);
    }
}
// pass 2: the conservative update with the c̄ dilation term (advection.jl:83),
// fill-report partials, cleanWisp! (VOFutil.jl:127-137) and BCf!
template <class T>
__global__ void __launch_bounds__(NT) k_vof_update(SweepArgs<T> a, const T *__restrict__ face) {
    const Grid &g = a.g;
    int p[3];
    bool act = thread_cell(g, a.t, p);
    double fv = 0;
    long fi = 0;
    bool fok = false, fnan = false;
    if (act) {
        int d = a.d;
        long I = lin(g, p), sd = g.st[d];
        const T *u1 = a.u1, *u2 = a.u2;
        T ff0 = face[I], ff1 = face[I + sd];
        T dud = u1[I + sd] - u1[I], du0d = u2[I + sd] - u2[I];
        T fnew = a.fin[I] + ((ff0 - ff1) + (((T)a.cb[I] * (dud + du0d)) * a.dt) / 2);
        fnan = fnew != fnew;
        fok = !fnan;
        fv = (double)fnew;
        fi = I;
        T fc = fnew < a.ctol ? T(0) : (fnew > 1 - a.ctol ? T(1) : fnew);
        write_images<T>(g, a.fout, p, fc, 0, 0, T(0));
        if (a.fmid) write_images<T>(g, a.fmid, p, (fc + a.fn[I]) / 2, 0, 0, T(0));
    }
    FillPart r = block_fill_cells(fv, fi, fok, fnan);
    if (threadIdx.x == 0) a.part[blockIdx.x] = r;
}

__device__ inline FillPart fill_identity() {
    FillPart acc;
    acc.vmax = -1e300;
    acc.vmin = 1e300;
    acc.imax = acc.imin = 1L << 60;
    acc.nan = 0;
    acc.pad = 0;
    return acc;
}
// first stage of the fill reduction: FILL_R2 blocks fold the per-block
// partials of a sweep (fill_combine is order-independent: ties go to the
// smallest index)
constexpr int FILL_R2 = 256;
__global__ void __launch_bounds__(NT) k_fill_reduce(const FillPart *part, long nblk, FillPart *out) {
    FillPart acc = fill_identity();
    for (long b = (long)blockIdx.x * NT + threadIdx.x; b < nblk; b += (long)gridDim.x * NT) fill_combine(acc, part[b]);
    FillPart r = block_fill(acc);
    if (threadIdx.x == 0) out[blockIdx.x] = r;
}
// reportFillError (advection.jl:145-189) on the reduced partials
template <class T>
__global__ void __launch_bounds__(NT) k_fill_finalize(Grid g, const FillPart *part, long nblk, const T *u1,
                                                      const T *u2, double tol, StepState *st) {
    FillPart acc = fill_identity();
    for (long b = threadIdx.x; b < nblk; b += NT) fill_combine(acc, part[b]);
    FillPart r = block_fill(acc);
    if (threadIdx.x != 0) return;
    int s = 0;
    if (r.nan) s = 1;
    else {
        auto divabs = [&](const T *u, long I) {
            T sdiv = 0;
            for (int i = 0; i < g.D; ++i) sdiv += u[g.ns * i + I + g.st[i]] - u[g.ns * i + I];
            return fabs(sdiv);
        };
        if (r.vmax - 1 > tol) {
            T du0 = divabs(u1, r.imax), du = divabs(u2, r.imax);
            s = (du0 + du > T(10) || (du0 + du) != (du0 + du)) ? 2 : 3;
        }
        if (s != 2 && -r.vmin > tol) {
            T du0 = divabs(u1, r.imin), du = divabs(u2, r.imin);
            int s2 = (du0 + du > T(10) || (du0 + du) != (du0 + du)) ? 2 : 3;
            s = s2 == 2 ? 2 : (s ? s : s2);
        }
    }
    if (s == 3) st->nwarn += 1;
    int old = st->status;
    int worst = (old == 1 || s == 1) ? 1 : ((old == 2 || s == 2) ? 2 : (old > s ? old : s));
    if (worst != old && (s == 1 || s == 2 || old == 0)) {
        st->maxf = r.vmax;
        st->minf = r.vmin;
        st->imax = r.imax;
        st->imin = r.imin;
    }
    st->status = worst;
}

// ---------------------------------------------------------------------------
// CMOM momentum sweep along d with the SynDRoM blend (flow.jl:20-57, :212-241)
// ---------------------------------------------------------------------------
template <class T> struct MomArgs {
    Grid g;
    Tiling t;
    const T *ru_in;
    T *ru_out;
    const T *fpre;       // f before this sweep's VOF update (Φ), with ghosts
    const T *ruf;        // mass flux ×inv(δt), component d (BC! applied on read)
    const int8_t *cb;
    const T *u1, *u2;    // velocity component d (dilation)
    const T *uold;       // uOld (vector)
    T *us;               // uStar (vector, every cell incl. ghosts): the reference's n̂ alias
    int d, lim;
    T dt, lr;
    T ubc[3];
};

template <class T> __device__ inline T mom_ustar(const MomArgs<T> &a, int i, const int q[3]) {
    const Grid &g = a.g;
    // exitBC: BC! leaves uStar's exit face at what the reference's scratch r
    // (Flow.f) holds there: ρu2u! does not reach it and the forcing left 0
    // (g_x = 0) or the previous sweep's fill!(r, 0) (flow.jl:197, :215)
    if (exit_face(g, i, q)) return T(0);
    if (vel_dirichlet(g, i, q)) return a.ubc[i];
    long K = lin_vel(g, i, q);
    return a.ru_in[g.ns * i + K] / linp((a.fpre[K] + a.fpre[K - g.st[i]]) / 2, a.lr);
}
template <class T> __device__ inline T mom_drho(const MomArgs<T> &a, int i, const int q[3]) {
    const Grid &g = a.g;
    int qq[3];
    if (!face_map(g, i, q, qq)) return T(1);  // dρ = ones(...) is never written there (cVOF.jl:198)
    long K = lin(g, qq);
    return (a.fpre[K] + a.fpre[K - g.st[i]]) / 2;
}
// uStar at every cell of the ghost-inclusive grid (flow.jl:196-206 with the
// BC! images): computed once per sweep, read by every face that needs it
template <class T>
__global__ void __launch_bounds__(NT) k_ustar(MomArgs<T> a) {
    const Grid &g = a.g;
    long K = (long)blockIdx.x * NT + threadIdx.x;
    if (K >= g.ns) return;
    int q[3];
    ghost_coord(g, K, q);
    for (int i = 0; i < g.D; ++i) a.us[g.ns * i + K] = mom_ustar(a, i, q);
}
// the sweep with the axis d and the dimension as compile-time constants
template <int DIR, class T> __device__ inline T mom_ruf(const MomArgs<T> &a, const int q[3]) {
    const Grid &g = a.g;
    // exitBC: ρuf keeps the computed mass flux of the exit face (flow.jl:207)
    if (DIR == 0 && exit_face(g, 0, q))
        return a.ruf[(long)q[0] + g.st[1] * map_scalar(g, 1, q[1]) + g.st[2] * map_scalar(g, 2, q[2])];
    if (!g.per[DIR] && (q[DIR] == 0 || q[DIR] == 1 || q[DIR] == g.Ng[DIR] - 1)) return a.ubc[DIR];
    return a.ruf[lin_scalar(g, q[0], q[1], q[2])];
}
template <int DIR, class T> __device__ inline T mom_dila(const MomArgs<T> &a, const int q[3]) {
    const Grid &g = a.g;
    long K = lin_scalar(g, q[0], q[1], q[2]);
    long s = g.st[DIR];
    return (linp((T)a.cb[K], a.lr) * ((a.u1[K + s] - a.u1[K]) + (a.u2[K + s] - a.u2[K]))) / 2;
}
template <int DIR, class T>
__device__ inline T mom_phiq(const MomArgs<T> &a, int i, const int q[3], T Psi, T uu, T cc, T dd) {
    int c[3] = {q[0], q[1], q[2]};
    if (Psi > 0) c[DIR] -= 1;
    T vI = cc, vd = limiter(a.lim, uu, cc, dd);
    T va = 2 * vI - vd;
    T mOut = fabs(Psi) * a.dt;
    T mOld = linp(mom_drho(a, i, c), a.lr);
    if (mOut > mOld) return Psi * vI;
    T l2 = fabs(mOut) / mOld, l1 = 1 - l2;
    T vb = l2 * va + l1 * vd;
    return (Psi * (vb + vd)) / 2;
}
template <int DIR, class T> __device__ inline T mom_face(const MomArgs<T> &a, int i, const int qin[3]) {
    const Grid &g = a.g;
    int n = g.Ng[DIR];
    int q[3] = {qin[0], qin[1], qin[2]};
    if (g.per[DIR] && q[DIR] == n - 1) q[DIR] = 1;  // flow.jl:241 reuses the lower-boundary flux
    int qm[3] = {q[0], q[1], q[2]};
    qm[i] -= 1;
    T Psi = (mom_ruf<DIR>(a, q) + mom_ruf<DIR>(a, qm)) / 2;
    const T *__restrict__ us = a.us + g.ns * i;
    long Q = lin(g, q), sd = g.st[DIR];
    auto u = [&](int off) { return us[Q + off * sd]; };
    // the limiter's (upwind, centre, downwind) triple for the sign of Psi,
    // selected first so that a wave whose lanes differ in sign evaluates the
    // face value once
    T uu, cc, dd;
    if (q[DIR] == 1 && !g.per[DIR]) {  // ϕuL
        T c0 = u(0), cm = u(-1);
        bool pos = Psi > 0;
        uu = pos ? 2 * cm - c0 : u(1);
        cc = pos ? cm : c0;
        dd = pos ? c0 : cm;
    } else if (q[DIR] == 1) {  // ϕuP: the far-upwind value is the periodic image of I-2δ
        bool pos = Psi > 0;
        uu = pos ? u(n - 4) : u(1);
        cc = pos ? u(-1) : u(0);
        dd = pos ? u(0) : u(-1);
    } else if (q[DIR] == n - 1) {  // ϕuR
        T c0 = u(0), cm = u(-1);
        bool neg = Psi < 0;
        uu = neg ? 2 * c0 - cm : u(-2);
        cc = neg ? c0 : cm;
        dd = neg ? cm : c0;
    } else {
        bool pos = Psi > 0;
        uu = pos ? u(-2) : u(1);
        cc = pos ? u(-1) : u(0);
        dd = pos ? u(0) : u(-1);
    }
    return mom_phiq<DIR>(a, i, q, Psi, uu, cc, dd);
}
// the momentum sweep along DIR in two passes.  Face pass: the SynDRoM face
// flux of every component at every face along DIR, once (each cell's lower
// face; the last cell also its upper face), into `face` (components last)
template <int D, int DIR, class T>
__global__ void __launch_bounds__(NT) k_mom_face(MomArgs<T> a, T *__restrict__ face) {
    const Grid &g = a.g;
    int p[3];
    if (DIR == 0 && (long)blockIdx.x >= a.t.nblk) {
        // the blocks past the tiling: the upper x-face of each x-row
        if (!top_face_x(g, (long)(blockIdx.x - a.t.nblk) * NT + threadIdx.x, p)) return;
        long It = lin(g, p);
#pragma unroll
        for (int i = 0; i < D; ++i) face[g.ns * i + It] = mom_face<DIR>(a, i, p);
        return;
    }
    if (!thread_cell(g, a.t, p)) return;
    long I = lin(g, p), sd = g.st[DIR];
    bool top = DIR != 0 && p[DIR] == g.Ng[DIR] - 2;
    int pu[3] = {p[0], p[1], p[2]};
    pu[DIR] += 1;
    // every component's face value before any store, so that the loads of
    // the components overlap (a store could alias them)
    T fv[D];
#pragma unroll
    for (int i = 0; i < D; ++i) fv[i] = mom_face<DIR>(a, i, p);
#pragma unroll
    for (int i = 0; i < D; ++i) face[g.ns * i + I] = fv[i];
    if (top) {
#pragma unroll
        for (int i = 0; i < D; ++i) fv[i] = mom_face<DIR>(a, i, pu);
#pragma unroll
        for (int i = 0; i < D; ++i) face[g.ns * i + I + sd] = fv[i];
    }
}
// cell pass: the two-pass difference of the face fluxes, the uOld dilation
// term and ρu += δt·r (flow.jl:216-231)
template <int D, int DIR, class T>
__global__ void __launch_bounds__(NT) k_mom_update(MomArgs<T> a, const T *__restrict__ face) {
    const Grid &g = a.g;
    int p[3];
    if (!thread_cell(g, a.t, p)) return;
    long I = lin(g, p), sd = g.st[DIR];
    T dl0 = mom_dila<DIR>(a, p);
    T rn[D];  // every component before any store (a store could alias the loads)
#pragma unroll
    for (int i = 0; i < D; ++i) {
        T flo = face[g.ns * i + I], fup = face[g.ns * i + I + sd];
        int pm[3] = {p[0], p[1], p[2]};
        pm[i] -= 1;
        T r = (T(0) + flo) - fup;
        r += a.uold[g.ns * i + I] * ((dl0 + mom_dila<DIR>(a, pm)) / 2);
        long o = g.ns * i + I;
        rn[i] = a.ru_in[o] + a.dt * r;
    }
#pragma unroll
    for (int i = 0; i < D; ++i) a.ru_out[g.ns * i + I] = rn[i];
}
// ---------------------------------------------------------------------------
// viscous stress (flow.jl:113-153) fused with the RK update (flow.jl:244-252)
// and the Poisson coefficients (flow.jl:254-259)
// ---------------------------------------------------------------------------
template <class T> struct ViscArgs {
    Grid g;
    Tiling t;
    const T *u;       // velocity the stress is taken from (ghosts valid)
    const T *fv;      // f for the face fractions
    const T *ff;      // the face fractions fFace (every cell incl. ghosts): the reference's n̂ alias
    T lm, mu, lr;
    T ubc[3];
    int has_mu;
};
// face fraction fFace[q, c] as left by f2face!/BCv! into n̂ (VOFutil.jl:229-234)
template <class T> __device__ inline T visc_ff(const ViscArgs<T> &a, int c, const int q[3]) {
    const Grid &g = a.g;
    int qq[3];
    // n̂ there still holds uStar (flow.jl:206): u_BC, or 0 on the exit face under exitBC
    if (!face_map(g, c, q, qq)) return exit_face(g, c, q) ? T(0) : a.ubc[c];
    long K = lin(g, qq);
    return (a.fv[K] + a.fv[K - g.st[c]]) / 2;
}
// f2face! + BCv! into n̂ (VOFutil.jl:229-234): fFace at every cell of the
// ghost-inclusive grid, once per use
template <class T>
__global__ void __launch_bounds__(NT) k_fface(ViscArgs<T> a, T *ff) {
    const Grid &g = a.g;
    long K = (long)blockIdx.x * NT + threadIdx.x;
    if (K >= g.ns) return;
    int q[3];
    ghost_coord(g, K, q);
    for (int c = 0; c < g.D; ++c) ff[g.ns * c + K] = visc_ff(a, c, q);
}
// viscF(i, j, Q) (flow.jl:144, getμ VOFutil.jl:186-191) on loaded values:
// f1..f4 the face fractions ff_i[Q-δj], ff_i[Q], ff_j[Q-δi], ff_j[Q]; then
// u_i[Q], u_i[Q-δj], u_j[Q], u_j[Q-δi]
template <class T>
__device__ inline T viscF_v(const ViscArgs<T> &a, T f1, T f2, T f3, T f4, T ui, T uim, T uj, T ujm) {
    T s = (((f1 + f2) + f3) + f4) / 4;
    T fr = a.lr < T(1) ? tmin(tmin(tmin(f1, f2), f3), f4) : tmax(tmax(tmax(f1, f2), f3), f4);
    T mu = a.mu * tmin(linp(s, a.lm), (s > T(0.5) ? T(1) : a.lm / a.lr) * linp(fr, a.lr));
    return mu * ((ui - uim) + (uj - ujm));
}
// the viscous force of every component of cell p (visc!, flow.jl:119-138):
// F_i = Σ_j -viscF(i,j,p) + viscF(i,j,p+δj), with each value loaded once.
// The D² viscF at p share the face fractions and velocities of p and its
// lower neighbours; viscF(i,j,p+δj) reads p's values again at its lower
// j-side (q wraps to the first interior cell on a periodic axis, flow.jl:153).
template <int D, class T> __device__ inline void visc_force_all(const ViscArgs<T> &a, const int p[3], T F[D]) {
    if (!a.has_mu) {
#pragma unroll
        for (int i = 0; i < D; ++i) F[i] = T(0);
        return;
    }
    const Grid &g = a.g;
    const T *__restrict__ ff = a.ff;
    const T *__restrict__ u = a.u;
    const long ns = g.ns, I = lin(g, p);
    T ffc[D], uc[D], ffm[D][D], um[D][D];  // [c][k]: component c at I - δk
#pragma unroll
    for (int c = 0; c < D; ++c) {
        ffc[c] = ff[ns * c + I];
        uc[c] = u[ns * c + I];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            ffm[c][k] = ff[ns * c + I - g.st[k]];
            um[c][k] = u[ns * c + I - g.st[k]];
        }
    }
#pragma unroll
    for (int i = 0; i < D; ++i) F[i] = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        // q = p + δj, wrapped to the first interior cell on a periodic axis (flow.jl:153)
        bool wrap = g.per[j] && p[j] + 1 == g.Ng[j] - 1;
        long Q = wrap ? I - (long)(g.Ng[j] - 3) * g.st[j] : I + g.st[j];
        T ffq[D], uq[D], ffjm[D], ujm[D];  // at Q: every component; component j at Q - δi
#pragma unroll
        for (int c = 0; c < D; ++c) {
            ffq[c] = ff[ns * c + Q];
            uq[c] = u[ns * c + Q];
        }
        T ffb[D], ub[D];  // every component at Q - δj (p itself unless wrapped)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            ffb[c] = wrap ? ff[ns * c + Q - g.st[j]] : ffc[c];
            ub[c] = wrap ? u[ns * c + Q - g.st[j]] : uc[c];
        }
#pragma unroll
        for (int i = 0; i < D; ++i) {
            ffjm[i] = i == j ? ffb[j] : ff[ns * j + Q - g.st[i]];
            ujm[i] = i == j ? ub[j] : u[ns * j + Q - g.st[i]];
        }
#pragma unroll
        for (int i = 0; i < D; ++i) {
            T lo = viscF_v(a, ffm[i][j], ffc[i], ffm[j][i], ffc[j], uc[i], um[i][j], uc[j], um[j][i]);
            T hi = viscF_v(a, ffb[i], ffq[i], ffjm[i], ffq[j], uq[i], ub[i], uq[j], ujm[i]);
            F[i] += -lo;
            F[i] -= -hi;
        }
    }
}

// ---------------------------------------------------------------------------
// diagnostics (metrics.jl:15-50) integrated over the interior, in FP64:
// Σρke, Σρu_i, Σρgh, Σ enstrophy (ω by backward differences at the cell's
// lower edges, which EnsI averages), Σf.  Seven block partials per block.
// ---------------------------------------------------------------------------
template <int D, class T> __device__ inline double vort_edge(const Grid &g, const T *u, long I, int a) {
    if (D == 2) {
        return (double)((u[g.ns + I] - u[g.ns + I - g.st[0]]) - (u[I] - u[I - g.st[1]]));
    }
    int j = (a + 1) % 3, k = (a + 2) % 3;
    return (double)((u[g.ns * k + I] - u[g.ns * k + I - g.st[j]]) - (u[g.ns * j + I] - u[g.ns * j + I - g.st[k]]));
}
// the metrics.jl integrands of interior cell p added into acc[0..6]
template <int D, class T>
__device__ inline void metrics_cell(const Grid &g, const T *u, const T *f, T lr, const double U[3], const double gv[3],
                                    const double wl[3], const int p[3], double acc[7]) {
    long I = lin(g, p);
    double rho = (double)linp(f[I], lr);
    double ke = 0, pe = 0;
    for (int i = 0; i < D; ++i) {
        const T *ui = u + g.ns * i;
        double a = (double)ui[I] - U[i], b = (double)ui[I + g.st[i]] - U[i];
        ke += (a * a + b * b) * rho;
        acc[1 + i] += 0.5 * (((double)ui[I] + (double)ui[I + g.st[i]]) - 2 * U[i]) * rho;
        pe += gv[i] * ((p[i] - 0.5) - wl[i]);
    }
    acc[0] += 0.25 * ke;
    acc[4] += -rho * pe;
    double en = 0;
    if (D == 2) {
        double a = vort_edge<D>(g, u, I, 0), b = vort_edge<D>(g, u, I + g.st[0], 0),
               c = vort_edge<D>(g, u, I + g.st[1], 0), d = vort_edge<D>(g, u, I + g.st[0] + g.st[1], 0);
        en = ((a * a + b * b) + c * c) + d * d;
    } else {
        for (int i = 0; i < 3; ++i) {
            int ix = (i + 1) % 3, iy = (i + 2) % 3;
            double a = vort_edge<D>(g, u, I, i), b = vort_edge<D>(g, u, I + g.st[ix], i),
                   c = vort_edge<D>(g, u, I + g.st[iy], i), d = vort_edge<D>(g, u, I + g.st[ix] + g.st[iy], i);
            en += ((a * a + b * b) + c * c) + d * d;
        }
    }
    acc[5] += 0.5 * 0.25 * en;
    acc[6] += (double)f[I];
}
template <int D, class T>
__global__ void __launch_bounds__(NT) k_metrics(Grid g, Tiling t, const T *u, const T *f, T lr, double U0, double U1,
                                                double U2, double g0, double g1, double g2, double w0, double w1,
                                                double w2, double *part) {
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    int p[3];
    if (thread_cell(g, t, p)) {
        double U[3] = {U0, U1, U2}, gv[3] = {g0, g1, g2}, wl[3] = {w0, w1, w2};
        metrics_cell<D, T>(g, u, f, lr, U, gv, wl, p, acc);
    }
    for (int q = 0; q < 7; ++q) {
        double s = block_sum(acc[q]);
        if (threadIdx.x == 0) part[(long)blockIdx.x * 7 + q] = s;
    }
}

// ---------------------------------------------------------------------------
// surface tension (surfaceTension.jl:7-124), per momentum direction d
// ---------------------------------------------------------------------------
template <class T> __device__ inline bool contain_if(T v) { return v > T(0) && v < T(1); }  // VOFutil.jl:144
// getPopinetHeightAdaptive(I, f, i, monotonic=true) about cell p (0-based,
// ghost-inclusive) along the signed axis i (1-based): to the air side, then to
// the water side; a walk stops at the array's end (validCI, util.jl:82)
template <class T> __device__ inline T popinet_height(const Grid &g, const T *f, const int p[3], int i) {
    int ax = abs(i) - 1, s = i > 0 ? 1 : -1;
    long st = g.st[ax] * s;
    long I0 = lin(g, p);
    T fnow = f[I0], H = fnow - T(0.5);
    bool fin = fnow < T(1);
    int c = p[ax];
    long In = I0;
    while (!fin || contain_if(fnow)) {
        c += s;
        In += st;
        if (c < 0 || c > g.Ng[ax] - 1) break;
        T v = f[In];
        fnow = v > fnow ? T(0) : v;
        H += fnow;
        if (contain_if(fnow)) fin = true;
    }
    fnow = f[I0];
    fin = fnow > T(0);
    c = p[ax];
    In = I0;
    while (!fin || contain_if(fnow)) {
        c -= s;
        In -= st;
        if (c < 0 || c > g.Ng[ax] - 1) break;
        T v = f[In];
        fnow = v < fnow ? T(1) : v;
        H += fnow - 1;
        if (contain_if(fnow)) fin = true;
    }
    return H;
}
template <class T> __device__ inline T root1p5(T a) { return sqrt((a * a) * a); }  // surfaceTension.jl:89
// getCurvature (surfaceTension.jl:35-80) with getXdir / getXYdir (util.jl:62-63)
template <int D, class T> __device__ inline T curvature(const Grid &g, const T *f, const int p[3], int i) {
    int s = i > 0 ? 1 : -1, ai = abs(i);
    if (D == 2) {
        int ix = ai == 1 ? -2 * s : s, ax = abs(ix) - 1, sx = ix > 0 ? 1 : -1;
        T H[3];
        for (int u = -1; u <= 1; ++u) {
            int q[3] = {p[0], p[1], p[2]};
            q[ax] += u * sx;
            H[u + 1] = popinet_height(g, f, q, i);
        }
        T Hx = (H[2] - H[0]) / 2;
        T Hxx = (H[2] + H[0]) - 2 * H[1];
        return Hxx / root1p5(1 + Hx * Hx);
    }
    int ix = s * (ai % 3 + 1), iy = (ai + 1) % 3 + 1;
    int axx = abs(ix) - 1, sx = ix > 0 ? 1 : -1;
    T H[3][3], flt = T(0.2);
    for (int u = -1; u <= 1; ++u)
        for (int v = -1; v <= 1; ++v) {
            int q[3] = {p[0], p[1], p[2]};
            q[axx] += u * sx;
            q[iy - 1] += v;
            H[u + 1][v + 1] = popinet_height(g, f, q, i);
        }
    T Hx = (H[2][1] - H[0][1]) / 2, Hy = (H[1][2] - H[1][0]) / 2;
    T Hxx = (((H[2][1] + H[0][1]) - 2 * H[1][1]) + ((H[2][0] + H[0][0]) - 2 * H[1][0]) * flt +
             ((H[2][2] + H[0][2]) - 2 * H[1][2]) * flt) / (1 + 2 * flt);
    T Hyy = (((H[1][2] + H[1][0]) - 2 * H[1][1]) + ((H[0][2] + H[0][0]) - 2 * H[0][1]) * flt +
             ((H[2][2] + H[2][0]) - 2 * H[2][1]) * flt) / (1 + 2 * flt);
    T Hxy = (((H[2][2] + H[0][0]) - H[2][0]) - H[0][2]) / 4;
    return ((Hxx * (1 + Hy * Hy) + Hyy * (1 + Hx * Hx)) - ((2 * Hxy) * Hx) * Hy) / root1p5((1 + Hx * Hx) + Hy * Hy);
}
// fbuffer[I] = ϕ(d, I, f) = (f[I] + f[I-δd]) / 2 on the interior
template <class T> __global__ void __launch_bounds__(NT) k_st_fbuf(Grid g, Tiling t, const T *f, T *fb, int d) {
    int p[3];
    if (!thread_cell(g, t, p)) return;
    long I = lin(g, p);
    fb[I] = (f[I] + f[I - g.st[d]]) / 2;
}
// BCf!(d, fbuffer) on the ghost planes of axis j (VOFutil.jl:76-89), axis by
// axis as the reference loops; the upper plane of axis d, which BCf!(d, ·)
// does not write, is set to 0
template <class T> __global__ void __launch_bounds__(NT) k_st_bc(Grid g, T *fb, int d, int j) {
    long n = g.ns / g.Ng[j];
    long q = (long)blockIdx.x * NT + threadIdx.x;
    if (q >= n) return;
    int c[3];
    long r = q;
    for (int a = 0; a < 3; ++a) {
        if (a == j) continue;
        c[a] = (int)(r % g.Ng[a]);
        r /= g.Ng[a];
    }
    c[j] = 0;
    long lo = lin(g, c), st = g.st[j], N = g.Ng[j];
    long hi = lo + (N - 1) * st;
    if (g.per[j]) {
        fb[lo] = fb[lo + (N - 2) * st];
        fb[hi] = fb[lo + st];
    } else if (j == d) {
        fb[lo] = fb[lo + 2 * st];
        fb[hi] = T(0);
    } else {
        fb[lo] = fb[lo + st];
        fb[hi] = fb[hi - st];
    }
}
// forcing[I,d] (+)= η κ(I, fbuffer, majorDir) (-∂_d f) at interface cells of
// fbuffer, 0 elsewhere (surfaceTension.jl:7-19)
template <int D, class T>
__global__ void __launch_bounds__(NT) k_surften(Grid g, Tiling t, const T *fb, const T *f, T eta, int d, T *out,
                                                int acc) {
    int p[3];
    if (!thread_cell(g, t, p)) return;
    long I = lin(g, p);
    T st = 0;
    bool hit = contain_if(fb[I]);
    if (hit) {
        T n[3];
        normal_wy(g, fb, I, n);
        int im = argabsmax(g.D, n);
        int i = (signbit(n[im]) ? -1 : 1) * (im + 1);
        st = eta * curvature<D>(g, fb, p, i) * -(f[I] - f[I - g.st[d]]);
    }
    long o = g.ns * d + I;
    if (!acc) out[o] = st;
    else if (hit) out[o] = out[o] + st;
}

template <class T> struct UpdArgs {
    ViscArgs<T> v;
    const T *u0;      // Flow.u⁰ (for ρu⁰ = u⁰ρ(f_cur), flow.jl:76/:99)
    const T *fcur;    // cVOF.f
    const T *fupd;    // f of the update and of the coefficients
    const T *ru;      // ρu after advection
    T *uout;
    T *L;             // μ₀
    const T *st;      // surface tension forcing (vector), or nullptr when η == nothing
    T dt, w, g[3];
};
template <int D, class T>
__global__ void __launch_bounds__(NT) k_visc_update(UpdArgs<T> a) {
    const Grid &g = a.v.g;
    int p[3];
    if (!thread_cell(g, a.v.t, p)) return;
    long I = lin(g, p);
    T wT = a.w, ac = T(1) / wT - 1;
    // every component's stress divergence first: its loads then share one
    // another's addresses, as no store intervenes
    T Fv[D];
    visc_force_all<D>(a.v, p, Fv);
    T un[D], rh[D];  // every component before any store (a store could alias the loads)
#pragma unroll
    for (int i = 0; i < D; ++i) {
        long o = g.ns * i + I;
        T F = Fv[i];
        if (a.st) F = F + a.st[o];
        T ru0 = a.u0[o] * linp((a.fcur[I] + a.fcur[I - g.st[i]]) / 2, a.v.lr);
        T run = ((ac * ru0 + a.ru[o]) + F * a.dt) * wT;
        T rho = linp((a.fupd[I] + a.fupd[I - g.st[i]]) / 2, a.v.lr);
        T u1 = run / rho;
        un[i] = u1 + (a.dt * wT) * a.g[i];
        rh[i] = rho;
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
        // exitBC: updateU!'s axpy! (flow.jl:251) is the only update of the exit face
        T xv = i == 0 && g.exitx ? a.v.u[I + g.st[0]] + (a.dt * wT) * a.g[0] : T(0);
        write_images<T>(g, a.uout + g.ns * i, p, un[i], 2, i, a.v.ubc[i], xv);
        write_images<T>(g, a.L + g.ns * i, p, T(1) / rh[i], 1, i, T(0));
    }
}
// viscSurfTenρu! alone (stage entry point): Flow.f = visc
template <int D, class T>
__global__ void __launch_bounds__(NT) k_visc_force(ViscArgs<T> a, T *F) {
    const Grid &g = a.g;
    int p[3];
    if (!thread_cell(g, a.t, p)) return;
    long I = lin(g, p);
    T Fv[D];
    visc_force_all<D>(a, p, Fv);
#pragma unroll
    for (int i = 0; i < D; ++i) F[g.ns * i + I] = Fv[i];
}
// updateL! alone (stage entry point)
template <class T>
__global__ void __launch_bounds__(NT) k_update_L(Grid g, Tiling t, const T *f, T *L, T lr) {
    int p[3];
    if (!thread_cell(g, t, p)) return;
    long I = lin(g, p);
    for (int i = 0; i < g.D; ++i)
        write_images<T>(g, L + g.ns * i, p, T(1) / linp((f[I] + f[I - g.st[i]]) / 2, lr), 1, i, T(0));
}

// z = div(u); (flow.jl:345)
template <class T>
__global__ void __launch_bounds__(NT) k_div(Grid g, Tiling t, const T *u, T *z) {
    int p[3];
    if (!thread_cell(g, t, p)) return;
    long I = lin(g, p);
    T s = 0;
    for (int i = 0; i < g.D; ++i) s += u[g.ns * i + I + g.st[i]] - u[g.ns * i + I];
    z[I] = s;
}
template <class T>
__global__ void k_scale(T *x, long n, T s) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) x[q] = x[q] * s;
}
template <class T>
__global__ void k_fill(T *x, long n, T v) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) x[q] = v;
}
template <class T> __device__ inline T xper(const Grid &g, const T *x, int p0, int p1, int p2) {
    return x[(long)map_per(g, 0, p0) + g.st[1] * map_per(g, 1, p1) + g.st[2] * map_per(g, 2, p2)];
}
template <class T>
__global__ void __launch_bounds__(NT) k_set_diag(Grid g, Tiling t, const T *L, T *Dg, T *iD) {
    int p[3];
    if (!thread_cell(g, t, p)) return;
    long I = lin(g, p);
    T s = 0;
    for (int d = 0; d < g.D; ++d) s -= (L[g.ns * d + I] + L[g.ns * d + I + g.st[d]]);
    Dg[I] = s;
    iD[I] = (s * s) < 2 * eps_of<T>() ? T(0) : T(1) / s;
}
template <class T>
__global__ void __launch_bounds__(NT) k_restrictL(Grid gc, Tiling t, Grid gf, const T *Lf, T *Lc) {
    int p[3];
    if (!thread_cell(gc, t, p)) return;
    for (int ii = 0; ii < gc.D; ++ii) {
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = a < gc.D ? 2 * p[a] - 1 : 0;
            hi[a] = a < gc.D ? (a == ii ? 2 * p[a] - 1 : 2 * p[a]) : 0;
        }
        T s = 0;
        for (int c2 = lo[2]; c2 <= hi[2]; ++c2)
            for (int c1 = lo[1]; c1 <= hi[1]; ++c1)
                for (int c0 = lo[0]; c0 <= hi[0]; ++c0) s += Lf[gf.ns * ii + c0 + gf.st[1] * c1 + gf.st[2] * c2];
        write_images<T>(gc, Lc + gc.ns * ii, p, (T)(0.5 * (double)s), 1, ii, T(0));
    }
}
__global__ void __launch_bounds__(NT) k_fin_sum(const double *part, long nblk, double *out) {
    double v = 0;
    for (long b = threadIdx.x; b < nblk; b += NT) v += part[b];
    double s = block_sum(v);
    if (threadIdx.x == 0) *out = s;
}
// myproject!: u -= L ∂x ; BC!(u)  (flow.jl:331-333, :82)
template <class T>
__global__ void __launch_bounds__(NT) k_project(Grid g, Tiling t, T *u, const T *L, const T *x, T ubc0, T ubc1, T ubc2) {
    int p[3];
    if (!thread_cell(g, t, p)) return;
    long I = lin(g, p);
    T ub[3] = {ubc0, ubc1, ubc2};
    for (int i = 0; i < g.D; ++i) {
        int q[3] = {p[0], p[1], p[2]};
        q[i] -= 1;
        T dx = x[I] - xper(g, x, q[0], q[1], q[2]);
        T un = u[g.ns * i + I] - L[g.ns * i + I] * dx;
        T xv = i == 0 && g.exitx ? u[I + g.st[0]] : T(0);  // exitBC: the exit face is left alone
        write_images<T>(g, u + g.ns * i, p, un, 2, i, ub[i], xv);
    }
}
// the two CFL maxima, Σf (each thread adds its blocks in order, then a fixed
// tree), min f and max f of the k_cfl partials
__global__ void __launch_bounds__(NT) k_fin_cfl(const double *part, long nblk, double *out) {
    double a = 0, b = 0, s = 0, mn = 1e300, mx = -1e300;
    for (long q = threadIdx.x; q < nblk; q += NT) {
        const double *o = part + 5 * q;
        a = o[0] > a ? o[0] : a;
        b = o[1] > b ? o[1] : b;
        s += o[2];
        mn = o[3] < mn ? o[3] : mn;
        mx = o[4] > mx ? o[4] : mx;
    }
    double ma = block_max(a);
    __syncthreads();
    double mb = block_max(b);
    __syncthreads();
    double ss = block_sum(s);
    __syncthreads();
    double mmx = block_max(mx);
    __syncthreads();
    double mmn = -block_max(-mn);
    if (threadIdx.x == 0) {
        out[0] = ma;
        out[1] = mb;
        out[2] = ss;
        out[3] = mmn;
        out[4] = mmx;
    }
}
// f⁰ = (f⁰ + f)/2 over every cell (flow.jl:74)
template <class T>
__global__ void k_avg(T *a, const T *b, long n) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) a[q] = (a[q] + b[q]) / 2;
}
// applyVOF! from SDF samples (VOFutil.jl:203-225) + cleanWisp! + BCf!
template <class T>
__global__ void __launch_bounds__(NT) k_apply_vof(Grid g, Tiling t, const T *sdf, T *f, T *al, T *nh, T ctol) {
    int p[3];
    if (!thread_cell(g, t, p)) return;
    long I = lin(g, p);
    long q = (long)(p[0] - 1) + (long)(g.Ng[0] - 2) * ((g.D > 1 ? p[1] - 1 : 0) + (long)(g.D > 1 ? g.Ng[1] - 2 : 1) * (g.D > 2 ? p[2] - 1 : 0));
    const T *s = sdf + q * (1 + 2 * g.D);
    T sumN = 0, sumN2 = 0, n[3];
    for (int d = 0; d < g.D; ++d) {
        T dd = s[2 * d + 1] - s[2 * d + 2];
        nh[g.ns * d + I] = dd;
        n[d] = dd;
        sumN += dd;
        sumN2 += dd * dd;
    }
    T a = sumN / 2 - sqrt(sumN2) * s[0];
    al[I] = a;
    T fv = getVolumeFraction(g.D, n, a);
    fv = fv < ctol ? T(0) : (fv > 1 - ctol ? T(1) : fv);
    write_images<T>(g, f, p, fv, 0, 0, T(0));
}
// reconstructInterface! materialised (stage entry point; BCVOF! on periodic sides)
template <class T>
__global__ void __launch_bounds__(NT) k_reconstruct(Grid g, Tiling t, const T *f, T *al, T *nh, int scheme) {
    int p[3];
    if (!thread_cell(g, t, p)) return;
    long I = lin(g, p);
    T n[3] = {0, 0, 0}, a = 0;
    if (!fullorempty(f[I])) {
        normal_scheme(scheme, g, f, I, n);
        a = getIntercept(g.D, n, f[I]);
    }
    // periodic images only (non-periodic ghosts of n̂/α are zero)
    int alt[3][2], na[3];
    for (int ax = 0; ax < 3; ++ax) {
        na[ax] = 1;
        alt[ax][0] = p[ax];
        if (ax < g.D && g.per[ax]) {
            if (p[ax] == 1) { alt[ax][1] = g.Ng[ax] - 1; na[ax] = 2; }
            else if (p[ax] == g.Ng[ax] - 2) { alt[ax][1] = 0; na[ax] = 2; }
        }
    }
    for (int i2 = 0; i2 < na[2]; ++i2)
        for (int i1 = 0; i1 < na[1]; ++i1)
            for (int i0 = 0; i0 < na[0]; ++i0) {
                long J = (long)alt[0][i0] + g.st[1] * alt[1][i1] + g.st[2] * alt[2][i2];
                al[J] = a;
                for (int d = 0; d < g.D; ++d) nh[g.ns * d + J] = n[d];
            }
}

// ---------------------------------------------------------------------------
// level-set reinitialisation (redistaning.jl)
// ---------------------------------------------------------------------------
// LevelSet(sim) (redistaning.jl:14-25): ϕ = 2f - 1 on every entry, ghosts
// included, and ϕini = ϕ
template <class T> __global__ void __launch_bounds__(NT) k_ls_init(long ns, const T *f, T *phi, T *ini) {
    long q = (long)blockIdx.x * NT + threadIdx.x;
    if (q >= ns) return;
    T v = 2 * f[q] - 1;
    phi[q] = v;
    ini[q] = v;
}
// redistaning.jl:110 minmod
template <class T> __device__ inline T ls_minmod(T a, T b) { return fabs(a) <= fabs(b) ? a : b; }
// redistaning.jl:122-143 𝛁ϕᵢ²: second-order ENO (minmod-limited) one-sided
// derivatives of ϕ[I-2..I+2], the upwind one for the sign s squared, else 0
template <class T> __device__ inline T ls_grad2(T a, T b, T c, T d, T e, T s) {
    T dp = d - c, dm = c - b;
    T d2p = (e + c) - 2 * d, d2o = (d + b) - 2 * c, d2m = (c + a) - 2 * b;
    T dR = dp - ls_minmod(d2p, d2o) / 2;
    T dL = dm + ls_minmod(d2o, d2m) / 2;
    T wR = dR * s, wL = dL * s;
    if (wR < 0 && (wR + wL) < 0) return dR * dR;
    if (wL > 0 && (wR + wL) > 0) return dL * dL;
    return T(0);
}
// One stage of redistaning! (redistaning.jl:28-31, :67-108, BCf! of
// VOFutil.jl:64-75):  L = ϕini (1 - |∇ϕ|) with |∇ϕ|² summed axis by axis
// (the first and last interior cells of an axis take the mirrored (Neumann)
// or wrapped (periodic) point two cells out), then
// out = α ϕ⁰ + (1-α)(ϕ + dτ L), written with its ghost images.  ϕ and out are
// different buffers; ϕ⁰ is read at I only and may be `out`.
template <int D, class T>
__global__ void __launch_bounds__(NT) k_ls_stage(Grid g, Tiling t, const T *phi, const T *phi0, const T *ini, T *out,
                                                 T dtau, T alpha) {
    int p[3];
    if (!thread_cell(g, t, p)) return;
    long I = lin(g, p);
    T si = ini[I];
    T s = si > 0 ? T(1) : (si < 0 ? T(-1) : si);
    T c = phi[I];
    T acc = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        long st = g.st[i], w = (long)(g.Ng[i] - 4) * st;
        int q = p[i];
        T b = phi[I - st], d = phi[I + st];
        T a = q == 1 ? (g.per[i] ? phi[I + w] : b) : phi[I - 2 * st];
        T e = q == g.Ng[i] - 2 ? (g.per[i] ? phi[I - w] : d) : phi[I + 2 * st];
        acc = acc + ls_grad2(a, b, c, d, e, s);
    }
    T L = si * (T(1) - sqrt(acc));
    T v = alpha * phi0[I] + (T(1) - alpha) * (c + dtau * L);
    write_images<T>(g, out, p, v, 0, 0, T(0));
}

}  // namespace ia
