// ia_device.h -- device-side building blocks for the gfx950 two-phase step.
//
// Coordinates inside kernels are 0-based and ghost-inclusive: p[a] in
// [0, Ng[a]), interior 1..Ng-2 (the reference's 2..Ng-1).  Arithmetic in
// every function follows the reference expression by expression (and the
// same order of operations), so that FP32 results equal the FP32 CPU oracle
// up to the last-ulp differences of the transcendental functions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ia {

struct Grid {
    int D;
    int Ng[3];    // ghost-inclusive extents (1 on an absent axis)
    int per[3];   // periodic flags
    long st[3];   // strides
    long ns;      // cells
    int exitx;    // WaterLily exitBC: the step's BC! keeps u_x on the exit face (x = Ng[0]-1)
};

// ---------------------------------------------------------------------------
// launch geometry: one thread per interior cell; 1-D grid of tiles whose
// block id is swizzled so that each XCD (8 on MI355X, blocks dispatched
// round-robin) works on one contiguous eighth of the domain, keeping stencil
// neighbours in the same L2.
// ---------------------------------------------------------------------------
constexpr int TX = 64, TY = 4;
struct Tiling {
    int nt[3];   // tiles per axis over the interior
    int tz;      // cells per tile along z (1)
    long nblk;
};
__device__ inline long swizzle_block(long b, long nblk) {
    const long nx = 8;
    long per = nblk / nx;
    if (per == 0) return b;
    long full = per * nx;
    if (b >= full) return b;
    return (b % nx) * per + b / nx;
}
// interior cell of this thread (returns false when outside the interior)
__device__ inline bool thread_cell(const Grid &g, const Tiling &t, int p[3]) {
    long b = swizzle_block((long)blockIdx.x, t.nblk);
    long bx = b % t.nt[0];
    long r = b / t.nt[0];
    long by = r % t.nt[1];
    long bz = r / t.nt[1];
    int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    p[0] = 1 + (int)bx * TX + tx;
    p[1] = g.D > 1 ? 1 + (int)by * TY + ty : 0;
    p[2] = g.D > 2 ? 1 + (int)bz : 0;
    bool ok = p[0] <= g.Ng[0] - 2;
    if (g.D > 1) ok = ok && p[1] <= g.Ng[1] - 2;
    if (g.D > 2) ok = ok && p[2] <= g.Ng[2] - 2;
    return ok;
}
__device__ inline long lin(const Grid &g, const int p[3]) {
    return (long)p[0] + g.st[1] * p[1] + g.st[2] * p[2];
}
// upper x-face q of the interior (one per x-row, rows in y-then-z order):
// the cell index of that face, x = Ng[0]-1.  A face pass along x computes
// these in blocks of their own, since inside the tiling they fall on one lane
// of a wavefront and would serialise it.
__device__ inline bool top_face_x(const Grid &g, long q, int p[3]) {
    long n1 = g.D > 1 ? g.Ng[1] - 2 : 1, n2 = g.D > 2 ? g.Ng[2] - 2 : 1;
    if (q >= n1 * n2) return false;
    p[0] = g.Ng[0] - 1;
    p[1] = g.D > 1 ? 1 + (int)(q % n1) : 0;
    p[2] = g.D > 2 ? 1 + (int)(q / n1) : 0;
    return true;
}
__host__ inline long top_face_x_blocks(const Grid &g) {
    long n1 = g.D > 1 ? g.Ng[1] - 2 : 1, n2 = g.D > 2 ? g.Ng[2] - 2 : 1;
    return (n1 * n2 + TX * TY - 1) / (TX * TY);
}

// ---------------------------------------------------------------------------
// boundary maps.  Reading through these equals reading the reference's array
// after its BC routine ran.
// ---------------------------------------------------------------------------
// BCf! / perBC! axis map for cell-centred scalars (VOFutil.jl:64-74)
__device__ inline int map_scalar(const Grid &g, int a, int q) {
    int n = g.Ng[a];
    if (a >= g.D) return q;
    if (q == 0) return g.per[a] ? n - 2 : 1;
    if (q == n - 1) return g.per[a] ? 1 : n - 2;
    return q;
}
__device__ inline long lin_scalar(const Grid &g, int p0, int p1, int p2) {
    return (long)map_scalar(g, 0, p0) + g.st[1] * map_scalar(g, 1, p1) + g.st[2] * map_scalar(g, 2, p2);
}
// periodic-only map (perBC! on Poisson iterates)
__device__ inline int map_per(const Grid &g, int a, int q) {
    int n = g.Ng[a];
    if (a >= g.D || !g.per[a]) return q;
    if (q == 0) return n - 2;
    if (q == n - 1) return 1;
    return q;
}
// WaterLily BC!(a, A, exitBC, perdir) on component c: true -> Dirichlet A[c]
__device__ inline bool vel_dirichlet(const Grid &g, int c, const int p[3]) {
    int n = g.Ng[c];
    return !g.per[c] && (p[c] == 0 || p[c] == 1 || (p[c] == n - 1 && !(g.exitx && c == 0)));
}
// the exit face of u_x under exitBC (left alone by BC!)
__device__ inline bool exit_face(const Grid &g, int c, const int p[3]) {
    return g.exitx && c == 0 && p[0] == g.Ng[0] - 1;
}
__device__ inline long lin_vel(const Grid &g, int c, const int p[3]) {
    int q[3];
    for (int a = 0; a < 3; ++a) q[a] = map_scalar(g, a, p[a]);
    return (long)q[0] + g.st[1] * q[1] + g.st[2] * q[2];
}
// BCv! on a face-averaged field, component c (VOFutil.jl:91-104): returns
// false when the slot is one f2face!/BCv! never writes (p_c = Ng-1 on a
// non-periodic axis), whose stale value the caller supplies.
__device__ inline bool face_map(const Grid &g, int c, const int p[3], int q[3]) {
    for (int a = 0; a < 3; ++a) {
        int n = g.Ng[a];
        int v = p[a];
        if (a < g.D) {
            if (g.per[a]) {
                if (v == 0) v = n - 2;
                else if (v == n - 1) v = 1;
            } else if (a == c) {
                if (v == n - 1) return false;
                if (v == 0) v = 2;
            } else {
                if (v == 0) v = 1;
                else if (v == n - 1) v = n - 2;
            }
        }
        q[a] = v;
    }
    return true;
}

// write an interior value and every ghost image of it.  kind 0: scalar
// (BCf!); kind 1: velocity-like component c (BC!(a, A, false) with Dirichlet
// value dval); kind 2: the step's BC!(u, u_BC, exitBC): as kind 1, except
// that under exitBC the exit face of u_x (x = Ng[0]-1) takes xval, the value
// BC! leaves there, and its Neumann images in y and z copy it.
template <class T>
__device__ inline void write_images(const Grid &g, T *a, const int p[3], T val, int kind, int c, T dval,
                                    T xval = T(0)) {
    const bool xexit = kind == 2 && c == 0 && g.exitx && !g.per[0];
    if (kind == 2) kind = 1;
    int alt[3][2], na[3];
    bool dir[3][2];
    for (int ax = 0; ax < 3; ++ax) {
        na[ax] = 1;
        alt[ax][0] = p[ax];
        dir[ax][0] = false;
        if (ax >= g.D) continue;
        int n = g.Ng[ax];
        bool vel_axis = kind == 1 && ax == c && !g.per[ax];
        if (vel_axis && p[ax] == 1) dir[ax][0] = true;
        int img = -1;
        if (g.per[ax]) {
            if (p[ax] == 1) img = n - 1;
            else if (p[ax] == n - 2) img = 0;
        } else {
            if (p[ax] == 1) img = 0;
            else if (p[ax] == n - 2) img = n - 1;
        }
        if (img >= 0) {
            alt[ax][1] = img;
            dir[ax][1] = vel_axis && !(xexit && ax == 0 && img == n - 1);
            na[ax] = 2;
        }
        // a 1-cell interior (n == 3) has both images
        if (n == 3 && !g.per[ax]) {
            // p == 1 == n-2: images 0 and 2; handled by the second pass below
        }
    }
    for (int i2 = 0; i2 < na[2]; ++i2)
        for (int i1 = 0; i1 < na[1]; ++i1)
            for (int i0 = 0; i0 < na[0]; ++i0) {
                bool d = dir[0][i0] || dir[1][i1] || dir[2][i2];
                bool xe = xexit && i0 == 1 && alt[0][1] == g.Ng[0] - 1;
                long idx = (long)alt[0][i0] + g.st[1] * alt[1][i1] + g.st[2] * alt[2][i2];
                a[idx] = d ? dval : (xe ? xval : val);
            }
}

// ---------------------------------------------------------------------------
// scalar helpers with Julia semantics
// ---------------------------------------------------------------------------
template <class T> __device__ inline T tmin(T a, T b) { return b < a ? b : a; }
template <class T> __device__ inline T tmax(T a, T b) { return b > a ? b : a; }
template <class T> __device__ inline T tsign(T x) { return x > 0 ? T(1) : (x < 0 ? T(-1) : x); }
template <class T> __device__ inline T linp(T f, T lam) { return lam + (T(1) - lam) * f; }  // VOFutil.jl:166
template <class T> __device__ inline bool fullorempty(T f) { return f == T(0) || f == T(1); }
template <class T> __device__ inline T median3(T a, T b, T c) { return tmax(tmin(a, b), tmin(tmax(a, b), c)); }
template <class T> __device__ inline T eps_of();
template <> __device__ inline float eps_of<float>() { return 1.1920928955078125e-07f; }
template <> __device__ inline double eps_of<double>() { return 2.220446049250313e-16; }

// ---------------------------------------------------------------------------
// PLIC (PLIC.jl, Scardovelli & Zaleski 2000)
// ---------------------------------------------------------------------------
template <class T> __device__ inline void sort2(T &a, T &b) {
    if (!(a < b)) { T t = a; a = b; b = t; }
}
template <class T> __device__ inline void sort3(T &a, T &b, T &c) {
    T t;
    if (a > c) { t = a; a = c; c = t; }
    if (a > b) { t = a; a = b; b = t; }
    if (b > c) { t = b; b = c; c = t; }
}
template <class T> __device__ inline T proot(T c0, T c1, T c2, T c3) {  // PLIC.jl:165-176
    T a0 = c0 / c3, a1 = c1 / c3, a2 = c2 / c3;
    T p0 = a1 / 3 - (a2 * a2) / 9;
    T q0 = (a1 * a2 - 3 * a0) / 6 - (a2 * a2 * a2) / 27;
    T a = q0 / sqrt(-(p0 * p0 * p0));
    T t = acos((a * a) <= T(1) ? a : T(0)) / 3;
    return sqrt(-p0) * (sqrt(T(3)) * sin(t) - cos(t)) - a2 / 3;
}
template <class T> __device__ inline T a2f2(T m1, T m2, T a) {  // PLIC.jl:92
    return a < m1 ? (a * a) / ((2 * m1) * m2) : (a - m1 / 2) / m2;
}
template <class T> __device__ inline T a2f3(T m1, T m2, T m3, T a) {  // PLIC.jl:93-107
    T m12 = m1 + m2;
    if (a < m1) return (a * a * a) / (((6 * m1) * m2) * m3);
    if (a < m2) return (a * (a - m1)) / ((2 * m2) * m3) + ((m2 == T(0) ? T(1) : m1 / m2) * m1) / (6 * m3);
    if (a < tmin(m3, m12))
        return ((a * a) * (3 * m12 - a) + (m1 * m1) * (m1 - 3 * a) + (m2 * m2) * (m2 - 3 * a)) / (((6 * m1) * m2) * m3);
    if (m3 < m12)
        return ((a * a) * (3 - 2 * a) + (m1 * m1) * (m1 - 3 * a) + (m2 * m2) * (m2 - 3 * a) + (m3 * m3) * (m3 - 3 * a)) /
               (((6 * m1) * m2) * m3);
    return (2 * a - m12) / (2 * m3);
}
template <class T> __device__ inline T f2a2(T m1, T m2, T v) {  // PLIC.jl:117
    return v < m1 / (2 * m2) ? sqrt(((2 * m1) * m2) * v) : m2 * v + m1 / 2;
}
template <class T> __device__ inline T f2a3(T m1, T m2, T m3, T v) {  // PLIC.jl:118-149
    T m12 = m1 + m2;
    T p = ((6 * m1) * m2) * m3;
    T v1 = ((m2 == T(0) ? T(1) : m1 / m2) * m1) / (6 * m3);
    T v2 = v1 + (m2 - m1) / (2 * m3);
    T v3 = m3 < m12 ? ((m3 * m3) * (3 * m12 - m3) + (m1 * m1) * (m1 - 3 * m3) + (m2 * m2) * (m2 - 3 * m3)) / p
                    : m12 / (2 * m3);
    if (v < v1) return cbrt(p * v);
    if (v < v2) return (m1 + sqrt(m1 * m1 + ((8 * m2) * m3) * (v - v1))) / 2;
    if (v < v3) {
        T c0 = (m1 * m1 * m1 + m2 * m2 * m2) - p * v;
        T c1 = -3 * (m1 * m1 + m2 * m2);
        T c2 = 3 * m12;
        return proot(c0, c1, c2, T(-1));
    }
    if (m3 < m12) {
        T c0 = ((m1 * m1 * m1 + m2 * m2 * m2) + m3 * m3 * m3) - p * v;
        T c1 = -3 * ((m1 * m1 + m2 * m2) + m3 * m3);
        return proot(c0, c1, T(3), T(-2));
    }
    return m3 * v + m12 / 2;
}
template <class T> __device__ inline T getIntercept(int D, const T *n, T g) {  // PLIC.jl:20-39
    T a;
    if (D == 2) {
        T t = fabs(n[0]) + fabs(n[1]);
        if (g != T(0.5)) {
            T m1 = fabs(n[0]) / t, m2 = fabs(n[1]) / t;
            sort2(m1, m2);
            a = f2a2(m1, m2, g < T(0.5) ? g : 1 - g);
        } else
            a = T(0.5);
        return ((g < T(0.5) ? a : 1 - a) * t + tmin(n[0], T(0))) + tmin(n[1], T(0));
    }
    T t = (fabs(n[0]) + fabs(n[1])) + fabs(n[2]);
    if (g != T(0.5)) {
        T m1 = fabs(n[0]) / t, m2 = fabs(n[1]) / t, m3 = fabs(n[2]) / t;
        sort3(m1, m2, m3);
        a = f2a3(m1, m2, m3, g < T(0.5) ? g : 1 - g);
    } else
        a = T(0.5);
    return (((g < T(0.5) ? a : 1 - a) * t + tmin(n[0], T(0))) + tmin(n[1], T(0))) + tmin(n[2], T(0));
}
template <class T> __device__ inline T getVolumeFraction(int D, const T *n, T b) {  // PLIC.jl:59-82
    if (D == 2) {
        T t = fabs(n[0]) + fabs(n[1]);
        T a = ((b - tmin(n[0], T(0))) - tmin(n[1], T(0))) / t;
        if (a <= T(0) || a == T(0.5) || a >= T(1)) return tmin(tmax(a, T(0)), T(1));
        T m1 = fabs(n[0]) / t, m2 = fabs(n[1]) / t;
        sort2(m1, m2);
        T v = a2f2(m1, m2, a < T(0.5) ? a : 1 - a);
        return a < T(0.5) ? v : 1 - v;
    }
    T t = (fabs(n[0]) + fabs(n[1])) + fabs(n[2]);
    T a = (((b - tmin(n[0], T(0))) - tmin(n[1], T(0))) - tmin(n[2], T(0))) / t;
    if (a <= T(0) || a == T(0.5) || a >= T(1)) return tmin(tmax(a, T(0)), T(1));
    T m1 = fabs(n[0]) / t, m2 = fabs(n[1]) / t, m3 = fabs(n[2]) / t;
    sort3(m1, m2, m3);
    T v = a2f3(m1, m2, m3, a < T(0.5) ? a : 1 - a);
    return a < T(0.5) ? v : 1 - v;
}

// ---------------------------------------------------------------------------
// interface normals (normalEstimation.jl)
// ---------------------------------------------------------------------------
template <class T> __device__ inline T h3(const Grid &g, const T *f, long I, int s) {  // VOFutil.jl:158
    return (f[I] + f[I - g.st[s]]) + f[I + g.st[s]];
}
template <class T> __device__ inline int argabsmax(int D, const T *n) {  // util.jl:19-31 (0-based)
    T mx = 0;
    int im = 0;
    for (int i = 0; i < D; ++i) {
        T cur = n[i] * n[i];
        if (cur > mx) { mx = cur; im = i; }
    }
    return im;
}
// normalEstimation.jl:75-97 (Column) ; I is a cell whose 3^D block is valid
template <class T> __device__ inline void normal_column(const Grid &g, const T *f, long I, T *n) {
    for (int d = 0; d < g.D; ++d) n[d] = f[I - g.st[d]] - f[I + g.st[d]];
    int dom = argabsmax(g.D, n);
    for (int d = 0; d < g.D; ++d) {
        if (d == dom) {
            T s = tsign(n[d]);
            n[d] = s == T(0) ? T(1) : s;
            continue;
        }
        T hl = h3(g, f, I - g.st[d], dom), hr = h3(g, f, I + g.st[d], dom);
        n[d] = (hl - hr) / 2;
    }
}
// normalEstimation.jl:105-154 (WH, the default).  Float64 literals of the
// Julia source stay double here.
template <class T> __device__ inline void normal_wh(const Grid &g, const T *f, long I, T *n) {
    normal_column(g, f, I, n);
    int im = argabsmax(g.D, n);
    int sdom = signbit(n[im]) ? -1 : 1;
    T an = fabs(n[im]);
    an = an == T(0) ? T(1) : an;
    for (int d = 0; d < g.D; ++d) n[d] /= an;
    for (int d = 0; d < g.D; ++d) {
        if (d == im) continue;
        T slope = fabs(n[d]);
        int scur = signbit(n[d]) ? -1 : 1;
        long off = scur * g.st[d];
        T hl = h3(g, f, I - off, im), hc = h3(g, f, I, im), hr = h3(g, f, I + off, im);
        T sh = (hl + hc) + hr;
        double s45 = 4.5 * (double)slope;
        if (s45 <= (double)sh && (double)sh <= 9 - s45) continue;
        long offd = sdom * g.st[im];
        double r1 = (double)(T(1) / (2 * slope));
        double mn = s45 < r1 ? s45 : r1;
        if ((double)sh < s45) {
            T wb = h3(g, f, I - offd, d);
            if ((double)wb > mn) n[d] = (T)copysign(((double)hl - 0.5) / ((double)wb - 0.5), (double)scur);
        }
        if ((double)sh > 9 - s45) {
            T wt = h3(g, f, I + offd, d);
            if ((double)wt < 3 - mn) n[d] = (T)copysign((2.5 - (double)hr) / (2.5 - (double)wt), (double)scur);
        }
    }
}
// normal_wh of a 3D cell from its 3x3x3 block, loaded in one round instead
// of WH's three dependent rounds (the central differences, then the column
// heights along the dominant axis, then the row widths).  Every value and
// operation is the one normal_wh uses, in the same order, so the result is
// identical; the axis choices select among column sums held in registers.
template <class T> __device__ inline T sel3(T v0, T v1, T v2, int k) { return k == 0 ? v0 : (k == 1 ? v1 : v2); }
template <class T> __device__ inline void normal_wh_block3(const Grid &g, const T *f, long I, T *n) {
    T b[3][3][3];  // b[dz+1][dy+1][dx+1] = f at I + (dx, dy, dz)
#pragma unroll
    for (int z = 0; z < 3; ++z)
#pragma unroll
        for (int y = 0; y < 3; ++y)
#pragma unroll
            for (int x = 0; x < 3; ++x) b[z][y][x] = f[I + (x - 1) + g.st[1] * (y - 1) + g.st[2] * (z - 1)];
    // Hm[d][s] = h3 at I - e_d along s, Hp[d][s] = h3 at I + e_d along s (d != s), Cs[s] = h3 at I along s
    T Hm[3][3], Hp[3][3], Cs[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const int es0 = s == 0, es1 = s == 1, es2 = s == 2;
        Cs[s] = (b[1][1][1] + b[1 - es2][1 - es1][1 - es0]) + b[1 + es2][1 + es1][1 + es0];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (d == s) {
                Hm[d][s] = Hp[d][s] = T(0);
                continue;
            }
            const int o0 = d == 0, o1 = d == 1, o2 = d == 2;
            Hm[d][s] = (b[1 - o2][1 - o1][1 - o0] + b[1 - o2 - es2][1 - o1 - es1][1 - o0 - es0]) +
                       b[1 - o2 + es2][1 - o1 + es1][1 - o0 + es0];
            Hp[d][s] = (b[1 + o2][1 + o1][1 + o0] + b[1 + o2 - es2][1 + o1 - es1][1 + o0 - es0]) +
                       b[1 + o2 + es2][1 + o1 + es1][1 + o0 + es0];
        }
    }
    // normal_column
    n[0] = b[1][1][0] - b[1][1][2];
    n[1] = b[1][0][1] - b[1][2][1];
    n[2] = b[0][1][1] - b[2][1][1];
    int dom = argabsmax(3, n);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d == dom) {
            T sg = tsign(n[d]);
            n[d] = sg == T(0) ? T(1) : sg;
            continue;
        }
        T hl = sel3(Hm[d][0], Hm[d][1], Hm[d][2], dom), hr = sel3(Hp[d][0], Hp[d][1], Hp[d][2], dom);
        n[d] = (hl - hr) / 2;
    }
    // normal_wh
    int im = argabsmax(3, n);
    T nim = sel3(n[0], n[1], n[2], im);
    int sdom = signbit(nim) ? -1 : 1;
    T an = fabs(nim);
    an = an == T(0) ? T(1) : an;
#pragma unroll
    for (int d = 0; d < 3; ++d) n[d] /= an;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d == im) continue;
        T slope = fabs(n[d]);
        int scur = signbit(n[d]) ? -1 : 1;
        T hm = sel3(Hm[d][0], Hm[d][1], Hm[d][2], im), hp = sel3(Hp[d][0], Hp[d][1], Hp[d][2], im);
        T hl = scur > 0 ? hm : hp, hc = sel3(Cs[0], Cs[1], Cs[2], im), hr = scur > 0 ? hp : hm;
        T sh = (hl + hc) + hr;
        double s45 = 4.5 * (double)slope;
        if (s45 <= (double)sh && (double)sh <= 9 - s45) continue;
        double r1 = (double)(T(1) / (2 * slope));
        double mn = s45 < r1 ? s45 : r1;
        // columns along d at I -+ sdom e_im
        T cm = sel3(Hm[0][d], Hm[1][d], Hm[2][d], im), cp = sel3(Hp[0][d], Hp[1][d], Hp[2][d], im);
        if ((double)sh < s45) {
            T wb = sdom > 0 ? cm : cp;
            if ((double)wb > mn) n[d] = (T)copysign(((double)hl - 0.5) / ((double)wb - 0.5), (double)scur);
        }
        if ((double)sh > 9 - s45) {
            T wt = sdom > 0 ? cp : cm;
            if ((double)wt < 3 - mn) n[d] = (T)copysign((2.5 - (double)hr) / (2.5 - (double)wt), (double)scur);
        }
    }
}
// normalEstimation.jl:35-68 (WY)
template <class T> __device__ inline void normal_wy(const Grid &g, const T *f, long I, T *n) {
    for (int d = 0; d < g.D; ++d) n[d] = f[I - g.st[d]] - f[I + g.st[d]];
    int dom = argabsmax(g.D, n);
    for (int d = 0; d < g.D; ++d) {
        if (d == dom) {
            T s = tsign(n[d]);
            n[d] = s == T(0) ? T(1) : s;
            continue;
        }
        T hl = h3(g, f, I - g.st[d], dom), hc = h3(g, f, I, dom), hr = h3(g, f, I + g.st[d], dom);
        T v = (hl - hr) / 2;
        if (fabs(v) > T(0.5)) v = ((double)v * ((double)hc - 1.5) >= 0.0) ? hc - hr : hl - hc;
        n[d] = v;
    }
}
// normalEstimation.jl:248-257 YoungSum: the 2^(D-1) cells II ∈ I-δxy:I and, for
// each, III ∈ II:II+δxy (δxy = 1 on every axis but d), in column-major order
template <class T> __device__ inline T young_sum(const Grid &g, const T *f, long I, int d) {
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int a = 0; a < g.D; ++a)
        if (a != d) lo[a] = -1;
    T s = 0;
    for (int c2 = lo[2]; c2 <= hi[2]; ++c2)
        for (int c1 = lo[1]; c1 <= hi[1]; ++c1)
            for (int c0 = lo[0]; c0 <= hi[0]; ++c0) {
                long II = I + c0 + g.st[1] * c1 + g.st[2] * c2;
                int h2[3] = {0, 0, 0};
                for (int a = 0; a < g.D; ++a)
                    if (a != d) h2[a] = 1;
                for (int e2 = 0; e2 <= h2[2]; ++e2)
                    for (int e1 = 0; e1 <= h2[1]; ++e1)
                        for (int e0 = 0; e0 <= h2[0]; ++e0) s += f[II + e0 + g.st[1] * e1 + g.st[2] * e2];
            }
    return s;
}
// normalEstimation.jl:232-247 (Youngs)
template <class T> __device__ inline void normal_youngs(const Grid &g, const T *f, long I, T *n) {
    T a = 0;
    for (int d = 0; d < g.D; ++d) {
        n[d] = (T)((double)(young_sum(g, f, I - g.st[d], d) - young_sum(g, f, I + g.st[d], d)) * 0.5);
        a += fabs(n[d]);
    }
    for (int d = 0; d < g.D; ++d) n[d] = a == T(0) ? (T)(1.0 / g.D) : n[d] / a;
}
// normalEstimation.jl:210-224 (CCi with dominant axis dc; signs from the Youngs normal ny)
template <class T> __device__ inline void normal_cci(const Grid &g, const T *f, long I, const T *ny, int dc, T *out) {
    T s = 0;
    for (int d = 0; d < g.D; ++d) {
        if (d == dc) {
            T sg = tsign(ny[d]);
            out[d] = sg == T(0) ? T(1) : sg;
        } else {
            T hu = h3(g, f, I + g.st[d], dc), hd = h3(g, f, I - g.st[d], dc);
            out[d] = -(hu - hd) / 2;
        }
    }
    for (int d = 0; d < g.D; ++d) s += fabs(out[d]);
    for (int d = 0; d < g.D; ++d) out[d] /= s;
}
// normalEstimation.jl:185-202 (MYC): Youngs, replaced by the CCi normal of the
// axis picked by the scan below when that one's dominant component is smaller
template <class T> __device__ inline void normal_myc(const Grid &g, const T *f, long I, T *n) {
    normal_youngs(g, f, I, n);
    T mx = 0;
    for (int i = 0; i < g.D; ++i) mx = fabs(n[i]) > mx ? fabs(n[i]) : mx;
    T curm0 = 0, cc[3];
    int ccz = 0;
    for (int iz = 0; iz < g.D; ++iz) {
        normal_cci(g, f, I, n, iz, cc);
        if (fabs(cc[iz]) > curm0) ccz = iz;
        curm0 = fabs(cc[iz]);
    }
    normal_cci(g, f, I, n, ccz, cc);
    if (fabs(cc[ccz]) < mx)
        for (int i = 0; i < g.D; ++i) n[i] = cc[i];
}
// normalEstimation.jl:161-165 (PCD) and :172-176 (SLIC: the dominant PCD axis only)
template <class T> __device__ inline void normal_pcd(const Grid &g, const T *f, long I, T *n) {
    for (int d = 0; d < g.D; ++d) n[d] = f[I - g.st[d]] - f[I + g.st[d]];
}
template <class T> __device__ inline void normal_slic(const Grid &g, const T *f, long I, T *n) {
    normal_pcd(g, f, I, n);
    int dom = argabsmax(g.D, n);
    for (int i = 0; i < g.D; ++i) n[i] = i == dom ? tsign(n[i]) : T(0);
}
template <class T> __device__ inline void normal_scheme(int scheme, const Grid &g, const T *f, long I, T *n) {
    if (scheme == 1) normal_column(g, f, I, n);
    else if (scheme == 2) normal_wy(g, f, I, n);
    else if (scheme == 3) normal_myc(g, f, I, n);
    else if (scheme == 4) normal_pcd(g, f, I, n);
    else if (scheme == 5) normal_youngs(g, f, I, n);
    else if (scheme == 6) normal_slic(g, f, I, n);
    else normal_wh(g, f, I, n);
}

// ---------------------------------------------------------------------------
// limiters (flow.jl:5-15)
// ---------------------------------------------------------------------------
template <class T> __device__ inline T limiter(int lim, T u, T c, T d) {
    if (lim == 1) return median3((3 * c - u) / 2, c, (c + d) / 2);  // minmod
    if (lim == 2) return c;                                         // upwind
    if (lim == 3) {                                                 // vanAlbada1
        T al = c - u, be = d - c;
        T w = (al == be && al == T(0)) ? T(0) : (al + be) / (al * al + be * be);
        return c + (tmax(al * be, T(0)) * w) / 2;
    }
    if (lim == 4) {  // superbee = Sweby(γ = 2)
        T s = tsign(d - u), gm = 2;
        if (c <= tmin(u, d) || c >= tmax(u, d)) return c;
        return c + (s * tmax(tmax(T(0), tmin((s * gm) * (c - u), s * (d - c))), tmin(s * (c - u), (s * gm) * (d - c)))) / 2;
    }
    if (lim == 5 || lim == 6 || lim == 7) {  // TVDcen, TVDdown, Sweby(γ = 1.5)
        T s = tsign(d - u);
        if (c <= tmin(u, d) || c >= tmax(u, d)) return c;
        if (lim == 5) return c + s * tmin(s * (c - u), (s * (d - c)) / 2);
        if (lim == 6) return c + s * tmin(s * (c - u), s * (d - c));
        T gm = T(1.5);
        return c + (s * tmax(tmax(T(0), tmin((s * gm) * (c - u), s * (d - c))), tmin(s * (c - u), (s * gm) * (d - c)))) / 2;
    }
    return median3(((7 * c + d) - 2 * u) / 6, c, median3(2 * c - u, c, d));  // Koren
}

}  // namespace ia
