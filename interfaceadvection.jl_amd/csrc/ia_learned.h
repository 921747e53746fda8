// ia_learned.h -- the learned Poisson correction: a five-layer 3x3x3
// convolutional network (3 -> 16 -> 16 -> 16 -> 16 -> 2 channels) evaluated
// with BF16 matrix-core (MFMA) products and FP32 accumulation, trained on
// the device from the solver's own residual.  It has no counterpart in the
// reference: it proposes a correction to the pressure warm start, which is
// kept only when it lowers the FP64 residual; the flexible CG then solves to
// the requested tolerance either way.
//
// Tensors are channels-last BF16 over the ghost-inclusive level-0 grid,
// 16 channels per voxel (unused channels hold zero); ghost voxels are never
// written and stay zero, which is the zero padding of the convolution.
// Weights are W[layer][tap][ci][co] in FP32 (tap = (dz+1)*9 + (dy+1)*3 + dx+1),
// mirrored into BF16 MFMA operand order for the forward and the transposed
// convolutions.
#pragma once
#include "ia_poisson.h"

namespace ia {

constexpr int CNN_C = 16;                         // channels of every tensor
constexpr int CNN_L = 5;                          // layers
constexpr int CNN_T = 27;                         // taps
constexpr int CNN_W = CNN_T * CNN_C * CNN_C;      // weights per layer
constexpr int CNN_NPAR = CNN_L * (CNN_W + CNN_C); // weights + biases
constexpr int CNN_IN = 3, CNN_OUT = 2;            // real input / output channels

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
// gfx950's v_mfma_f32_16x16x32_bf16 contracts 32 = 2 taps × 16 input channels
// per product: the 27 taps go in CNN_P pairs, the last pair's second tap
// with zero weights.  Lane l holds, of the pair's K = 32, the eight values
// k = 8·(l/16) .. +7 of A (row l%16) and of B (column l%16): tap 2p + (l/32),
// channels 8·((l/16)&1) .. +7.  (The hardware contracts A's and B's k in
// the same order, so the product sums exactly these 32 terms.)
constexpr int CNN_P = (CNN_T + 1) / 2;
__device__ inline bf16x8 ld_bf16x8(const unsigned short *p) { return *(const bf16x8 *)p; }
__device__ inline int pair_tap(int p, int kg) { return min(2 * p + (kg >> 1), CNN_T - 1); }

__device__ inline unsigned short f2bf(float f) {  // round to nearest even
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// one 16-voxel row segment (x0 .. x0+15 at y, z) per tile
struct CnnTiles {
    Grid g;
    int tx;      // tiles per row = N0 / 16
    long ntile;  // tx * N1 * N2
};
__device__ inline long tile_base(const CnnTiles &t, long tile) {
    long r = tile / t.tx;
    int xt = (int)(tile % t.tx);
    int y = 1 + (int)(r % (t.g.Ng[1] - 2));
    int z = 1 + (int)(r / (t.g.Ng[1] - 2));
    return (long)(1 + 16 * xt) + t.g.st[1] * y + t.g.st[2] * z;
}
__device__ inline long tap_off(const Grid &g, int t) {
    int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    return (long)dx + g.st[1] * dy + g.st[2] * dz;
}

// the ReLU derivative of the backward pass, fused into the transposed
// convolution's store: dH = dY ⊙ (H > 0) with H the layer's input (mask)
__device__ inline void cnn_mask(const unsigned short *__restrict__ mask, long o, unsigned short v[4]) {
    if (!mask) return;
    uint2 h = *(const uint2 *)(mask + o);
    unsigned short hv[4] = {(unsigned short)(h.x & 0xffffu), (unsigned short)(h.x >> 16),
                            (unsigned short)(h.y & 0xffffu), (unsigned short)(h.y >> 16)};
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (!(bf2f(hv[q]) > 0.f)) v[q] = 0;
}

// Y = act(conv(X, W) + b) on interior voxels (then masked: see cnn_mask).  Per wave and tile, the 27
// taps are CNN_P MFMA 16x16x32 products D[co][vox] += A[co][k] B[k][vox]
// over a tap pair, A the pair's weights (lane l: co = l%16) and B the input
// rows (lane l: voxel l%16, eight channels of its tap, one 16-byte load);
// lane l then holds channels 4(l/16)..+3 of voxel l%16 and stores them as
// 8 bytes.
__global__ void __launch_bounds__(256) k_cnn_conv(CnnTiles t, const unsigned short *__restrict__ X,
                                                  unsigned short *__restrict__ Y,
                                                  const bf16x8 *__restrict__ Wf,  // [CNN_P][64]
                                                  const float *__restrict__ bias, int relu,
                                                  const unsigned short *__restrict__ mask) {
    const int lane = threadIdx.x & 63;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long nwave = ((long)gridDim.x * blockDim.x) >> 6;
    const int j = lane & 15, kg = lane >> 4, h8 = 8 * (kg & 1);
    bf16x8 a[CNN_P];
#pragma unroll
    for (int p = 0; p < CNN_P; ++p) a[p] = Wf[p * 64 + lane];
    float b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) b[q] = bias[4 * kg + q];
    long off[CNN_P];
#pragma unroll
    for (int p = 0; p < CNN_P; ++p) off[p] = tap_off(t.g, pair_tap(p, kg));
    for (long tile = wave; tile < t.ntile; tile += nwave) {
        long V = tile_base(t, tile) + j;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int p = 0; p < CNN_P; ++p) {
            bf16x8 bx = ld_bf16x8(X + (V + off[p]) * CNN_C + h8);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], bx, acc, 0, 0, 0);
        }
        unsigned short o[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float v = acc[q] + b[q];
            if (relu && v < 0.f) v = 0.f;
            o[q] = f2bf(v);
        }
        cnn_mask(mask, V * CNN_C + 4 * kg, o);
        *(uint2 *)(Y + V * CNN_C + 4 * kg) = make_uint2(o[0] | ((unsigned)o[1] << 16), o[2] | ((unsigned)o[3] << 16));
    }
}

// The same convolution with the input staged through LDS: a block owns a
// 16 (x) by 4 (y) column of voxels and walks CNN_ZB planes in z.  A ring of
// four LDS planes holds the input rows y-1..y+4, x-1..x+16 of planes z-1,
// z, z+1 and the prefetched z+2; every tap pair's MFMA operand is one
// 16-byte LDS read, and each input plane is read from memory once per block.
constexpr int CNN_ZB = 32;
struct CnnBlocks {
    Grid g;
    int tx, ty, tz;  // x tiles of 16, y tiles of 4, z tiles of CNN_ZB
};
__global__ void __launch_bounds__(256) k_cnn_conv_lds(CnnBlocks t, const unsigned short *__restrict__ X,
                                                      unsigned short *__restrict__ Y,
                                                      const bf16x8 *__restrict__ Wf, const float *__restrict__ bias,
                                                      int relu, const unsigned short *__restrict__ mask) {
    constexpr int PX = 18, PY = 6, PV = PX * PY;  // voxels per staged plane
    __shared__ __align__(16) unsigned short pl[4][PV * CNN_C];
    const Grid &g = t.g;
    long b = blockIdx.x;
    int bx = (int)(b % t.tx), by = (int)((b / t.tx) % t.ty), bz = (int)(b / ((long)t.tx * t.ty));
    int x0 = 1 + 16 * bx, y0 = 1 + 4 * by, zlo = 1 + CNN_ZB * bz;
    int zhi = min(zlo + CNN_ZB - 1, g.Ng[2] - 2);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j = lane & 15, kg = lane >> 4, h8 = 8 * (kg & 1);
    // staging: thread s < 216 moves 16 bytes (half a voxel's channels) of a plane
    const int s = threadIdx.x;
    const bool stg = s < 2 * PV;
    const int sv = s >> 1, sh = s & 1;
    const long sI = stg ? (long)(x0 - 1 + sv % PX) + g.st[1] * (y0 - 1 + sv / PX) : 0;
    auto fetch = [&](int z) -> uint4 {
        if (!stg) return make_uint4(0u, 0u, 0u, 0u);
        return *(const uint4 *)(X + (sI + g.st[2] * z) * CNN_C + 8 * sh);
    };
    auto put = [&](int slot, uint4 v) {
        if (stg) *(uint4 *)(&pl[slot][sv * CNN_C + 8 * sh]) = v;
    };
    bf16x8 a[CNN_P];
#pragma unroll
    for (int p = 0; p < CNN_P; ++p) a[p] = Wf[p * 64 + lane];
    // each lane's tap of every pair: the (x, y) offset in the plane and the ring slot shift
    int loff[CNN_P], lz[CNN_P];
#pragma unroll
    for (int p = 0; p < CNN_P; ++p) {
        int q = pair_tap(p, kg);
        int dx = q % 3 - 1, dy = (q / 3) % 3 - 1;
        loff[p] = ((w + 1 + dy) * PX + (j + 1 + dx)) * CNN_C + h8;
        lz[p] = q / 9 - 1;
    }
    float bb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) bb[q] = bias[4 * kg + q];
    put((zlo - 1) & 3, fetch(zlo - 1));
    put(zlo & 3, fetch(zlo));
    put((zlo + 1) & 3, fetch(zlo + 1));
    for (int z = zlo; z <= zhi; ++z) {
        __syncthreads();
        uint4 nxt = make_uint4(0u, 0u, 0u, 0u);
        if (z + 2 <= zhi + 1) nxt = fetch(z + 2);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int p = 0; p < CNN_P; ++p) {
            bf16x8 bx = ld_bf16x8(&pl[(z + lz[p]) & 3][loff[p]]);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[p], bx, acc, 0, 0, 0);
        }
        unsigned short o[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float v = acc[q] + bb[q];
            if (relu && v < 0.f) v = 0.f;
            o[q] = f2bf(v);
        }
        long V = (long)(x0 + j) + g.st[1] * (y0 + w) + g.st[2] * z;
        cnn_mask(mask, V * CNN_C + 4 * kg, o);
        *(uint2 *)(Y + V * CNN_C + 4 * kg) = make_uint2(o[0] | ((unsigned)o[1] << 16), o[2] | ((unsigned)o[3] << 16));
        if (z + 2 <= zhi + 1) put((z + 2) & 3, nxt);
    }
}

// MFMA operand order of one layer's weights, per tap pair p: forward
// A[co][ci] (lane l: co = l%16, tap 2p + l/32, ci = 8((l/16)&1) .. +7), and
// the transposed convolution's A[ci][co] of the flipped tap; the missing
// 28th tap is zero
__global__ void k_cnn_frag(const float *__restrict__ W, bf16x8 *__restrict__ Ff, bf16x8 *__restrict__ Fb) {
    int p = blockIdx.x, lane = threadIdx.x;  // CNN_P blocks of 64
    int j = lane & 15, kg = lane >> 4, h8 = 8 * (kg & 1);
    int t = 2 * p + (kg >> 1);
    union {
        unsigned short u[8];
        bf16x8 v;
    } f, r;
    for (int e = 0; e < 8; ++e) {
        bool ok = t < CNN_T;
        f.u[e] = ok ? f2bf(W[(t * CNN_C + h8 + e) * CNN_C + j]) : 0;                  // W[t][ci=h8+e][co=j]
        r.u[e] = ok ? f2bf(W[((CNN_T - 1 - t) * CNN_C + j) * CNN_C + h8 + e]) : 0;    // W[26-t][ci=j][co=h8+e]
    }
    Ff[p * 64 + lane] = f.v;
    Fb[p * 64 + lane] = r.v;
}

// weight and bias gradient partials of one layer:
// gW[t][ci][co] = Σ_v X[v+off_t][ci] dY[v][co], gb[co] = Σ_v dY[v][co].
// Each tap is a 16x16 product over voxels, v_mfma_f32_16x16x32_bf16 with
// A[ci][v] = X[v+off_t][ci] and B[v][co] = dY[v][co] over 32 voxels (one
// x-row segment).  The segment's nine input rows (y-1..y+1, z-1..z+1, x-1 ..
// x+32) and its output gradients are staged in LDS channel-major, the input
// rows once per x-shift (-1, 0, +1), so that a lane's eight voxels of one
// channel are one aligned 16-byte read.  Wave w accumulates taps w, w+4, ...;
// wave 3's last slot multiplies dY by ones, which is the bias gradient.  A
// product of two BF16 values is exact in FP32; the sums run in FP32 per
// block, then over the blocks in FP64 (k_cnn_wgrad_sum).  The next
// segment's loads are issued before the current one's products.
constexpr int WG_SEG = 32, WG_ROW = 40, WG_K = 7;  // voxels per segment, LDS row (padded), slots per wave
struct CnnSegs {
    Grid g;
    int sx;      // segments per x-row: ceil(N0 / 32)
    long nseg;   // sx * N1 * N2
};
__device__ inline void wg_unpack(uint4 v, unsigned short u[8]) {
    unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        u[2 * k] = (unsigned short)(w[k] & 0xffffu);
        u[2 * k + 1] = (unsigned short)(w[k] >> 16);
    }
}
__global__ void __launch_bounds__(256) k_cnn_wgrad(CnnSegs t, const unsigned short *__restrict__ X,
                                                   const unsigned short *__restrict__ dY,
                                                   float *__restrict__ part /* [nblk][27*256 + 16] */) {
    __shared__ __align__(16) unsigned short xt[9][3][CNN_C][WG_ROW];
    __shared__ __align__(16) unsigned short dyt[CNN_C][WG_ROW];
    constexpr int NST = 9 * (WG_SEG + 2) * 2;  // staged 16-byte pieces of the input rows
    constexpr int NIT = (NST + 255) / 256;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m = lane & 15, kg = lane >> 4;
    const Grid &g = t.g;
    const int nx = g.Ng[0] - 2;
    f32x4 acc[WG_K];
#pragma unroll
    for (int k = 0; k < WG_K; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};
    const unsigned short one[8] = {0x3F80, 0x3F80, 0x3F80, 0x3F80, 0x3F80, 0x3F80, 0x3F80, 0x3F80};  // BF16 1.0
    const bf16x8 ones = ld_bf16x8(one);
    uint4 xv[NIT], dv = make_uint4(0u, 0u, 0u, 0u);
    auto fetch = [&](long s) {
        long r = s / t.sx;
        int x0 = 1 + WG_SEG * (int)(s % t.sx);
        int y = 1 + (int)(r % (g.Ng[1] - 2)), z = 1 + (int)(r / (g.Ng[1] - 2));
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            int e = threadIdx.x + 256 * it;
            xv[it] = make_uint4(0u, 0u, 0u, 0u);
            if (e < NST) {
                int h = e & 1, vv = (e >> 1) % (WG_SEG + 2), row = (e >> 1) / (WG_SEG + 2);
                int x = x0 - 1 + vv;
                if (x <= g.Ng[0] - 1) {
                    long V = (long)x + g.st[1] * (y + row % 3 - 1) + g.st[2] * (z + row / 3 - 1);
                    xv[it] = *(const uint4 *)(X + V * CNN_C + 8 * h);
                }
            }
        }
        dv = make_uint4(0u, 0u, 0u, 0u);
        if (threadIdx.x < 2 * WG_SEG) {
            int h = threadIdx.x & 1, v = threadIdx.x >> 1;
            if (x0 + v <= nx) {
                long V = (long)(x0 + v) + g.st[1] * y + g.st[2] * z;
                dv = *(const uint4 *)(dY + V * CNN_C + 8 * h);
            }
        }
    };
    long s = blockIdx.x;
    if (s < t.nseg) fetch(s);
    for (; s < t.nseg; s += gridDim.x) {
        __syncthreads();  // the previous segment's products have read the LDS
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            int e = threadIdx.x + 256 * it;
            if (e < NST) {
                int h = e & 1, vv = (e >> 1) % (WG_SEG + 2), row = (e >> 1) / (WG_SEG + 2);
                unsigned short u[8];
                wg_unpack(xv[it], u);
#pragma unroll
                for (int sh = 0; sh < 3; ++sh) {
                    int v = vv - sh;  // shift sh holds X at x0 + v + sh - 1
                    if (v >= 0 && v < WG_SEG) {
#pragma unroll
                        for (int j = 0; j < 8; ++j) xt[row][sh][8 * h + j][v] = u[j];
                    }
                }
            }
        }
        if (threadIdx.x < 2 * WG_SEG) {
            int h = threadIdx.x & 1, v = threadIdx.x >> 1;
            unsigned short u[8];
            wg_unpack(dv, u);
#pragma unroll
            for (int j = 0; j < 8; ++j) dyt[8 * h + j][v] = u[j];
        }
        __syncthreads();
        if (s + gridDim.x < t.nseg) fetch(s + gridDim.x);
        const bf16x8 b = ld_bf16x8(&dyt[m][8 * kg]);
#pragma unroll
        for (int k = 0; k < WG_K; ++k) {
            int q = w + 4 * k;  // tap q: row q / 3 = (dz+1)*3 + (dy+1), shift q % 3 = dx + 1
            if (q < CNN_T) {
                const bf16x8 a = ld_bf16x8(&xt[q / 3][q % 3][m][8 * kg]);
                acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[k], 0, 0, 0);
            } else if (q == CNN_T) {
                acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ones, b, acc[k], 0, 0, 0);
            }
        }
    }
    // lane l holds D[ci = 4(l/16) + j][co = l%16]
    float *P = part + (long)blockIdx.x * (CNN_W + CNN_C);
#pragma unroll
    for (int k = 0; k < WG_K; ++k) {
        int q = w + 4 * k;
        if (q < CNN_T) {
#pragma unroll
            for (int j = 0; j < 4; ++j) P[(q * CNN_C + 4 * kg + j) * CNN_C + m] = acc[k][j];
        } else if (q == CNN_T && kg == 0) {
            P[CNN_W + m] = acc[k][0];  // every row of the ones product is Σ_v dY[v][co]
        }
    }
}
// fixed-order sum of the per-block partials: a block owns 16 outputs, its
// thread (o, k) sums blocks k, k+16, ... of output o in FP64, and thread
// (o, 0) adds the 16 group sums in order
__global__ void __launch_bounds__(256) k_cnn_wgrad_sum(const float *__restrict__ part, long nblk,
                                                       float *__restrict__ gW, float *__restrict__ gb) {
    __shared__ double sh[16][17];
    const int o = threadIdx.x & 15, k = threadIdx.x >> 4;
    const int i = blockIdx.x * 16 + o;
    double s = 0;
    if (i < CNN_W + CNN_C)
        for (long b = k; b < nblk; b += 16) s += part[b * (CNN_W + CNN_C) + i];
    sh[k][o] = s;
    __syncthreads();
    if (k == 0 && i < CNN_W + CNN_C) {
        double t = 0;
        for (int q = 0; q < 16; ++q) t += sh[q][o];
        if (i < CNN_W) gW[i] = (float)t;
        else gb[i - CNN_W] = (float)t;
    }
}

// Adam on every parameter; frozen entries (inputs that are always zero,
// outputs that are never used) keep a zero gradient and stay zero
__global__ void k_cnn_adam(float *__restrict__ w, const float *__restrict__ gw, float *__restrict__ m,
                           float *__restrict__ v, int n, float lr, float b1, float b2, float c1, float c2,
                           float eps) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = gw[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------------
// solver side: features, applying the correction, the loss gradient
// ---------------------------------------------------------------------------
// RMS of the interior of a FP64 field, from partials of Σa²
template <class T> struct CnnSolve {
    FcgV<T> f;
    CnnTiles t;
    const T *fv;  // volume fraction
};
// features: ch0 = f, ch1 = z/s_z, ch2 = r/s_r (r = z - A x0), other channels 0
template <class T>
__global__ void __launch_bounds__(256) k_cnn_features(CnnSolve<T> a, unsigned short *__restrict__ X0,
                                                      const double *part_zz, const double *part_rr) {
    double sz = sqrt(block_total(part_zz, a.f.t.nblk) / (double)a.f.ncell);
    double sr = sqrt(block_total(part_rr, a.f.t.nblk) / (double)a.f.ncell);
    sz = sz > 0 ? sz : 1.0;
    sr = sr > 0 ? sr : 1.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.f.S->sload = sr;
    COL_LOOP(a.f, {
        unsigned short o[CNN_C];
        for (int c = 0; c < CNN_C; ++c) o[c] = 0;
        o[0] = f2bf((float)a.fv[I]);
        o[1] = f2bf((float)(a.f.z[I] / sz));
        o[2] = f2bf((float)(a.f.r[I] / sr));
        uint4 *d = (uint4 *)(X0 + I * CNN_C);
        d[0] = make_uint4(o[0] | ((unsigned)o[1] << 16), o[2] | ((unsigned)o[3] << 16), o[4] | ((unsigned)o[5] << 16),
                          o[6] | ((unsigned)o[7] << 16));
        d[1] = make_uint4(o[8] | ((unsigned)o[9] << 16), o[10] | ((unsigned)o[11] << 16),
                          o[12] | ((unsigned)o[13] << 16), o[14] | ((unsigned)o[15] << 16));
    })
}
// the 2x2x2 block average of channel 1 at voxel I (the coarse correction);
// block-averaging is symmetric, so the same operator serves the backward pass
__device__ inline float blockavg(const Grid &g, const unsigned short *Y, int ch, const int p[3]) {
    int b0 = ((p[0] - 1) & ~1) + 1, b1 = ((p[1] - 1) & ~1) + 1, b2 = ((p[2] - 1) & ~1) + 1;
    float s = 0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx)
                s += bf2f(Y[((long)(b0 + dx) + g.st[1] * (b1 + dy) + g.st[2] * (b2 + dz)) * CNN_C + ch]);
    return s * 0.125f;
}
// x_keep = x; x += iD (2/3 r + s_r (out0 + avg(out1))): the proposed warm
// start is a damped Jacobi step on the residual r = z - A x plus the
// network's correction of it, in the same Jacobi-scaled units
template <class T>
__global__ void __launch_bounds__(256) k_cnn_apply(CnnSolve<T> a, const unsigned short *__restrict__ Y,
                                                   double *__restrict__ xkeep) {
    double sr = a.f.S->sload;
    COL_LOOP(a.f, {
        float d = bf2f(Y[I * CNN_C]) + blockavg(a.f.g, Y, 1, p);
        double id = (double)a.f.iD[I];
        xkeep[I] = a.f.x[I];
        a.f.x[I] += id * ((2.0 / 3.0) * a.f.r[I] + sr * (double)d);
    })
}
// restore the warm start (the proposal did not lower the residual)
template <class T>
__global__ void __launch_bounds__(256) k_cnn_restore(CnnSolve<T> a, const double *__restrict__ xkeep) {
    COL_LOOP(a.f, { a.f.x[I] = xkeep[I]; })
}
// loss ‖r1‖²/‖r0‖² with r1 = z - A x'; d loss/d x' = -2 A r1 / ‖r0‖² (A is
// symmetric); the output gradient is that times s_r iD (channel 0) and the
// block average of it (channel 1)
template <class T>
__global__ void __launch_bounds__(256) k_cnn_loss_grad(CnnSolve<T> a, float *__restrict__ gtmp, double r0) {
    double sc = -2.0 * a.f.S->sload / (r0 > 0 ? r0 : 1.0);
    const Grid &g = a.f.g;
    COL_LOOP(a.f, {
        double Ar = g.D == 3 ? mult64_d<3>(g, p, I, a.f.L, a.f.r) : mult64_d<2>(g, p, I, a.f.L, a.f.r);
        gtmp[I] = (float)(sc * (double)a.f.iD[I] * Ar);
    })
}
template <class T>
__global__ void __launch_bounds__(256) k_cnn_loss_grad2(CnnSolve<T> a, unsigned short *__restrict__ dY,
                                                        const float *__restrict__ gtmp) {
    const Grid &g = a.f.g;
    COL_LOOP(a.f, {
        int b0 = ((p[0] - 1) & ~1) + 1, b1 = ((p[1] - 1) & ~1) + 1, b2 = ((p[2] - 1) & ~1) + 1;
        float s = 0;
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) s += gtmp[(long)(b0 + dx) + g.st[1] * (b1 + dy) + g.st[2] * (b2 + dz)];
        unsigned short o[CNN_C];
        for (int c = 0; c < CNN_C; ++c) o[c] = 0;
        o[0] = f2bf(gtmp[I]);
        o[1] = f2bf(s * 0.125f);
        uint4 *d = (uint4 *)(dY + I * CNN_C);
        d[0] = make_uint4(o[0] | ((unsigned)o[1] << 16), 0u, 0u, 0u);
        d[1] = make_uint4(0u, 0u, 0u, 0u);
    })
}

// test entry points: interior planes (x fastest, channel planes last) <-> tensor
__global__ void __launch_bounds__(256) k_cnn_pack(Grid g, const float *__restrict__ src, int nch,
                                                  unsigned short *__restrict__ X) {
    long n0 = g.Ng[0] - 2, n1 = g.Ng[1] - 2, n2 = g.Ng[2] - 2, n = n0 * n1 * n2;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long x = i % n0, y = (i / n0) % n1, z = i / (n0 * n1);
    long V = (x + 1) + g.st[1] * (y + 1) + g.st[2] * (z + 1);
    for (int c = 0; c < CNN_C; ++c) X[V * CNN_C + c] = c < nch ? f2bf(src[c * n + i]) : (unsigned short)0;
}
__global__ void __launch_bounds__(256) k_cnn_unpack(Grid g, const unsigned short *__restrict__ X, int nch,
                                                    float *__restrict__ dst) {
    long n0 = g.Ng[0] - 2, n1 = g.Ng[1] - 2, n2 = g.Ng[2] - 2, n = n0 * n1 * n2;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long x = i % n0, y = (i / n0) % n1, z = i / (n0 * n1);
    long V = (x + 1) + g.st[1] * (y + 1) + g.st[2] * (z + 1);
    for (int c = 0; c < nch; ++c) dst[c * n + i] = bf2f(X[V * CNN_C + c]);
}

}  // namespace ia
