// Immersed-boundary coupling on gfx950.
//
// Spreading: ibm_st1 = ibm_st1a (zero) + ibm_st1b (4x4x4 cosine-delta
// spread) + ibm_st1c (fold of halo contributions onto periodic images)
// (ibm.cu:27-131, 214-369).  Interpolation + advection: ibm_st2a
// (ibm.cu:372-639).
//
// The stencil of a vertex is selected exactly as the reference does in its
// padded frame (xtr = X + (nx_p-1) dx/2, first node floor(xtr/dx)-1, nodes at
// ix_p*dx) and then mapped to the unpadded lattice by ix = ix_p - 2 with
// periodic wrapping, which is what the halo fold achieves; nodes beyond the
// walls are dropped (spread) or contribute zero velocity (interpolation).
// Vertex positions and stencil selection are FP64 in every precision mode;
// weights, lattice fields (u, F) and nodal forces use the arithmetic type T.
//
// Both directions work on lattice tiles of 16x8x8 nodes.  After each change
// of the positions the vertices are binned by the tile of their first stencil
// node (counting sort with wave-aggregated counters).
//   - Spreading: one work group per tile takes the vertices of its own bin
//     and of its 7 lower neighbours, keeps those whose stencil meets the tile
//     (per-wave queue), and accumulates their contributions into the tile in
//     LDS, one lane per vertex, with 64-bit integer LDS atomics on fixed-point
//     values (no global atomics).  The power-of-two scale per tile is fixed
//     by max |f| of the call and the tile's candidate count, so sums cannot
//     overflow and do not depend on the order of the additions (bitwise
//     reproducible); 64-bit integer LDS atomics also run about 4x faster
//     than FP32 LDS atomics here.  The tile then writes all of its nodes,
//     zeros included, which replaces the zeroing pass of ibm_st1a.
//   - Interpolation: one work group per tile stages the velocity of the tile
//     plus a 3/4-node halo in LDS once and runs the four frozen-field RK
//     stages of its vertices from LDS, one lane per vertex.
#include "common.hpp"

namespace lbmibm {
namespace {

struct IbmGeom {
    int nx, ny, nz;     // unpadded
    int nxp, nyp, nzp;  // padded
    int walls;          // 1: y beyond [0,ny) is wall
    double dx;
};

struct TileGeom {
    int ntx, nty, ntz, ntiles;
};

__device__ __forceinline__ bool wrap_index(int &i, int n) {
    if (i < 0) i += n;
    else if (i >= n) i -= n;
    return (unsigned)i < (unsigned)n;
}

__device__ __forceinline__ int pmod(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}

// sin and cos of (pi/2) t for t in [0, 1).  FP32: Taylor series of both to
// degree 13 / 12 in x = (pi/2) t (truncation below 1e-9 on [0, pi/2), so the
// result is within a few ulp of the correctly rounded value), evaluated by
// Horner in x^2 with no range reduction.  FP64: the library sincos.
__device__ __forceinline__ void sincos_quarter(float t, float *s, float *c) {
    const float x = 1.57079632679489662f * t, x2 = x * x;
    float ps = -1.0f / 6227020800.0f;
    ps = ps * x2 + 1.0f / 39916800.0f;
    ps = ps * x2 - 1.0f / 362880.0f;
    ps = ps * x2 + 1.0f / 5040.0f;
    ps = ps * x2 - 1.0f / 120.0f;
    ps = ps * x2 + 1.0f / 6.0f;
    // sin x = x (1 - x^2 (1/6 - x^2 (1/120 - ...)))
    *s = x - x * x2 * ps;
    float pc = 1.0f / 479001600.0f;
    pc = pc * x2 - 1.0f / 3628800.0f;
    pc = pc * x2 + 1.0f / 40320.0f;
    pc = pc * x2 - 1.0f / 720.0f;
    pc = pc * x2 + 1.0f / 24.0f;
    pc = pc * x2 - 0.5f;
    *c = 1.0f + x2 * pc;
}
__device__ __forceinline__ void sincos_quarter(double t, double *s, double *c) { sincos(M_PI / 2 * t, s, c); }

// Weights 1 + cos(pi |x_node - xtr| / (2h)) of the 4 nodes from
// floor(xtr/h)-1 (ibm.cu:103-121).  With t = frac(xtr/h) the distances are
// 1+t, t, 1-t, 2-t, so the four cosines are -sin, cos, sin, -cos of pi t/2:
// one sincos of an argument in [0, pi/2) per axis, evaluated in W.
// xtr / h; h == 1 (lattice units) skips the FP64 division, with the same result
__device__ __forceinline__ double over_h(double xtr, double h) { return h == 1.0 ? xtr : xtr / h; }

template <typename W>
__device__ __forceinline__ void stencil(double xtr, double h, int &first, W w[4], bool unit_spacing) {
    const double q = unit_spacing ? xtr : over_h(xtr, h);
    const double fl = floor(q);
    first = (int)fl - 1;
    const W t = (W)(q - fl);
    W sn, cs;
    sincos_quarter(t, &sn, &cs);
    w[0] = W(1) - sn;
    w[1] = W(1) + cs;
    w[2] = W(1) + sn;
    w[3] = W(1) - cs;
}

// First stencil node (padded frame, as stencil()) and fraction t = frac(xtr/h).
template <typename W>
__device__ __forceinline__ void stencil_first(double xtr, double h, int &first, W &t) {
    const double q = over_h(xtr, h);
    const double fl = floor(q);
    first = (int)fl - 1;
    t = (W)(q - fl);
}

// Weights of the 4 nodes from the fraction t (as stencil()).
template <typename W> __device__ __forceinline__ void stencil_weights(W t, W w[4]) {
    W sn, cs;
    sincos_quarter(t, &sn, &cs);
    w[0] = W(1) - sn;
    w[1] = W(1) + cs;
    w[2] = W(1) + sn;
    w[3] = W(1) - cs;
}

// Unpadded first stencil nodes (unwrapped) and weights at a position.
template <typename W>
__device__ __forceinline__ void stencil3(const IbmGeom &g, double x, double y, double z, bool unit, int &x1, int &y1,
                                         int &z1, W wx[4], W wy[4], W wz[4]) {
    if (!unit) {
        stencil<W>(x + (double)(g.nxp - 1) * g.dx / 2.0, g.dx, x1, wx, false);
        stencil<W>(y + (double)(g.nyp - 1) * g.dx / 2.0, g.dx, y1, wy, false);
        stencil<W>(z + (double)(g.nzp - 1) * g.dx / 2.0, g.dx, z1, wz, false);
    } else {  // RK2 / Euler branches use unit spacing (ibm.cu:524-551, 581-608)
        stencil<W>(x + (double)(g.nxp - 1) / 2.0, g.dx, x1, wx, true);
        stencil<W>(y + (double)(g.nyp - 1) / 2.0, g.dx, y1, wy, true);
        stencil<W>(z + (double)(g.nzp - 1) / 2.0, g.dx, z1, wz, true);
    }
    x1 -= 2;
    y1 -= 2;
    z1 -= 2;
}

// Mask of the 4 stencil nodes first..first+3 (periodic when wrap) that fall
// in [o, o+len) of an axis of n nodes; bit k for node first+k.
__device__ __forceinline__ unsigned axis_mask(int first, int o, int len, int n, bool wrap) {
    unsigned m = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int a = first + k;
        if (wrap) {
            if (a < 0) a += n;
            else if (a >= n) a -= n;
        }
        if (a >= o && a < o + len && a >= 0 && a < n) m |= 1u << k;
    }
    return m;
}

// Tiles along one axis whose nodes can hold the first stencil node of a
// vertex reaching [o, o+len): the tile itself and the tiles of nodes o-3..o-1
// (periodic when wrap).  Returns the number of distinct tiles (1..4).
__device__ __forceinline__ int source_tiles(int t, int o, int n, int tile, bool wrap, int out[4]) {
    int cnt = 1;
    out[0] = t;
    for (int d = 1; d <= 3; d++) {
        int a = o - d;
        if (a < 0) {
            if (!wrap) break;
            a = ((a % n) + n) % n;
        }
        const int s = a / tile;
        bool seen = false;
        for (int i = 0; i < cnt; i++) seen |= out[i] == s;
        if (!seen) out[cnt++] = s;
    }
    return cnt;
}

// ---- binning ---------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void k_bin_count(const double *__restrict__ xs, const T *__restrict__ cf, int nv,
                                                   BinGeom bg, int *__restrict__ home,
                                                   int *__restrict__ tile_count, float *__restrict__ blk_fmax) {
    __shared__ float shm[4];
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = v < nv;
    int t = 0;
    float fm = 0.0f;
    if (on) {
        t = home_tile(bg, xs[v], xs[nv + v], xs[2 * (size_t)nv + v]);
        home[v] = t;
        fm = fmaxf(fabsf((float)cf[v]), fmaxf(fabsf((float)cf[nv + v]), fabsf((float)cf[2 * (size_t)nv + v])));
        fm *= 1.0001f;  // margin for the float rounding of |f|
    }
    (void)wave_aggregated_add(tile_count, t, on);
    // per-block max |f| (one store per block; k_tile_scan reduces them)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) fm = fmaxf(fm, __shfl_xor(fm, o));
    if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = fm;
    __syncthreads();
    if (threadIdx.x == 0) blk_fmax[blockIdx.x] = fmaxf(fmaxf(shm[0], shm[1]), fmaxf(shm[2], shm[3]));
}

// Heavy-first order of the tiles: a counting sort on a 256-level key of the
// per-tile load, largest first.  The tile kernels take their tiles in this
// order, so the dispatcher starts the longest work groups first and fills the
// tail with light ones (longest-processing-time-first scheduling); the order
// within a key level is arbitrary and does not affect any result.
__device__ void heavy_first_order(const int *load, int n, int *hist, int *order) {
    __shared__ int maxload;
    if (threadIdx.x == 0) maxload = 0;
    for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    int m = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) m = max(m, load[i]);
    atomicMax(&maxload, m);
    __syncthreads();
    int shift = 0;
    while ((maxload >> shift) > 255) shift++;
    for (int i = threadIdx.x; i < n; i += blockDim.x) atomicAdd(&hist[255 - (load[i] >> shift)], 1);
    __syncthreads();
    if (threadIdx.x == 0) {  // exclusive scan of 256 bins
        int acc = 0;
        for (int k = 0; k < 256; k++) {
            const int c = hist[k];
            hist[k] = acc;
            acc += c;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) order[atomicAdd(&hist[255 - (load[i] >> shift)], 1)] = i;
    __syncthreads();
}

__global__ __launch_bounds__(1024) void k_tile_scan(const int *__restrict__ tile_count, int *__restrict__ tile_off,
                                                    int *__restrict__ tile_cur, int ntiles,
                                                    const float *__restrict__ blk_fmax, int nblk,
                                                    unsigned *__restrict__ absmax) {
    // exclusive scan over tiles in one work group, 1024 entries per round,
    // and the maximum of the per-block force maxima
    __shared__ int sh[1024];
    __shared__ float shm[1024];
    __shared__ int carry;
    float fm = 0.0f;
    for (int i = threadIdx.x; i < nblk; i += 1024) fm = fmaxf(fm, blk_fmax[i]);
    shm[threadIdx.x] = fm;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (threadIdx.x < o) shm[threadIdx.x] = fmaxf(shm[threadIdx.x], shm[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *absmax = __float_as_uint(shm[0]);
    for (int base = 0; base < ntiles; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < ntiles ? tile_count[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const int t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < ntiles) {
            tile_off[i] = carry + sh[threadIdx.x] - v;
            tile_cur[i] = 0;
        }
        __syncthreads();
        if (threadIdx.x == 1023) carry += sh[1023];
        __syncthreads();
    }
}

// The heavy-first tile orders of the spreading (by candidate count: own bin
// and the bins below on each axis) and of the interpolation (by own count).
// They only steer the dispatch order, and tile loads change slowly with the
// cells, so they are refreshed on every TILE_ORDER_PERIOD-th binning (a call
// counter on the device, so captured graphs refresh them too); any
// permutation is a valid order, the initial one is the identity.
constexpr int TILE_ORDER_PERIOD = 16;

__global__ __launch_bounds__(1024) void k_tile_order(const int *__restrict__ tile_count, int ntiles, IbmGeom g,
                                                     TileGeom tg, int *__restrict__ tile_cand,
                                                     int *__restrict__ order_spread, int *__restrict__ order_interp,
                                                     int *__restrict__ calls) {
    __shared__ int sh[256];
    __shared__ int call;
    if (threadIdx.x == 0) call = (*calls)++;
    __syncthreads();
    if (call % TILE_ORDER_PERIOD != 0) return;
    for (int t = threadIdx.x; t < ntiles; t += 1024) {
        const int tx = t % tg.ntx, ty = (t / tg.ntx) % tg.nty, tz = t / (tg.ntx * tg.nty);
        int sxs[4], sys[4], szs[4];
        const int nsx = source_tiles(tx, tx * TILE_X, g.nx, TILE_X, true, sxs);
        const int nsy = source_tiles(ty, ty * TILE_Y, g.ny, TILE_Y, !g.walls, sys);
        const int nsz = source_tiles(tz, tz * TILE_Z, g.nz, TILE_Z, true, szs);
        int c = 0;
        for (int iz = 0; iz < nsz; iz++)
            for (int iy = 0; iy < nsy; iy++)
                for (int ix = 0; ix < nsx; ix++) c += tile_count[sxs[ix] + tg.ntx * (sys[iy] + tg.nty * szs[iz])];
        tile_cand[t] = c;
    }
    __syncthreads();
    heavy_first_order(tile_cand, ntiles, sh, order_spread);
    heavy_first_order(tile_count, ntiles, sh, order_interp);
}

// Spreading data of a vertex, stored in bin order next to its first stencil
// node (node0: x and z wrapped, y + 3): the stencil fractions and the force.
template <typename T> struct SpreadRec {
    T t[3];
    T f[3];
};

template <typename T>
__global__ __launch_bounds__(256) void k_bin_scatter(const double *__restrict__ xs, const T *__restrict__ cf,
                                                     const int *__restrict__ home, int nv, IbmGeom g,
                                                     const int *__restrict__ tile_off, int *__restrict__ tile_cur,
                                                     int *__restrict__ home_list, short4 *__restrict__ node0,
                                                     SpreadRec<T> *__restrict__ rec) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = v < nv;
    const int t = on ? home[v] : 0;
    const int r = wave_aggregated_add(tile_cur, t, on);
    if (on) {
        const int pos = tile_off[t] + r;
        home_list[pos] = v;
        int x1, y1, z1;
        SpreadRec<T> sr;
        stencil_first<T>(xs[v] + (double)(g.nxp - 1) * g.dx / 2.0, g.dx, x1, sr.t[0]);
        stencil_first<T>(xs[nv + v] + (double)(g.nyp - 1) * g.dx / 2.0, g.dx, y1, sr.t[1]);
        stencil_first<T>(xs[2 * (size_t)nv + v] + (double)(g.nzp - 1) * g.dx / 2.0, g.dx, z1, sr.t[2]);
        x1 -= 2; y1 -= 2; z1 -= 2;  // unpadded
        int xw = x1, zw = z1;
        wrap_index(xw, g.nx);
        wrap_index(zw, g.nz);
        xw = min(max(xw, -30000), 30000);
        zw = min(max(zw, -30000), 30000);
        node0[pos] = make_short4((short)xw, (short)min(max(y1 + 3, -30000), 30000), (short)zw, 0);
        // the forces read in vertex order (coalesced), written in bin order
        sr.f[0] = cf[v]; sr.f[1] = cf[nv + v]; sr.f[2] = cf[2 * (size_t)nv + v];
        rec[pos] = sr;
    }
}

// The force part of the spreading records, in bin order, and max |f|, when
// the bins of unchanged positions are reused for new forces (a gather of cf).
template <typename T>
__global__ __launch_bounds__(256) void k_bin_forces(const int *__restrict__ home_list, const T *__restrict__ cf, int nv,
                                                    SpreadRec<T> *__restrict__ rec, unsigned *__restrict__ absmax) {
    __shared__ float shm[4];
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    float fm = 0.0f;
    if (pos < nv) {
        const int v = home_list[pos];
        const T f0 = cf[v], f1 = cf[nv + v], f2 = cf[2 * (size_t)nv + v];
        rec[pos].f[0] = f0; rec[pos].f[1] = f1; rec[pos].f[2] = f2;
        fm = fmaxf(fabsf((float)f0), fmaxf(fabsf((float)f1), fabsf((float)f2)));
        fm *= 1.0001f;  // margin for the float rounding of |f|
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) fm = fmaxf(fm, __shfl_xor(fm, o));
    if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = fm;
    __syncthreads();
    // non-negative floats order as their bit patterns: an integer maximum
    if (threadIdx.x == 0) atomicMax(absmax, __float_as_uint(fmaxf(fmaxf(shm[0], shm[1]), fmaxf(shm[2], shm[3]))));
}

__global__ void k_zero_int(int *__restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---- spreading ---------------------------------------------------------------

// LDS accumulator layout of a 16x8x8 tile with row strides 20 and 176: the
// 64 (kx,ky,kz) offsets of one stencil fall on 32 distinct 8-byte bank slots.
constexpr int ACC_SX = 20, ACC_SXY = 176, ACC_N = ACC_SXY * TILE_Z;

// Accumulation into the tile in 64-bit fixed point with a power-of-two scale,
// so the integer sums do not depend on the order of the additions.
//   FP32: the scale keeps |sum| below 2^47; the product w*(f*scale) is formed
//   by one FP64 FMA with 1.5*2^52 added, which leaves round(w*f*scale) in the
//   low mantissa bits (valid for |value| < 2^51), see below.
//   FP64: the scale keeps |sum| below 2^62 and llrint converts.
template <typename T> struct FixedPoint;
// FP32: t = fma(w, f scale, 1.5 2^52) holds k = round(w f scale) as
// bits(t) = 0x4338 << 48 + k for |k| < 2^51.  The constant part is a multiple
// of 2^48, so the raw bits are summed as they are and the low 48 bits of the
// sum are the sum of the k modulo 2^48: exact whenever the final sum is below
// 2^47 in magnitude, which the per-tile scale guarantees (kBits = 47), however
// the partial sums wrap.  No per-term subtraction.
template <> struct FixedPoint<float> {
    static constexpr int kBits = 47;
    double f[3];
    __device__ __forceinline__ FixedPoint(float f0, float f1, float f2, double scale) {
        f[0] = (double)f0 * scale; f[1] = (double)f1 * scale; f[2] = (double)f2 * scale;
    }
    __device__ __forceinline__ void add(unsigned long long *acc, int li, float w, int stride = ACC_N) const {
        const double wd = (double)w;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double t = fma(wd, f[d], 6755399441055744.0);  // + 1.5 * 2^52
            atomicAdd(&acc[d * stride + li], (unsigned long long)__double_as_longlong(t));
        }
    }
    __device__ __forceinline__ static long long value(unsigned long long a) {
        return (long long)(a << 16) >> 16;  // sign-extend the low 48 bits
    }
};
template <> struct FixedPoint<double> {
    static constexpr int kBits = 62;
    double f[3], scale;
    __device__ __forceinline__ FixedPoint(double f0, double f1, double f2, double s) : scale(s) {
        f[0] = f0; f[1] = f1; f[2] = f2;
    }
    __device__ __forceinline__ void add(unsigned long long *acc, int li, double w, int stride = ACC_N) const {
#pragma unroll
        for (int d = 0; d < 3; d++) atomicAdd(&acc[d * stride + li], (unsigned long long)llrint(w * f[d] * scale));
    }
    __device__ __forceinline__ static long long value(unsigned long long a) { return (long long)a; }
};

template <typename T>
__device__ __forceinline__ void spread_vertex(const SpreadRec<T> &sr, short4 n0, const IbmGeom &g, int ox, int oy,
                                              int oz, int lx, int ly, int lz, double scale, T w64,
                                              unsigned long long *acc) {
    const int x1 = n0.x, y1 = n0.y - 3, z1 = n0.z;
    const unsigned mx = axis_mask(x1, ox, lx, g.nx, true);
    const unsigned my = axis_mask(y1, oy, ly, g.ny, !g.walls);
    const unsigned mz = axis_mask(z1, oz, lz, g.nz, true);
    if (!mx || !my || !mz) return;
    T wx[4], wy[4], wz[4];
    stencil_weights<T>(sr.t[0], wx);
    stencil_weights<T>(sr.t[1], wy);
    stencil_weights<T>(sr.t[2], wz);
    const FixedPoint<T> fp(sr.f[0], sr.f[1], sr.f[2], scale);
    // tile-relative offsets of the stencil nodes on each axis
    int ax[4], by[4], cz[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int a = x1 + k, b = y1 + k, c = z1 + k;
        wrap_index(a, g.nx);
        wrap_index(c, g.nz);
        if (!g.walls) wrap_index(b, g.ny);
        ax[k] = a - ox;
        by[k] = ACC_SX * (b - oy);
        cz[k] = ACC_SXY * (c - oz);
    }
#pragma unroll
    for (int kz = 0; kz < 4; kz++) {
        if (!((mz >> kz) & 1u)) continue;
#pragma unroll
        for (int ky = 0; ky < 4; ky++) {
            if (!((my >> ky) & 1u)) continue;
            const T wyz = wy[ky] * wz[kz] * w64;
            const int row = by[ky] + cz[kz];
#pragma unroll
            for (int kx = 0; kx < 4; kx++) {
                if (!((mx >> kx) & 1u)) continue;
                fp.add(acc, row + ax[kx], wx[kx] * wyz);
            }
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_spread_tiles(const SpreadRec<T> *__restrict__ rec,
                                                      const short4 *__restrict__ node0,
                                                      const int *__restrict__ tile_count,
                                                      const int *__restrict__ tile_off, T *__restrict__ F,
                                                      IbmGeom g, TileGeom tg, const unsigned *__restrict__ absmax,
                                                      const int *__restrict__ order) {
    __shared__ unsigned long long acc[3 * ACC_N];
    __shared__ int queue[4][128];
    const int tile = order[blockIdx.x];
    const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
    const int ox = tx * TILE_X, oy = ty * TILE_Y, oz = tz * TILE_Z;
    const int lx = min(TILE_X, g.nx - ox), ly = min(TILE_Y, g.ny - oy), lz = min(TILE_Z, g.nz - oz);
    for (int k = threadIdx.x; k < 3 * ACC_N; k += blockDim.x) acc[k] = 0ull;
    // source bins: tiles holding first stencil nodes that reach this tile
    int sxs[4], sys[4], szs[4];
    const int nsx = source_tiles(tx, ox, g.nx, TILE_X, true, sxs);
    const int nsy = source_tiles(ty, oy, g.ny, TILE_Y, !g.walls, sys);
    const int nsz = source_tiles(tz, oz, g.nz, TILE_Z, true, szs);
    // fixed-point scale: at most `total` vertices reach one node, each with a
    // weight at most 1/(8 dx^3), so |sum| <= total |f|max / (8 dx^3), which
    // the power of two keeps below 2^kBits / scale
    long long total = 0;
    for (int iz = 0; iz < nsz; iz++)
        for (int iy = 0; iy < nsy; iy++)
            for (int ix = 0; ix < nsx; ix++) total += tile_count[sxs[ix] + tg.ntx * (sys[iy] + tg.nty * szs[iz])];
    const double fmax = (double)__uint_as_float(*absmax);
    double scale = 1.0;
    if (fmax > 0.0 && total > 0) {
        const double bound = (double)total * fmax / (8.0 * g.dx * g.dx * g.dx);
        scale = ldexp(1.0, FixedPoint<T>::kBits - 1 - ilogb(bound));
    }
    const T w64 = (T)(1.0 / (64.0 * g.dx * g.dx * g.dx));
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int *q = queue[wave];
    int qn = 0;
    for (int s = 0; s < nsx * nsy * nsz; s++) {
        const int bin = sxs[s % nsx] + tg.ntx * (sys[(s / nsx) % nsy] + tg.nty * szs[s / (nsx * nsy)]);
        const int cnt = tile_count[bin], off = tile_off[bin];
        // the bin read as 64 columns of R rows, lane l taking column l: the
        // lanes of a wave get vertices far apart in the bin, whose stencils
        // rarely share LDS addresses (64-bit LDS atomics to one address
        // serialise, and neighbouring vertices in a bin are close in space)
        const int R = (cnt + 63) >> 6;
        for (int r = wave; r < R; r += 4) {
            const int i = lane * R + r;
            int v = -1;  // position in bin order
            bool touch = false;
            if (i < cnt) {
                v = off + i;
                const short4 n0 = node0[v];
                touch = axis_mask(n0.x, ox, lx, g.nx, true) && axis_mask(n0.y - 3, oy, ly, g.ny, !g.walls) &&
                        axis_mask(n0.z, oz, lz, g.nz, true);
            }
            const unsigned long long m = __ballot(touch);
            if (touch) q[qn + __popcll(m & ((1ull << lane) - 1ull))] = v;
            qn += __popcll(m);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (qn >= 64) {
                spread_vertex<T>(rec[q[lane]], node0[q[lane]], g, ox, oy, oz, lx, ly, lz, scale, w64, acc);
                const int rest = qn - 64;
                const int mv = lane < rest ? q[64 + lane] : 0;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (lane < rest) q[lane] = mv;
                qn = rest;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
    }
    if (lane < qn) spread_vertex<T>(rec[q[lane]], node0[q[lane]], g, ox, oy, oz, lx, ly, lz, scale, w64, acc);
    __syncthreads();
    const double inv_scale = 1.0 / scale;
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    for (int k = threadIdx.x; k < TILE_X * TILE_Y * TILE_Z; k += blockDim.x) {
        const int la = k % TILE_X, lb = (k / TILE_X) % TILE_Y, lc = k / (TILE_X * TILE_Y);
        if (la >= lx || lb >= ly || lc >= lz) continue;
        const size_t nidx = (size_t)(ox + la) + (size_t)g.nx * ((size_t)(oy + lb) + (size_t)g.ny * (oz + lc));
        const int li = la + ACC_SX * lb + ACC_SXY * lc;
#pragma unroll
        for (int d = 0; d < 3; d++)
            F[d * N + nidx] = (T)((double)FixedPoint<T>::value(acc[d * ACC_N + li]) * inv_scale);
    }
}

// Spreading, form 1.  One work group per tile takes its home vertices only
// and adds each one's whole 4x4x4 stencil into an LDS box of the tile plus
// the 3 nodes above it on each axis (its stencils cannot leave that box), in
// the fixed point of form 0, with no per-node masks; the box is then written
// to a staging array, and k_spread_gather sums for each lattice node the box
// entries that map to it.  Form 0 visits a vertex once per tile its stencil
// touches (2.2 tiles on average with 16x8x8 tiles), each visit walking the 64
// nodes under masks.  Lane l of wave w takes column NW l + w of the bin read
// as WG columns, so that the lanes of a wave hold vertices far apart.
constexpr int SB_X = TILE_X + 3, SB_Y = TILE_Y + 3, SB_Z = TILE_Z + 3, SB_N = SB_X * SB_Y * SB_Z;
constexpr int SPREAD_BOX_WG = 512;
constexpr int SB_MAXC = 4;  // box entries per lattice coordinate and axis (k_spread_gather)

template <typename T>
__global__ __launch_bounds__(SPREAD_BOX_WG) void k_spread_box(const SpreadRec<T> *__restrict__ rec,
                                                              const short4 *__restrict__ node0,
                                                              const int *__restrict__ tile_count,
                                                              const int *__restrict__ tile_off, T *__restrict__ S,
                                                              IbmGeom g, TileGeom tg,
                                                              const int *__restrict__ order) {
    __shared__ unsigned long long acc[3 * SB_N];
    __shared__ float wmax[SPREAD_BOX_WG / 64];
    const int tile = order[blockIdx.x];
    const int cnt = tile_count[tile], off = tile_off[tile];
    if (cnt == 0) return;  // whole work group leaves together; the gather skips empty tiles
    const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
    const int ox = tx * TILE_X, oy = ty * TILE_Y, oz = tz * TILE_Z;
    for (int k = threadIdx.x; k < 3 * SB_N; k += blockDim.x) acc[k] = 0ull;
    constexpr int NW = SPREAD_BOX_WG / 64;
    const int R = (cnt + SPREAD_BOX_WG - 1) / SPREAD_BOX_WG;
    const int col = NW * (threadIdx.x & 63) + (threadIdx.x >> 6);
    // a box node takes at most one term per home vertex, each at most
    // |f|max / (8 dx^3) with |f|max over the tile's own vertices (a maximum,
    // so independent of the order of the records): the power of two keeps
    // the sums below 2^kBits
    float fm = 0.0f;
    for (int r = 0; r < R; r++) {
        const int i = col * R + r;
        if (i >= cnt) break;
        const SpreadRec<T> sr = rec[off + i];
        fm = fmaxf(fm, fmaxf(fabsf((float)sr.f[0]), fmaxf(fabsf((float)sr.f[1]), fabsf((float)sr.f[2]))));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) fm = fmaxf(fm, __shfl_xor(fm, o));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = fm;
    __syncthreads();
    fm = wmax[0];
#pragma unroll
    for (int w = 1; w < NW; w++) fm = fmaxf(fm, wmax[w]);
    // the float rounding of a double |f| may round down
    const double fmax = (double)fm * (1.0 + 0x1p-20);
    double scale = 1.0;
    if (fmax > 0.0) {
        const double bound = (double)cnt * fmax / (8.0 * g.dx * g.dx * g.dx);
        scale = ldexp(1.0, FixedPoint<T>::kBits - 1 - ilogb(bound));
    }
    const T w64 = (T)(1.0 / (64.0 * g.dx * g.dx * g.dx));
    for (int r = 0; r < R; r++) {
        const int i = col * R + r;
        if (i >= cnt) break;
        const short4 n0 = node0[off + i];
        const SpreadRec<T> sr = rec[off + i];
        // box coordinates of the first stencil node (x, z wrapped at binning;
        // y raw: with walls it may lie below the lattice in the first tile row)
        const int bx = n0.x - ox, bz = n0.z - oz;
        int y1 = n0.y - 3;
        if (!g.walls) wrap_index(y1, g.ny);
        const int by = y1 - oy;
        if (bx < 0 || bx >= TILE_X || bz < 0 || bz >= TILE_Z || by < -3 || by >= TILE_Y) continue;  // not a valid home
        T wx[4], wy[4], wz[4];
        stencil_weights<T>(sr.t[0], wx);
        stencil_weights<T>(sr.t[1], wy);
        stencil_weights<T>(sr.t[2], wz);
        const FixedPoint<T> fp(sr.f[0], sr.f[1], sr.f[2], scale);
#pragma unroll
        for (int kz = 0; kz < 4; kz++) {
#pragma unroll
            for (int ky = 0; ky < 4; ky++) {
                if (by + ky < 0) continue;  // below the wall
                const T wyz = wy[ky] * wz[kz] * w64;
                const int row = bx + SB_X * ((by + ky) + SB_Y * (bz + kz));
#pragma unroll
                for (int kx = 0; kx < 4; kx++) fp.add(acc, row + kx, wx[kx] * wyz, SB_N);
            }
        }
    }
    __syncthreads();
    const double inv_scale = 1.0 / scale;
    T *St = S + (size_t)tile * 3 * SB_N;
    for (int k = threadIdx.x; k < SB_N; k += blockDim.x)
#pragma unroll
        for (int d = 0; d < 3; d++)
            St[d * SB_N + k] = (T)((double)FixedPoint<T>::value(acc[d * SB_N + k]) * inv_scale);
}

// Sums, per lattice node, the box entries of k_spread_box that map to it, in
// a fixed order (deterministic).  cx/cy/cz list for each coordinate of an axis
// up to SB_MAXC entries (tile index along the axis) * 32 + box coordinate,
// padded with -1 (built once on the host).
template <typename T>
__global__ __launch_bounds__(256) void k_spread_gather(const T *__restrict__ S, const int *__restrict__ tile_count,
                                                       const int *__restrict__ cx, const int *__restrict__ cy,
                                                       const int *__restrict__ cz, T *__restrict__ F, IbmGeom g,
                                                       TileGeom tg) {
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    // each XCD gathers one contiguous eighth of the lattice: the box rows
    // of neighbouring y rows share cache lines, which are then fetched into
    // one L2 instead of into the L2s of the XCDs the rows' work groups land on
    const int nblocks = (int)((N + 255) / 256);
    const int lb = xcd_block(nblocks);
    if (lb >= nblocks) return;
    const size_t n = (size_t)lb * 256 + threadIdx.x;
    if (n >= N) return;
    const int x = (int)(n % g.nx), y = (int)((n / g.nx) % g.ny), z = (int)(n / ((size_t)g.nx * g.ny));
    // all table entries first, then per (z, y) entry the tile counts and box
    // values of every x entry, so that independent loads are in flight
    // together; the sums keep the order (z, y, x entry)
    int ex[SB_MAXC], ey[SB_MAXC], ez[SB_MAXC];
#pragma unroll
    for (int i = 0; i < SB_MAXC; i++) {
        ex[i] = cx[SB_MAXC * x + i];
        ey[i] = cy[SB_MAXC * y + i];
        ez[i] = cz[SB_MAXC * z + i];
    }
    T a0 = T(0), a1 = T(0), a2 = T(0);
#pragma unroll
    for (int iz = 0; iz < SB_MAXC; iz++) {
        if (ez[iz] < 0) break;
#pragma unroll
        for (int iy = 0; iy < SB_MAXC; iy++) {
            if (ey[iy] < 0) break;
            const int tyz = tg.ntx * ((ey[iy] >> 5) + tg.nty * (ez[iz] >> 5));
            const int byz = SB_X * ((ey[iy] & 31) + SB_Y * (ez[iz] & 31));
            int cnt[SB_MAXC];
#pragma unroll
            for (int ix = 0; ix < SB_MAXC; ix++) cnt[ix] = ex[ix] >= 0 ? tile_count[(ex[ix] >> 5) + tyz] : 0;
            T v0[SB_MAXC], v1[SB_MAXC], v2[SB_MAXC];
#pragma unroll
            for (int ix = 0; ix < SB_MAXC; ix++) {
                v0[ix] = v1[ix] = v2[ix] = T(0);
                if (cnt[ix] > 0) {
                    const size_t b = (size_t)((ex[ix] >> 5) + tyz) * 3 * SB_N + (ex[ix] & 31) + byz;
                    v0[ix] = S[b];
                    v1[ix] = S[b + SB_N];
                    v2[ix] = S[b + 2 * SB_N];
                }
            }
#pragma unroll
            for (int ix = 0; ix < SB_MAXC; ix++)
                if (cnt[ix] > 0) { a0 += v0[ix]; a1 += v1[ix]; a2 += v2[ix]; }
        }
    }
    F[n] = a0;
    F[N + n] = a1;
    F[2 * N + n] = a2;
}

// ---- interpolation + advection ------------------------------------------------

// Velocity box staged per tile: the tile plus 1 node below and 4 above on
// every axis.  The first stencil node of a home vertex lies in the tile, an
// RK stage moves it by less than one node in practice (|u| M dt << dx), and
// the stencil spans 4 nodes; stages that leave the box read global memory.
constexpr int RX = TILE_X + 5, RY = TILE_Y + 5, RZ = TILE_Z + 5, RN = RX * RY * RZ;
constexpr int RLO = 1;  // nodes of the box below the tile
constexpr int INTERP_WG = 512;  // work-group size of k_interp_tiles

// Out-of-line and not unrolled: it is the rare path, and inlining it into
// every RK stage would multiply the kernel's code size.
template <typename T>
__device__ __forceinline__ void interp_global(const T *__restrict__ U, size_t N, const IbmGeom &g, double x,
                                           double y, double z, bool unit, T &ur, T &vr, T &wr) {
    int x1, y1, z1;
    T wx[4], wy[4], wz[4];
    stencil3<T>(g, x, y, z, unit, x1, y1, z1, wx, wy, wz);
    ur = vr = wr = T(0);
#pragma unroll 1
    for (int kz = 0; kz < 4; kz++) {
        int iz = z1 + kz;
        if (!wrap_index(iz, g.nz)) continue;
#pragma unroll 1
        for (int ky = 0; ky < 4; ky++) {
            int iy = y1 + ky;
            if (g.walls) {
                if (iy < 0 || iy >= g.ny) continue;
            } else if (!wrap_index(iy, g.ny)) continue;
#pragma unroll 1
            for (int kx = 0; kx < 4; kx++) {
                int ix = x1 + kx;
                if (!wrap_index(ix, g.nx)) continue;
                const T dlt = wx[kx] * wy[ky] * wz[kz] * T(1.0 / 64.0);
                const size_t n = (size_t)ix + (size_t)g.nx * ((size_t)iy + (size_t)g.ny * iz);
                ur += dlt * U[n];
                vr += dlt * U[N + n];
                wr += dlt * U[2 * N + n];
            }
        }
    }
}

// Interpolation from the staged box when the stencil lies inside it.
template <typename T>
__device__ __forceinline__ void interp_tile(const T *__restrict__ U, size_t N, const T *__restrict__ Us,
                                            const IbmGeom &g, int bx, int by, int bz, double x, double y, double z,
                                            bool unit, T &ur, T &vr, T &wr) {
    int x1, y1, z1;
    T wx[4], wy[4], wz[4];
    stencil3<T>(g, x, y, z, unit, x1, y1, z1, wx, wy, wz);
    int rx = x1 - bx, rz = z1 - bz, ry = y1 - by;
    if (rx < 0) rx += g.nx; else if (rx >= g.nx) rx -= g.nx;
    if (rz < 0) rz += g.nz; else if (rz >= g.nz) rz -= g.nz;
    if (!g.walls) { if (ry < 0) ry += g.ny; else if (ry >= g.ny) ry -= g.ny; }
    if (rx < 0 || rx + 3 >= RX || ry < 0 || ry + 3 >= RY || rz < 0 || rz + 3 >= RZ) {
        interp_global<T>(U, N, g, x, y, z, unit, ur, vr, wr);
        return;
    }
    ur = vr = wr = T(0);
#pragma unroll
    for (int kz = 0; kz < 4; kz++) {
        const T wz64 = wz[kz] * T(1.0 / 64.0);
#pragma unroll
        for (int ky = 0; ky < 4; ky++) {
            const T wyz = wy[ky] * wz64;
            const int row = rx + RX * ((ry + ky) + RY * (rz + kz));
#pragma unroll
            for (int kx = 0; kx < 4; kx++) {
                const T dlt = wx[kx] * wyz;
                ur += dlt * Us[row + kx];
                vr += dlt * Us[RN + row + kx];
                wr += dlt * Us[2 * RN + row + kx];
            }
        }
    }
}

struct AdvectArgs {
    double dt, M;
    double lx, ly, lz;
    int rk;
    int periodic_y;
};

// Four consecutive values of one stencil row, read with one (unaligned)
// wide load.
template <typename T> struct Row4 {
    T v[4];
} __attribute__((aligned(sizeof(T))));

// Interpolation straight from the velocity field: one wide load per stencil
// row and component, the same products and summation order as interp_tile.
// A row that wraps in x is read node by node.
template <typename T>
__device__ __forceinline__ void interp_rows(const T *__restrict__ U, size_t N, const IbmGeom &g, double x,
                                            double y, double z, bool unit, T &ur, T &vr, T &wr) {
    int x1, y1, z1;
    T wx[4], wy[4], wz[4];
    stencil3<T>(g, x, y, z, unit, x1, y1, z1, wx, wy, wz);
    const bool whole = x1 >= 0 && x1 + 3 < g.nx;
    int ixw[4];
    bool okx[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        ixw[k] = x1 + k;
        okx[k] = wrap_index(ixw[k], g.nx);
    }
    ur = vr = wr = T(0);
#pragma unroll
    for (int kz = 0; kz < 4; kz++) {
        int iz = z1 + kz;
        if (!wrap_index(iz, g.nz)) continue;
        const T wz64 = wz[kz] * T(1.0 / 64.0);
#pragma unroll
        for (int ky = 0; ky < 4; ky++) {
            int iy = y1 + ky;
            if (g.walls) {
                if (iy < 0 || iy >= g.ny) continue;
            } else if (!wrap_index(iy, g.ny)) continue;
            const T wyz = wy[ky] * wz64;
            const size_t rowb = (size_t)g.nx * ((size_t)iy + (size_t)g.ny * iz);
            T ru[4], rv[4], rw[4];
            if (whole) {
                const Row4<T> a = *(const Row4<T> *)(U + rowb + x1);
                const Row4<T> b = *(const Row4<T> *)(U + N + rowb + x1);
                const Row4<T> c = *(const Row4<T> *)(U + 2 * N + rowb + x1);
#pragma unroll
                for (int kx = 0; kx < 4; kx++) { ru[kx] = a.v[kx]; rv[kx] = b.v[kx]; rw[kx] = c.v[kx]; }
            } else {
#pragma unroll
                for (int kx = 0; kx < 4; kx++) {
                    const size_t n = rowb + (okx[kx] ? ixw[kx] : 0);
                    ru[kx] = okx[kx] ? U[n] : T(0);
                    rv[kx] = okx[kx] ? U[N + n] : T(0);
                    rw[kx] = okx[kx] ? U[2 * N + n] : T(0);
                }
            }
#pragma unroll
            for (int kx = 0; kx < 4; kx++) {
                const T dlt = wx[kx] * wyz;
                ur += dlt * ru[kx];
                vr += dlt * rv[kx];
                wr += dlt * rw[kx];
            }
        }
    }
}

// The RK stages of vertex j (ibm.cu:430-517, 524-551, 581-608) with the
// interpolation ip(x, y, z, unit, u, v, w), then the periodic wrap of the new
// position (ibm.cu:616-632) and its stores.  xin and xo may be the same array
// (each vertex is read before it is written, by the same thread).
template <typename T, typename Interp>
__device__ __forceinline__ void advect_vertex(int j, const double *xin, double *xo, T *__restrict__ uvw,
                                              double *__restrict__ xd, int nv, const AdvectArgs &a, Interp ip) {
    const double M = a.M, dt = a.dt;
    const bool unit = a.rk != 4;
    const double xn = xin[j], yn = xin[nv + j], zn = xin[2 * (size_t)nv + j];
    // stage s is taken at X + k_{s-1} (M dt/2, M dt/2, M dt) for RK4 and
    // X + k_1 M for RK2, each offset evaluated as (k M) dt / 2 in that order;
    // the stage sum is (k1 + 2 k2 + 2 k3 + k4)/6, (k1 + k2)/2, k1.  One loop
    // keeps a single inlined copy of the interpolation.
    const int ns = a.rk == 4 ? 4 : (a.rk == 2 ? 2 : 1);
    T ku = T(0), kv = T(0), kw = T(0), su = T(0), sv = T(0), sw = T(0);
#pragma unroll 1
    for (int s = 0; s < ns; s++) {
        double ox = 0.0, oy = 0.0, oz = 0.0;
        if (s > 0) {
            ox = (double)ku * M; oy = (double)kv * M; oz = (double)kw * M;
            if (a.rk == 4) {
                ox *= dt; oy *= dt; oz *= dt;
                if (s < 3) { ox /= 2.0; oy /= 2.0; oz /= 2.0; }
            }
        }
        ip(xn + ox, yn + oy, zn + oz, unit, ku, kv, kw);
        const T wt = (a.rk == 4 && (s == 1 || s == 2)) ? T(2) : T(1);
        su += wt * ku; sv += wt * kv; sw += wt * kw;
    }
    const T den = a.rk == 4 ? T(6) : (a.rk == 2 ? T(2) : T(1));
    const T um = su / den, vm = sv / den, wm = sw / den;
    double xr = xn + (double)um * M * dt, yr = yn + (double)vm * M * dt, zr = zn + (double)wm * M * dt;
    if (xr < -a.lx / 2.0) xr += a.lx;
    if (xr > a.lx / 2.0) xr -= a.lx;
    if (zr < -a.lz / 2.0) zr += a.lz;
    if (zr > a.lz / 2.0) zr -= a.lz;
    if (a.periodic_y) {
        if (yr < -a.ly / 2.0) yr += a.ly;
        if (yr > a.ly / 2.0) yr -= a.ly;
    }
    if (uvw) { uvw[j] = um; uvw[nv + j] = vm; uvw[2 * (size_t)nv + j] = wm; }
    xo[j] = xr; xo[nv + j] = yr; xo[2 * (size_t)nv + j] = zr;
    xd[3 * (size_t)j] = xr; xd[3 * (size_t)j + 1] = yr; xd[3 * (size_t)j + 2] = zr;
}

// Interpolation form 0: one work group per tile stages the tile's velocity
// plus halo in LDS once and runs its home vertices from there.
template <typename T>
__global__ __launch_bounds__(INTERP_WG) void k_interp_tiles(const T *__restrict__ U, const double *xin, double *xo,
                                                      T *__restrict__ uvw, double *__restrict__ xd,
                                                      const int *__restrict__ tile_count,
                                                      const int *__restrict__ tile_off,
                                                      const int *__restrict__ home_list, int nv, IbmGeom g,
                                                      TileGeom tg, AdvectArgs a, const int *__restrict__ order) {
    extern __shared__ double dyn[];
    T *Us = (T *)dyn;  // [3][RZ][RY][RX]
    const int tile = order[blockIdx.x];
    const int cnt = tile_count[tile], off = tile_off[tile];
    if (cnt == 0) return;  // whole work group leaves together
    const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
    const int bx = tx * TILE_X - RLO, by = ty * TILE_Y - RLO, bz = tz * TILE_Z - RLO;
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    for (int k = threadIdx.x; k < RN; k += blockDim.x) {
        const int rx = k % RX, ry = (k / RX) % RY, rz = k / (RX * RY);
        // the box may span an axis more than once on small lattices
        const int ix = pmod(bx + rx, g.nx), iz = pmod(bz + rz, g.nz);
        int iy = by + ry;
        bool ok = true;
        if (g.walls) ok = iy >= 0 && iy < g.ny;
        else iy = pmod(iy, g.ny);
        T u = T(0), v = T(0), w = T(0);
        if (ok) {
            const size_t n = (size_t)ix + (size_t)g.nx * ((size_t)iy + (size_t)g.ny * iz);
            u = U[n]; v = U[N + n]; w = U[2 * N + n];
        }
        Us[k] = u; Us[RN + k] = v; Us[2 * RN + k] = w;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        advect_vertex<T>(home_list[off + i], xin, xo, uvw, xd, nv, a,
                         [&](double x, double y, double z, bool unit, T &u, T &v, T &w) {
                             interp_tile<T>(U, N, Us, g, bx, by, bz, x, y, z, unit, u, v, w);
                         });
    }
}

// Interpolation form 1: one thread per vertex in bin order (neighbouring
// lanes read neighbouring rows), rows read from the field through the caches.
template <typename T>
__global__ __launch_bounds__(256) void k_interp_vert(const T *__restrict__ U, const double *xin, double *xo,
                                                     T *__restrict__ uvw, double *__restrict__ xd,
                                                     const int *__restrict__ home_list, int nv, IbmGeom g,
                                                     AdvectArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    advect_vertex<T>(home_list[i], xin, xo, uvw, xd, nv, a,
                     [&](double x, double y, double z, bool unit, T &u, T &v, T &w) {
                         interp_rows<T>(U, N, g, x, y, z, unit, u, v, w);
                     });
}

// ---- adjoints (vector-Jacobian products) ------------------------------------
//
// The delta weights are C^1 in the position: with t = frac(xtr/h) the four
// weights are (1 - sin, 1 + cos, 1 + sin, 1 - cos)(pi t/2), and their
// derivatives (pi/2h) (-cos, -sin, cos, sin)(pi t/2) agree at t = 0 and 1
// with those of the neighbouring stencil, so the stencil switch adds no term.
// The spread adjoint is a gather over each vertex's 4x4x4 stencil in global
// memory (one thread per vertex).  The advection adjoint gathers for the
// position terms the same way, and adds the velocity adjoint to the nodes
// per tile in LDS (k_advect_adj_scatter).

template <typename W>
__device__ __forceinline__ void stencil_d(double xtr, double h, int &first, W w[4], W dw[4]) {
    const double q = over_h(xtr, h);
    const double fl = floor(q);
    first = (int)fl - 1;
    const W t = (W)(q - fl);
    W sn, cs;
    sincos_quarter(t, &sn, &cs);
    w[0] = W(1) - sn;
    w[1] = W(1) + cs;
    w[2] = W(1) + sn;
    w[3] = W(1) - cs;
    const W k = (W)(M_PI / (2.0 * h));
    dw[0] = -k * cs;
    dw[1] = -k * sn;
    dw[2] = k * cs;
    dw[3] = k * sn;
}

// Stencil of a position as stencil3 (spacing dx) with weight derivatives.
// a[k] for a runtime k by selects (a runtime index would put a in scratch)
template <typename T> __device__ __forceinline__ T sel4(const T (&a)[4], int k) {
    return k == 0 ? a[0] : k == 1 ? a[1] : k == 2 ? a[2] : a[3];
}

template <typename W> struct StencilD {
    int x1, y1, z1;
    W wx[4], wy[4], wz[4], dx[4], dy[4], dz[4];
    __device__ __forceinline__ StencilD(const IbmGeom &g, double x, double y, double z) {
        stencil_d<W>(x + (double)(g.nxp - 1) * g.dx / 2.0, g.dx, x1, wx, dx);
        stencil_d<W>(y + (double)(g.nyp - 1) * g.dx / 2.0, g.dx, y1, wy, dy);
        stencil_d<W>(z + (double)(g.nzp - 1) * g.dx / 2.0, g.dx, z1, wz, dz);
        x1 -= 2;
        y1 -= 2;
        z1 -= 2;
    }
    // visits every stencil node inside the lattice: f(node index, weight
    // product, its gradient with respect to the position, stencil offsets)
    template <typename Fn> __device__ __forceinline__ void each(const IbmGeom &g, Fn f) const {
#pragma unroll
        for (int kz = 0; kz < 4; kz++) {
            int iz = z1 + kz;
            if (!wrap_index(iz, g.nz)) continue;
#pragma unroll
            for (int ky = 0; ky < 4; ky++) {
                int iy = y1 + ky;
                if (g.walls) {
                    if (iy < 0 || iy >= g.ny) continue;
                } else if (!wrap_index(iy, g.ny)) continue;
#pragma unroll
                for (int kx = 0; kx < 4; kx++) {
                    int ix = x1 + kx;
                    if (!wrap_index(ix, g.nx)) continue;
                    const size_t n = (size_t)ix + (size_t)g.nx * ((size_t)iy + (size_t)g.ny * iz);
                    const W w = wx[kx] * wy[ky] * wz[kz];
                    const W gx = dx[kx] * wy[ky] * wz[kz], gy = wx[kx] * dy[ky] * wz[kz], gz = wx[kx] * wy[ky] * dz[kz];
                    f(n, w, gx, gy, gz, kx, ky, kz);
                }
            }
        }
    }
    // the 16 nodes of z-slice kz only (same products as each)
    template <typename Fn> __device__ __forceinline__ void slice(const IbmGeom &g, int kz, Fn f) const {
        int iz = z1 + kz;
        if (!wrap_index(iz, g.nz)) return;
        const W wzk = sel4(wz, kz), dzk = sel4(dz, kz);
#pragma unroll
        for (int ky = 0; ky < 4; ky++) {
            int iy = y1 + ky;
            if (g.walls) {
                if (iy < 0 || iy >= g.ny) continue;
            } else if (!wrap_index(iy, g.ny)) continue;
#pragma unroll
            for (int kx = 0; kx < 4; kx++) {
                int ix = x1 + kx;
                if (!wrap_index(ix, g.nx)) continue;
                const size_t n = (size_t)ix + (size_t)g.nx * ((size_t)iy + (size_t)g.ny * iz);
                const W w = wx[kx] * wy[ky] * wzk;
                const W gx = dx[kx] * wy[ky] * wzk, gy = wx[kx] * dy[ky] * wzk, gz = wx[kx] * wy[ky] * dzk;
                f(n, w, gx, gy, gz);
            }
        }
    }
};

// Sum over the 4 lanes of an aligned quad; every lane gets the same bits
// ((a + b) + (c + d) in each, addition being commutative).
template <typename T> __device__ __forceinline__ T quad_sum(T v) {
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    return v;
}

// Adjoint of ibm_st1 at the current positions and nodal forces: given Fa (the
// adjoint of the lattice force), cfa = sum D Fa is the adjoint of the nodal
// force and xa += sum (Fa . f) grad D that of the position, D = phi^3/dx^3.
template <typename T>
__global__ __launch_bounds__(256) void k_spread_adj(const double *__restrict__ xs, const T *__restrict__ cf,
                                                    const T *__restrict__ Fa, T *__restrict__ cfa,
                                                    double *__restrict__ xa, int nv, IbmGeom g) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nv) return;
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    const StencilD<T> st(g, xs[j], xs[nv + j], xs[2 * (size_t)nv + j]);
    const T f0 = cf[j], f1 = cf[nv + j], f2 = cf[2 * (size_t)nv + j];
    const T w64 = (T)(1.0 / (64.0 * g.dx * g.dx * g.dx));
    T b0 = T(0), b1 = T(0), b2 = T(0), x0 = T(0), x1 = T(0), x2 = T(0);
    st.each(g, [&](size_t n, T w, T gx, T gy, T gz, int, int, int) {
        const T a0 = Fa[n], a1 = Fa[N + n], a2 = Fa[2 * N + n];
        b0 += w * a0; b1 += w * a1; b2 += w * a2;
        const T s = a0 * f0 + a1 * f1 + a2 * f2;
        x0 += s * gx; x1 += s * gy; x2 += s * gz;
    });
    cfa[j] = b0 * w64; cfa[nv + j] = b1 * w64; cfa[2 * (size_t)nv + j] = b2 * w64;
    xa[j] += (double)(x0 * w64);
    xa[nv + j] += (double)(x1 * w64);
    xa[2 * (size_t)nv + j] += (double)(x2 * w64);
}

// Adjoint of ibm_st2 with RK4 at the current positions and velocity field:
// xa holds the adjoint of the advected positions on entry and that of the
// current positions on exit; the adjoint of the velocity field is added to
// ua.  The stage positions are recomputed; stage s contributes
// ua += phi^3(X_s) kb_s and X_s-bar = sum (u . kb_s) grad phi^3, which feeds
// the stage before it (ibm.cu:430-517: X2 = X + h/2 k1, X3 = X + h/2 k2,
// X4 = X + h k3, X' = X + h (k1 + 2 k2 + 2 k3 + k4)/6, h = M dt).
// Two kernels.  k_advect_adj_vert, four lanes per vertex in bin order, does
// the gathers (the stage positions, X-bar) and writes each stage's position
// and kb_s to a record; k_advect_adj_scatter, one work group per tile over
// its home vertices, adds the velocity adjoint of the records.

template <typename T> struct AdvRec {
    double p[3];  // stage position
    T k[3];       // kb_s
};

// Four lanes per vertex (bin order): lane z of the quad sums the z-th slice
// of every stencil, and quad_sum completes the sums in all four lanes, so
// the four follow the same stage chain.  One lane per vertex left most of the
// machine idle on the small cell counts the gradient is taken on.
template <typename T>
__global__ __launch_bounds__(256) void k_advect_adj_vert(const double *__restrict__ xs, const T *__restrict__ U,
                                                         double *__restrict__ xa, const int *__restrict__ home_list,
                                                         int nv, IbmGeom g, double h, AdvRec<T> *__restrict__ rec,
                                                         unsigned *__restrict__ kbmax) {
    __shared__ float shm[4];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t >> 2, sub = t & 3;  // vertex in bin order, stencil z-slice
    float km = 0.0f;
    if (i < nv) {  // the same for the four lanes of a quad
        const int j = home_list[i];
        const size_t N = (size_t)g.nx * g.ny * g.nz;
        const T q64 = T(1.0 / 64.0);
        const double X[3] = {xs[j], xs[nv + j], xs[2 * (size_t)nv + j]};
        auto interp = [&](const double p[3], double k[3]) {
            const StencilD<T> st(g, p[0], p[1], p[2]);
            T u0 = T(0), u1 = T(0), u2 = T(0);
            st.slice(g, sub, [&](size_t n, T w, T, T, T) {
                u0 += w * U[n]; u1 += w * U[N + n]; u2 += w * U[2 * N + n];
            });
            k[0] = (double)(quad_sum(u0) * q64);
            k[1] = (double)(quad_sum(u1) * q64);
            k[2] = (double)(quad_sum(u2) * q64);
        };
        double P[4][3], k[3];
#pragma unroll
        for (int d = 0; d < 3; d++) P[0][d] = X[d];
        interp(P[0], k);
#pragma unroll
        for (int d = 0; d < 3; d++) P[1][d] = X[d] + k[d] * h / 2.0;
        interp(P[1], k);
#pragma unroll
        for (int d = 0; d < 3; d++) P[2][d] = X[d] + k[d] * h / 2.0;
        interp(P[2], k);
#pragma unroll
        for (int d = 0; d < 3; d++) P[3][d] = X[d] + k[d] * h;
        const double xb[3] = {xa[j], xa[nv + j], xa[2 * (size_t)nv + j]};
        double acc[3] = {xb[0], xb[1], xb[2]};
        double kb[4][3];
        const double cw[4] = {h / 6.0, h / 3.0, h / 3.0, h / 6.0};
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int d = 0; d < 3; d++) kb[s][d] = cw[s] * xb[d];
        const double back[4] = {0.0, h / 2.0, h / 2.0, h};  // X_s = X + back[s] k_{s-1}
#pragma unroll
        for (int s = 3; s >= 0; s--) {  // kb_s is final here (stage s+1 came first)
            const T k0 = (T)kb[s][0], k1 = (T)kb[s][1], k2 = (T)kb[s][2];
            if (sub == s) {  // the quad's lanes write one record each
                AdvRec<T> r;
                r.p[0] = P[s][0]; r.p[1] = P[s][1]; r.p[2] = P[s][2];
                r.k[0] = k0; r.k[1] = k1; r.k[2] = k2;
                rec[4 * (size_t)i + s] = r;
            }
            km += fmaxf(fabsf((float)k0), fmaxf(fabsf((float)k1), fabsf((float)k2)));
            const StencilD<T> st(g, P[s][0], P[s][1], P[s][2]);
            T g0 = T(0), g1 = T(0), g2 = T(0);
            st.slice(g, sub, [&](size_t n, T, T gx, T gy, T gz) {
                const T sdot = U[n] * k0 + U[N + n] * k1 + U[2 * N + n] * k2;
                g0 += sdot * gx; g1 += sdot * gy; g2 += sdot * gz;
            });
            const double xs_[3] = {(double)(quad_sum(g0) * q64), (double)(quad_sum(g1) * q64),
                                   (double)(quad_sum(g2) * q64)};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                acc[d] += xs_[d];
                if (s > 0) kb[s - 1][d] += back[s] * xs_[d];
            }
        }
        if (sub == 0) {
            xa[j] = acc[0];
            xa[nv + j] = acc[1];
            xa[2 * (size_t)nv + j] = acc[2];
        }
        km *= 1.0001f;  // margin for the float rounding
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) km = fmaxf(km, __shfl_xor(km, o));
    if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = km;
    __syncthreads();
    // non-negative floats order as their bit patterns: an integer maximum
    if (threadIdx.x == 0) atomicMax(kbmax, __float_as_uint(fmaxf(fmaxf(shm[0], shm[1]), fmaxf(shm[2], shm[3]))));
}

// One work group per tile takes its home vertices (the bins of the
// interpolation) and adds ua += sum_s phi^3(X_s) kb_s / 64 into an LDS box
// of the tile plus its halo (the staging box of k_interp_tiles), then adds
// the box to ua with one global atomic per node; stencils that leave the box
// add to ua directly.  The four stage stencils of a vertex nearly always
// start at the same node (the stages move it by h |u| << 1): then the four
// terms are summed per node and added once.  Lane l of wave w takes column
// 8l + w of the bin read as 512 columns, so the lanes of a wave hold vertices
// far apart, whose nodes rarely coincide (LDS atomics to one address
// serialise).  FP32 accumulates in 64-bit fixed point, as the spreading
// (raw bits of fma(v, scale, 1.5 2^52), whose low 48 bits sum exactly), with
// the scale from the bound cnt max_v sum_s |kb_s| / 8 of a node's sum (a
// weight product is at most 8/64); FP64 uses FP64 LDS atomics.
constexpr int ADV_WG = 512;  // work-group size of k_advect_adj_scatter

template <typename T>
__global__ __launch_bounds__(ADV_WG) void k_advect_adj_scatter(const AdvRec<T> *__restrict__ rec, T *__restrict__ ua,
                                                            const int *__restrict__ tile_count,
                                                            const int *__restrict__ tile_off, IbmGeom g,
                                                            TileGeom tg, const unsigned *__restrict__ kbmax,
                                                            const int *__restrict__ order) {
    extern __shared__ double dyn[];
    constexpr bool FIX = sizeof(T) == 4;
    unsigned long long *Ub = (unsigned long long *)dyn;  // [3][RZ][RY][RX]
    double *Ud = dyn;
    const int tile = order[blockIdx.x];
    const int cnt = tile_count[tile], off = tile_off[tile];
    if (cnt == 0) return;  // whole work group leaves together
    const int tx = tile % tg.ntx, ty = (tile / tg.ntx) % tg.nty, tz = tile / (tg.ntx * tg.nty);
    const int bx = tx * TILE_X - RLO, by = ty * TILE_Y - RLO, bz = tz * TILE_Z - RLO;
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    double scale = 1.0;
    if (FIX) {
        const double km = (double)__uint_as_float(*kbmax);
        if (km > 0.0) scale = ldexp(1.0, 46 - ilogb((double)cnt * km / 8.0));
    }
    for (int k = threadIdx.x; k < 3 * RN; k += blockDim.x) Ub[k] = 0ull;
    __syncthreads();
    const T q64 = T(1.0 / 64.0);
    constexpr int NW = ADV_WG / 64;
    const int R = (cnt + ADV_WG - 1) / ADV_WG;
    const int col = NW * (threadIdx.x & 63) + (threadIdx.x >> 6);
    for (int r = 0; r < R; r++) {
        const int i = col * R + r;
        if (i >= cnt) break;
        int x1[4], y1[4], z1[4];
        T wx[4][4], wy[4][4], wz[4][4], kk[4][3];
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const AdvRec<T> rr = rec[4 * (size_t)(off + i) + s];
            stencil3<T>(g, rr.p[0], rr.p[1], rr.p[2], false, x1[s], y1[s], z1[s], wx[s], wy[s], wz[s]);
            kk[s][0] = rr.k[0] * q64; kk[s][1] = rr.k[1] * q64; kk[s][2] = rr.k[2] * q64;
        }
        bool same = true;
#pragma unroll
        for (int s = 1; s < 4; s++) same = same && x1[s] == x1[0] && y1[s] == y1[0] && z1[s] == z1[0];
        // one walk of the common stencil with all four stages, or one walk per
        // stage with the other stages masked out
        const int ns = same ? 1 : 4;
        for (int it = 0; it < ns; it++) {
            T m[4];
            int sx = x1[0], sy = y1[0], sz = z1[0];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                m[s] = (same || s == it) ? T(1) : T(0);
                if (!same && s == it) { sx = x1[s]; sy = y1[s]; sz = z1[s]; }
            }
            // box coordinates of the first stencil node (as interp_tile)
            int rx = sx - bx, ry = sy - by, rz = sz - bz;
            if (rx < 0) rx += g.nx; else if (rx >= g.nx) rx -= g.nx;
            if (rz < 0) rz += g.nz; else if (rz >= g.nz) rz -= g.nz;
            if (!g.walls) { if (ry < 0) ry += g.ny; else if (ry >= g.ny) ry -= g.ny; }
            const bool inbox = rx >= 0 && rx + 3 < RX && ry >= 0 && ry + 3 < RY && rz >= 0 && rz + 3 < RZ;
#pragma unroll 1
            for (int kz = 0; kz < 4; kz++) {
                int iz = sz + kz;
                if (!wrap_index(iz, g.nz)) continue;
                T cz[4][3];  // mask * kb_s/64 * wz_s
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    const T w = m[s] * sel4(wz[s], kz);
                    cz[s][0] = w * kk[s][0]; cz[s][1] = w * kk[s][1]; cz[s][2] = w * kk[s][2];
                }
#pragma unroll
                for (int ky = 0; ky < 4; ky++) {
                    int iy = sy + ky;
                    if (g.walls) {
                        if (iy < 0 || iy >= g.ny) continue;
                    } else if (!wrap_index(iy, g.ny)) continue;
                    T cy[4][3];
#pragma unroll
                    for (int s = 0; s < 4; s++)
#pragma unroll
                        for (int d = 0; d < 3; d++) cy[s][d] = wy[s][ky] * cz[s][d];
#pragma unroll
                    for (int kx = 0; kx < 4; kx++) {
                        int ix = sx + kx;
                        if (!wrap_index(ix, g.nx)) continue;
                        const size_t n = (size_t)ix + (size_t)g.nx * ((size_t)iy + (size_t)g.ny * iz);
                        const int bi = (rx + kx) + RX * ((ry + ky) + RY * (rz + kz));
                        T v0 = T(0), v1 = T(0), v2 = T(0);
#pragma unroll
                        for (int s = 0; s < 4; s++) {
                            v0 += wx[s][kx] * cy[s][0]; v1 += wx[s][kx] * cy[s][1]; v2 += wx[s][kx] * cy[s][2];
                        }
                        // into the box when the stencil lies in it, else into ua (written
                        // out here: the LDS atomics must see the LDS pointer itself)
                        if (inbox) {
                            if (FIX) {
                                const double big = 6755399441055744.0;  // 1.5 2^52
                                atomicAdd(Ub + bi,
                                          (unsigned long long)__double_as_longlong(fma((double)v0, scale, big)));
                                atomicAdd(Ub + RN + bi,
                                          (unsigned long long)__double_as_longlong(fma((double)v1, scale, big)));
                                atomicAdd(Ub + 2 * RN + bi,
                                          (unsigned long long)__double_as_longlong(fma((double)v2, scale, big)));
                            } else {
                                atomicAdd(Ud + bi, (double)v0);
                                atomicAdd(Ud + RN + bi, (double)v1);
                                atomicAdd(Ud + 2 * RN + bi, (double)v2);
                            }
                        } else {
                            atomicAdd(&ua[n], v0);
                            atomicAdd(&ua[N + n], v1);
                            atomicAdd(&ua[2 * N + n], v2);
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    // the box into ua (the box may wrap onto itself on small lattices: atomics)
    const double inv_scale = 1.0 / scale;
    for (int k = threadIdx.x; k < RN; k += blockDim.x) {
        auto val = [&](int i) {
            return FIX ? (T)((double)((long long)(Ub[i] << 16) >> 16) * inv_scale) : (T)Ud[i];
        };
        const T a0 = val(k), a1 = val(RN + k), a2 = val(2 * RN + k);
        if (a0 == T(0) && a1 == T(0) && a2 == T(0)) continue;
        const int rx = k % RX, ry = (k / RX) % RY, rz = k / (RX * RY);
        const int ix = pmod(bx + rx, g.nx), iz = pmod(bz + rz, g.nz);
        int iy = by + ry;
        if (g.walls) {
            if (iy < 0 || iy >= g.ny) continue;
        } else {
            iy = pmod(iy, g.ny);
        }
        const size_t n = (size_t)ix + (size_t)g.nx * ((size_t)iy + (size_t)g.ny * iz);
        atomicAdd(&ua[n], a0);
        atomicAdd(&ua[N + n], a1);
        atomicAdd(&ua[2 * N + n], a2);
    }
}

IbmGeom geom_of(const lbmibm_ctx *ctx) {
    IbmGeom g;
    g.nx = ctx->nx; g.ny = ctx->ny; g.nz = ctx->nz;
    g.nxp = ctx->nxp; g.nyp = ctx->nyp; g.nzp = ctx->nzp;
    g.walls = ctx->p.periodic_y == 0;
    g.dx = ctx->p.dx;
    return g;
}

BinGeom bin_geom(const lbmibm_ctx *ctx) {
    BinGeom b;
    b.nx = ctx->nx; b.ny = ctx->ny; b.nz = ctx->nz;
    b.nxp = ctx->nxp; b.nyp = ctx->nyp; b.nzp = ctx->nzp;
    b.walls = ctx->p.periodic_y == 0;
    b.ntx = ctx->ntx; b.nty = ctx->nty;
    b.dx = ctx->p.dx;
    return b;
}

// The offsets and orders of the tiles from their counts (k_tile_scan with no
// force maxima, k_tile_order).
lbmibm_status tile_offsets(lbmibm_ctx *ctx, hipStream_t stream, int nblk_fmax) {
    const IbmGeom g = geom_of(ctx);
    TileGeom tg{ctx->ntx, ctx->nty, ctx->ntz, ctx->ntiles};
    hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, stream, ctx->tile_count, ctx->tile_off, ctx->tile_cur,
                       ctx->ntiles, ctx->blk_fmax, nblk_fmax, ctx->absmax);
    LBMIBM_LAUNCHED(ctx, "k_tile_scan");
    hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(1024), 0, stream, ctx->tile_count, ctx->ntiles, g, tg,
                       ctx->tile_cand, ctx->order_spread, ctx->order_interp, ctx->order_calls);
    LBMIBM_LAUNCHED(ctx, "k_tile_order");
    return LBMIBM_OK;
}

// The scatter of the binning: home list, first stencil nodes and spreading
// records (fractions, current forces) in bin order.
template <typename T> lbmibm_status bin_scatter(lbmibm_ctx *ctx, hipStream_t stream) {
    hipLaunchKernelGGL(k_bin_scatter<T>, dim3(ctx->nblk_bin), dim3(256), 0, stream, (const double *)ctx->xs,
                       (const T *)spread_force(ctx), ctx->home, ctx->nv, geom_of(ctx), ctx->tile_off, ctx->tile_cur,
                       ctx->home_list, ctx->node0, (SpreadRec<T> *)ctx->srec);
    LBMIBM_LAUNCHED(ctx, "k_bin_scatter");
    return LBMIBM_OK;
}

// Binning of the vertices by the tile of their first stencil node, with the
// spreading records of the current nodal forces and max |f|.
template <typename T> lbmibm_status bin_impl(lbmibm_ctx *ctx) {
    hipStream_t stream = ctx->stream;
    const int nv = ctx->nv;
    const hipEvent_t kt = kt_begin(ctx, stream);
    hipLaunchKernelGGL(k_zero_int, dim3((ctx->ntiles + 255) / 256), dim3(256), 0, stream, ctx->tile_count,
                       ctx->ntiles);
    LBMIBM_LAUNCHED(ctx, "k_zero_int");
    hipLaunchKernelGGL(k_bin_count<T>, dim3(ctx->nblk_bin), dim3(256), 0, stream, (const double *)ctx->xs,
                       (const T *)spread_force(ctx), nv, bin_geom(ctx), ctx->home, ctx->tile_count, ctx->blk_fmax);
    LBMIBM_LAUNCHED(ctx, "k_bin_count");
    lbmibm_status s = tile_offsets(ctx, stream, ctx->nblk_bin);
    if (s == LBMIBM_OK) s = bin_scatter<T>(ctx, stream);
    if (s != LBMIBM_OK) return s;
    kt_end(ctx, stream, "binning", kt);
    ctx->bins_valid = true;
    ctx->counts_ready = false;
    return LBMIBM_OK;
}

// The spreading of the binned records into F with the given form.
template <typename T> lbmibm_status spread_launch(lbmibm_ctx *ctx, int form) {
    TileGeom tg{ctx->ntx, ctx->nty, ctx->ntz, ctx->ntiles};
    const IbmGeom g = geom_of(ctx);
    if (form == 1) {
        LBMIBM_KT(ctx, ctx->stream, "k_spread_box",
                  hipLaunchKernelGGL(k_spread_box<T>, dim3(ctx->ntiles), dim3(SPREAD_BOX_WG), 0, ctx->stream,
                                     (const SpreadRec<T> *)ctx->srec, ctx->node0, ctx->tile_count, ctx->tile_off,
                                     (T *)ctx->sbox, g, tg, ctx->order_interp));
        LBMIBM_LAUNCHED(ctx, "k_spread_box");
        LBMIBM_KT(ctx, ctx->stream, "k_spread_gather",
                  hipLaunchKernelGGL(k_spread_gather<T>, dim3((unsigned)xcd_grid((int)((ctx->N + 255) / 256))), dim3(256), 0,
                                     ctx->stream, (const T *)ctx->sbox, ctx->tile_count, ctx->sb_cand,
                                     ctx->sb_cand + SB_MAXC * ctx->nx, ctx->sb_cand + SB_MAXC * (ctx->nx + ctx->ny),
                                     (T *)ctx->F, g, tg));
        LBMIBM_LAUNCHED(ctx, "k_spread_gather");
        return LBMIBM_OK;
    }
    LBMIBM_KT(ctx, ctx->stream, "k_spread_tiles",
              hipLaunchKernelGGL(k_spread_tiles<T>, dim3(ctx->ntiles), dim3(256), 0, ctx->stream,
                                 (const SpreadRec<T> *)ctx->srec, ctx->node0, ctx->tile_count, ctx->tile_off,
                                 (T *)ctx->F, g, tg, ctx->absmax, ctx->order_spread));
    LBMIBM_LAUNCHED(ctx, "k_spread_tiles");
    return LBMIBM_OK;
}

template <typename T> lbmibm_status spread_impl(lbmibm_ctx *ctx) {
    if (ctx->nv == 0) return lattice_fill(ctx, ctx->F, 3 * (size_t)ctx->N, 0.0, false);
    if (!ctx->bins_valid && ctx->counts_ready && ctx->spread_form == 1) {
        // counted with the volume correction: only the scatter is left (form 1
        // scales each tile by its own max |f| and needs no global maximum)
        const hipEvent_t kt = kt_begin(ctx, ctx->stream);
        lbmibm_status s = bin_scatter<T>(ctx, ctx->stream);
        if (s != LBMIBM_OK) return s;
        kt_end(ctx, ctx->stream, "binning", kt);
        ctx->bins_valid = true;
        ctx->counts_ready = false;
    } else if (!ctx->bins_valid) {
        lbmibm_status s = bin_impl<T>(ctx);
        if (s != LBMIBM_OK) return s;
    } else {  // same positions, new forces: refresh the force part of the records
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, ctx->stream, (int *)ctx->absmax, 1);
        LBMIBM_LAUNCHED(ctx, "k_zero_int");
        hipLaunchKernelGGL(k_bin_forces<T>, dim3(ctx->nblk_bin), dim3(256), 0, ctx->stream, ctx->home_list,
                           (const T *)spread_force(ctx), ctx->nv, (SpreadRec<T> *)ctx->srec, ctx->absmax);
        LBMIBM_LAUNCHED(ctx, "k_bin_forces");
    }
    return spread_launch<T>(ctx, ctx->spread_form);
}

// One interpolation + advection of all vertices with the given form, from
// xin into xo / xd (uvw may be null).
template <typename T>
lbmibm_status interp_launch(lbmibm_ctx *ctx, int form, const double *xin, double *xo, T *uvw, double *xd) {
    AdvectArgs a;
    a.dt = ctx->p.dt; a.M = ctx->p.m;
    a.lx = ctx->nx * ctx->p.dx; a.ly = ctx->ny * ctx->p.dx; a.lz = ctx->nz * ctx->p.dx;
    a.rk = ctx->p.runge_kutta;
    a.periodic_y = ctx->p.periodic_y;
    if (form == 1) {
        LBMIBM_KT(ctx, ctx->stream, "k_interp_vert",
                  hipLaunchKernelGGL(k_interp_vert<T>, dim3((ctx->nv + 255) / 256), dim3(256), 0, ctx->stream,
                                     (const T *)ctx->U, xin, xo, uvw, xd, ctx->home_list, ctx->nv, geom_of(ctx), a));
        LBMIBM_LAUNCHED(ctx, "k_interp_vert");
        return LBMIBM_OK;
    }
    TileGeom tg{ctx->ntx, ctx->nty, ctx->ntz, ctx->ntiles};
    const size_t lds = 3 * (size_t)RN * sizeof(T);
    LBMIBM_KT(ctx, ctx->stream, "k_interp_tiles",
              hipLaunchKernelGGL(k_interp_tiles<T>, dim3(ctx->ntiles), dim3(INTERP_WG), lds, ctx->stream,
                                 (const T *)ctx->U, xin, xo, uvw, xd, ctx->tile_count, ctx->tile_off, ctx->home_list,
                                 ctx->nv, geom_of(ctx), tg, a, ctx->order_interp));
    LBMIBM_LAUNCHED(ctx, "k_interp_tiles");
    return LBMIBM_OK;
}

template <typename T> lbmibm_status interp_impl(lbmibm_ctx *ctx) {
    if (ctx->nv == 0) return LBMIBM_OK;
    if (!ctx->bins_valid) {
        lbmibm_status s = bin_impl<T>(ctx);
        if (s != LBMIBM_OK) return s;
    }
    double *xo = ctx->xs_out ? ctx->xs_out : (double *)ctx->xs;
    lbmibm_status s = interp_launch<T>(ctx, ctx->interp_form, (const double *)ctx->xs, xo,
                                       ctx->p.store_macros ? (T *)ctx->uvw : (T *)nullptr, (double *)ctx->xd);
    invalidate_bins(ctx);  // the positions moved
    return s;
}

// Times the interpolation forms on the current positions and velocity field,
// writing to scratch arrays (the vertices do not move), and keeps the faster.
template <typename T> lbmibm_status interp_autotune_impl(lbmibm_ctx *ctx) {
    if (!ctx->bins_valid) {
        lbmibm_status s = bin_impl<T>(ctx);
        if (s != LBMIBM_OK) return s;
    }
    double *xo = nullptr, *xd = nullptr;
    const size_t bytes = 3 * (size_t)ctx->nv * sizeof(double);
    LBMIBM_HIP(ctx, hipMalloc((void **)&xo, bytes));
    if (hipMalloc((void **)&xd, bytes) != hipSuccess) {
        (void)hipFree(xo);
        return LBMIBM_ERR_HIP;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    lbmibm_status s = LBMIBM_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) s = LBMIBM_ERR_HIP;
    static const char *names[2] = {"tiles in LDS", "rows from memory"};
    char line[160];
    float best = 0.0f;
    for (int f = 0; f < 2 && s == LBMIBM_OK; f++) {
        s = interp_launch<T>(ctx, f, (const double *)ctx->xs, xo, (T *)nullptr, xd);  // warm-up
        if (s != LBMIBM_OK) break;
        if (hipEventRecord(e0, ctx->stream) != hipSuccess) { s = LBMIBM_ERR_HIP; break; }
        for (int r = 0; r < 3 && s == LBMIBM_OK; r++)
            s = interp_launch<T>(ctx, f, (const double *)ctx->xs, xo, (T *)nullptr, xd);
        float ms = 0.0f;
        if (s == LBMIBM_OK && (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                               hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            s = LBMIBM_ERR_HIP;
        if (s != LBMIBM_OK) break;
        ms /= 3.0f;
        snprintf(line, sizeof line, "interp form %s: %.4f ms\n", names[f], ms);
        ctx->tune_report += line;
        if (f == 0 || ms < best) { best = ms; ctx->interp_form = f; }
    }
    if (s == LBMIBM_OK) {
        snprintf(line, sizeof line, "interp: using %s\n", names[ctx->interp_form]);
        ctx->tune_report += line;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(xo);
    (void)hipFree(xd);
    return s;
}

template <typename T> lbmibm_status spread_adj_impl(lbmibm_ctx *ctx) {
    if (ctx->nv == 0) return LBMIBM_OK;
    hipLaunchKernelGGL(k_spread_adj<T>, dim3((ctx->nv + 255) / 256), dim3(256), 0, ctx->stream,
                       (const double *)ctx->xs, (const T *)ctx->cf, (const T *)ctx->adj_F, (T *)ctx->adj_cf,
                       ctx->adj_x, ctx->nv, geom_of(ctx));
    LBMIBM_LAUNCHED(ctx, "k_spread_adj");
    return LBMIBM_OK;
}

template <typename T> lbmibm_status advect_adj_impl(lbmibm_ctx *ctx) {
    if (ctx->nv == 0) return LBMIBM_OK;
    if (!ctx->bins_valid) {
        lbmibm_status s = bin_impl<T>(ctx);
        if (s != LBMIBM_OK) return s;
    }
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, ctx->stream, (int *)ctx->adj_xmax, 1);
    LBMIBM_LAUNCHED(ctx, "k_zero_int");
    const IbmGeom g = geom_of(ctx);
    hipLaunchKernelGGL(k_advect_adj_vert<T>, dim3((unsigned)((4 * (size_t)ctx->nv + 255) / 256)), dim3(256), 0,
                       ctx->stream,
                       (const double *)ctx->xs, (const T *)ctx->U, ctx->adj_x, ctx->home_list, ctx->nv, g,
                       ctx->p.m * ctx->p.dt, (AdvRec<T> *)ctx->adj_rec, ctx->adj_xmax);
    LBMIBM_LAUNCHED(ctx, "k_advect_adj_vert");
    TileGeom tg{ctx->ntx, ctx->nty, ctx->ntz, ctx->ntiles};
    const size_t lds = 3 * (size_t)RN * sizeof(unsigned long long);
    hipLaunchKernelGGL(k_advect_adj_scatter<T>, dim3(ctx->ntiles), dim3(ADV_WG), lds, ctx->stream,
                       (const AdvRec<T> *)ctx->adj_rec, (T *)ctx->adj_u, ctx->tile_count, ctx->tile_off, g, tg,
                       ctx->adj_xmax, ctx->order_interp);
    LBMIBM_LAUNCHED(ctx, "k_advect_adj_scatter");
    return LBMIBM_OK;
}

}  // namespace

lbmibm_status ibm_spread_adjoint(lbmibm_ctx *ctx) {
    return ctx->prec == LBMIBM_F64 ? spread_adj_impl<double>(ctx) : spread_adj_impl<float>(ctx);
}

lbmibm_status ibm_interp_adjoint(lbmibm_ctx *ctx) {
    if (ctx->p.runge_kutta != 4) {
        ctx->err = "adjoint: the advection adjoint supports RUNGE_KUTTA = 4";
        return LBMIBM_ERR_STATE;
    }
    return ctx->prec == LBMIBM_F64 ? advect_adj_impl<double>(ctx) : advect_adj_impl<float>(ctx);
}

lbmibm_status ibm_setup(lbmibm_ctx *ctx) {
    const int lds64 = (int)(3 * (size_t)RN * sizeof(double));
    const int lds32 = (int)(3 * (size_t)RN * sizeof(float));
    LBMIBM_HIP(ctx, hipFuncSetAttribute((const void *)k_interp_tiles<double>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, lds64));
    LBMIBM_HIP(ctx, hipFuncSetAttribute((const void *)k_interp_tiles<float>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, lds32));
    LBMIBM_HIP(ctx, hipFuncSetAttribute((const void *)k_advect_adj_scatter<double>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, lds64));
    LBMIBM_HIP(ctx, hipFuncSetAttribute((const void *)k_advect_adj_scatter<float>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, lds64));
    return LBMIBM_OK;
}

lbmibm_status ibm_count_begin(lbmibm_ctx *ctx, hipStream_t stream) {
    hipLaunchKernelGGL(k_zero_int, dim3((ctx->ntiles + 255) / 256), dim3(256), 0, stream, ctx->tile_count,
                       ctx->ntiles);
    LBMIBM_LAUNCHED(ctx, "k_zero_int");
    return LBMIBM_OK;
}

lbmibm_status ibm_count_end(lbmibm_ctx *ctx, hipStream_t stream) { return tile_offsets(ctx, stream, 0); }

lbmibm_status ibm_spread(lbmibm_ctx *ctx) {
    lbmibm_status s = ctx->prec == LBMIBM_F64 ? spread_impl<double>(ctx) : spread_impl<float>(ctx);
    if (s == LBMIBM_OK) ctx->force_zero = (ctx->nv == 0);
    return s;
}

// Host tables of spread form 1: for each coordinate of each axis, the
// (tile along the axis, box coordinate) pairs whose box node lands on it,
// tiles ascending; false when a coordinate has more than SB_MAXC (lattices
// much smaller than a box) or the tile counts do not fit the packing.
static bool spread_box_tables(const lbmibm_ctx *ctx, std::vector<int> &out) {
    const int n[3] = {ctx->nx, ctx->ny, ctx->nz}, nt[3] = {ctx->ntx, ctx->nty, ctx->ntz};
    const int tile[3] = {TILE_X, TILE_Y, TILE_Z}, R[3] = {SB_X, SB_Y, SB_Z};
    const bool wrap[3] = {true, ctx->p.periodic_y != 0, true};
    out.assign((size_t)SB_MAXC * (n[0] + n[1] + n[2]), -1);
    size_t base = 0;
    for (int a = 0; a < 3; a++) {
        if (nt[a] >= (1 << 25)) return false;
        for (int t = 0; t < nt[a]; t++)
            for (int k = 0; k < R[a]; k++) {
                int cc = t * tile[a] + k;
                if (cc >= n[a]) {
                    if (!wrap[a]) continue;
                    cc %= n[a];
                }
                int *e = &out[base + (size_t)SB_MAXC * cc];
                int slot = 0;
                while (slot < SB_MAXC && e[slot] >= 0) slot++;
                if (slot == SB_MAXC) return false;
                e[slot] = t * 32 + k;
            }
        base += (size_t)SB_MAXC * n[a];
    }
    return true;
}

// Allocates the staging boxes and uploads the tables of spread form 1;
// *ok = false where the form does not apply.
static lbmibm_status spread_box_setup(lbmibm_ctx *ctx, bool *ok) {
    *ok = false;
    std::vector<int> tab;
    if (!spread_box_tables(ctx, tab)) return LBMIBM_OK;
    if (!ctx->sb_cand) {
        lbmibm_status s = dev_alloc(ctx, (void **)&ctx->sb_cand, tab.size() * sizeof(int));
        if (s != LBMIBM_OK) return s;
        LBMIBM_HIP(ctx, hipMemcpy(ctx->sb_cand, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    if (!ctx->sbox) {
        lbmibm_status s = dev_alloc(ctx, &ctx->sbox, (size_t)ctx->ntiles * 3 * SB_N * ctx->szT);
        if (s != LBMIBM_OK) return s;
    }
    *ok = true;
    return LBMIBM_OK;
}

// Times the spreading forms on the binned initial positions (the force
// field they write is overwritten by the first step) and keeps the faster.
template <typename T> lbmibm_status spread_autotune_impl(lbmibm_ctx *ctx) {
    if (!ctx->bins_valid) {
        lbmibm_status s = bin_impl<T>(ctx);
        if (s != LBMIBM_OK) return s;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    lbmibm_status s = LBMIBM_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) s = LBMIBM_ERR_HIP;
    static const char *names[2] = {"source bins per tile", "home boxes + gather"};
    char line[160];
    float best = 0.0f;
    for (int f = 0; f < 2 && s == LBMIBM_OK; f++) {
        s = spread_launch<T>(ctx, f);  // warm-up
        if (s != LBMIBM_OK) break;
        if (hipEventRecord(e0, ctx->stream) != hipSuccess) { s = LBMIBM_ERR_HIP; break; }
        for (int r = 0; r < 3 && s == LBMIBM_OK; r++) s = spread_launch<T>(ctx, f);
        float ms = 0.0f;
        if (s == LBMIBM_OK && (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                               hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            s = LBMIBM_ERR_HIP;
        if (s != LBMIBM_OK) break;
        ms /= 3.0f;
        snprintf(line, sizeof line, "spread form %s: %.4f ms\n", names[f], ms);
        ctx->tune_report += line;
        if (f == 0 || ms < best) { best = ms; ctx->spread_form = f; }
    }
    if (s == LBMIBM_OK) {
        snprintf(line, sizeof line, "spread: using %s\n", names[ctx->spread_form]);
        ctx->tune_report += line;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return s;
}

// a form fixed in the environment: "0" .. last into *form, -1 when unset
static lbmibm_status env_form(const char *name, int *form, char last = '1') {
    *form = -1;
    const char *v = getenv(name);
    if (!v || !v[0]) return LBMIBM_OK;
    if (v[0] < '0' || v[0] > last || v[1]) return LBMIBM_ERR_ARG;
    *form = v[0] - '0';
    return LBMIBM_OK;
}

lbmibm_status ibm_autotune(lbmibm_ctx *ctx) {
    ctx->interp_form = 0;
    ctx->spread_form = 0;
    if (ctx->nv == 0) return LBMIBM_OK;
    // LBMIBM_INTERP_FORM / LBMIBM_SPREAD_FORM = 0|1 fix a form;
    // LBMIBM_AUTOTUNE=0 keeps forms 0
    int fi = -1, fs = -1;
    lbmibm_status s = env_form("LBMIBM_INTERP_FORM", &fi);
    if (s == LBMIBM_OK) s = env_form("LBMIBM_SPREAD_FORM", &fs);
    if (s != LBMIBM_OK) return s;
    bool box_ok = false;
    s = spread_box_setup(ctx, &box_ok);
    if (s != LBMIBM_OK) return s;
    if (fs == 1 && !box_ok) return LBMIBM_ERR_ARG;  // the lattice is too small for the boxes
    const char *env = getenv("LBMIBM_AUTOTUNE");
    const bool tune = !(env && env[0] == '0');
    if (fi >= 0) {
        ctx->interp_form = fi;
        ctx->tune_report += "interp form fixed by LBMIBM_INTERP_FORM\n";
    } else if (tune) {
        s = ctx->prec == LBMIBM_F64 ? interp_autotune_impl<double>(ctx) : interp_autotune_impl<float>(ctx);
        if (s != LBMIBM_OK) return s;
    }
    if (fs >= 0) {
        ctx->spread_form = fs;
        ctx->tune_report += "spread form fixed by LBMIBM_SPREAD_FORM\n";
    } else if (tune && box_ok) {
        s = ctx->prec == LBMIBM_F64 ? spread_autotune_impl<double>(ctx) : spread_autotune_impl<float>(ctx);
        if (s != LBMIBM_OK) return s;
    }
    return LBMIBM_OK;
}

lbmibm_status ibm_interp(lbmibm_ctx *ctx) {
    return ctx->prec == LBMIBM_F64 ? interp_impl<double>(ctx) : interp_impl<float>(ctx);
}

}  // namespace lbmibm
