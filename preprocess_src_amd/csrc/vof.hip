// VOF indicator of the cell interiors: front tracking (fronttrack.cu:274-724)
// with its red-black SOR Poisson solve (matrix_solver.cu:112-216), and the
// THINC/WLIC advection (thincwlic.cu:12-372).  The slab (walls in y,
// periodic x and z) or the tube (walls outside the disc, periodic z), FP64
// throughout.
//
// Front tracking works on the reference's padded lattice (nx+4)(ny+4)(nz+4)
// so that its halo fold, its BUFF and WALL rows and its padding rows are the
// reference's own:
//   - the elements' inward area vectors are spread with the cosine delta by
//     64-bit fixed-point atomics (integer sums do not depend on the order, so
//     the field is deterministic; one unit is 2^-44 of 1/dx);
//   - the x/z halo sums are folded onto their COMP images and copied back;
//   - b = div G / (-6/dx^2) on COMP nodes;
//   - the SOR runs matrix-free on the Jacobi-scaled rows: COMP x_i - (1/6)
//     sum of the 6 neighbours = b_i, BUFF x_i - x_image = 0, WALL x_i = 0,
//     and the expand2pow2(n) - n identity rows of the padding, whose common
//     value is carried as one scalar.  Per colour, the COMP and WALL rows go
//     first and the BUFF rows then read the updated image (the reference
//     updates them in one kernel, with a read of a same-colour image that
//     races with its update; this is one of the orders it can produce).
//     The convergence test sqrt(r.r / b.b) < RESIDUAL runs on the device
//     after every iteration; later kernels of a batch return at once, so the
//     host reads the status once per batch.
// THINC/WLIC works on the unpadded lattice: halo values are the periodic
// images, which is what the reference's BUFF copies hold, and the wall
// layers hold vf = 0.
#include <algorithm>
#include <climits>

#include "common.hpp"

namespace lbmibm {
namespace {

constexpr double FT_UNIT = 17592186044416.0;  // 2^44 fixed-point units per 1/dx
constexpr int VWG = 256;
constexpr int FT_MAX_ATTEMPTS = 10;
constexpr int SOR_CHUNKS = 4;  // x-pair chunks of VWG per work-group of k_sor_rows
constexpr int CHECK_WG = 1024;

struct FtGeom {
    int nx, ny, nz;  // padded sizes
    long long np;
    double dx, lx, lz;  // spacing, physical LX, LZ
    int nxh;            // (nx + 1) / 2: nodes of one colour per row
    long long half;     // nxh * ny * nz: length of one colour's array
    const unsigned *tube_col;  // the tube's column words (bit 0: inside), or null for the slab
};

enum { FT_COMP = 0, FT_BUFF = 1, FT_WALL = 2 };

// bc class of a padded node: the slab (walls beyond the y range, BUFF in
// the x/z halo) or the tube (walls outside the disc, BUFF in the z halo),
// cdo_init (init1.cu:164-178, 217-229)
__device__ __forceinline__ int ft_kind(const FtGeom &g, int ix, int iy, int iz) {
    if (g.tube_col) {
        const int ux = ix - 2, uy = iy - 2, nxu = g.nx - 4, nyu = g.ny - 4;
        if (ux < 0 || ux >= nxu || uy < 0 || uy >= nyu || !(g.tube_col[ux + nxu * uy] & 1u)) return FT_WALL;
        return iz >= 2 && iz <= g.nz - 3 ? FT_COMP : FT_BUFF;
    }
    if (iy < 2 || iy > g.ny - 3) return FT_WALL;
    return ix >= 2 && ix <= g.nx - 3 && iz >= 2 && iz <= g.nz - 3 ? FT_COMP : FT_BUFF;
}

// The SOR iterate and right-hand side are stored by colour: node (ix, iy,
// iz) of colour c = (ix + iy + iz) & 1 sits at c * half + (ix >> 1) +
// nxh * (iy + ny * iz), so a colour's sweep reads and writes contiguous
// stretches and its neighbours (all of the other colour) likewise.
__device__ __forceinline__ long long packed(const FtGeom &g, int ix, int iy, int iz) {
    return (long long)((ix + iy + iz) & 1) * g.half + (ix >> 1) + (long long)g.nxh * (iy + (long long)g.ny * iz);
}

struct SorCtl {
    int done;  // 0 running, 1 converged, 2 ITRMAX reached
    int iter;
    int conv_iter;
    int pad;
    double xpad;  // value of every padding row
    double bb;    // b.b
    double err;
};

__device__ double block_sum(double v, double *lds) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += lds[k];
    return s;
}

template <typename V> __global__ void k_vfill(V *__restrict__ p, size_t n, V v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// front_tracking_source2 (fronttrack.cu:336-441): 64 lanes per element, one
// per node of the 4x4x4 box starting at floor(.)-1 (the box is 3 wide where
// the centroid sits on a node: ceil = floor).  A work group takes FT_CHUNK
// consecutive elements (FT_CHUNK / 4 rounds of 4 elements) and sums their
// terms into an LDS box around the chunk, then adds the box's nonzero nodes
// to G: one global atomic per touched node and component instead of one per
// term.  The sums are integers, so G does not depend on that grouping.  A
// chunk whose stencils span more than the box (elements of two cells, or a
// cell across a periodic face) adds its terms to G directly.
constexpr int FT_CHUNK = 64;
constexpr int FTB = 12, FTB_N = FTB * FTB * FTB;  // LDS box edge (nodes)

struct FtTerm {
    int ix, iy, iz;  // padded node; ix < 0: no term
    int fx, fy, fz;  // the first node floor(.)-1 of the element's 4x4x4 box
    long long v[3];
};

// The term of lane k of element e (as front_tracking_source2); VALUES false:
// the node only.
template <bool VALUES>
__device__ __forceinline__ FtTerm ft_term(const double *__restrict__ xs, int nv, const int *__restrict__ ele,
                                          long long e, int k, const FtGeom &g) {
    FtTerm out;
    out.ix = -1;
    const int kx = k & 3, ky = (k >> 2) & 3, kz = k >> 4;
    const double *X = xs, *Y = xs + nv, *Z = xs + 2 * (size_t)nv;
    const int j1 = ele[3 * e], j2 = ele[3 * e + 1], j3 = ele[3 * e + 2];
    const double LX = g.lx, LZ = g.lz, dx = g.dx;
    double tx1 = X[j1], tx2 = X[j2], tx3 = X[j3];
    double ty1 = Y[j1], ty2 = Y[j2], ty3 = Y[j3];
    double tz1 = Z[j1], tz2 = Z[j2], tz3 = Z[j3];
    double xx1 = tx2 - tx1, yy1 = ty2 - ty1, zz1 = tz2 - tz1;
    if (xx1 > LX * 0.5) { xx1 -= LX; tx1 += LX; }
    else if (xx1 < -LX * 0.5) { xx1 += LX; tx2 += LX; }
    if (zz1 > LZ * 0.5) { zz1 -= LZ; tz1 += LZ; }
    else if (zz1 < -LZ * 0.5) { zz1 += LZ; tz2 += LZ; }
    double xx2 = tx3 - tx1, yy2 = ty3 - ty1, zz2 = tz3 - tz1;
    if (xx2 > LX * 0.5) { xx2 -= LX; tx1 += LX; }
    else if (xx2 < -LX * 0.5) { xx2 += LX; tx3 += LX; }
    if (zz2 > LZ * 0.5) { zz2 -= LZ; tz1 += LZ; }
    else if (zz2 < -LZ * 0.5) { zz2 += LZ; tz3 += LZ; }
    const double xtr = (tx1 + tx2 + tx3) / 3.0 + (double)(g.nx - 1) * dx / 2.0;
    const double ytr = (ty1 + ty2 + ty3) / 3.0 + (double)(g.ny - 1) * dx / 2.0;
    const double ztr = (tz1 + tz2 + tz3) / 3.0 + (double)(g.nz - 1) * dx / 2.0;
    out.fx = (int)floor(xtr / dx) - 1;
    out.fy = (int)floor(ytr / dx) - 1;
    out.fz = (int)floor(ztr / dx) - 1;
    const int ix = out.fx + kx, iy = out.fy + ky, iz = out.fz + kz;
    if (ix > (int)ceil(xtr / dx) + 1 || iy > (int)ceil(ytr / dx) + 1 || iz > (int)ceil(ztr / dx) + 1) return out;
    if (ix < 0 || ix >= g.nx || iy < 0 || iy >= g.ny || iz < 0 || iz >= g.nz) return out;
    out.ix = ix; out.iy = iy; out.iz = iz;
    if (!VALUES) return out;
    double nvx = -(yy1 * zz2 - zz1 * yy2);
    double nvy = -(zz1 * xx2 - xx1 * zz2);
    double nvz = -(xx1 * yy2 - yy1 * xx2);
    const double nn = sqrt(nvx * nvx + nvy * nvy + nvz * nvz);
    const double ds = nn / 2.0;
    nvx /= nn; nvy /= nn; nvz /= nn;
    const double xcos = 1.0 + cos(M_PI * ((double)ix * dx - xtr) / (2.0 * dx));
    const double ycos = 1.0 + cos(M_PI * ((double)iy * dx - ytr) / (2.0 * dx));
    const double zcos = 1.0 + cos(M_PI * ((double)iz * dx - ztr) / (2.0 * dx));
    const double dlt = xcos * ycos * zcos / (64.0 * dx * dx * dx);
    const double s = dx * FT_UNIT;
    out.v[0] = __double2ll_rn(dlt * nvx * ds * s);
    out.v[1] = __double2ll_rn(dlt * nvy * ds * s);
    out.v[2] = __double2ll_rn(dlt * nvz * ds * s);
    return out;
}

__global__ __launch_bounds__(VWG) void k_ft_spread(const double *__restrict__ xs, int nv, const int *__restrict__ ele,
                                                   int ne, FtGeom g, unsigned long long *__restrict__ G) {
    __shared__ unsigned long long box[3 * FTB_N];
    __shared__ int lo[3], hi[3];
    const long long e0 = (long long)blockIdx.x * FT_CHUNK;
    const int k = threadIdx.x & 63, sub = threadIdx.x >> 6;  // 4 elements per round
    if (threadIdx.x < 3) { lo[threadIdx.x] = INT_MAX; hi[threadIdx.x] = INT_MIN; }
    for (int i = threadIdx.x; i < 3 * FTB_N; i += VWG) box[i] = 0ull;
    __syncthreads();
    // the extent of the chunk's terms: the 4x4x4 box from floor(.)-1 of each
    // element's centroid, one element per lane of the first wave
    if (threadIdx.x < 64) {
        int l[3] = {INT_MAX, INT_MAX, INT_MAX}, h[3] = {INT_MIN, INT_MIN, INT_MIN};
        const long long e = e0 + threadIdx.x;
        if (threadIdx.x < FT_CHUNK && e < ne) {
            const FtTerm t = ft_term<false>(xs, nv, ele, e, 0, g);  // lane 0: the box's first node
            l[0] = t.fx; l[1] = t.fy; l[2] = t.fz;
            h[0] = t.fx + 3; h[1] = t.fy + 3; h[2] = t.fz + 3;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                l[d] = min(l[d], __shfl_xor(l[d], o));
                h[d] = max(h[d], __shfl_xor(h[d], o));
            }
        if (threadIdx.x == 0) {
            // only nodes inside the padded lattice take terms
            lo[0] = max(l[0], 0); lo[1] = max(l[1], 0); lo[2] = max(l[2], 0);
            hi[0] = min(h[0], g.nx - 1); hi[1] = min(h[1], g.ny - 1); hi[2] = min(h[2], g.nz - 1);
        }
    }
    __syncthreads();
    if (lo[0] > hi[0]) return;  // no term in the chunk (uniform)
    const int ox = lo[0], oy = lo[1], oz = lo[2];
    const int ex = hi[0] - ox + 1, ey = hi[1] - oy + 1, ez = hi[2] - oz + 1;
    const bool local = ex <= FTB && ey <= FTB && ez <= FTB;
    for (int r = 0; r < FT_CHUNK / 4; r++) {
        const long long e = e0 + 4 * r + sub;
        if (e >= ne) break;
        const FtTerm t = ft_term<true>(xs, nv, ele, e, k, g);
        if (t.ix < 0) continue;
        if (local) {
            const int b = (t.ix - ox) + FTB * ((t.iy - oy) + FTB * (t.iz - oz));
            for (int d = 0; d < 3; d++) atomicAdd(&box[d * FTB_N + b], (unsigned long long)t.v[d]);
        } else {
            const long long i = t.ix + (long long)g.nx * (t.iy + (long long)g.ny * t.iz);
            for (int d = 0; d < 3; d++) atomicAdd(G + d * g.np + i, (unsigned long long)t.v[d]);
        }
    }
    if (!local) return;  // uniform
    __syncthreads();
    for (int b = threadIdx.x; b < ex * ey * ez; b += VWG) {
        const int bx = b % ex, by = (b / ex) % ey, bz = b / (ex * ey);
        const int lb = bx + FTB * (by + FTB * bz);
        const long long i = (ox + bx) + (long long)g.nx * ((oy + by) + (long long)g.ny * (oz + bz));
        for (int d = 0; d < 3; d++) {
            const unsigned long long v = box[d * FTB_N + lb];
            if (v) atomicAdd(G + d * g.np + i, v);
        }
    }
}

__device__ __forceinline__ bool ft_comp(const FtGeom &g, int ix, int iy, int iz) {
    return ft_kind(g, ix, iy, iz) == FT_COMP;
}

// front_tracking_source3 (fronttrack.cu:444-499) in exact integer sums.
__global__ __launch_bounds__(VWG) void k_ft_fold(FtGeom g, unsigned long long *__restrict__ G) {
    const long long i = (long long)blockIdx.x * VWG + threadIdx.x;
    if (i >= g.np) return;
    const int ix = (int)(i % g.nx), iy = (int)((i / g.nx) % g.ny), iz = (int)(i / ((long long)g.nx * g.ny));
    if (!ft_comp(g, ix, iy, iz)) return;
    // the tube folds z only (fronttrack.cu:466-482)
    const bool xl = !g.tube_col && ix <= 3, xh = !g.tube_col && ix >= g.nx - 4, zl = iz <= 3, zh = iz >= g.nz - 4;
    if (!(xl || xh || zl || zh)) return;
    const long long sx = g.nx - 4, sz = (long long)g.nx * g.ny * (g.nz - 4);
    const long long ox = xl ? sx : -sx, oz = zl ? sz : -sz;
    long long img[3] = {-1, -1, -1};
    if ((xl || xh) && (zl || zh)) { img[0] = i + ox; img[1] = i + oz; img[2] = i + ox + oz; }
    else if (xl || xh) img[0] = i + ox;
    else img[0] = i + oz;
    for (int c = 0; c < 3; c++) {
        unsigned long long *Gc = G + c * g.np;
        unsigned long long v = Gc[i];
        for (int k = 0; k < 3; k++)
            if (img[k] >= 0) v += Gc[img[k]];
        Gc[i] = v;
        for (int k = 0; k < 3; k++)
            if (img[k] >= 0) Gc[img[k]] = v;
    }
}

// front_tracking_source4 + jacobi_source (fronttrack.cu:502-556, 314-333);
// per-block sums of b^2.
__global__ __launch_bounds__(VWG) void k_ft_rhs(FtGeom g, const unsigned long long *__restrict__ G,
                                                double *__restrict__ b, double *__restrict__ part) {
    __shared__ double lds[VWG / 64];
    const long long i = (long long)blockIdx.x * VWG + threadIdx.x;
    double bi = 0.0;
    if (i < g.np) {
        const int ix = (int)(i % g.nx), iy = (int)((i / g.nx) % g.ny), iz = (int)(i / ((long long)g.nx * g.ny));
        if (ft_comp(g, ix, iy, iz)) {
            const double u = 1.0 / (g.dx * FT_UNIT);
            const long long nx = g.nx, nxy = (long long)g.nx * g.ny;
            auto v = [&](int c, long long j) { return (double)(long long)G[c * g.np + j] * u; };
            const double dg = (v(0, i + 1) - v(0, i - 1) + v(1, i + nx) - v(1, i - nx) + v(2, i + nxy) - v(2, i - nxy)) /
                              (2.0 * g.dx);
            bi = dg / (-6.0 / (g.dx * g.dx));
        }
        b[packed(g, ix, iy, iz)] = bi;
    }
    const double s = block_sum(bi * bi, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_ft_start(SorCtl *__restrict__ ctl, const double *__restrict__ part, int nparts, double cnt) {
    __shared__ double lds[VWG / 64];
    double s = 0.0;
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) s += part[k];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) {
        ctl->done = 0;
        ctl->iter = 0;
        ctl->conv_iter = 0;
        ctl->xpad = cnt;
        ctl->bb = s;
        ctl->err = 0.0;
    }
}

// One colour of the COMP and WALL rows.  One lane per x-pair of a row: the
// lane takes the node of colour C, so a wave reads and writes a contiguous
// stretch of the row.  A work-group takes SOR_CHUNKS consecutive stretches
// and the work-groups of one XCD a contiguous range of planes (their
// neighbour planes stay in that XCD's L2); one partial sum of r^2 per
// work-group.
template <int C>
__global__ __launch_bounds__(VWG) void k_sor_rows(FtGeom g, double *__restrict__ x, const double *__restrict__ b,
                                                  const SorCtl *__restrict__ ctl, double *__restrict__ part, double a,
                                                  int nblocks) {
#pragma clang fp contract(off)
    __shared__ double lds[VWG / 64];
    if (ctl->done) return;
    const int lb = xcd_block(nblocks);
    if (lb >= nblocks) return;
    const unsigned nxh = (unsigned)(g.nx + 1) >> 1, nyu = (unsigned)g.ny;
    double r2 = 0.0;
    for (int k = 0; k < SOR_CHUNKS; k++) {
    const unsigned t = ((unsigned)lb * SOR_CHUNKS + k) * VWG + threadIdx.x;
    const unsigned row = t / nxh;
    if (row < nyu * (unsigned)g.nz) {
        const unsigned izu = row / nyu;
        const int iy = (int)(row - izu * nyu), iz = (int)izu;
        const int ix = 2 * (int)(t - row * nxh) + ((C + iy + iz) & 1);
        if (ix < g.nx) {
            // own colour C at k + nxh*row; the neighbours, of colour 1-C, at
            // their own x >> 1 in rows row, row -+ 1 (y), row -+ ny (z)
            const unsigned k = t - row * nxh;
            double *xc = x + (size_t)C * g.half;
            const double *xo = x + (size_t)(1 - C) * g.half;
            const unsigned i = k + nxh * row;
            const int kind = ft_kind(g, ix, iy, iz);
            const bool wall = kind == FT_WALL, comp = kind == FT_COMP;
            if (comp || wall) {
                double sum, bi = 0.0;
                if (comp) {
                    const unsigned rowz = nxh * (unsigned)g.ny;
                    const unsigned xm = ((unsigned)ix - 1) >> 1, xp = ((unsigned)ix + 1) >> 1;
                    const unsigned r0 = nxh * row;
                    sum = 0.0;
                    sum += a * xo[k + r0 - rowz];
                    sum += a * xo[k + r0 - nxh];
                    sum += a * xo[xm + r0];
                    sum += 1.0 * xc[i];
                    sum += a * xo[xp + r0];
                    sum += a * xo[k + r0 + nxh];
                    sum += a * xo[k + r0 + rowz];
                    bi = b[(size_t)C * g.half + i];
                } else {
                    sum = 1.0 * xc[i];
                }
                const double r = bi - sum;
                xc[i] = xc[i] + 1.3 * r;
                r2 += r * r;
            }
        }
    }
    }
    const double s = block_sum(r2, lds);
    if (threadIdx.x == 0) part[lb] = s;
}

// One colour of the BUFF rows x_i - x_image = 0: the x halo (all z) then
// the z halo (COMP x), y within the COMP range.
template <int C>
__global__ __launch_bounds__(VWG) void k_sor_buff(FtGeom g, double *__restrict__ x, const SorCtl *__restrict__ ctl,
                                                  double *__restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double lds[VWG / 64];
    if (ctl->done) return;
    const long long t = (long long)blockIdx.x * VWG + threadIdx.x;
    const long long ny4 = g.ny - 4, nx4 = g.nx - 4;
    const long long n1 = 4 * ny4 * g.nz, n2 = 4 * nx4 * ny4;
    double r2 = 0.0;
    if (t < n1 + n2) {
        int ix, iy, iz;
        if (t < n1) {
            const int a = (int)(t & 3);
            const long long rest = t >> 2;
            ix = a < 2 ? a : g.nx - 4 + a;
            iy = 2 + (int)(rest % ny4);
            iz = (int)(rest / ny4);
        } else {
            const long long u = t - n1;
            const int a = (int)(u & 3);
            const long long rest = u >> 2;
            iz = a < 2 ? a : g.nz - 4 + a;
            ix = 2 + (int)(rest % nx4);
            iy = 2 + (int)(rest / nx4);
        }
        if (((ix + iy + iz) & 1) == C && ft_kind(g, ix, iy, iz) == FT_BUFF) {
            const int jx = ix <= 1 ? ix + (g.nx - 4) : (ix >= g.nx - 2 ? ix - (g.nx - 4) : ix);
            const int jz = iz <= 1 ? iz + (g.nz - 4) : (iz >= g.nz - 2 ? iz - (g.nz - 4) : iz);
            const long long i = packed(g, ix, iy, iz), j = packed(g, jx, iy, jz);
            const double r = 0.0 - (x[i] - x[j]);
            x[i] = x[i] + 1.3 * r;
            r2 = r * r;
        }
    }
    const double s = block_sum(r2, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Residual norm of one iteration (the padding rows, all equal, contribute
// npad x_pad^2) and the stop test of matrix_solver.cu:198-205.
__global__ __launch_bounds__(CHECK_WG) void k_sor_check(SorCtl *__restrict__ ctl, const double *__restrict__ part,
                                                         int nparts, double npad, int itrmax, double residual) {
#pragma clang fp contract(off)
    __shared__ double lds[CHECK_WG / 64];
    if (ctl->done) return;
    double s = 0.0;
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) s += part[k];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) {
        const double xp = ctl->xpad;
        const double rr = s + npad * (xp * xp);
        ctl->xpad = xp + 1.3 * (0.0 - 1.0 * xp);
        const int it = ctl->iter + 1;
        ctl->iter = it;
        const double err = sqrt(rr / ctl->bb);
        ctl->err = err;
        if (err < residual) {
            ctl->done = 1;
            ctl->conv_iter = it;
        } else if (it >= itrmax) {
            ctl->done = 2;
        }
    }
}

__global__ __launch_bounds__(VWG) void k_ft_extract(FtGeom g, const double *__restrict__ x, double *__restrict__ vf,
                                                    int nx, int ny, long long N) {
    const long long j = (long long)blockIdx.x * VWG + threadIdx.x;
    if (j >= N) return;
    const int ix = (int)(j % nx), iy = (int)((j / nx) % ny), iz = (int)(j / ((long long)nx * ny));
    vf[j] = x[packed(g, ix + 2, iy + 2, iz + 2)];
}

// ---- THINC/WLIC ---------------------------------------------------------

struct VofDims {
    int nx, ny, nz;  // unpadded
    long long N;
    int m;
    double dt, dx;
    const unsigned *tube_col;  // the tube's column words, or null for the slab
};

// a wall node of the unpadded frame (beyond the slab's y range, or outside
// the tube's disc)
__device__ __forceinline__ bool vof_wall(const VofDims &d, int x, int y) {
    if (d.tube_col) return x < 0 || x >= d.nx || y < 0 || y >= d.ny || !(d.tube_col[x + d.nx * y] & 1u);
    return y < 0 || y >= d.ny;
}

__device__ __forceinline__ int wrapi(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// thincwlic.cu:187-197 (expf / coshf of float arguments, as there)
__device__ double thinc_flux(double c, double alpha, double uc, int isgn, int m, double dt, double dx) {
#pragma clang fp contract(off)
    const double beta = 2.0;
    const double a1 = (double)expf((float)(beta * (2.0 * c - 1.0) / alpha));
    const double a3 = (double)expf((float)beta);
    const double xc = 0.5 / beta * log((a3 * a3 - a1 * a3) / (a1 * a3 - 1.0));
    const double a4 = (double)coshf((float)(beta * ((double)isgn - uc * (double)m * dt / dx - xc)));
    const double a5 = (double)coshf((float)(beta * ((double)isgn - xc)));
    return 0.5 * (uc * (double)m * dt - alpha * dx / beta * log(a4 / a5));
}

// wlic1d4md_st1 (thincwlic.cu:108-214) of sweep D: the flux through the
// face below each node along D (the reference's flux[i]), from vin.
template <typename T, int D>
__global__ __launch_bounds__(VWG) void k_thinc_flux(VofDims d, const double *__restrict__ vin,
                                                    double *__restrict__ flux, const T *__restrict__ u) {
#pragma clang fp contract(off)
    const unsigned j = blockIdx.x * VWG + threadIdx.x;
    if (j >= (unsigned)d.N) return;
    const unsigned r = j / (unsigned)d.nx, izu = r / (unsigned)d.ny;
    const int ix = (int)(j - r * (unsigned)d.nx), iy = (int)(r - izu * (unsigned)d.ny), iz = (int)izu;
    auto V = [&](int x, int y, int z) -> double {
        if (vof_wall(d, d.tube_col ? x : 0, y)) return 0.0;  // wall nodes hold vf = 0
        x = wrapi(x, d.nx);
        z = wrapi(z, d.nz);
        return vin[(unsigned)x + (unsigned)d.nx * ((unsigned)y + (unsigned)d.ny * (unsigned)z)];
    };
    auto Uv = [&](int x, int y, int z) -> double {
        x = wrapi(x, d.nx);
        z = wrapi(z, d.nz);
        return (double)u[(unsigned)x + (unsigned)d.nx * ((unsigned)y + (unsigned)d.ny * (unsigned)z)];
    };
    constexpr int ex = D == 0, ey = D == 1, ez = D == 2;
    // flux through the face between (bx,by,bz) - e_D and (bx,by,bz)
    auto face = [&](int bx, int by, int bz) -> double {
        const int ax = bx - ex, ay = by - ey, az = bz - ez;
        if (d.tube_col) {  // a wall node on either side (thincwlic.cu:205)
            if (vof_wall(d, bx, by) || vof_wall(d, ax, ay)) return 0.0;
        } else if (D == 1 && (by == 0 || by == d.ny)) {
            return 0.0;
        }
        const double uc = (Uv(bx, by, bz) + Uv(ax, ay, az)) / 2.0;
        const int isgn = uc <= 0.0 ? 0 : 1;
        const int px = isgn ? ax : bx, py = isgn ? ay : by, pz = isgn ? az : bz;
        const double c = V(px, py, pz);
        if (c >= 1.0 - 1.0e-6 || c <= 1.0e-6) return c * uc * (double)d.m * d.dt;
        // normal3d_st1/st2 at the upwind node: 3x3x3 neighbourhood
        double v[3][3][3];
#pragma unroll
        for (int c2 = 0; c2 < 3; c2++)
#pragma unroll
            for (int b2 = 0; b2 < 3; b2++)
#pragma unroll
                for (int a2 = 0; a2 < 3; a2++) v[c2][b2][a2] = V(px + a2 - 1, py + b2 - 1, pz + c2 - 1);
        // normal3d_st1/st2 (thincwlic.cu:12-106): the 2x2x2-cube differences
        // averaged over the 8 cubes around the node are the central
        // difference along the axis weighted (1,2,1) x (1,2,1) across it,
        // over 32 dx; summed in that form (9 differences per axis instead of
        // 64 terms; the same value up to rounding)
        auto n1 = [&](int dir) -> double {
            double acc = 0.0;
#pragma unroll
            for (int b2 = 0; b2 < 3; b2++)
#pragma unroll
                for (int c2 = 0; c2 < 3; c2++) {
                    const double w = (double)((b2 == 1 ? 2 : 1) * (c2 == 1 ? 2 : 1));
                    double diff;
                    if (dir == 0) diff = v[c2][b2][2] - v[c2][b2][0];
                    else if (dir == 1) diff = v[c2][2][b2] - v[c2][0][b2];
                    else diff = v[2][c2][b2] - v[0][c2][b2];
                    acc += w * diff;
                }
            return acc / (32.0 * d.dx);
        };
        const double n1x = n1(0), n1y = n1(1), n1z = n1(2);
        const double nn = fabs(n1x) + fabs(n1y) + fabs(n1z);
        if (nn <= 1.0e-6) return c * uc * (double)d.m * d.dt;
        const double alpha = v[1 - ez][1 - ey][1 - ex] <= v[1 + ez][1 + ey][1 + ex] ? 1.0 : -1.0;
        const double f = thinc_flux(c, alpha, uc, isgn, d.m, d.dt, d.dx);
        const double weight = fabs(D == 0 ? n1x : (D == 1 ? n1y : n1z)) / nn;
        return f * weight + c * uc * (double)d.m * d.dt * (1.0 - weight);
    };
    flux[j] = face(ix, iy, iz);
}

// wlic1d4md_st2 (thincwlic.cu:217-250) of sweep D; the face above the last
// row along y is a wall face (flux 0), along x and z it is the periodic
// image's lower face.
template <int D>
__global__ __launch_bounds__(VWG) void k_thinc_update(VofDims d, const double *__restrict__ vin,
                                                      const double *__restrict__ flux, double *__restrict__ vout) {
#pragma clang fp contract(off)
    const unsigned j = blockIdx.x * VWG + threadIdx.x;
    if (j >= (unsigned)d.N) return;
    const unsigned r = j / (unsigned)d.nx, izu = r / (unsigned)d.ny;
    const int ix = (int)(j - r * (unsigned)d.nx), iy = (int)(r - izu * (unsigned)d.ny), iz = (int)izu;
    const unsigned nxy = (unsigned)d.nx * (unsigned)d.ny;
    if (d.tube_col && vof_wall(d, ix, iy)) {  // outside the tube: vf = 0
        vout[j] = 0.0;
        return;
    }
    double hi;
    if (D == 0) hi = flux[ix == d.nx - 1 ? j + 1 - (unsigned)d.nx : j + 1];
    else if (D == 1) hi = iy == d.ny - 1 ? 0.0 : flux[j + (unsigned)d.nx];
    else hi = flux[iz == d.nz - 1 ? j - nxy * (unsigned)(d.nz - 1) : j + nxy];
    vout[j] = vin[j] - (hi - flux[j]) / d.dx;
}

FtGeom ft_geom(const lbmibm_ctx *ctx) {
    FtGeom g;
    g.nx = ctx->nxp; g.ny = ctx->nyp; g.nz = ctx->nzp;
    g.np = ctx->np;
    g.dx = ctx->p.dx;
    g.lx = ctx->p.lx * ctx->p.dx;
    g.lz = ctx->p.lz * ctx->p.dx;
    g.nxh = (g.nx + 1) / 2;
    g.half = (long long)g.nxh * g.ny * g.nz;
    g.tube_col = ctx->p.tube_radius > 0.0 ? (const unsigned *)ctx->tube_col : nullptr;
    return g;
}

inline unsigned blocks_of(long long n) { return (unsigned)((n + VWG - 1) / VWG); }

long long sor_row_threads(const FtGeom &g) { return (long long)((g.nx + 1) / 2) * g.ny * g.nz; }
int sor_row_blocks(const FtGeom &g) {
    return (int)((sor_row_threads(g) + (long long)VWG * SOR_CHUNKS - 1) / ((long long)VWG * SOR_CHUNKS));
}
long long sor_buff_threads(const FtGeom &g) {
    return 4LL * (g.ny - 4) * g.nz + 4LL * (g.nx - 4) * (g.ny - 4);
}

int expand_pow2(long long n) {  // etc.cu:425-446
    int cnt = 0;
    while (n > 1) {
        if (n % 2 == 1) n++;
        n /= 2;
        cnt++;
    }
    return cnt;
}

}  // namespace

lbmibm_status vof_alloc(lbmibm_ctx *ctx) {
    if (ctx->vof[0]) return LBMIBM_OK;
    const FtGeom g = ft_geom(ctx);
    const long long nb_rows = sor_row_blocks(g), nb_buff = blocks_of(sor_buff_threads(g));
    ctx->ft_nparts = (int)(2 * (nb_rows + nb_buff));
    const long long nparts = std::max<long long>(ctx->ft_nparts, blocks_of(g.np));
    lbmibm_status s = dev_alloc(ctx, (void **)&ctx->vof[0], ctx->N * sizeof(double));
    if (s == LBMIBM_OK) s = dev_alloc(ctx, (void **)&ctx->vof[1], ctx->N * sizeof(double));
    if (s == LBMIBM_OK) s = dev_alloc(ctx, (void **)&ctx->ft_g, 3 * g.np * sizeof(unsigned long long));
    if (s == LBMIBM_OK) s = dev_alloc(ctx, (void **)&ctx->ft_x, 2 * g.half * sizeof(double));
    if (s == LBMIBM_OK) s = dev_alloc(ctx, (void **)&ctx->ft_b, 2 * g.half * sizeof(double));
    if (s == LBMIBM_OK) s = dev_alloc(ctx, (void **)&ctx->ft_part, nparts * sizeof(double));
    if (s == LBMIBM_OK) s = dev_alloc(ctx, &ctx->ft_ctl, sizeof(SorCtl));
    if (s != LBMIBM_OK) return s;
    for (int k = 0; k < 2; k++) {
        hipLaunchKernelGGL(k_vfill<double>, dim3(blocks_of(ctx->N)), dim3(VWG), 0, ctx->stream, ctx->vof[k],
                           (size_t)ctx->N, 0.0);
        LBMIBM_LAUNCHED(ctx, "k_vfill");
    }
    ctx->vof_cur = 0;
    return LBMIBM_OK;
}

lbmibm_status vof_front_tracking(lbmibm_ctx *ctx, int *iterations, int *attempts) {
    if (iterations) *iterations = 0;
    if (attempts) *attempts = 0;
    lbmibm_status s = vof_alloc(ctx);
    if (s != LBMIBM_OK) return s;
    const FtGeom g = ft_geom(ctx);
    hipStream_t st = ctx->stream;
    unsigned long long *G = ctx->ft_g;
    hipLaunchKernelGGL(k_vfill<unsigned long long>, dim3(blocks_of(3 * g.np)), dim3(VWG), 0, st, G,
                       (size_t)(3 * g.np), 0ull);
    LBMIBM_LAUNCHED(ctx, "k_vfill");
    if (ctx->nt > 0) {
        hipLaunchKernelGGL(k_ft_spread, dim3((unsigned)((ctx->nt + FT_CHUNK - 1) / FT_CHUNK)), dim3(VWG), 0, st,
                           (const double *)ctx->xs,
                           ctx->nv, ctx->ele, ctx->nt, g, G);
        LBMIBM_LAUNCHED(ctx, "k_ft_spread");
    }
    hipLaunchKernelGGL(k_ft_fold, dim3(blocks_of(g.np)), dim3(VWG), 0, st, g, G);
    LBMIBM_LAUNCHED(ctx, "k_ft_fold");
    hipLaunchKernelGGL(k_ft_rhs, dim3(blocks_of(g.np)), dim3(VWG), 0, st, g, (const unsigned long long *)G, ctx->ft_b,
                       ctx->ft_part);
    LBMIBM_LAUNCHED(ctx, "k_ft_rhs");
    // per-block sums of b^2 are in ft_part until k_ft_start adds them up
    const int nb_b = (int)blocks_of(g.np);
    const unsigned nb_rows = sor_row_blocks(g), nb_buff = blocks_of(sor_buff_threads(g));
    const unsigned grid_rows = xcd_grid(nb_rows);
    double *part = ctx->ft_part;
    double *p_r0 = part, *p_b0 = part + nb_rows, *p_r1 = part + nb_rows + nb_buff, *p_b1 = p_r1 + nb_rows;
    SorCtl *ctl = (SorCtl *)ctx->ft_ctl;
    const double a = (1.0 / (g.dx * g.dx)) / (-6.0 / (g.dx * g.dx));  // fronttrack.cu:97-110, 294-311
    const long long mn = 1LL << expand_pow2(g.np);
    const double npad = (double)(mn - g.np);
    // b.b of the first start, kept for the restarts (b does not change)
    hipLaunchKernelGGL(k_ft_start, dim3(1), dim3(VWG), 0, st, ctl, (const double *)part, nb_b, 0.0);
    LBMIBM_LAUNCHED(ctx, "k_ft_start");
    SorCtl h{};
    LBMIBM_HIP(ctx, hipMemcpyAsync(&h, ctl, sizeof(SorCtl), hipMemcpyDeviceToHost, st));
    LBMIBM_HIP(ctx, hipStreamSynchronize(st));
    const double bb = h.bb;
    int cnt = 0;
    for (;;) {
        hipLaunchKernelGGL(k_vfill<double>, dim3(blocks_of(2 * g.half)), dim3(VWG), 0, st, ctx->ft_x,
                           (size_t)(2 * g.half), (double)cnt);
        LBMIBM_LAUNCHED(ctx, "k_vfill");
        h = SorCtl{};
        h.xpad = (double)cnt;
        h.bb = bb;
        LBMIBM_HIP(ctx, hipMemcpyAsync(ctl, &h, sizeof(SorCtl), hipMemcpyHostToDevice, st));
        int batch = 16;
        int issued = 0;
        for (;;) {
            for (int k = 0; k < batch; k++) {
                hipLaunchKernelGGL(k_sor_rows<0>, dim3(grid_rows), dim3(VWG), 0, st, g, ctx->ft_x,
                                   (const double *)ctx->ft_b, (const SorCtl *)ctl, p_r0, a, (int)nb_rows);
                hipLaunchKernelGGL(k_sor_buff<0>, dim3(nb_buff), dim3(VWG), 0, st, g, ctx->ft_x, (const SorCtl *)ctl,
                                   p_b0);
                hipLaunchKernelGGL(k_sor_rows<1>, dim3(grid_rows), dim3(VWG), 0, st, g, ctx->ft_x,
                                   (const double *)ctx->ft_b, (const SorCtl *)ctl, p_r1, a, (int)nb_rows);
                hipLaunchKernelGGL(k_sor_buff<1>, dim3(nb_buff), dim3(VWG), 0, st, g, ctx->ft_x, (const SorCtl *)ctl,
                                   p_b1);
                hipLaunchKernelGGL(k_sor_check, dim3(1), dim3(CHECK_WG), 0, st, ctl, (const double *)part,
                                   (int)(2 * (nb_rows + nb_buff)), npad, ctx->p.sor_itrmax, ctx->p.sor_residual);
            }
            LBMIBM_LAUNCHED(ctx, "k_sor_*");
            issued += batch;
            LBMIBM_HIP(ctx, hipMemcpyAsync(&h, ctl, sizeof(SorCtl), hipMemcpyDeviceToHost, st));
            LBMIBM_HIP(ctx, hipStreamSynchronize(st));
            if (h.done) break;
            if (issued > ctx->p.sor_itrmax + 1) {  // the device test stops by ITRMAX
                ctx->err = "front tracking: SOR status did not settle";
                return LBMIBM_ERR_STATE;
            }
            batch = std::min(batch * 2, 256);
        }
        cnt++;
        if (cnt > FT_MAX_ATTEMPTS) {  // fronttrack.cu:718: error(5) on the eleventh start
            if (attempts) *attempts = cnt;
            ctx->err = "front tracking: the red-black SOR did not converge from 10 starts (error 5)";
            return LBMIBM_ERR_VOF;
        }
        if (h.done == 1) break;
    }
    hipLaunchKernelGGL(k_ft_extract, dim3(blocks_of(ctx->N)), dim3(VWG), 0, st, g, (const double *)ctx->ft_x,
                       ctx->vof[ctx->vof_cur], ctx->nx, ctx->ny, ctx->N);
    LBMIBM_LAUNCHED(ctx, "k_ft_extract");
    if (iterations) *iterations = h.conv_iter;
    if (attempts) *attempts = cnt;
    return LBMIBM_OK;
}

lbmibm_status vof_thinc(lbmibm_ctx *ctx) {
    lbmibm_status s = vof_alloc(ctx);
    if (s != LBMIBM_OK) return s;
    VofDims d{ctx->nx, ctx->ny, ctx->nz, ctx->N, ctx->p.m, ctx->p.dt, ctx->p.dx,
              ctx->p.tube_radius > 0.0 ? (const unsigned *)ctx->tube_col : nullptr};
    const unsigned nb = blocks_of(ctx->N);
    const size_t N = (size_t)ctx->N;
    auto sweep = [&](auto kflux, auto kupd, auto u) {
        const double *vin = ctx->vof[ctx->vof_cur];
        hipLaunchKernelGGL(kflux, dim3(nb), dim3(VWG), 0, ctx->stream, d, vin, ctx->ft_x, u);
        hipLaunchKernelGGL(kupd, dim3(nb), dim3(VWG), 0, ctx->stream, d, vin, (const double *)ctx->ft_x,
                           ctx->vof[ctx->vof_cur ^ 1]);
        ctx->vof_cur ^= 1;
    };
    // the flux of a sweep lives in the front-tracking iterate's array
    // (np >= N doubles), which is free between front-tracking calls
    if (ctx->prec == LBMIBM_F64) {
        const double *U = (const double *)ctx->U;
        sweep(k_thinc_flux<double, 0>, k_thinc_update<0>, U);
        sweep(k_thinc_flux<double, 1>, k_thinc_update<1>, U + N);
        sweep(k_thinc_flux<double, 2>, k_thinc_update<2>, U + 2 * N);
    } else {
        const float *U = (const float *)ctx->U;
        sweep(k_thinc_flux<float, 0>, k_thinc_update<0>, U);
        sweep(k_thinc_flux<float, 1>, k_thinc_update<1>, U + N);
        sweep(k_thinc_flux<float, 2>, k_thinc_update<2>, U + 2 * N);
    }
    LBMIBM_LAUNCHED(ctx, "k_thinc");
    return LBMIBM_OK;
}

}  // namespace lbmibm
