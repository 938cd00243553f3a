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
// The delta factorises, so each vertex evaluates 4 cosines per axis (12)
// instead of 3 per stencil node (192).  Vertex positions, stencil selection
// and weights are FP64 in every precision mode; the lattice fields (u, F)
// and the nodal forces are in the arithmetic type T.
#include "common.hpp"

namespace lbmibm {
namespace {

struct IbmGeom {
    int nx, ny, nz;     // unpadded
    int nxp, nyp, nzp;  // padded
    int walls;          // 1: y beyond [0,ny) is wall
    double dx;
};

__device__ __forceinline__ bool wrap_index(int &i, int n) {
    if (i < 0) i += n;
    else if (i >= n) i -= n;
    return (unsigned)i < (unsigned)n;
}

template <typename T>
__device__ __forceinline__ int stencil(T xtr, T h, int &first, T w[4], bool unit_spacing) {
    // weights 1 + cos(pi |x_node - xtr| / (2h)) of the 4 nodes from floor(xtr/h)-1
    const T s = unit_spacing ? T(1) : h;
    first = (int)floor(xtr / s) - 1;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const T r = fabs((T)(first + k) * s - xtr);
        w[k] = T(1) + cos(T(M_PI) * r / (T(2) * s));
    }
    return first;
}

template <typename T>
__global__ __launch_bounds__(256) void k_spread(const double *__restrict__ xs, const T *__restrict__ cf,
                                                T *__restrict__ F, int nv, IbmGeom g) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nv) return;
    const double dx = g.dx;
    const double X = xs[j], Y = xs[nv + j], Z = xs[2 * (size_t)nv + j];
    const T fx = cf[j], fy = cf[nv + j], fz = cf[2 * nv + j];
    int x1, y1, z1;
    double wx[4], wy[4], wz[4];
    stencil<double>(X + (double)(g.nxp - 1) * dx / 2.0, dx, x1, wx, false);
    stencil<double>(Y + (double)(g.nyp - 1) * dx / 2.0, dx, y1, wy, false);
    stencil<double>(Z + (double)(g.nzp - 1) * dx / 2.0, dx, z1, wz, false);
    const double inv = 1.0 / (64.0 * dx * dx * dx);
    const size_t N = (size_t)g.nx * g.ny * g.nz;
#pragma unroll
    for (int kz = 0; kz < 4; kz++) {
        int iz = z1 + kz - 2;
        if (!wrap_index(iz, g.nz)) continue;
#pragma unroll
        for (int ky = 0; ky < 4; ky++) {
            int iy = y1 + ky - 2;
            if (g.walls) {
                if (iy < 0 || iy >= g.ny) continue;
            } else if (!wrap_index(iy, g.ny)) continue;
            const double wyz = wy[ky] * wz[kz];
#pragma unroll
            for (int kx = 0; kx < 4; kx++) {
                int ix = x1 + kx - 2;
                if (!wrap_index(ix, g.nx)) continue;
                const T dlt = (T)(wx[kx] * wyz * inv);
                const size_t n = (size_t)ix + (size_t)g.nx * ((size_t)iy + (size_t)g.ny * iz);
                unsafeAtomicAdd(F + n, dlt * fx);
                unsafeAtomicAdd(F + N + n, dlt * fy);
                unsafeAtomicAdd(F + 2 * N + n, dlt * fz);
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ void interp(const T *__restrict__ U, size_t N, const IbmGeom &g, double x, double y,
                                       double z, bool rk4, T &ur, T &vr, T &wr) {
    const double dx = g.dx;
    int x1, y1, z1;
    double wx[4], wy[4], wz[4];
    if (rk4) {
        stencil<double>(x + (double)(g.nxp - 1) * dx / 2.0, dx, x1, wx, false);
        stencil<double>(y + (double)(g.nyp - 1) * dx / 2.0, dx, y1, wy, false);
        stencil<double>(z + (double)(g.nzp - 1) * dx / 2.0, dx, z1, wz, false);
    } else {  // RK2 / Euler branches use unit spacing (ibm.cu:524-551, 581-608)
        stencil<double>(x + (double)(g.nxp - 1) / 2.0, dx, x1, wx, true);
        stencil<double>(y + (double)(g.nyp - 1) / 2.0, dx, y1, wy, true);
        stencil<double>(z + (double)(g.nzp - 1) / 2.0, dx, z1, wz, true);
    }
    ur = vr = wr = T(0);
#pragma unroll
    for (int kz = 0; kz < 4; kz++) {
        int iz = z1 + kz - 2;
        if (!wrap_index(iz, g.nz)) continue;
#pragma unroll
        for (int ky = 0; ky < 4; ky++) {
            int iy = y1 + ky - 2;
            if (g.walls) {
                if (iy < 0 || iy >= g.ny) continue;
            } else if (!wrap_index(iy, g.ny)) continue;
#pragma unroll
            for (int kx = 0; kx < 4; kx++) {
                int ix = x1 + kx - 2;
                if (!wrap_index(ix, g.nx)) continue;
                const T dlt = (T)(wx[kx] * wy[ky] * wz[kz] / 64.0);
                const size_t n = (size_t)ix + (size_t)g.nx * ((size_t)iy + (size_t)g.ny * iz);
                ur += dlt * U[n];
                vr += dlt * U[N + n];
                wr += dlt * U[2 * N + n];
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

template <typename T>
__global__ __launch_bounds__(256) void k_interp_advect(const T *__restrict__ U, double *__restrict__ xs,
                                                       T *__restrict__ uvw, double *__restrict__ xd, int nv,
                                                       IbmGeom g, AdvectArgs a) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nv) return;
    const size_t N = (size_t)g.nx * g.ny * g.nz;
    const double xn = xs[j], yn = xs[nv + j], zn = xs[2 * (size_t)nv + j];
    const double M = a.M, dt = a.dt;
    T um, vm, wm;
    if (a.rk == 4) {
        T u1, v1, w1, u2, v2, w2, u3, v3, w3, u4, v4, w4;
        interp<T>(U, N, g, xn, yn, zn, true, u1, v1, w1);
        interp<T>(U, N, g, xn + (double)u1 * M * dt / 2.0, yn + (double)v1 * M * dt / 2.0,
                  zn + (double)w1 * M * dt / 2.0, true, u2, v2, w2);
        interp<T>(U, N, g, xn + (double)u2 * M * dt / 2.0, yn + (double)v2 * M * dt / 2.0,
                  zn + (double)w2 * M * dt / 2.0, true, u3, v3, w3);
        interp<T>(U, N, g, xn + (double)u3 * M * dt, yn + (double)v3 * M * dt, zn + (double)w3 * M * dt, true, u4,
                  v4, w4);
        um = (u1 + T(2) * u2 + T(2) * u3 + u4) / T(6);
        vm = (v1 + T(2) * v2 + T(2) * v3 + v4) / T(6);
        wm = (w1 + T(2) * w2 + T(2) * w3 + w4) / T(6);
    } else if (a.rk == 2) {
        T u1, v1, w1, u2, v2, w2;
        interp<T>(U, N, g, xn, yn, zn, false, u1, v1, w1);
        interp<T>(U, N, g, xn + (double)u1 * M, yn + (double)v1 * M, zn + (double)w1 * M, false, u2, v2, w2);
        um = (u1 + u2) / T(2);
        vm = (v1 + v2) / T(2);
        wm = (w1 + w2) / T(2);
    } else {
        interp<T>(U, N, g, xn, yn, zn, false, um, vm, wm);
    }
    double xr = xn + (double)um * M * dt, yr = yn + (double)vm * M * dt, zr = zn + (double)wm * M * dt;
    const double lx = a.lx, ly = a.ly, lz = a.lz;
    if (xr < -lx / 2.0) xr += lx;
    if (xr > lx / 2.0) xr -= lx;
    if (zr < -lz / 2.0) zr += lz;
    if (zr > lz / 2.0) zr -= lz;
    if (a.periodic_y) {
        if (yr < -ly / 2.0) yr += ly;
        if (yr > ly / 2.0) yr -= ly;
    }
    uvw[j] = um; uvw[nv + j] = vm; uvw[2 * nv + j] = wm;
    xs[j] = xr; xs[nv + j] = yr; xs[2 * (size_t)nv + j] = zr;
    xd[3 * j] = xr; xd[3 * j + 1] = yr; xd[3 * j + 2] = zr;
}

IbmGeom geom_of(const lbmibm_ctx *ctx) {
    IbmGeom g;
    g.nx = ctx->nx; g.ny = ctx->ny; g.nz = ctx->nz;
    g.nxp = ctx->nxp; g.nyp = ctx->nyp; g.nzp = ctx->nzp;
    g.walls = ctx->p.periodic_y == 0;
    g.dx = ctx->p.dx;
    return g;
}

template <typename T> lbmibm_status spread_impl(lbmibm_ctx *ctx) {
    lbmibm_status s = lattice_fill(ctx, ctx->F, 3 * (size_t)ctx->N, 0.0, false);
    if (s != LBMIBM_OK) return s;
    if (ctx->nv == 0) return LBMIBM_OK;
    hipLaunchKernelGGL(k_spread<T>, dim3((ctx->nv + 255) / 256), dim3(256), 0, ctx->stream, (const double *)ctx->xs,
                       (const T *)ctx->cf, (T *)ctx->F, ctx->nv, geom_of(ctx));
    LBMIBM_LAUNCHED(ctx, "k_spread");
    return LBMIBM_OK;
}

template <typename T> lbmibm_status interp_impl(lbmibm_ctx *ctx) {
    if (ctx->nv == 0) return LBMIBM_OK;
    AdvectArgs a;
    a.dt = ctx->p.dt; a.M = ctx->p.m;
    a.lx = ctx->nx * ctx->p.dx; a.ly = ctx->ny * ctx->p.dx; a.lz = ctx->nz * ctx->p.dx;
    a.rk = ctx->p.runge_kutta;
    a.periodic_y = ctx->p.periodic_y;
    hipLaunchKernelGGL(k_interp_advect<T>, dim3((ctx->nv + 255) / 256), dim3(256), 0, ctx->stream,
                       (const T *)ctx->U, (double *)ctx->xs, (T *)ctx->uvw, (double *)ctx->xd, ctx->nv, geom_of(ctx), a);
    LBMIBM_LAUNCHED(ctx, "k_interp_advect");
    return LBMIBM_OK;
}

}  // namespace

lbmibm_status ibm_spread(lbmibm_ctx *ctx) {
    lbmibm_status s = ctx->prec == LBMIBM_F64 ? spread_impl<double>(ctx) : spread_impl<float>(ctx);
    if (s == LBMIBM_OK) ctx->force_zero = (ctx->nv == 0);
    return s;
}

lbmibm_status ibm_interp(lbmibm_ctx *ctx) {
    return ctx->prec == LBMIBM_F64 ? interp_impl<double>(ctx) : interp_impl<float>(ctx);
}

}  // namespace lbmibm
