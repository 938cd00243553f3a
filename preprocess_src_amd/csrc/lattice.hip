// D3Q19 stream-collide for gfx950.
//
// Reference: lbm_CapsuleSHEARFLOW / lbm_CapsuleCHANNELFLOW / lbm_CHANNELFLOW
// (lbm.cu:481-1312) plus lbm_*_Boundary and lbm_AVERAGE (lbm.cu:1316-1460),
// host launcher lbm_d3q19 (lbm.cu:1463-1579).
//
// Layout: structure-of-arrays f[q][N] over the unpadded fluid lattice
// (i = ix + nx*(iy + ny*iz)); the reference's two halo layers and its bc mask
// are replaced by index wrapping (periodic x, z and optionally y) and the
// analytic wall rows iy = -1 and iy = ny (half-way bounce-back at
// y = +-LY/2).  The buffers hold POST-collision populations: one launch
// pulls f_q(x) = f*_q(x - c_q) (or f*_opp(q)(x) across a wall), forms the
// moments, collides and stores f*(x) at x.  Every population is read once
// and written once, both fully coalesced; the reference's post-stream array
// is the pull-gather of this state (lattice_get_fstar + host gather).
//
// Storage: deviations g = f - w_q rho0 (rho0 = 1), in FP64, FP32, or FP16
// as g/w_q; arithmetic in T (double or float).
#include "common.hpp"

namespace lbmibm {
namespace {

__host__ __device__ constexpr int cx(int q) {
    return q == 1 || q == 7 || q == 10 || q == 15 || q == 17 ? 1 : (q == 2 || q == 8 || q == 9 || q == 16 || q == 18 ? -1 : 0);
}
__host__ __device__ constexpr int cy(int q) {
    return q == 3 || q == 7 || q == 9 || q == 11 || q == 14 ? 1 : (q == 4 || q == 8 || q == 10 || q == 12 || q == 13 ? -1 : 0);
}
__host__ __device__ constexpr int cz(int q) {
    return q == 5 || q == 11 || q == 13 || q == 15 || q == 18 ? 1 : (q == 6 || q == 12 || q == 14 || q == 16 || q == 17 ? -1 : 0);
}
__host__ __device__ constexpr int opp(int q) { return q == 0 ? 0 : q + ((q % 2) << 1) - 1; }
template <typename T> __host__ __device__ constexpr T wq(int q) {
    return q == 0 ? T(1) / T(3) : (q <= 6 ? T(1) / T(18) : T(1) / T(36));
}
template <typename T> __host__ __device__ constexpr T inv_wq(int q) { return q == 0 ? T(3) : (q <= 6 ? T(18) : T(36)); }

template <typename S, typename T> struct Pop {
    __device__ static T ld(const S *p, int) { return (T)(*p); }
    __device__ static S st(T g, int) { return (S)g; }
};
template <typename T> struct Pop<__half, T> {
    __device__ static T ld(const __half *p, int q) { return (T)__half2float(*p) * wq<T>(q); }
    __device__ static __half st(T g, int q) { return __float2half((float)(g * inv_wq<T>(q))); }
};

struct LbmDims {
    unsigned N;
    int nx, ny, nz;
    int nblocks;
};

// Geometry of k_lbm: periodic box, slab (half-way bounce-back walls at the
// y ends) or the tube (nodes outside the disc are walls; links that cross
// the disc's edge are interpolated from the node's own populations,
// lbm_cylinder.cu:61-85, 193-235).
// GEOM_SHEAR is the slab with walls sliding along z (lbm.cu:951-970).
enum { GEOM_PERIODIC = 0, GEOM_SLAB = 1, GEOM_TUBE = 2, GEOM_SHEAR = 3 };

// BounceBackLinearInterpolation_DEV (lbm_cylinder.cu:61-85) with f1 the
// half-way bounced value, f2 the value streamed in along the link and f3 the
// node's own value going the other way: 2r f1 + (1 - 2r) f2 for r < 1/2,
// f1/(2r) + (1 - 1/(2r)) f3 for 1/2 <= r <= 1, f1 otherwise, r being the
// wall's fraction of the link from the node.  The table holds s = 2r for
// r < 1/2, s = -1/(2r) for 1/2 <= r <= 1 and s = -2 for a link without an
// interpolation, so that no lane divides: every wave of a tube row holds an
// edge node and runs this for the union of its lanes' cut links.
template <typename T> __device__ __forceinline__ T tube_interp(double s, T f1, T f2, T f3) {
    if (s < -1.0) return f1;
    const T a = (T)fabs(s);
    return a * f1 + (T(1) - a) * (s >= 0.0 ? f2 : f3);
}

// S: population storage, T: the lattice fields (F, H, U, macros), A: the
// arithmetic of the collision.  FP32 storage collides in FP64: the kernel is
// bound by its HBM traffic, which S and T set, and FP64 arithmetic removes
// the drift that FP32 rounding of the collision accumulates over thousands
// of steps (about 1e-8 of the velocity per step on a smooth flow), leaving
// the rounding of the stored deviations, which does not drift.
template <typename S, typename T, typename A, bool HASF, bool MACROS, int GEOM>
__global__ __launch_bounds__(WG) void k_lbm(const S *__restrict__ fin, S *__restrict__ fout, LbmDims d,
                                            const T *__restrict__ F, const T *__restrict__ H, T *__restrict__ U,
                                            T *__restrict__ Ui, T *__restrict__ rho_out, T *__restrict__ rhom, A om0,
                                            A inv_cr, A pex, A pey, A pez, A B, const DevStep *__restrict__ st,
                                            int m, const double *__restrict__ vf, A l1m1, A df,
                                            const unsigned *__restrict__ tube_col,
                                            const double *__restrict__ tube_r, T *__restrict__ mw, A mw_lo,
                                            A mw_hi) {
    constexpr bool WALLS = GEOM == GEOM_SLAB || GEOM == GEOM_SHEAR, MOVING = GEOM == GEOM_SHEAR;
    const int lb = xcd_block(d.nblocks);
    if (lb >= d.nblocks) return;
    const unsigned i = (unsigned)lb * WG + threadIdx.x;
    if (i >= d.N) return;
    const unsigned nx = d.nx, ny = d.ny, nz = d.nz;
    const size_t N = d.N;
    const unsigned ix = i % nx, r = i / nx, iy = r % ny, iz = r / ny;
    const unsigned xm = ix == 0 ? nx - 1 : ix - 1, xp = ix == nx - 1 ? 0 : ix + 1;
    const unsigned zm = iz == 0 ? nz - 1 : iz - 1, zp = iz == nz - 1 ? 0 : iz + 1;
    bool wall_lo = false, wall_hi = false;
    unsigned ym, yp;
    if (WALLS) {
        wall_lo = iy == 0;       // sources at iy-1 are wall nodes
        wall_hi = iy == ny - 1;  // sources at iy+1 are wall nodes
        ym = iy - 1;
        yp = iy + 1;
    } else {
        ym = iy == 0 ? ny - 1 : iy - 1;
        yp = iy == ny - 1 ? 0 : iy + 1;
    }

    // tube: bit 0 of the column word says the column is inside the disc,
    // bit q that the pull source of direction q is a wall node
    unsigned cut = 0;
    if (GEOM == GEOM_TUBE) {
        cut = tube_col[ix + nx * iy];
        if (!(cut & 1u)) return;  // a wall node: never updated
    }

    A g[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const unsigned sx = cx(q) == 1 ? xm : (cx(q) == -1 ? xp : ix);
        const unsigned sy = cy(q) == 1 ? ym : (cy(q) == -1 ? yp : iy);
        const unsigned sz = cz(q) == 1 ? zm : (cz(q) == -1 ? zp : iz);
        const bool bb = (WALLS && ((cy(q) == 1 && wall_lo) || (cy(q) == -1 && wall_hi))) ||
                        (GEOM == GEOM_TUBE && q > 0 && ((cut >> q) & 1u));
        const size_t src = bb ? (size_t)opp(q) * N + i : (size_t)q * N + (sx + nx * (sy + ny * sz));
        g[q] = Pop<S, A>::ld(fin + src, bb ? opp(q) : q);
    }
    // sliding walls: Ladd's term on the links bounced at a wall moving along
    // z, 6 w_q rho Cr (c_q . u_w) with the rho and u_w of the step that
    // collided this node (lbm.cu:951-970), kept per wall-row node in mw
    const size_t wcol = (size_t)ix + (size_t)nx * iz;
    if (MOVING && wall_lo) {
        const A m = (A)mw[wcol];
        g[11] += wq<A>(11) * m;  // c = (0, 1, 1)
        g[14] -= wq<A>(14) * m;  // c = (0, 1, -1)
    }
    if (MOVING && wall_hi) {
        const A m = (A)mw[(size_t)nx * nz + wcol];
        g[12] -= wq<A>(12) * m;  // c = (0, -1, -1)
        g[13] += wq<A>(13) * m;  // c = (0, -1, 1)
    }
    if (GEOM == GEOM_TUBE && tube_r && (cut & ~1u)) {
        // g[q] of a cut link holds the half-way value (f1); f2 is the value
        // pulled along the opposite direction, or the node's own f_q where
        // that source is a wall too; f3 is the node's own f_q.  Only values
        // of uncut directions are read as f2, so the in-place update is safe.
        // all loads first (each under its lane mask), then the arithmetic,
        // so that a wave waits once rather than once per cut direction
        // (the node's own values are held as stored and converted at use)
        const unsigned rt = (ix + nx * iy) * (unsigned)Q;  // 32-bit offset into the column table
        S own[Q];
        double sq[Q];
#pragma unroll
        for (int q = 1; q < Q; q++) {
            own[q] = S(0);
            sq[q] = -2.0;
            if ((cut >> q) & 1u) {
                own[q] = fin[(size_t)q * N + i];
                sq[q] = tube_r[rt + q];
            }
        }
#pragma unroll
        for (int q = 1; q < Q; q++) {
            if (!((cut >> q) & 1u)) continue;
            const A f3 = Pop<S, A>::ld(&own[q], q);
            const A f2 = ((cut >> opp(q)) & 1u) ? f3 : g[opp(q)];
            g[q] = tube_interp<A>(sq[q], g[q], f2, f3);
        }
    }

    A drho = A(0), jx = A(0), jy = A(0), jz = A(0);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        drho += g[q];
        if (cx(q)) jx += cx(q) * g[q];
        if (cy(q)) jy += cy(q) * g[q];
        if (cz(q)) jz += cz(q) * g[q];
    }
    const A rho = A(1) + drho;
    const A amp = (A)st->amp;
    if (MOVING) {
        const A uw = (A)st->wall;
        if (wall_lo) mw[wcol] = (T)(mw_lo * rho * uw);
        if (wall_hi) mw[(size_t)nx * nz + wcol] = (T)(mw_hi * rho * uw);
    }
    A ax = pex * amp, ay = pey * amp, az = pez * amp;
    if (HASF) {
        ax += B * (A)F[i];
        ay += B * (A)F[N + i];
        az += B * (A)F[2 * N + i];
    }
    if (H) {  // per-node body-force field, same convention as P e
        ax += (A)H[i];
        ay += (A)H[N + i];
        az += (A)H[2 * N + i];
    }
    // relaxation rate with the VOF indicator (lbm.cu:1131), uniform without
    const A om = vf ? A(1) / (A(3) * (A(1) + l1m1 * (A)vf[i]) * df + A(0.5)) : om0;
    const A irho = A(1) / rho;
    // U = (j + G/2)/rho with the Guo force G = 2 rho (P e + B F) (lbm.cu:1116-1118)
    const A Ux = jx * irho + ax, Uy = jy * irho + ay, Uz = jz * irho + az;
    const A uu = Ux * Ux + Uy * Uy + Uz * Uz;
    const A Ua = Ux * ax + Uy * ay + Uz * az;
    const A kf = (A(1) - om * A(0.5)) * A(2) * rho;

#pragma unroll
    for (int q = 0; q < Q; q++) {
        const A cu = (cx(q) ? cx(q) * Ux : A(0)) + (cy(q) ? cy(q) * Uy : A(0)) + (cz(q) ? cz(q) * Uz : A(0));
        const A ca = (cx(q) ? cx(q) * ax : A(0)) + (cy(q) ? cy(q) * ay : A(0)) + (cz(q) ? cz(q) * az : A(0));
        const A geq = wq<A>(q) * (drho + rho * (A(3) * cu + A(4.5) * cu * cu - A(1.5) * uu));
        const A src = wq<A>(q) * kf * (A(3) * ca * (A(1) + A(3) * cu) - A(3) * Ua);
        const A gs = g[q] + om * (geq - g[q]) + src;
        fout[(size_t)q * N + i] = Pop<S, A>::st(gs, q);
    }

    const int iter = st->iter;
    const bool reset = ((iter - 1) % m) == 0, last = (iter % m) == 0;
    const T ux = (T)(Ux * inv_cr), uy = (T)(Uy * inv_cr), uz = (T)(Uz * inv_cr);
    if (m == 1) {
        U[i] = ux;
        U[N + i] = uy;
        U[2 * N + i] = uz;
    } else {
        // running sum reset at (iter-1)%M==0 and divided by M at iter%M==0
        // (lbm.cu:1097-1102, 1120-1127, 1455-1457)
        T a0 = reset ? ux : U[i] + ux, a1 = reset ? uy : U[N + i] + uy, a2 = reset ? uz : U[2 * N + i] + uz;
        if (last) { a0 /= (T)m; a1 /= (T)m; a2 /= (T)m; }
        U[i] = a0;
        U[N + i] = a1;
        U[2 * N + i] = a2;
    }
    if (MACROS) {
        rho_out[i] = (T)rho;
        Ui[i] = ux;
        Ui[N + i] = uy;
        Ui[2 * N + i] = uz;
        rhom[i] = reset ? (T)rho : rhom[i] + (T)rho;  // dm is never divided (lbm.cu:1455-1457)
    }
}

// Vector-Jacobian product of one k_lbm step with M = 1, taken at the state
// the step would run from (pre-step populations fin, lattice force F, body
// force field H).  Given the adjoint of the post-collision populations ga_out
// and of the velocity output ua (u = U/Cr), it forms the adjoint of the pulled
// pre-collision populations at every node and writes each to the location
// the forward pulled it from, so ga_in (the adjoint of fin) is written once
// per entry, without atomics.  The adjoint of the acceleration a = P e amp +
// B F + H gives Fa = B abar and is added to Ha.  Derivatives of the BGK
// equilibrium and the Guo source as written in k_lbm (drho, j, U = j/rho + a,
// kf = (2 - om) rho).
template <typename S, typename T, bool WALLS>
__global__ __launch_bounds__(WG) void k_lbm_adj(const S *__restrict__ fin, const T *__restrict__ ga_out,
                                                const T *__restrict__ ua, T *__restrict__ ga_in, LbmDims d,
                                                const T *__restrict__ F, const T *__restrict__ H,
                                                T *__restrict__ Fa, double *__restrict__ Ha, T om, T inv_cr,
                                                T pex, T pey, T pez, T B, const DevStep *__restrict__ st) {
    const int lb = xcd_block(d.nblocks);
    if (lb >= d.nblocks) return;
    const unsigned i = (unsigned)lb * WG + threadIdx.x;
    if (i >= d.N) return;
    const unsigned nx = d.nx, ny = d.ny, nz = d.nz;
    const size_t N = d.N;
    const unsigned ix = i % nx, r = i / nx, iy = r % ny, iz = r / ny;
    const unsigned xm = ix == 0 ? nx - 1 : ix - 1, xp = ix == nx - 1 ? 0 : ix + 1;
    const unsigned zm = iz == 0 ? nz - 1 : iz - 1, zp = iz == nz - 1 ? 0 : iz + 1;
    bool wall_lo = false, wall_hi = false;
    unsigned ym, yp;
    if (WALLS) {
        wall_lo = iy == 0;
        wall_hi = iy == ny - 1;
        ym = iy - 1;
        yp = iy + 1;
    } else {
        ym = iy == 0 ? ny - 1 : iy - 1;
        yp = iy == ny - 1 ? 0 : iy + 1;
    }
    T g[Q];
    size_t src[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const unsigned sx = cx(q) == 1 ? xm : (cx(q) == -1 ? xp : ix);
        const unsigned sy = cy(q) == 1 ? ym : (cy(q) == -1 ? yp : iy);
        const unsigned sz = cz(q) == 1 ? zm : (cz(q) == -1 ? zp : iz);
        const bool bb = WALLS && ((cy(q) == 1 && wall_lo) || (cy(q) == -1 && wall_hi));
        src[q] = bb ? (size_t)opp(q) * N + i : (size_t)q * N + (sx + nx * (sy + ny * sz));
        g[q] = Pop<S, T>::ld(fin + src[q], bb ? opp(q) : q);
    }
    T drho = T(0), jx = T(0), jy = T(0), jz = T(0);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        drho += g[q];
        if (cx(q)) jx += cx(q) * g[q];
        if (cy(q)) jy += cy(q) * g[q];
        if (cz(q)) jz += cz(q) * g[q];
    }
    const T rho = T(1) + drho;
    const T amp = (T)st->amp;
    T ax = pex * amp, ay = pey * amp, az = pez * amp;
    if (F) {
        ax += B * F[i];
        ay += B * F[N + i];
        az += B * F[2 * N + i];
    }
    if (H) {
        ax += H[i];
        ay += H[N + i];
        az += H[2 * N + i];
    }
    const T irho = T(1) / rho;
    const T Ux = jx * irho + ax, Uy = jy * irho + ay, Uz = jz * irho + az;
    const T uu = Ux * Ux + Uy * Uy + Uz * Uz;
    const T Ua = Ux * ax + Uy * ay + Uz * az;
    const T kf = (T(2) - om) * rho;
    // adjoints: of drho, U (from the velocity output first), a
    T drb = T(0);
    T Ubx = ua[i] * inv_cr, Uby = ua[N + i] * inv_cr, Ubz = ua[2 * N + i] * inv_cr;
    T abx = T(0), aby = T(0), abz = T(0);
    T gb[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const T go = ga_out[(size_t)q * N + i];
        gb[q] = (T(1) - om) * go;
        const T cu = (cx(q) ? cx(q) * Ux : T(0)) + (cy(q) ? cy(q) * Uy : T(0)) + (cz(q) ? cz(q) * Uz : T(0));
        const T ca = (cx(q) ? cx(q) * ax : T(0)) + (cy(q) ? cy(q) * ay : T(0)) + (cz(q) ? cz(q) * az : T(0));
        const T w = wq<T>(q);
        const T eb = om * go * w;  // adjoint of geq / w
        const T sb = go * w;       // adjoint of src / w
        const T Sq = T(3) * ca * (T(1) + T(3) * cu) - T(3) * Ua;
        drb += eb * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * uu) + sb * (T(2) - om) * Sq;
        // geq: rho (3 c + 9 cu c - 3 U); src: kf (9 ca c - 3 a) for U, kf (3 c (1 + 3 cu) - 3 U) for a
        const T kc = eb * rho * (T(3) + T(9) * cu) + sb * kf * T(9) * ca;
        const T ka = sb * kf * T(3) * (T(1) + T(3) * cu);
        if (cx(q)) { Ubx += cx(q) * kc; abx += cx(q) * ka; }
        if (cy(q)) { Uby += cy(q) * kc; aby += cy(q) * ka; }
        if (cz(q)) { Ubz += cz(q) * kc; abz += cz(q) * ka; }
        const T ku = eb * rho * T(3), kv = sb * kf * T(3);
        Ubx -= ku * Ux + kv * ax; Uby -= ku * Uy + kv * ay; Ubz -= ku * Uz + kv * az;
        abx -= kv * Ux; aby -= kv * Uy; abz -= kv * Uz;
    }
    // U = j / rho + a
    abx += Ubx; aby += Uby; abz += Ubz;
    const T jbx = Ubx * irho, jby = Uby * irho, jbz = Ubz * irho;
    drb -= (jbx * jx + jby * jy + jbz * jz) * irho;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        T v = gb[q] + drb;
        if (cx(q)) v += cx(q) * jbx;
        if (cy(q)) v += cy(q) * jby;
        if (cz(q)) v += cz(q) * jbz;
        ga_in[src[q]] = v;
    }
    Fa[i] = B * abx;
    Fa[N + i] = B * aby;
    Fa[2 * N + i] = B * abz;
    if (Ha) {
        Ha[i] += (double)abx;
        Ha[N + i] += (double)aby;
        Ha[2 * N + i] += (double)abz;
    }
}

template <typename S, typename T>
__global__ void k_from_fstar(const double *__restrict__ fs, S *__restrict__ f, unsigned N) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
#pragma unroll
    for (int q = 0; q < Q; q++) f[(size_t)q * N + i] = Pop<S, T>::st((T)(fs[(size_t)q * N + i] - wq<double>(q)), q);
}

template <typename S, typename T>
__global__ void k_to_fstar(const S *__restrict__ f, double *__restrict__ fs, unsigned N) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
#pragma unroll
    for (int q = 0; q < Q; q++) fs[(size_t)q * N + i] = (double)Pop<S, T>::ld(f + (size_t)q * N + i, q) + wq<double>(q);
}

template <typename V> __global__ void k_fill(V *__restrict__ p, size_t n, V v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename S, typename T, typename A> lbmibm_status step_impl(lbmibm_ctx *ctx) {
    const lbmibm_params &p = ctx->p;
    LbmDims d;
    d.N = (unsigned)ctx->N;
    d.nx = ctx->nx; d.ny = ctx->ny; d.nz = ctx->nz;
    d.nblocks = (int)((ctx->N + WG - 1) / WG);
    const int grid = xcd_grid(d.nblocks);
    const double om = 1.0 / (3.0 * p.df + 0.5);  // lbm.cu:1131 with vf = 0
    if (p.vof_period > 0 && !ctx->vof[0]) {
        ctx->err = "the VOF indicator is not allocated";
        return LBMIBM_ERR_STATE;
    }
    const S *fin = (const S *)ctx->f[ctx->cur];
    S *fout = (S *)ctx->f[ctx->cur ^ 1];
    const bool hasf = !ctx->force_zero, macros = p.store_macros != 0, walls = p.periodic_y == 0;
    auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(WG), 0, ctx->stream, fin, fout, d, (const T *)ctx->F,
                           (const T *)ctx->Hb, (T *)ctx->U, (T *)ctx->Ui, (T *)ctx->rho, (T *)ctx->rhom, (A)om, (A)(1.0 / p.cr),
                           (A)(p.p * p.ex), (A)(p.p * p.ey), (A)(p.p * p.ez), (A)p.b, ctx->step, p.m,
                           (const double *)(p.vof_period > 0 ? ctx->vof[ctx->vof_cur] : nullptr), (A)(p.lambda1 - 1.0),
                           (A)p.df, (const unsigned *)ctx->tube_col, (const double *)ctx->tube_r, (T *)ctx->mw,
                           (A)(p.wall_motion == 1 ? -6.0 * p.cr : 0.0), (A)(6.0 * p.cr));
    };
    const int geom = p.tube_radius > 0.0 ? GEOM_TUBE : (walls ? (p.wall_motion ? GEOM_SHEAR : GEOM_SLAB) : GEOM_PERIODIC);
#define LBM_CASE(HF, MC, G) \
    if (hasf == HF && macros == MC && geom == G) launch(k_lbm<S, T, A, HF, MC, G>);
#define LBM_GEOMS(HF, MC) LBM_CASE(HF, MC, GEOM_PERIODIC) LBM_CASE(HF, MC, GEOM_SLAB) LBM_CASE(HF, MC, GEOM_TUBE) \
    LBM_CASE(HF, MC, GEOM_SHEAR)
    LBMIBM_KT(ctx, ctx->stream, "k_lbm",
              LBM_GEOMS(true, true) LBM_GEOMS(true, false) LBM_GEOMS(false, true) LBM_GEOMS(false, false));
#undef LBM_GEOMS
#undef LBM_CASE
    LBMIBM_LAUNCHED(ctx, "k_lbm");
    ctx->cur ^= 1;  // pointer swap of lbm.cu:1575
    return LBMIBM_OK;
}

// Adjoint of the lattice step from the current state into ga_in (separate
// from ga_out); the forward state is left unchanged.
template <typename S, typename T> lbmibm_status adj_step_impl(lbmibm_ctx *ctx, const void *ga_out, void *ga_in) {
    const lbmibm_params &p = ctx->p;
    LbmDims d;
    d.N = (unsigned)ctx->N;
    d.nx = ctx->nx; d.ny = ctx->ny; d.nz = ctx->nz;
    d.nblocks = (int)((ctx->N + WG - 1) / WG);
    const int grid = xcd_grid(d.nblocks);
    const double om = 1.0 / (3.0 * p.df + 0.5);
    const S *fin = (const S *)ctx->f[ctx->cur];
    const T *F = ctx->force_zero ? nullptr : (const T *)ctx->F;
    auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid), dim3(WG), 0, ctx->stream, fin, (const T *)ga_out, (const T *)ctx->adj_u,
                           (T *)ga_in, d, F, (const T *)ctx->Hb, (T *)ctx->adj_F, ctx->grad_H, (T)om,
                           (T)(1.0 / p.cr), (T)(p.p * p.ex), (T)(p.p * p.ey), (T)(p.p * p.ez), (T)p.b, ctx->step);
    };
    if (p.periodic_y == 0) launch(k_lbm_adj<S, T, true>);
    else launch(k_lbm_adj<S, T, false>);
    LBMIBM_LAUNCHED(ctx, "k_lbm_adj");
    return LBMIBM_OK;
}

template <typename S, typename T> lbmibm_status from_fstar_impl(lbmibm_ctx *ctx, const double *dev_fs) {
    const unsigned N = (unsigned)ctx->N;
    hipLaunchKernelGGL((k_from_fstar<S, T>), dim3((N + 255) / 256), dim3(256), 0, ctx->stream, dev_fs,
                       (S *)ctx->f[ctx->cur], N);
    LBMIBM_LAUNCHED(ctx, "k_from_fstar");
    return LBMIBM_OK;
}

template <typename S, typename T> lbmibm_status to_fstar_impl(lbmibm_ctx *ctx, double *dev_fs) {
    const unsigned N = (unsigned)ctx->N;
    hipLaunchKernelGGL((k_to_fstar<S, T>), dim3((N + 255) / 256), dim3(256), 0, ctx->stream,
                       (const S *)ctx->f[ctx->cur], dev_fs, N);
    LBMIBM_LAUNCHED(ctx, "k_to_fstar");
    return LBMIBM_OK;
}

}  // namespace

lbmibm_status lattice_step(lbmibm_ctx *ctx) {
    switch (ctx->prec) {
        case LBMIBM_F64: return step_impl<double, double, double>(ctx);
        case LBMIBM_F32:
            return ctx->p.fp32_collide ? step_impl<float, float, float>(ctx) : step_impl<float, float, double>(ctx);
        case LBMIBM_F16: return step_impl<__half, float, float>(ctx);
    }
    return LBMIBM_ERR_ARG;
}

lbmibm_status lattice_adjoint(lbmibm_ctx *ctx, const void *ga_out, void *ga_in) {
    switch (ctx->prec) {
        case LBMIBM_F64: return adj_step_impl<double, double>(ctx, ga_out, ga_in);
        case LBMIBM_F32: return adj_step_impl<float, float>(ctx, ga_out, ga_in);
        default: ctx->err = "adjoint: FP16 population storage has no adjoint"; return LBMIBM_ERR_STATE;
    }
}

lbmibm_status lattice_set_state_from_fstar(lbmibm_ctx *ctx, const double *host_fs) {
    double *dev = nullptr;
    const size_t bytes = sizeof(double) * Q * (size_t)ctx->N;
    LBMIBM_HIP(ctx, hipMalloc(&dev, bytes));
    lbmibm_status s = LBMIBM_OK;
    if (hipMemcpy(dev, host_fs, bytes, hipMemcpyHostToDevice) != hipSuccess) s = LBMIBM_ERR_HIP;
    if (s == LBMIBM_OK) {
        switch (ctx->prec) {
            case LBMIBM_F64: s = from_fstar_impl<double, double>(ctx, dev); break;
            case LBMIBM_F32: s = from_fstar_impl<float, float>(ctx, dev); break;
            case LBMIBM_F16: s = from_fstar_impl<__half, float>(ctx, dev); break;
        }
    }
    // the uploaded state carries its wall terms already: the next pull adds none
    if (s == LBMIBM_OK && ctx->mw) s = lattice_fill(ctx, ctx->mw, 2 * (size_t)ctx->nx * ctx->nz, 0.0, false);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess && s == LBMIBM_OK) s = LBMIBM_ERR_HIP;
    (void)hipFree(dev);
    if (s != LBMIBM_OK && ctx->err.empty()) ctx->err = "population upload failed";
    return s;
}

lbmibm_status lattice_get_fstar(lbmibm_ctx *ctx, double *host_fs) {
    double *dev = nullptr;
    const size_t bytes = sizeof(double) * Q * (size_t)ctx->N;
    LBMIBM_HIP(ctx, hipMalloc(&dev, bytes));
    lbmibm_status s = LBMIBM_OK;
    switch (ctx->prec) {
        case LBMIBM_F64: s = to_fstar_impl<double, double>(ctx, dev); break;
        case LBMIBM_F32: s = to_fstar_impl<float, float>(ctx, dev); break;
        case LBMIBM_F16: s = to_fstar_impl<__half, float>(ctx, dev); break;
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess && s == LBMIBM_OK) s = LBMIBM_ERR_HIP;
    if (s == LBMIBM_OK && hipMemcpy(host_fs, dev, bytes, hipMemcpyDeviceToHost) != hipSuccess) s = LBMIBM_ERR_HIP;
    (void)hipFree(dev);
    if (s != LBMIBM_OK && ctx->err.empty()) ctx->err = "population download failed";
    return s;
}

lbmibm_status lattice_fill(lbmibm_ctx *ctx, void *ptr, size_t count, double value, bool storage_type) {
    if (!ptr || count == 0) return LBMIBM_OK;
    const dim3 grid((unsigned)((count + 255) / 256));
    const size_t sz = storage_type ? ctx->szS : ctx->szT;
    if (sz == 8) hipLaunchKernelGGL(k_fill<double>, grid, dim3(256), 0, ctx->stream, (double *)ptr, count, value);
    else if (sz == 4) hipLaunchKernelGGL(k_fill<float>, grid, dim3(256), 0, ctx->stream, (float *)ptr, count, (float)value);
    else hipLaunchKernelGGL(k_fill<__half>, grid, dim3(256), 0, ctx->stream, (__half *)ptr, count, __float2half((float)value));
    LBMIBM_LAUNCHED(ctx, "k_fill");
    return LBMIBM_OK;
}

}  // namespace lbmibm
