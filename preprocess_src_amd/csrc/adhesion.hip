// Cytoadhesion of cell knobs to wall receptors (adhesion.cu), for the slab
// (CapsuleSHEARFLOW) with the linear spring or WLC law: receptors on the wall
// y = -LY/2 on an nxw x nzw grid (wall_init, adhesion.cu:441-458), knobs
// with springs_per_knob springs, Bell-model attach / detach
// (adhesion.cu:513-615) and the spring force that the spreading adds to the
// membrane force (ibm_st1bb, ibm.cu:133-212).
//
// The reference leaves two things to the hardware; both are fixed here:
//   - random numbers: splitmix64 of (seed, iteration, spring id), one draw
//     per spring per call, mapped to (0, 1] (curand seeded from clock() in
//     the reference);
//   - a receptor wanted by several springs in one call goes to the lowest
//     spring id (vertex * springs_per_knob + spring), the others trying
//     their next candidate, in rounds (an atomicOr race in the reference);
//     the bids are 64-bit keys stamped with the call and round.
// Detachments go first; a spring that detaches does not re-attach in the
// same call (one draw decides either branch in the reference).
#include <algorithm>
#include <climits>
#include <vector>

#include "common.hpp"

namespace lbmibm {
namespace {

constexpr int AWG = 256;

struct AdhDev {
    double rc, k_ad, l0, k_on0, k_off0, chi_beta, kbt, shear_rate, gs;
    int limit, maxn, ns, nxw, nzw;
    double dxw, dzw, lx, ly, lz, dx, dtm;  // dtm = dt M
    unsigned long long seed;
    int nv;
    int wlc;                                      // 0 LINEAR_SPRING, 1 WLC
    double lcont, f_wlc, ra, alpha1, alpha2;
    int ring;                                     // the tube's receptor ring
    double radius;
    const double *ring_xy;                        // [nxw][2]: R cos, R sin of each column (host libm)
    double reach;                                 // angular half-width of the columns within LIMIT dx
};

// Extension measure and force of the Bell rates: LINEAR_SPRING
// (adhesion.cu:562-563) or WLC with LL = Lcont dx (adhesion.cu:564-566).
__device__ __forceinline__ double bell_extension(const AdhDev &a, double r) {
#pragma clang fp contract(off)
    if (a.wlc) {
        const double LL = a.lcont * a.dx;
        return fmax(0.0, 0.25 / ((1.0 - r / LL) * (1.0 - r / LL)) + r / LL - 0.25);
    }
    return fmax(0.0, (r - a.l0 * a.dx) * a.rc);
}
__device__ __forceinline__ double bell_force(const AdhDev &a, double dr) {
    return a.wlc ? fabs(a.f_wlc * dr) : fabs(a.k_ad * dr);
}

__host__ __device__ __forceinline__ double adh_uniform(unsigned long long seed, long long iter, long long sid) {
    // splitmix64 (Steele, Lea, Flood 2014) of the key, mapped to (0, 1]
    unsigned long long z = seed ^ ((unsigned long long)iter * 0x9E3779B97F4A7C15ull) ^
                           ((unsigned long long)sid * 0xD1B54A32D192ED03ull);
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)((z >> 11) + 1) * (1.0 / 9007199254740992.0);
}

// A bid of spring sid in round `round` of the call numbered `call` (a
// device counter that every call advances): later rounds and calls outrank
// every older bid, so claims need no clearing, and within a round the
// lowest spring id is the largest key.
__device__ __forceinline__ unsigned long long bid_key(unsigned long long call, int round, int maxn, int sid) {
    const unsigned long long stamp = call * (unsigned long long)(maxn + 2) + (unsigned)round + 1u;
    return (stamp << 32) | (unsigned long long)(0xFFFFFFFFu - (unsigned)sid);
}

// Advances the call counter. Stamps must stay below 2^32 (the key keeps 32
// bits of stamp): when the next call's largest stamp would not, every claim
// is cleared and the counter restarts at 0, whose stamps outrank a cleared
// claim. With MAXN = 100 that happens once in about 42 million calls.
__global__ void __launch_bounds__(256) k_adh_tick(unsigned long long *__restrict__ calls,
                                                  unsigned long long *__restrict__ claim, long long nrec, int maxn) {
    const unsigned long long next = calls[0] + 1ull;
    const bool wrap = (next + 1ull) * (unsigned long long)(maxn + 2) > 0xFFFFFFFFull;
    if (wrap)
        for (long long i = threadIdx.x; i < nrec; i += blockDim.x) claim[i] = 0ull;
    __syncthreads();
    if (threadIdx.x == 0) calls[0] = wrap ? 0ull : next;
}

// Receptor id = iz * nxw + ix on the slab's wall y = -LY/2 (adhesion.cu:
// 452-457) or on the tube's ring at angle ix * dxw (adhesion.cu:475-490).
__device__ __forceinline__ void receptor(const AdhDev &a, int id, double &x, double &y, double &z) {
    const int iz = id / a.nxw, ix = id - iz * a.nxw;
    if (a.ring) {
        x = a.ring_xy[2 * ix];
        y = a.ring_xy[2 * ix + 1];
    } else {
        x = (ix - a.nxw / 2) * a.dxw;
        y = -0.5 * a.ly;
    }
    z = (iz - a.nzw / 2) * a.dzw;
}

__device__ __forceinline__ double image_shift(double d, double L, double xi, double xj) {
    if (fabs(d) > L / 2.0) {  // adhesion.cu:656-668
        if (xi > xj) d += L;
        else if (xi < xj) d -= L;
    }
    return d;
}

__device__ __forceinline__ double p_on(const AdhDev &a, double r, double &dr) {
#pragma clang fp contract(off)
    dr = bell_extension(a, r);  // adhesion.cu:562-573
    const double f_ad = bell_force(a, dr);
    const double k_on = a.k_on0 / a.shear_rate * exp(f_ad * (a.chi_beta - 0.5 * dr) / a.kbt);
    return 1.0 - exp(-k_on * a.dtm);
}

// AdhesiveSpringForce_GPU (adhesion.cu:617-706) and the spreading force
// cf + f_adh of ibm_st1bb.
template <typename T>
__global__ __launch_bounds__(AWG) void k_adh_force(AdhDev a, const double *__restrict__ xs,
                                                   const T *__restrict__ cf, const unsigned char *__restrict__ knob,
                                                   const int *__restrict__ attach, double *__restrict__ f_adh,
                                                   T *__restrict__ fsum) {
#pragma clang fp contract(off)
    const int v = blockIdx.x * AWG + threadIdx.x;
    if (v >= a.nv) return;
    double f[3] = {0.0, 0.0, 0.0};
    if (knob[v]) {
        const double xi = xs[v], yi = xs[a.nv + v], zi = xs[2 * (size_t)a.nv + v];
        // WLC: the repulsive term keeps the previous spring's value where
        // |dy| >= L0 dx, as the reference's variable does (adhesion.cu:677-690)
        double repulf = 0.0;
        for (int s = 0; s < a.ns; s++) {
            const int id = attach[v * a.ns + s];
            if (id == -1) continue;
            double xj, yj, zj;
            receptor(a, id, xj, yj, zj);
            const double dx = a.ring ? xj - xi : image_shift(xj - xi, a.lx, xi, xj), dy = yj - yi;
            const double dz = image_shift(zj - zi, a.lz, zi, zj);
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (a.wlc) {
                const double LL = a.lcont * a.dx;
                const double dr = 0.25 / ((1.0 - r / LL) * (1.0 - r / LL)) + r / LL - 0.25;
                if (fabs(dy) < a.l0 * a.dx) {
                    const double ex = exp(-a.alpha2 * fabs(dy));
                    repulf = a.alpha1 / (a.gs * a.ra) * a.alpha2 * ex / (1.0 - ex) * dy / fabs(dy);
                }
                f[0] += a.f_wlc / (a.gs * a.ra) * dr * dx / r;
                f[1] += a.f_wlc / (a.gs * a.ra) * dr * dy / r + repulf;
                f[2] += a.f_wlc / (a.gs * a.ra) * dr * dz / r;
                continue;
            }
            const double dr = r - a.l0 * a.dx;
            f[0] += a.k_ad / a.gs * dr * dx / r;
            f[1] += a.k_ad / a.gs * dr * dy / r;
            f[2] += a.k_ad / a.gs * dr * dz / r;
        }
    }
    for (int c = 0; c < 3; c++) {
        f_adh[3 * (size_t)v + c] = f[c];
        fsum[c * (size_t)a.nv + v] = (T)((double)cf[c * (size_t)a.nv + v] + f[c]);
    }
}

// StoreNeighWall_CUBE_GPU (adhesion.cu:109-206): receptors within LIMIT dx
// of a knob near the wall, nearest first, ties in scan order, at most MAXN.
__global__ __launch_bounds__(AWG) void k_adh_search(AdhDev a, const double *__restrict__ xs,
                                                    const unsigned char *__restrict__ knob, int *__restrict__ id_wall,
                                                    double *__restrict__ r_wall) {
#pragma clang fp contract(off)
    const int v = blockIdx.x * AWG + threadIdx.x;
    if (v >= a.nv || !knob[v]) return;  // other vertices keep the lists set up empty
    int *ids = id_wall + (size_t)v * a.maxn;
    double *rs = r_wall + (size_t)v * a.maxn;
    for (int k = 0; k < a.maxn; k++) { ids[k] = -1; rs[k] = 1.0e15; }
    const double xi = xs[v], yi = xs[a.nv + v], zi = xs[2 * (size_t)a.nv + v];
    if (a.ring ? sqrt(xi * xi + yi * yi) < a.radius - a.limit * a.dx  // adhesion.cu:248-249
               : sqrt(yi * yi) < 0.5 * a.ly - a.limit * a.dx)
        return;
    const int iz = (int)((zi - (-a.lz / 2.0)) / a.dzw);
    const int span = a.limit * (int)(a.dx / a.dzw);
    // Only columns within LIMIT dx can pass the range test: an x window on
    // the slab, an angular window around the knob on the ring (two ranges
    // when it wraps past angle 0). Scanning them in increasing column order,
    // as a scan of every column does, keeps the list and its ties.
    int ra0, ra1, rb0 = 0, rb1 = -1;
    if (a.ring) {
        double t = atan2(yi, xi);
        if (t < 0.0) t += 2.0 * M_PI;
        const int j0 = (int)floor((t - a.reach) / a.dxw) - 1, j1 = (int)ceil((t + a.reach) / a.dxw) + 1;
        if (j1 - j0 + 1 >= a.nxw) { ra0 = 0; ra1 = a.nxw - 1; }
        else if (j0 < 0) { ra0 = 0; ra1 = j1; rb0 = j0 + a.nxw; rb1 = a.nxw - 1; }
        else if (j1 >= a.nxw) { ra0 = 0; ra1 = j1 - a.nxw; rb0 = j0; rb1 = a.nxw - 1; }
        else { ra0 = j0; ra1 = j1; }
    } else {
        ra0 = max(0, (int)floor((xi - a.limit * a.dx) / a.dxw) + a.nxw / 2 - 1);
        ra1 = min(a.nxw - 1, (int)ceil((xi + a.limit * a.dx) / a.dxw) + a.nxw / 2 + 1);
    }
    int cnt = 0;
    for (int i = 0; i < 2 * span; i++) {
        int jz = iz + i - (span - 1);
        double zj = (jz - a.nzw / 2) * a.dzw;
        if (jz < 0) { jz += a.nzw; zj += a.lz; }
        else if (jz >= a.nzw) { jz -= a.nzw; zj -= a.lz; }
        for (int k = ra0, kend = ra1 + (rb1 - rb0 + 1); k <= kend; k++) {
            const int jx = k <= ra1 ? k : rb0 + (k - ra1 - 1);
            const int id = jz * a.nxw + jx;
            double xj, yj, zr;
            receptor(a, id, xj, yj, zr);
            const double r = sqrt((xj - xi) * (xj - xi) + (yj - yi) * (yj - yi) + (zj - zi) * (zj - zi));
            if (r > a.limit * a.dx) continue;
            int pos = cnt < a.maxn ? cnt : a.maxn;
            while (pos > 0 && rs[pos - 1] > r) pos--;
            if (pos >= a.maxn) continue;
            for (int k = (cnt < a.maxn ? cnt : a.maxn - 1); k > pos; k--) { ids[k] = ids[k - 1]; rs[k] = rs[k - 1]; }
            ids[pos] = id;
            rs[pos] = r;
            if (cnt < a.maxn) cnt++;
        }
    }
}

// Detachment of attached springs (adhesion.cu:584-612, z image only); the
// free springs become the candidates of the attachment rounds.
__global__ __launch_bounds__(AWG) void k_adh_detach(AdhDev a, const double *__restrict__ xs,
                                                    const unsigned char *__restrict__ knob, int *__restrict__ attach,
                                                    double *__restrict__ r_ij, int *__restrict__ is_attached,
                                                    int *__restrict__ pos, int *__restrict__ active,
                                                    const DevStep *__restrict__ st, int init,
                                                    const int *__restrict__ klist, int nk) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * AWG + threadIdx.x;
    if (t >= nk * a.ns) return;
    const int v = klist[t / a.ns], sid = v * a.ns + t % a.ns;
    pos[sid] = 0;
    active[sid] = 0;
    if (!knob[v]) return;
    const int id = attach[sid];
    if (id == -1) {
        active[sid] = 1;
        return;
    }
    if (init) return;
    double xj, yj, zj;
    receptor(a, id, xj, yj, zj);
    const double xi = xs[v], yi = xs[a.nv + v], zi = xs[2 * (size_t)a.nv + v];
    const double dx = xj - xi, dy = yj - yi, dz = image_shift(zj - zi, a.lz, zi, zj);
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double dr = bell_extension(a, r);
    const double f_ad = bell_force(a, dr);
    const double k_off = a.k_off0 / a.shear_rate * exp(f_ad * a.chi_beta / a.kbt);
    const double p_off = 1.0 - exp(-k_off * a.dtm);
    if (p_off <= adh_uniform(a.seed, st->iter, sid)) return;
    is_attached[id] = 0;
    attach[sid] = -1;
    r_ij[sid] = dr;
}

// One attachment round: each active spring proposes its next acceptable
// free candidate and bids for it with its id.
__global__ __launch_bounds__(AWG) void k_adh_propose(AdhDev a, const int *__restrict__ id_wall,
                                                     const double *__restrict__ r_wall,
                                                     const int *__restrict__ is_attached, int *__restrict__ pos,
                                                     int *__restrict__ active, int *__restrict__ prop,
                                                     unsigned long long *__restrict__ claim, const DevStep *__restrict__ st,
                                                     int init, int *__restrict__ nprop, int round,
                                                     const int *__restrict__ klist, int nk,
                                                     const unsigned long long *__restrict__ calls) {
    // after a round without proposals nothing changes any more
    if (round > 0 && nprop[round - 1] == 0) return;
    const int t = blockIdx.x * AWG + threadIdx.x;
    if (t >= nk * a.ns) return;
    const int v = klist[t / a.ns], sid = v * a.ns + t % a.ns;
    prop[sid] = -1;
    if (!active[sid]) return;
    const double u = init ? 0.0 : adh_uniform(a.seed, st->iter, sid);
    int c = pos[sid];
    for (; c < a.maxn; c++) {
        const int id = id_wall[(size_t)v * a.maxn + c];
        if (id == -1) continue;
        if (!init) {
            double dr;
            if (p_on(a, r_wall[(size_t)v * a.maxn + c], dr) <= u) continue;
        }
        if (is_attached[id]) continue;
        break;
    }
    if (c >= a.maxn) {
        active[sid] = 0;
        return;
    }
    pos[sid] = c;
    const int id = id_wall[(size_t)v * a.maxn + c];
    prop[sid] = id;
    atomicMax(claim + id, bid_key(*calls, round, a.maxn, sid));
    atomicAdd(nprop + round, 1);
}

__global__ __launch_bounds__(AWG) void k_adh_resolve(AdhDev a, const double *__restrict__ r_wall,
                                                     int *__restrict__ attach, double *__restrict__ r_ij,
                                                     int *__restrict__ is_attached, int *__restrict__ pos,
                                                     int *__restrict__ active, const int *__restrict__ prop,
                                                     const unsigned long long *__restrict__ claim, int init,
                                                     const int *__restrict__ nprop, int round,
                                                     const int *__restrict__ klist, int nk,
                                                     const unsigned long long *__restrict__ calls) {
    if (nprop[round] == 0) return;
    const int t = blockIdx.x * AWG + threadIdx.x;
    if (t >= nk * a.ns) return;
    const int sid = klist[t / a.ns] * a.ns + t % a.ns;
    const int id = prop[sid];
    if (id < 0) return;
    if (claim[id] == bid_key(*calls, round, a.maxn, sid)) {
        attach[sid] = id;
        is_attached[id] = 1;
        if (!init) {
            double dr;
            (void)p_on(a, r_wall[(size_t)(sid / a.ns) * a.maxn + pos[sid]], dr);
            r_ij[sid] = dr;
        }
        active[sid] = 0;
    } else {
        pos[sid] = pos[sid] + 1;
    }
}

template <typename V> __global__ void k_afill(V *__restrict__ p, size_t n, V v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

AdhDev adh_dev(const lbmibm_ctx *ctx) {
    const lbmibm_adhesion_params &h = ctx->adh;
    AdhDev a;
    a.rc = h.rc; a.k_ad = h.k_ad; a.l0 = h.l0; a.k_on0 = h.k_on0; a.k_off0 = h.k_off0;
    a.chi_beta = h.chi_beta; a.kbt = h.kbt; a.shear_rate = h.shear_rate; a.gs = h.gs;
    a.limit = h.limit; a.maxn = h.max_neighwall; a.ns = h.springs_per_knob;
    a.nxw = ctx->adh_nxw; a.nzw = ctx->adh_nzw; a.dxw = ctx->adh_dxw; a.dzw = ctx->adh_dzw;
    a.dx = ctx->p.dx;
    a.lx = ctx->p.lx * a.dx; a.ly = ctx->p.ly * a.dx; a.lz = ctx->p.lz * a.dx;
    a.dtm = ctx->p.dt * (double)ctx->p.m;
    a.seed = h.seed;
    a.nv = ctx->nv;
    a.wlc = h.wlc;
    a.lcont = h.lcont; a.f_wlc = h.f_wlc; a.ra = h.ra; a.alpha1 = h.alpha1; a.alpha2 = h.alpha2;
    a.ring = ctx->p.tube_radius > 0.0;
    a.radius = ctx->p.tube_radius;
    a.ring_xy = ctx->adh_ring;
    a.reach = 0.0;
    if (a.ring) {  // chord >= 2 sqrt(R rho) |sin(dtheta / 2)| with rho >= R - LIMIT dx for a searching knob
        const double L = h.limit * a.dx, R = a.radius, rho = std::max(R - L, 1.0e-300);
        const double arg = L / (2.0 * std::sqrt(R * rho));
        a.reach = arg >= 1.0 ? M_PI : 2.0 * std::asin(arg);
    }
    return a;
}

inline unsigned ablocks(long long n) { return n > 0 ? (unsigned)((n + AWG - 1) / AWG) : 1u; }

// Search, detachment and the attachment rounds (init: the nearest free
// receptors with no draw, AttachmentPointAtWall_GPU, adhesion.cu:278-314).
lbmibm_status adh_bell(lbmibm_ctx *ctx, int init) {
    const AdhDev a = adh_dev(ctx);
    hipStream_t st = ctx->stream;
    const long long nsp = (long long)ctx->adh_nk * a.ns;  // springs of the knobs
    const int *kl = ctx->adh_klist;
    const int nk = ctx->adh_nk;
    hipLaunchKernelGGL(k_adh_search, dim3(ablocks(a.nv)), dim3(AWG), 0, st, a, (const double *)ctx->xs,
                       (const unsigned char *)ctx->adh_knob, ctx->adh_idw, ctx->adh_rw);
    hipLaunchKernelGGL(k_adh_detach, dim3(ablocks(nsp)), dim3(AWG), 0, st, a, (const double *)ctx->xs,
                       (const unsigned char *)ctx->adh_knob, ctx->adh_attach, ctx->adh_rij, ctx->adh_isatt,
                       ctx->adh_pos, ctx->adh_active, (const DevStep *)ctx->step, init, kl, nk);
    hipLaunchKernelGGL(k_afill<int>, dim3(1), dim3(64), 0, st, ctx->adh_nprop, (size_t)(a.maxn + 1), 0);
    hipLaunchKernelGGL(k_adh_tick, dim3(1), dim3(256), 0, st, ctx->adh_calls, ctx->adh_claim,
                       (long long)a.nxw * a.nzw, a.maxn);
    for (int r = 0; r <= a.maxn; r++) {  // each round retires or advances every active spring
        hipLaunchKernelGGL(k_adh_propose, dim3(ablocks(nsp)), dim3(AWG), 0, st, a, (const int *)ctx->adh_idw,
                           (const double *)ctx->adh_rw, (const int *)ctx->adh_isatt, ctx->adh_pos, ctx->adh_active,
                           ctx->adh_prop, ctx->adh_claim, (const DevStep *)ctx->step, init, ctx->adh_nprop, r, kl,
                           nk, (const unsigned long long *)ctx->adh_calls);
        hipLaunchKernelGGL(k_adh_resolve, dim3(ablocks(nsp)), dim3(AWG), 0, st, a, (const double *)ctx->adh_rw,
                           ctx->adh_attach, ctx->adh_rij, ctx->adh_isatt, ctx->adh_pos, ctx->adh_active,
                           (const int *)ctx->adh_prop, (const unsigned long long *)ctx->adh_claim, init,
                           (const int *)ctx->adh_nprop, r, kl, nk, (const unsigned long long *)ctx->adh_calls);
    }
    LBMIBM_LAUNCHED(ctx, "k_adh_*");
    return LBMIBM_OK;
}

}  // namespace

// Adhesion() (adhesion.cu:772-790): the spring force of the current
// attachments (and the spreading force cf + f_adh), then the Bell update.
lbmibm_status adhesion_step(lbmibm_ctx *ctx) {
    const AdhDev a = adh_dev(ctx);
    if (ctx->prec == LBMIBM_F64)
        hipLaunchKernelGGL(k_adh_force<double>, dim3(ablocks(a.nv)), dim3(AWG), 0, ctx->stream, a,
                           (const double *)ctx->xs, (const double *)ctx->cf, (const unsigned char *)ctx->adh_knob,
                           (const int *)ctx->adh_attach, ctx->adh_fadh, (double *)ctx->adh_fsum);
    else
        hipLaunchKernelGGL(k_adh_force<float>, dim3(ablocks(a.nv)), dim3(AWG), 0, ctx->stream, a,
                           (const double *)ctx->xs, (const float *)ctx->cf, (const unsigned char *)ctx->adh_knob,
                           (const int *)ctx->adh_attach, ctx->adh_fadh, (float *)ctx->adh_fsum);
    LBMIBM_LAUNCHED(ctx, "k_adh_force");
    return adh_bell(ctx, 0);
}

lbmibm_status adhesion_setup(lbmibm_ctx *ctx, const lbmibm_adhesion_params *h, const unsigned char *is_knob) {
    ctx->adh = *h;
    // wall_init (adhesion.cu:441-458 slab, 467-480 tube ring)
    const double dx = ctx->p.dx, LX = ctx->p.lx * dx, LZ = ctx->p.lz * dx;
    const double dz = std::sqrt(1.0 / h->receptor_density) / h->rc;
    ctx->adh_nzw = (int)(LZ / dz + 1.0e-03);
    if (ctx->adh_nzw < 1) {
        ctx->err = "adhesion: receptor spacing larger than the box";
        return LBMIBM_ERR_ARG;
    }
    ctx->adh_dzw = LZ / (double)ctx->adh_nzw;
    if (ctx->p.tube_radius > 0.0) {
        ctx->adh_dxw = ctx->adh_dzw / ctx->p.tube_radius;  // dtheta
        ctx->adh_nxw = (int)(2.0 * M_PI / ctx->adh_dxw);   // ntheta
    } else {
        ctx->adh_nxw = (int)(LX / dz + 1.0e-03);
        ctx->adh_dxw = ctx->adh_nxw > 0 ? LX / (double)ctx->adh_nxw : 0.0;
    }
    if (ctx->adh_nxw < 1) {
        ctx->err = "adhesion: receptor spacing larger than the box";
        return LBMIBM_ERR_ARG;
    }
    const long long nrec = (long long)ctx->adh_nxw * ctx->adh_nzw, nv = ctx->nv;
    const long long nsp = nv * h->springs_per_knob, nn = nv * h->max_neighwall;
    // knobs: the caller's flags (the reference's ./input/icosa file) or
    // drawn with probability KNOB / (vertices of the cell)
    std::vector<int> voff(ctx->nc + 1);
    LBMIBM_HIP(ctx, hipMemcpy(voff.data(), ctx->voff, voff.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<unsigned char> knob((size_t)nv, 0);
    for (int c = 0; c < ctx->nc; c++)
        for (int v = voff[c]; v < voff[c + 1]; v++) {
            if (is_knob) {
                knob[v] = is_knob[v] ? 1 : 0;
            } else {  // iteration -1 is the knob draw's stream
                knob[v] = adh_uniform(h->seed, -1, v) < h->knob / (double)(voff[c + 1] - voff[c]) ? 1 : 0;
            }
        }
    std::vector<int> klist;
    for (long long v = 0; v < nv; v++)
        if (knob[v]) klist.push_back((int)v);
    ctx->adh_nk = (int)klist.size();
    // A second set-up may change every size (receptor count, MAXN, springs
    // per knob, knobs): the previous buffers, and the step graph recorded with
    // their addresses, are released once the device is done with them.
    lbmibm_status s = LBMIBM_OK;
    if (ctx->adh_knob) {
        LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
        ctx->graph = nullptr;
        ctx->graph_ready = false;
        ctx->adh_on = false;
        void **bufs[] = {&ctx->adh_knob, (void **)&ctx->adh_idw, (void **)&ctx->adh_rw, (void **)&ctx->adh_attach,
                         (void **)&ctx->adh_rij, (void **)&ctx->adh_pos, (void **)&ctx->adh_active,
                         (void **)&ctx->adh_prop, (void **)&ctx->adh_isatt, (void **)&ctx->adh_claim,
                         (void **)&ctx->adh_calls, (void **)&ctx->adh_fadh, (void **)&ctx->adh_nprop,
                         (void **)&ctx->adh_klist, &ctx->adh_fsum, (void **)&ctx->adh_ring};
        for (void **p : bufs) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
    }
    auto al = [&](void **p, size_t bytes) {
        if (s == LBMIBM_OK) s = dev_alloc(ctx, p, bytes);
    };
    al(&ctx->adh_knob, nv);
    al((void **)&ctx->adh_idw, nn * sizeof(int));
    al((void **)&ctx->adh_rw, nn * sizeof(double));
    al((void **)&ctx->adh_attach, nsp * sizeof(int));
    al((void **)&ctx->adh_rij, nsp * sizeof(double));
    al((void **)&ctx->adh_pos, nsp * sizeof(int));
    al((void **)&ctx->adh_active, nsp * sizeof(int));
    al((void **)&ctx->adh_prop, nsp * sizeof(int));
    al((void **)&ctx->adh_isatt, nrec * sizeof(int));
    al((void **)&ctx->adh_claim, nrec * sizeof(unsigned long long));
    al((void **)&ctx->adh_calls, sizeof(unsigned long long));
    al((void **)&ctx->adh_fadh, 3 * nv * sizeof(double));
    al((void **)&ctx->adh_nprop, (size_t)(h->max_neighwall + 1) * sizeof(int));
    al((void **)&ctx->adh_klist, std::max<size_t>(klist.size(), 1) * sizeof(int));
    al(&ctx->adh_fsum, 3 * nv * ctx->szT);
    std::vector<double> ring;
    if (ctx->p.tube_radius > 0.0) {  // the columns' coordinates with the host's libm, as wall_init computes them
        ring.resize(2 * (size_t)ctx->adh_nxw);
        for (int ix = 0; ix < ctx->adh_nxw; ix++) {
            ring[2 * ix] = ctx->p.tube_radius * std::cos(ix * ctx->adh_dxw);
            ring[2 * ix + 1] = ctx->p.tube_radius * std::sin(ix * ctx->adh_dxw);
        }
        al((void **)&ctx->adh_ring, ring.size() * sizeof(double));
    }
    if (s != LBMIBM_OK) return s;
    if (!ring.empty())
        LBMIBM_HIP(ctx, hipMemcpy(ctx->adh_ring, ring.data(), ring.size() * sizeof(double), hipMemcpyHostToDevice));
    LBMIBM_HIP(ctx, hipMemcpy(ctx->adh_knob, knob.data(), nv, hipMemcpyHostToDevice));
    if (!klist.empty())
        LBMIBM_HIP(ctx, hipMemcpy(ctx->adh_klist, klist.data(), klist.size() * sizeof(int), hipMemcpyHostToDevice));
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(k_afill<int>, dim3(ablocks(nsp)), dim3(AWG), 0, st, ctx->adh_attach, (size_t)nsp, -1);
    hipLaunchKernelGGL(k_afill<double>, dim3(ablocks(nsp)), dim3(AWG), 0, st, ctx->adh_rij, (size_t)nsp, -1.0e10);
    hipLaunchKernelGGL(k_afill<int>, dim3(ablocks(nrec)), dim3(AWG), 0, st, ctx->adh_isatt, (size_t)nrec, 0);
    hipLaunchKernelGGL(k_afill<unsigned long long>, dim3(ablocks(nrec)), dim3(AWG), 0, st, ctx->adh_claim,
                       (size_t)nrec, 0ull);
    hipLaunchKernelGGL(k_afill<unsigned long long>, dim3(1), dim3(64), 0, st, ctx->adh_calls, (size_t)1, 0ull);
    hipLaunchKernelGGL(k_afill<double>, dim3(ablocks(3 * nv)), dim3(AWG), 0, st, ctx->adh_fadh, (size_t)(3 * nv), 0.0);
    hipLaunchKernelGGL(k_afill<int>, dim3(ablocks(nn)), dim3(AWG), 0, st, ctx->adh_idw, (size_t)nn, -1);
    hipLaunchKernelGGL(k_afill<double>, dim3(ablocks(nn)), dim3(AWG), 0, st, ctx->adh_rw, (size_t)nn, 1.0e15);
    LBMIBM_LAUNCHED(ctx, "k_afill");
    ctx->adh_on = true;
    s = adh_bell(ctx, 1);
    if (s == LBMIBM_OK) LBMIBM_HIP(ctx, hipStreamSynchronize(st));
    return s;
}

}  // namespace lbmibm
