// Per-cell volume / area and volume correction on gfx950
// (volume_correction.cu:28-524).
//
// The reference iterates all cells together from the host: per pass it
// launches the element-normal kernel, performs 2 thrust reductions per cell
// and copies lambda to the device, up to 1000 passes.  A cell whose volume
// is within tolerance gets lambda = 0 and is never moved again, so each cell
// evolves independently.  Here a correction is four launches:
//   - k_vol_sums: the first pass's volume and area sums over chunks of
//     VOL_CH elements spread over the whole chip, with the positions unwrapped
//     around the previous centroid on the fly (CapsulePeriodicBefore);
//   - k_volume: one work group per cell; a cell whose first sums are within
//     tolerance (most of them, most steps) is done, the others unwrap their
//     positions and run their own passes to convergence;
//   - k_vol_center / k_vol_fin: the centroid sums over chunks of VOL_CH
//     vertices with the positions wrapped back into the box
//     (CapsulePeriodicAfter), then the per-cell centroid.
// Element terms are FP64 and every sum runs in a fixed order.  The
// reference's return value iter-1 is the largest per-cell pass count minus
// one, and it fails when that count reaches 1000.
#include "common.hpp"

namespace lbmibm {
namespace {

struct VolArgs {
    double lx, ly, lz;
    int py;
};

// The unit normal of element (a, b, c) and its volume and area terms
// (CalcEleNormal_GPU, volume_correction.cu:138-200).
__device__ __forceinline__ void elem_terms(const double a[3], const double b[3], const double c[3], double n[3],
                                           double &dv, double &ds) {
    const double v1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double v2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double cr[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                          v1[0] * v2[1] - v1[1] * v2[0]};
    const double nrm = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    n[0] = cr[0] / nrm;
    n[1] = cr[1] / nrm;
    n[2] = cr[2] / nrm;
    ds = nrm / 2.0;
    const double r0 = (a[0] + b[0] + c[0]) / 3.0, r1 = (a[1] + b[1] + c[1]) / 3.0, r2 = (a[2] + b[2] + c[2]) / 3.0;
    dv = 1.0 / 3.0 * (r0 * n[0] + r1 * n[1] + r2 * n[2]) * ds;
}

// The volume and area terms of element (a, b, c) without its unit normal:
// (r . n) ds / 3 = (r . cr) / 6 with cr = (b - a) x (c - a), |cr| = 2 ds; the
// same terms as elem_terms up to rounding.
__device__ __forceinline__ void elem_sums(const double a[3], const double b[3], const double c[3], double &dv,
                                          double &ds) {
    const double v1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double v2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    const double cr[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                          v1[0] * v2[1] - v1[1] * v2[0]};
    ds = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) / 2.0;
    dv = ((a[0] + b[0] + c[0]) * cr[0] + (a[1] + b[1] + c[1]) * cr[1] + (a[2] + b[2] + c[2]) * cr[2]) / 18.0;
}

// CapsulePeriodicBefore (:28-67) of one coordinate around the centroid g.
__device__ __forceinline__ double unwrap1(double x, double g, double l) {
    const double d = x - g;
    return d > l * 0.5 ? x - l : (d < -l * 0.5 ? x + l : x);
}

// CapsulePeriodicAfter (:70-94) of one coordinate.
__device__ __forceinline__ double rewrap1(double x, double l) {
    return x < -l / 2.0 ? x + l : (x > l / 2.0 ? x - l : x);
}

__device__ void cell_sums(const double *__restrict__ xs, int nv, const int *__restrict__ ele, int e0, int e1,
                          double *__restrict__ normal, WgReduce &red, double &V, double &A) {
    // CalcEleNormal_GPU + CalcSum_Thrust
    double pv = 0.0, pa = 0.0;
    for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        double p[3][3];
        for (int k = 0; k < 3; k++) {
            const int v = ele[3 * e + k];
            p[k][0] = xs[v];
            p[k][1] = xs[nv + v];
            p[k][2] = xs[2 * (size_t)nv + v];
        }
        double n[3], dv, ds;
        elem_terms(p[0], p[1], p[2], n, dv, ds);
        if (normal) {
            normal[3 * (size_t)e] = n[0];
            normal[3 * (size_t)e + 1] = n[1];
            normal[3 * (size_t)e + 2] = n[2];
        }
        pv += dv;
        pa += ds;
    }
    double va[2] = {pv, pa};
    red.sum<2>(va);
    V = va[0];
    A = va[1];
}

// CalcCenter (volume_correction.cu:310-352) of the positions as they are
// (volume_init).
__device__ void cell_center(const double *__restrict__ xs, int nv, int v0, int v1, WgReduce &red, const VolArgs &a,
                            double g[3]) {
    double s[3] = {0.0, 0.0, 0.0};
    for (int j = v0 + threadIdx.x; j < v1; j += blockDim.x) {
        s[0] += xs[j];
        s[1] += xs[nv + j];
        s[2] += xs[2 * (size_t)nv + j];
    }
    red.sum<3>(s);
    for (int d = 0; d < 3; d++) g[d] = s[d] / (double)(v1 - v0);
    g[0] = rewrap1(g[0], a.lx);
    g[2] = rewrap1(g[2], a.lz);
    if (a.py) g[1] = rewrap1(g[1], a.ly);
}

// The first pass's sums over one chunk of a cell's elements, with each vertex
// unwrapped around the cell's previous centroid as CapsulePeriodicBefore
// would store it.
__global__ __launch_bounds__(256) void k_vol_sums(const int4 *__restrict__ chunks, const double *__restrict__ xs,
                                                  int nv, const int *__restrict__ ele, const double *__restrict__ gx,
                                                  double2 *__restrict__ part, VolArgs a) {
    __shared__ double sh[128];
    WgReduce red{sh, 0};
    const int4 ch = chunks[blockIdx.x];
    const double g[3] = {gx[3 * ch.x], gx[3 * ch.x + 1], gx[3 * ch.x + 2]};
    double pv = 0.0, pa = 0.0;
    for (int e = ch.y + threadIdx.x; e < ch.y + ch.z; e += 256) {
        double p[3][3];
        for (int k = 0; k < 3; k++) {
            const int v = ele[3 * e + k];
            p[k][0] = unwrap1(xs[v], g[0], a.lx);
            p[k][1] = a.py ? unwrap1(xs[nv + v], g[1], a.ly) : xs[nv + v];
            p[k][2] = unwrap1(xs[2 * (size_t)nv + v], g[2], a.lz);
        }
        double dv, ds;
        elem_sums(p[0], p[1], p[2], dv, ds);
        pv += dv;
        pa += ds;
    }
    double va[2] = {pv, pa};
    red.sum<2>(va);
    if (threadIdx.x == 0) part[blockIdx.x] = make_double2(va[0], va[1]);
}

// mode 0: volume_init (:427-450), sums and centroid.  mode 1: the
// correction of the cells whose first sums (the chunk partials) are out of
// tolerance: CapsulePeriodicBefore stored, then passes to convergence
// (:354-424); need[c] records which cells ran them.
template <int NT>
__global__ __launch_bounds__(NT) void k_volume(const int *__restrict__ cells, int mode, double *__restrict__ xs, int nv,
                                               const int *__restrict__ voff, const int *__restrict__ eoff,
                                               const int *__restrict__ ele, const int *__restrict__ nle,
                                               const int *__restrict__ le, double *__restrict__ normal,
                                               double *__restrict__ vol0, double *__restrict__ vol,
                                               double *__restrict__ area, double *__restrict__ gx,
                                               int *__restrict__ passes, const int4 *__restrict__ crange,
                                               const double2 *__restrict__ epart, int *__restrict__ need,
                                               DevStatus *__restrict__ status, VolArgs a) {
    __shared__ double sh[128];
    WgReduce red{sh, 0};
    const int c = cells[blockIdx.x];
    const int v0 = voff[c], v1 = voff[c + 1], e0 = eoff[c], e1 = eoff[c + 1];
    double *X = xs, *Y = xs + nv, *Z = xs + 2 * (size_t)nv;
    double V = 0.0, A = 0.0;
    if (mode == 0) {
        cell_sums(xs, nv, ele, e0, e1, nullptr, red, V, A);
        if (threadIdx.x == 0) { vol0[c] = V; vol[c] = V; area[c] = A; }
        double g[3];
        cell_center(xs, nv, v0, v1, red, a, g);
        if (threadIdx.x == 0) { gx[3 * c] = g[0]; gx[3 * c + 1] = g[1]; gx[3 * c + 2] = g[2]; }
        return;
    }
    const double V0 = vol0[c];
    {
        // the first pass from the chunk sums, in chunk order
        const int4 r = crange[c];
        for (int k = 0; k < r.y; k++) {
            const double2 pr = epart[r.x + k];
            V += pr.x;
            A += pr.y;
        }
        if (fabs(V - V0) / V0 * 100.0 < 1.0e-3) {  // EPS (:15, :388-389)
            if (threadIdx.x == 0) {
                vol[c] = V;
                area[c] = A;
                passes[c] = 1;
                need[c] = 0;
                atomicMax(&status->vol_passes, 1);
            }
            return;
        }
    }
    // CapsulePeriodicBefore around the previous centroid (:28-67)
    const double g[3] = {gx[3 * c], gx[3 * c + 1], gx[3 * c + 2]};
    for (int j = v0 + threadIdx.x; j < v1; j += blockDim.x) {
        X[j] = unwrap1(X[j], g[0], a.lx);
        Z[j] = unwrap1(Z[j], g[2], a.lz);
        if (a.py) Y[j] = unwrap1(Y[j], g[1], a.ly);
    }
    __syncthreads();
    int pass = 1;
    for (; pass <= 1000; pass++) {
        cell_sums(xs, nv, ele, e0, e1, normal, red, V, A);  // its barrier orders normals and positions
        const double dV = fabs(V - V0) / V0 * 100.0;
        if (dV < 1.0e-3) break;
        const double lambda = 18.0 * (V - V0) / A;
        // VolumeCorrection_GPU (:97-135)
        for (int j = v0 + threadIdx.x; j < v1; j += blockDim.x) {
            double nn0 = 0.0, nn1 = 0.0, nn2 = 0.0;
            const int cnt = nle[j];
            for (int k = 0; k < cnt; k++) {
                const size_t e = le[6 * (size_t)j + k];
                nn0 += normal[3 * e] / cnt;
                nn1 += normal[3 * e + 1] / cnt;
                nn2 += normal[3 * e + 2] / cnt;
            }
            X[j] = X[j] + -lambda / 6.0 * nn0 * 0.1;
            Y[j] = Y[j] + -lambda / 6.0 * nn1 * 0.1;
            Z[j] = Z[j] + -lambda / 6.0 * nn2 * 0.1;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        vol[c] = V;
        area[c] = A;
        passes[c] = pass;  // 1001 = not converged within 1000 passes
        need[c] = 1;
        atomicMax(&status->vol_passes, pass);
    }
}

// CalcCenter's sums over one chunk of a cell's vertices, of the positions as
// CapsulePeriodicBefore left them (unwrapped here for the cells that ran no
// pass), and CapsulePeriodicAfter: each position wrapped back into the box.
// With `home` the final positions are also counted into the spreading tiles
// (home tile of each vertex, per-tile counts), as the binning would.
__global__ __launch_bounds__(256) void k_vol_center(const int4 *__restrict__ chunks, double *__restrict__ xs, int nv,
                                                    const double *__restrict__ gx, const int *__restrict__ need,
                                                    double4 *__restrict__ part, VolArgs a, BinGeom bg,
                                                    int *__restrict__ home, int *__restrict__ tile_count) {
    __shared__ double sh[128];
    WgReduce red{sh, 0};
    const int4 ch = chunks[blockIdx.x];
    const bool unwrap = need[ch.x] == 0;
    const double g[3] = {gx[3 * ch.x], gx[3 * ch.x + 1], gx[3 * ch.x + 2]};
    const double len[3] = {a.lx, a.ly, a.lz};
    double s[3] = {0.0, 0.0, 0.0};
    // a uniform trip count: the tile counters aggregate over whole waves
    for (int base = 0; base < ch.z; base += 256) {
        const int j = ch.y + base + (int)threadIdx.x;
        const bool on = base + (int)threadIdx.x < ch.z;
        double fin[3] = {0.0, 0.0, 0.0};
        if (on) {
            for (int d = 0; d < 3; d++) {
                double *P = xs + d * (size_t)nv + j;
                const double x0 = *P;
                const bool per = d != 1 || a.py;
                const double x = unwrap && per ? unwrap1(x0, g[d], len[d]) : x0;
                s[d] += x;
                const double w = per ? rewrap1(x, len[d]) : x;
                if (w != x0) *P = w;
                fin[d] = w;
            }
        }
        if (home) {
            const int t = on ? home_tile(bg, fin[0], fin[1], fin[2]) : 0;
            if (on) home[j] = t;
            (void)wave_aggregated_add(tile_count, t, on);
        }
    }
    red.sum<3>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = make_double4(s[0], s[1], s[2], 0.0);
}

// The centroid of each cell from its chunk sums, in chunk order, wrapped into
// the box.
__global__ __launch_bounds__(256) void k_vol_fin(const int4 *__restrict__ crange, const int *__restrict__ voff, int nc,
                                                 const double4 *__restrict__ part, double *__restrict__ gx,
                                                 VolArgs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int4 r = crange[c];
    double s[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < r.w; k++) {
        const double4 pr = part[r.z + k];
        s[0] += pr.x;
        s[1] += pr.y;
        s[2] += pr.z;
    }
    const double n = (double)(voff[c + 1] - voff[c]);
    gx[3 * c] = rewrap1(s[0] / n, a.lx);
    gx[3 * c + 1] = a.py ? rewrap1(s[1] / n, a.ly) : s[1] / n;
    gx[3 * c + 2] = rewrap1(s[2] / n, a.lz);
}

lbmibm_status run_impl(lbmibm_ctx *ctx, int mode, hipStream_t stream) {
    VolArgs a{ctx->nx * ctx->p.dx, ctx->ny * ctx->p.dx, ctx->nz * ctx->p.dx, ctx->p.periodic_y};
    if (mode == 1) {
        LBMIBM_KT(ctx, stream, "k_vol_sums",
                  hipLaunchKernelGGL(k_vol_sums, dim3(ctx->vol_ne), dim3(256), 0, stream, ctx->vol_echunks,
                                     (const double *)ctx->xs, ctx->nv, ctx->ele, ctx->gx, ctx->vol_epart, a));
        LBMIBM_LAUNCHED(ctx, "k_vol_sums");
    }
    // cells of up to VOL_SMALL vertices with 256 threads, larger ones with
    // 1024; the two classes touch disjoint cells, so on the main stream the
    // large one runs on the side stream (forked and joined with events)
    // concurrently with the other
    const bool fork = stream == ctx->stream && ctx->vol_n[0] > 0 && ctx->vol_n[1] > 0;
    hipStream_t big = stream;
    if (fork) {
        LBMIBM_HIP(ctx, hipEventRecord(ctx->ev_vfork, ctx->stream));
        LBMIBM_HIP(ctx, hipStreamWaitEvent(ctx->side, ctx->ev_vfork, 0));
        big = ctx->side;
    }
#define LBMIBM_VOLUME(NT, K, S)                                                                                    \
    hipLaunchKernelGGL(k_volume<NT>, dim3(ctx->vol_n[K]), dim3(NT), 0, S, ctx->vol_cells[K], mode,                  \
                       (double *)ctx->xs, ctx->nv, ctx->voff, ctx->eoff, ctx->ele, ctx->nle, ctx->le, ctx->normal,  \
                       ctx->vol0, ctx->vol, ctx->area, ctx->gx, ctx->vol_passes, ctx->vol_crange, ctx->vol_epart,   \
                       ctx->vol_need, ctx->status, a)
    if (ctx->vol_n[1] > 0) {
        LBMIBM_KT(ctx, big, "k_volume<1024>", LBMIBM_VOLUME(1024, 1, big));
        LBMIBM_LAUNCHED(ctx, "k_volume<1024>");
    }
    if (ctx->vol_n[0] > 0) {
        LBMIBM_KT(ctx, stream, "k_volume<256>", LBMIBM_VOLUME(256, 0, stream));
        LBMIBM_LAUNCHED(ctx, "k_volume<256>");
    }
#undef LBMIBM_VOLUME
    if (fork) {
        LBMIBM_HIP(ctx, hipEventRecord(ctx->ev_vjoin, ctx->side));
        LBMIBM_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_vjoin, 0));
    }
    if (mode == 1) {
        // the corrected positions are final until the next interpolation:
        // count them into the spreading tiles here (spread form 1, whose
        // scatter then needs no count of its own)
        const bool count = ctx->spread_form == 1 && ctx->nv > 0 && ctx->tile_count;
        if (count) {
            lbmibm_status s = ibm_count_begin(ctx, stream);
            if (s != LBMIBM_OK) return s;
        }
        const BinGeom bg{ctx->nx, ctx->ny, ctx->nz, ctx->nxp, ctx->nyp, ctx->nzp, ctx->p.periodic_y == 0,
                         ctx->ntx, ctx->nty, ctx->p.dx};
        LBMIBM_KT(ctx, stream, "k_vol_center",
                  hipLaunchKernelGGL(k_vol_center, dim3(ctx->vol_nvc), dim3(256), 0, stream, ctx->vol_vchunks,
                                     (double *)ctx->xs, ctx->nv, ctx->gx, ctx->vol_need, ctx->vol_vpart, a, bg,
                                     count ? ctx->home : nullptr, ctx->tile_count));
        LBMIBM_LAUNCHED(ctx, "k_vol_center");
        if (count) {
            lbmibm_status s = ibm_count_end(ctx, stream);
            if (s != LBMIBM_OK) return s;
            ctx->counts_ready = true;
        }
        hipLaunchKernelGGL(k_vol_fin, dim3((ctx->nc + 255) / 256), dim3(256), 0, stream, ctx->vol_crange, ctx->voff,
                           ctx->nc, ctx->vol_vpart, ctx->gx, a);
        LBMIBM_LAUNCHED(ctx, "k_vol_fin");
    }
    return LBMIBM_OK;
}

}  // namespace

lbmibm_status volume_run(lbmibm_ctx *ctx, int mode) { return volume_run_on(ctx, mode, ctx->stream); }

lbmibm_status volume_run_on(lbmibm_ctx *ctx, int mode, hipStream_t stream) {
    if (ctx->nc == 0) return LBMIBM_OK;
    if (mode == 1) invalidate_bins(ctx);  // the correction moves the vertices
    return run_impl(ctx, mode, stream);
}

}  // namespace lbmibm
