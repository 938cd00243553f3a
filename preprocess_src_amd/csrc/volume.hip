// Per-cell volume / area and volume correction on gfx950
// (volume_correction.cu:28-524).
//
// The reference iterates all cells together from the host: per pass it
// launches the element-normal kernel, performs 2 thrust reductions per cell
// and copies lambda to the device, up to 1000 passes.  A cell whose volume
// is within tolerance gets lambda = 0 and is never moved again, so each cell
// evolves independently; one work group per cell therefore runs its own
// passes to convergence on the device with FP64 element terms and FP64
// reductions in a fixed order.  The reference's return value iter-1 is the
// largest per-cell pass count minus one, and it fails when that count
// reaches 1000.
#include "common.hpp"

namespace lbmibm {
namespace {

struct VolArgs {
    double lx, ly, lz;
    int py;
};

__device__ void cell_sums(const double *__restrict__ xs, int nv, const int *__restrict__ ele, int e0, int e1,
                          double *__restrict__ normal, WgReduce &red, double &V, double &A) {
    // CalcEleNormal_GPU (volume_correction.cu:138-200) + CalcSum_Thrust
    double pv = 0.0, pa = 0.0;
    for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const int na = ele[3 * e], nb = ele[3 * e + 1], nc = ele[3 * e + 2];
        const double xa = xs[na], ya = xs[nv + na], za = xs[2 * (size_t)nv + na];
        const double xb = xs[nb], yb = xs[nv + nb], zb = xs[2 * (size_t)nv + nb];
        const double xc = xs[nc], yc = xs[nv + nc], zc = xs[2 * (size_t)nv + nc];
        const double v1[3] = {xb - xa, yb - ya, zb - za};
        const double v2[3] = {xc - xa, yc - ya, zc - za};
        const double cr[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                              v1[0] * v2[1] - v1[1] * v2[0]};
        const double nrm = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        const double n0 = cr[0] / nrm, n1 = cr[1] / nrm, n2 = cr[2] / nrm;
        if (normal) {
            normal[3 * (size_t)e] = n0;
            normal[3 * (size_t)e + 1] = n1;
            normal[3 * (size_t)e + 2] = n2;
        }
        const double ds = nrm / 2.0;
        const double r0 = (xa + xb + xc) / 3.0, r1 = (ya + yb + yc) / 3.0, r2 = (za + zb + zc) / 3.0;
        pv += 1.0 / 3.0 * (r0 * n0 + r1 * n1 + r2 * n2) * ds;
        pa += ds;
    }
    double va[2] = {pv, pa};
    red.sum<2>(va);
    V = va[0];
    A = va[1];
}

// CalcCenter (volume_correction.cu:310-352), and with `rewrap` also
// CapsulePeriodicAfter (:70-94) in the same pass: the centroid sums use the
// unwrapped positions, which are then wrapped back into the box.
__device__ void cell_center(double *__restrict__ xs, int nv, int v0, int v1, WgReduce &red, const VolArgs &a,
                            bool rewrap, double g[3]) {
    double *X = xs, *Y = xs + nv, *Z = xs + 2 * (size_t)nv;
    double s[3] = {0.0, 0.0, 0.0};
    for (int j = v0 + threadIdx.x; j < v1; j += blockDim.x) {
        double x = X[j], y = Y[j], z = Z[j];
        s[0] += x; s[1] += y; s[2] += z;
        if (rewrap) {
            if (x < -a.lx / 2.0) X[j] = x + a.lx; else if (x > a.lx / 2.0) X[j] = x - a.lx;
            if (z < -a.lz / 2.0) Z[j] = z + a.lz; else if (z > a.lz / 2.0) Z[j] = z - a.lz;
            if (a.py) { if (y < -a.ly / 2.0) Y[j] = y + a.ly; else if (y > a.ly / 2.0) Y[j] = y - a.ly; }
        }
    }
    red.sum<3>(s);
    for (int d = 0; d < 3; d++) g[d] = s[d] / (double)(v1 - v0);
    if (g[0] < -a.lx / 2.0) g[0] += a.lx; else if (g[0] > a.lx / 2.0) g[0] -= a.lx;
    if (g[2] < -a.lz / 2.0) g[2] += a.lz; else if (g[2] > a.lz / 2.0) g[2] -= a.lz;
    if (a.py) { if (g[1] < -a.ly / 2.0) g[1] += a.ly; else if (g[1] > a.ly / 2.0) g[1] -= a.ly; }
}

template <int NT>
__global__ __launch_bounds__(NT) void k_volume(const int *__restrict__ cells, int mode, double *__restrict__ xs, int nv,
                                               const int *__restrict__ voff,
                                                const int *__restrict__ eoff, const int *__restrict__ ele,
                                                const int *__restrict__ nle, const int *__restrict__ le,
                                                double *__restrict__ normal, double *__restrict__ vol0,
                                                double *__restrict__ vol, double *__restrict__ area,
                                                double *__restrict__ gx, int *__restrict__ passes,
                                                DevStatus *__restrict__ status, VolArgs a) {
    __shared__ double sh[128];
    WgReduce red{sh, 0};
    const int c = cells[blockIdx.x];
    const int v0 = voff[c], v1 = voff[c + 1], e0 = eoff[c], e1 = eoff[c + 1];
    double *X = xs, *Y = xs + nv, *Z = xs + 2 * (size_t)nv;
    double V = 0.0, A = 0.0;
    if (mode == 0) {  // volume_init (:427-450)
        cell_sums(xs, nv, ele, e0, e1, nullptr, red, V, A);
        if (threadIdx.x == 0) { vol0[c] = V; vol[c] = V; area[c] = A; }
    } else {
        // CapsulePeriodicBefore around the previous centroid (:28-67)
        const double gx0 = gx[3 * c], gy0 = gx[3 * c + 1], gz0 = gx[3 * c + 2];
        for (int j = v0 + threadIdx.x; j < v1; j += blockDim.x) {
            const double dxv = (double)X[j] - gx0;
            if (dxv > a.lx * 0.5) X[j] -= a.lx; else if (dxv < -a.lx * 0.5) X[j] += a.lx;
            const double dzv = (double)Z[j] - gz0;
            if (dzv > a.lz * 0.5) Z[j] -= a.lz; else if (dzv < -a.lz * 0.5) Z[j] += a.lz;
            if (a.py) {
                const double dyv = (double)Y[j] - gy0;
                if (dyv > a.ly * 0.5) Y[j] -= a.ly; else if (dyv < -a.ly * 0.5) Y[j] += a.ly;
            }
        }
        __syncthreads();
        const double V0 = vol0[c];
        int pass = 1;
        for (; pass <= 1000; pass++) {
            cell_sums(xs, nv, ele, e0, e1, normal, red, V, A);  // its barrier orders normals and positions
            const double dV = fabs(V - V0) / V0 * 100.0;
            if (dV < 1.0e-3) break;  // EPS (:15, :388-389)
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
            atomicMax(&status->vol_passes, pass);
        }
    }
    double g[3];
    cell_center(xs, nv, v0, v1, red, a, mode == 1, g);
    if (threadIdx.x == 0) { gx[3 * c] = g[0]; gx[3 * c + 1] = g[1]; gx[3 * c + 2] = g[2]; }
}

lbmibm_status run_impl(lbmibm_ctx *ctx, int mode, hipStream_t stream) {
    VolArgs a{ctx->nx * ctx->p.dx, ctx->ny * ctx->p.dx, ctx->nz * ctx->p.dx, ctx->p.periodic_y};
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
    if (ctx->vol_n[1] > 0) {
        LBMIBM_KT(ctx, big, "k_volume<1024>",
                  hipLaunchKernelGGL(k_volume<1024>, dim3(ctx->vol_n[1]), dim3(1024), 0, big, ctx->vol_cells[1], mode,
                                     (double *)ctx->xs, ctx->nv, ctx->voff, ctx->eoff, ctx->ele, ctx->nle, ctx->le,
                                     ctx->normal, ctx->vol0, ctx->vol, ctx->area, ctx->gx, ctx->vol_passes,
                                     ctx->status, a));
        LBMIBM_LAUNCHED(ctx, "k_volume<1024>");
    }
    if (ctx->vol_n[0] > 0) {
        LBMIBM_KT(ctx, stream, "k_volume<256>",
                  hipLaunchKernelGGL(k_volume<256>, dim3(ctx->vol_n[0]), dim3(256), 0, stream, ctx->vol_cells[0], mode,
                                     (double *)ctx->xs, ctx->nv, ctx->voff, ctx->eoff, ctx->ele, ctx->nle, ctx->le,
                                     ctx->normal, ctx->vol0, ctx->vol, ctx->area, ctx->gx, ctx->vol_passes,
                                     ctx->status, a));
        LBMIBM_LAUNCHED(ctx, "k_volume<256>");
    }
    if (fork) {
        LBMIBM_HIP(ctx, hipEventRecord(ctx->ev_vjoin, ctx->side));
        LBMIBM_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_vjoin, 0));
    }
    return LBMIBM_OK;
}

}  // namespace

lbmibm_status volume_run(lbmibm_ctx *ctx, int mode) { return volume_run_on(ctx, mode, ctx->stream); }

lbmibm_status volume_run_on(lbmibm_ctx *ctx, int mode, hipStream_t stream) {
    if (ctx->nc == 0) return LBMIBM_OK;
    if (mode == 1) ctx->bins_valid = false;  // the correction moves the vertices
    return run_impl(ctx, mode, stream);
}

}  // namespace lbmibm
