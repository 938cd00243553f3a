// C ABI of liblbmibm: context set-up (the reference's cdo_init / flu_init /
// fem_init / mem_init, init1.cu:13-1092, init2.cu:258-400), the five stage
// entry points that replace the reference host functions, the main.cu loop
// body, graph-captured multi-step runs, and host<->device exchange in the
// reference's padded / AoS layouts.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"

using namespace lbmibm;

namespace {

constexpr int CXh[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int CYh[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 0, 0, 0, 0};
constexpr int CZh[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
constexpr double WDh[Q] = {3, 18, 18, 18, 18, 18, 18, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36};
inline int opph(int q) { return q == 0 ? 0 : q + ((q % 2) << 1) - 1; }

struct StepParams {
    int force_modulated, flowstart;
    double dt, freq, wall_velocity;
};

__global__ void k_set_step(DevStep *s, int iter, StepParams p) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int it = iter >= 0 ? iter : s->iter + 1;
    double amp = 1.0;
    if (p.force_modulated == 1) amp = it > p.flowstart ? 1.0 + 0.5 * sin(M_PI * (p.dt * (double)it) / p.freq) : 0.0;
    else if (p.force_modulated == 2) amp = it > p.flowstart ? 1.0 : 0.0;  // lbm_cylinder.cu:292-293
    s->iter = it;
    s->amp = amp;
    s->wall = it < p.flowstart ? 0.0 : p.wall_velocity;  // sra (lbm.cu:755)
}

__global__ void k_reset_status(DevStatus *s) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        s->bicg_flag = 0;
        s->vol_passes = 0;
    }
}

template <typename T> __global__ void k_convert_from_double(const double *__restrict__ a, T *__restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (T)a[i];
}
template <typename T> __global__ void k_convert_to_double(const T *__restrict__ a, double *__restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

}  // namespace

namespace lbmibm {

lbmibm_status dev_alloc(lbmibm_ctx *ctx, void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return LBMIBM_OK;
    LBMIBM_HIP(ctx, hipMalloc(p, bytes));
    return LBMIBM_OK;
}

// Host double array -> device buffer of the arithmetic type.
lbmibm_status put_T(lbmibm_ctx *ctx, void *dst, const double *src, size_t n) {
    if (n == 0) return LBMIBM_OK;
    if (ctx->szT == 8) {
        LBMIBM_HIP(ctx, hipMemcpy(dst, src, n * 8, hipMemcpyHostToDevice));
        return LBMIBM_OK;
    }
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; i++) tmp[i] = (float)src[i];
    LBMIBM_HIP(ctx, hipMemcpy(dst, tmp.data(), n * 4, hipMemcpyHostToDevice));
    return LBMIBM_OK;
}

lbmibm_status get_T(lbmibm_ctx *ctx, double *dst, const void *src, size_t n) {
    if (n == 0) return LBMIBM_OK;
    LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->szT == 8) {
        LBMIBM_HIP(ctx, hipMemcpy(dst, src, n * 8, hipMemcpyDeviceToHost));
        return LBMIBM_OK;
    }
    std::vector<float> tmp(n);
    LBMIBM_HIP(ctx, hipMemcpy(tmp.data(), src, n * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) dst[i] = (double)tmp[i];
    return LBMIBM_OK;
}

// The tube mask of cdo_init (init1.cu:217-229) at fluid-frame column (ix,
// iy), which may lie in the halo: inside where sqrt(x^2 + y^2) <= R.
bool tube_inside(const lbmibm_ctx *ctx, int ix, int iy) {
    const double dx = ctx->p.dx;
    const double x = -ctx->p.lx * dx / 2.0 - 1.5 * dx + (double)(ix + 2) * dx;
    const double y = -ctx->p.ly * dx / 2.0 - 1.5 * dx + (double)(iy + 2) * dx;
    return std::sqrt(x * x + y * y) <= ctx->p.tube_radius;
}

// SetDistanceFromWall_GPU (lbm_cylinder.cu:147-186) for the link from
// column (ix, iy) along direction q (z taken at 0; the fraction does not
// depend on it): the wall's fraction of the link, or -1 where the reference
// records no crossing.
double tube_fraction(const lbmibm_ctx *ctx, int ix, int iy, int q) {
    const double dxl = ctx->p.dx, R = ctx->p.tube_radius;
    const double x_i = -ctx->p.lx * dxl / 2.0 - 1.5 * dxl + (double)(ix + 2) * dxl;
    const double y_i = -ctx->p.ly * dxl / 2.0 - 1.5 * dxl + (double)(iy + 2) * dxl;
    const double x_j = -ctx->p.lx * dxl / 2.0 - 1.5 * dxl + (double)(ix + 2 + CXh[q]) * dxl;
    const double y_j = -ctx->p.ly * dxl / 2.0 - 1.5 * dxl + (double)(iy + 2 + CYh[q]) * dxl;
    const double dx = x_j - x_i, dy = y_j - y_i, dz = (double)CZh[q] * dxl;
    const double r0 = std::sqrt(dx * dx + dy * dy + dz * dz);
    const double Det = (x_i * dx + y_i * dy) * (x_i * dx + y_i * dy) - (dx * dx + dy * dy) * (x_i * x_i + y_i * y_i - R * R);
    if (Det < 0.0) return -1.0;
    const double t1 = (-x_i * dx - y_i * dy + std::sqrt(Det)) / (dx * dx + dy * dy);
    const double t2 = (-x_i * dx - y_i * dy - std::sqrt(Det)) / (dx * dx + dy * dy);
    double t;
    if (t1 > 0.0 && t1 < 1.0) t = t1;
    else if (t2 > 0.0 && t2 < 1.0) t = t2;
    else return -1.0;
    const double ex = t * dx, ey = t * dy, ez = t * dz;
    return std::sqrt(ex * ex + ey * ey + ez * ez) / r0;
}

long long padded_index(const lbmibm_ctx *ctx, int ix, int iy, int iz) {
    return (long long)(ix + 2) + (long long)ctx->nxp * ((iy + 2) + (long long)ctx->nyp * (iz + 2));
}

// Source node of direction q at fluid node (ix,iy,iz) for the pull, or -1
// when it lies beyond a wall.
long long pull_source(const lbmibm_ctx *ctx, int ix, int iy, int iz, int q) {
    int sx = ix - CXh[q], sy = iy - CYh[q], sz = iz - CZh[q];
    if (ctx->p.tube_radius > 0.0) {
        if (!tube_inside(ctx, sx, sy)) return -1;
        return (long long)sx + (long long)ctx->nx * (sy + (long long)ctx->ny * ((sz + ctx->nz) % ctx->nz));
    }
    sx = (sx + ctx->nx) % ctx->nx;
    sz = (sz + ctx->nz) % ctx->nz;
    if (ctx->p.periodic_y) sy = (sy + ctx->ny) % ctx->ny;
    else if (sy < 0 || sy >= ctx->ny) return -1;
    return (long long)sx + (long long)ctx->nx * (sy + (long long)ctx->ny * sz);
}

long long push_target(const lbmibm_ctx *ctx, int ix, int iy, int iz, int q) {
    int tx = ix + CXh[q], ty = iy + CYh[q], tz = iz + CZh[q];
    if (ctx->p.tube_radius > 0.0) {
        if (!tube_inside(ctx, tx, ty)) return -1;
        return (long long)tx + (long long)ctx->nx * (ty + (long long)ctx->ny * ((tz + ctx->nz) % ctx->nz));
    }
    tx = (tx + ctx->nx) % ctx->nx;
    tz = (tz + ctx->nz) % ctx->nz;
    if (ctx->p.periodic_y) ty = (ty + ctx->ny) % ctx->ny;
    else if (ty < 0 || ty >= ctx->ny) return -1;
    return (long long)tx + (long long)ctx->nx * (ty + (long long)ctx->ny * tz);
}

// Periodic image in the padded frame (lbm.cu:1350-1353, 1410-1411), or -1
// for a wall node.
inline long long padded_image(const lbmibm_ctx *ctx, int px, int py, int pz, bool &is_comp) {
    auto w = [](int a, int n) { return a < 2 ? a + (n - 4) : (a > n - 3 ? a - (n - 4) : a); };
    if (ctx->p.tube_radius > 0.0) {  // only z is periodic; outside the disc is wall
        is_comp = false;
        if (!tube_inside(ctx, px - 2, py - 2)) return -1;
        is_comp = pz >= 2 && pz <= ctx->nzp - 3;
        return (long long)(px - 2) + (long long)ctx->nx * ((py - 2) + (long long)ctx->ny * (w(pz, ctx->nzp) - 2));
    }
    const bool iny = py >= 2 && py <= ctx->nyp - 3;
    is_comp = px >= 2 && px <= ctx->nxp - 3 && iny && pz >= 2 && pz <= ctx->nzp - 3;
    if (!iny && !ctx->p.periodic_y) return -1;
    const int ix = w(px, ctx->nxp) - 2, iy = w(py, ctx->nyp) - 2, iz = w(pz, ctx->nzp) - 2;
    return (long long)ix + (long long)ctx->nx * (iy + (long long)ctx->ny * iz);
}

lbmibm_status sync_check(lbmibm_ctx *ctx) {
    LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return LBMIBM_OK;
}

lbmibm_status read_status(lbmibm_ctx *ctx, DevStatus &s) {
    LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LBMIBM_HIP(ctx, hipMemcpy(&s, ctx->status, sizeof(DevStatus), hipMemcpyDeviceToHost));
    return LBMIBM_OK;
}

lbmibm_status check_flags(lbmibm_ctx *ctx, int *vol_iters) {
    DevStatus s;
    lbmibm_status r = read_status(ctx, s);
    if (r != LBMIBM_OK) return r;
    if (s.bicg_flag == 1) { ctx->err = "BiCGSTAB breakdown: (r, r^) = 0"; return LBMIBM_ERR_BICG_BREAKDOWN; }
    if (s.bicg_flag == 2) { ctx->err = "BiCGSTAB did not converge within ITRMAX"; return LBMIBM_ERR_BICG_NOCONV; }
    if (s.vol_passes >= 1000) { ctx->err = "volume correction failed (1000 passes)"; return LBMIBM_ERR_VOLUME; }
    if (vol_iters) *vol_iters = s.vol_passes > 0 ? s.vol_passes - 1 : 0;
    return LBMIBM_OK;
}

lbmibm_status reset_status(lbmibm_ctx *ctx) {
    hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(64), 0, ctx->stream, ctx->status);
    LBMIBM_LAUNCHED(ctx, "k_reset_status");
    return LBMIBM_OK;
}

lbmibm_status join_volume(lbmibm_ctx *ctx) {
    if (!ctx->vol_pending) return LBMIBM_OK;
    ctx->vol_pending = false;
    LBMIBM_HIP(ctx, hipEventRecord(ctx->ev_pjoin, ctx->vstream));
    LBMIBM_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_pjoin, 0));
    return LBMIBM_OK;
}

// One main.cu loop body on the stream, no host synchronisation.  With
// overlap_volume the volume correction is left running on vstream (joined by
// the next step before it reads the positions, or by join_volume).  The
// membrane force reads fem->xd, which the correction does not change
// (volume_correction.cu:453-524 moves cel->x only), so the two may overlap;
// bending reads cel->x, so with BR != 0 the correction is joined first.
lbmibm_status enqueue_step(lbmibm_ctx *ctx, int iter, bool device_advance, bool overlap_volume) {
    lbmibm_status s = set_step(ctx, device_advance ? -1 : iter, false);
    const int M = ctx->p.m;
    if (s == LBMIBM_OK && (iter - 1) % M == 0 && ctx->nc > 0) {
        if (ctx->p.br != 0.0) s = join_volume(ctx);
        if (s == LBMIBM_OK) s = membrane_forces(ctx);
        if (s == LBMIBM_OK) s = join_volume(ctx);
        if (s == LBMIBM_OK && ctx->adh_on) s = adhesion_step(ctx);  // Adhesion(), main.cu:84-86
        if (s == LBMIBM_OK) s = ibm_spread(ctx);
    }
    if (s == LBMIBM_OK) s = lattice_step(ctx);
    if (s == LBMIBM_OK && iter % M == 0 && ctx->nc > 0) {
        s = join_volume(ctx);
        if (s == LBMIBM_OK) s = ibm_interp(ctx);
        // THINC/WLIC, then front tracking every K steps, before the volume
        // correction (main.cu:93-97); front tracking waits for the host
        if (s == LBMIBM_OK && ctx->p.vof_period > 0) {
            s = vof_thinc(ctx);
            if (s == LBMIBM_OK && iter % ctx->p.vof_period == 0) s = vof_front_tracking(ctx, nullptr, nullptr);
        }
        if (s == LBMIBM_OK && overlap_volume && ctx->p.br == 0.0) {
            // the volume correction overlapping the next membrane force (the
            // binning of the corrected positions stays with the spreading: on
            // this stream it took CUs from the large-cell solve)
            LBMIBM_HIP(ctx, hipEventRecord(ctx->ev_pfork, ctx->stream));
            LBMIBM_HIP(ctx, hipStreamWaitEvent(ctx->vstream, ctx->ev_pfork, 0));
            if (ctx->p.volume_correction) s = volume_run_on(ctx, 1, ctx->vstream);
            ctx->vol_pending = true;
        } else if (s == LBMIBM_OK && ctx->p.volume_correction) {
            s = volume_run(ctx, 1);
        }
    }
    return s;
}

}  // namespace lbmibm

namespace {

void free_all(lbmibm_ctx *ctx) {
    void *ptrs[] = {ctx->f[0], ctx->f[1], ctx->F, ctx->U, ctx->Ui, ctx->rho, ctx->rhom, ctx->voff, ctx->eoff,
                    ctx->vcell, ctx->ele, ctx->ec, ctx->nln, ctx->ln, ctx->lo8, ctx->nle, ctx->le, ctx->xs,
                    ctx->xd, ctx->uvw, ctx->cf, ctx->c0, ctx->s0, ctx->nref, ctx->tt, ctx->rg, ctx->tp, ctx->qv, ctx->qp,
                    ctx->qi, ctx->coef, ctx->bv, ctx->rv, ctx->rh, ctx->pv, ctx->vv, ctx->tv, ctx->gscale,
                    ctx->normal, ctx->vol0, ctx->vol, ctx->area, ctx->gx, ctx->bicg_iters, ctx->vol_passes,
                    ctx->step, ctx->status, ctx->node0, ctx->home, ctx->home_list, ctx->tile_count,
                    ctx->tile_off, ctx->tile_cur, ctx->order_calls, ctx->blk_fmax, ctx->absmax, ctx->nbr, ctx->bicg_cells[0],
                    ctx->bicg_cells[1], ctx->bicg_cells[2], ctx->vol_cells[0], ctx->vol_cells[1], ctx->dsv,
                    ctx->vol_echunks, ctx->vol_vchunks, ctx->vol_crange, ctx->vol_epart, ctx->vol_vpart,
                    ctx->vol_need,
                    ctx->tile_cand, ctx->order_spread, ctx->order_interp, ctx->corner, ctx->cslot, ctx->epair,
                    ctx->asm_chunks, ctx->asm_elems, ctx->asm_list[0], ctx->asm_list[1], ctx->asm_list[2], ctx->eslot,
                    ctx->srec, ctx->onchip_cells, ctx->Hb, ctx->adj_f[0], ctx->adj_f[1], ctx->adj_u, ctx->adj_F,
                    ctx->adj_x, ctx->adj_cf, ctx->grad_H, ctx->grad_gs, ctx->adj_lam, ctx->adj_dsv,
                    ctx->adj_corner, ctx->adj_scratch, ctx->adj_cscale, ctx->adj_gpart, ctx->adj_xmax, ctx->adj_rec, ctx->sbox, ctx->sb_cand,
                    ctx->vof[0], ctx->vof[1], ctx->ft_g, ctx->ft_x, ctx->ft_b, ctx->ft_part, ctx->ft_ctl,
                    ctx->tube_col, ctx->tube_r, ctx->adh_knob, ctx->adh_idw, ctx->adh_attach, ctx->adh_pos,
                    ctx->adh_active, ctx->adh_prop, ctx->adh_isatt, ctx->adh_claim, ctx->adh_rw, ctx->adh_rij,
                    ctx->adh_fadh, ctx->adh_fsum, ctx->adh_nprop, ctx->adh_klist, ctx->adh_calls, ctx->adh_ring, ctx->mw};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (ctx->graph) (void)hipGraphExecDestroy(ctx->graph);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->ev_afork) (void)hipEventDestroy(ctx->ev_afork);
    if (ctx->ev_vfork) (void)hipEventDestroy(ctx->ev_vfork);
    if (ctx->ev_vjoin) (void)hipEventDestroy(ctx->ev_vjoin);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->ev_pfork) (void)hipEventDestroy(ctx->ev_pfork);
    if (ctx->ev_pjoin) (void)hipEventDestroy(ctx->ev_pjoin);
    if (ctx->vstream) (void)hipStreamDestroy(ctx->vstream);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
}

}  // namespace

namespace lbmibm {
lbmibm_status set_step(lbmibm_ctx *ctx, int iter, bool) {
    StepParams sp{ctx->p.force_modulated, ctx->p.flowstart, ctx->p.dt, ctx->p.freq,
                  ctx->p.wall_motion ? ctx->p.wall_velocity : 0.0};
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, ctx->stream, ctx->step, iter, sp);
    LBMIBM_LAUNCHED(ctx, "k_set_step");
    return LBMIBM_OK;
}
}  // namespace lbmibm

extern "C" {

void lbmibm_default_params(lbmibm_params *p, int lx, int ly, int lz) {
    std::memset(p, 0, sizeof(*p));
    p->lx = lx; p->ly = ly; p->lz = lz;
    p->periodic_y = 0;
    p->dx = 1.0; p->dt = 1.0; p->cr = 1.0;
    p->df = 0.1;  // tau = 0.8
    p->lambda1 = 1.0;
    p->p = 0.0; p->ex = 1.0; p->ey = 0.0; p->ez = 0.0;
    p->force_modulated = 0; p->flowstart = 0; p->freq = 1.0;
    p->b = 0.0; p->m = 1;
    p->c = 100.0; p->rgs = 1.0; p->br = 0.0;
    p->itrmax = 1000; p->residual = 1.0e-10; p->runge_kutta = 4; p->neohookean = 0;
    p->volume_correction = 1;
    p->precision = LBMIBM_F32;
    p->store_macros = 0;
    p->fileoutput = 0;
    p->vof_period = 0;
    p->sor_itrmax = 1000;
    p->sor_residual = 1.0e-10;
    p->tube_radius = 0.0;
    p->tube_staircase = 0;
    p->wall_motion = 0;
    p->wall_velocity = 0.0;
    p->bicg_warm_start = 0;
    p->residual_floor = 1.0e-7;
    p->fp32_collide = 0;
}

const char *lbmibm_last_error(const lbmibm_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

const char *lbmibm_autotune_report(const lbmibm_ctx *ctx) { return ctx ? ctx->tune_report.c_str() : ""; }

void lbmibm_sizes(const lbmibm_ctx *ctx, long long *n_padded, int *nv, int *nt, int *nc) {
    if (!ctx) return;
    if (n_padded) *n_padded = ctx->np;
    if (nv) *nv = ctx->nv;
    if (nt) *nt = ctx->nt;
    if (nc) *nc = ctx->nc;
}

lbmibm_status lbmibm_create(const lbmibm_params *p, const lbmibm_cells *cells, lbmibm_ctx **out) {
    if (!p || !out) return LBMIBM_ERR_ARG;
    *out = nullptr;
    if (p->lx < 4 || p->ly < 4 || p->lz < 4 || p->m < 1 || p->dx <= 0.0 || p->cr <= 0.0 ||
        (p->runge_kutta != 1 && p->runge_kutta != 2 && p->runge_kutta != 4) || p->precision < 0 ||
        p->precision > 2 || p->itrmax < 1 || (p->force_modulated && p->freq == 0.0) || p->bicg_warm_start < 0 ||
        p->bicg_warm_start > 2)
        return LBMIBM_ERR_ARG;
    // the viscosity contrast acts through the VOF indicator, which needs cells
    // and the slab's walls (CapsuleSHEARFLOW, fronttrack.cu:87-125)
    if (p->vof_period < 0 || (p->lambda1 != 1.0 && p->vof_period == 0)) return LBMIBM_ERR_ARG;
    // the tube's halo columns must be walls (cdo_init tests only r <= Lc/2)
    if (p->tube_radius < 0.0 || p->force_modulated < 0 || p->force_modulated > 2 ||
        (p->tube_radius > 0.0 && (p->periodic_y ||
                                  p->tube_radius > 0.5 * std::min(p->lx, p->ly) * p->dx - p->dx)))
        return LBMIBM_ERR_ARG;
    // sliding walls belong to the slab (CapsuleSHEARFLOW / SHEARFLOW)
    if (p->wall_motion < 0 || p->wall_motion > 2 || p->tube_staircase < 0 || p->tube_staircase > 1 ||
        (p->wall_motion != 0 && (p->periodic_y || p->tube_radius > 0.0)))
        return LBMIBM_ERR_ARG;
    if (p->vof_period > 0 && (p->periodic_y || !cells || cells->nc < 1 || p->sor_itrmax < 1 ||
                              !(p->sor_residual > 0.0) ||
                              (long long)(p->lx + 4) * (p->ly + 4) * (p->lz + 4) >= (1LL << 31)))
        return LBMIBM_ERR_ARG;  // the VOF kernels index the padded lattice in 32 bits
    const long long N = (long long)p->lx * p->ly * p->lz;
    if (N >= (1LL << 32)) return LBMIBM_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return LBMIBM_ERR_HIP;

    lbmibm_ctx *ctx = new (std::nothrow) lbmibm_ctx();
    if (!ctx) return LBMIBM_ERR_ARG;
    ctx->p = *p;
    ctx->nx = p->lx; ctx->ny = p->ly; ctx->nz = p->lz;
    ctx->N = N;
    ctx->nxp = p->lx + 4; ctx->nyp = p->ly + 4; ctx->nzp = p->lz + 4;
    ctx->np = (long long)ctx->nxp * ctx->nyp * ctx->nzp;
    ctx->prec = p->precision;
    ctx->szS = p->precision == LBMIBM_F64 ? 8 : (p->precision == LBMIBM_F32 ? 4 : 2);
    ctx->szT = p->precision == LBMIBM_F64 ? 8 : 4;
    auto fail = [&](lbmibm_status s) {
        free_all(ctx);
        delete ctx;
        return s;
    };
#define CK(x)                                    \
    do {                                         \
        lbmibm_status s_ = (x);                  \
        if (s_ != LBMIBM_OK) return fail(s_);    \
    } while (0)
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return fail(LBMIBM_ERR_HIP);
    if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess) return fail(LBMIBM_ERR_HIP);
    if (hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess) return fail(LBMIBM_ERR_HIP);
    if (hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess) return fail(LBMIBM_ERR_HIP);
    if (hipEventCreateWithFlags(&ctx->ev_afork, hipEventDisableTiming) != hipSuccess) return fail(LBMIBM_ERR_HIP);
    if (hipEventCreateWithFlags(&ctx->ev_vfork, hipEventDisableTiming) != hipSuccess) return fail(LBMIBM_ERR_HIP);
    if (hipEventCreateWithFlags(&ctx->ev_vjoin, hipEventDisableTiming) != hipSuccess) return fail(LBMIBM_ERR_HIP);
    if (hipStreamCreateWithFlags(&ctx->vstream, hipStreamNonBlocking) != hipSuccess) return fail(LBMIBM_ERR_HIP);
    if (hipEventCreateWithFlags(&ctx->ev_pfork, hipEventDisableTiming) != hipSuccess) return fail(LBMIBM_ERR_HIP);
    if (hipEventCreateWithFlags(&ctx->ev_pjoin, hipEventDisableTiming) != hipSuccess) return fail(LBMIBM_ERR_HIP);

    const size_t szS = ctx->szS, szT = ctx->szT;
    CK(dev_alloc(ctx, &ctx->f[0], (size_t)Q * N * szS));
    CK(dev_alloc(ctx, &ctx->f[1], (size_t)Q * N * szS));
    CK(dev_alloc(ctx, &ctx->F, 3 * (size_t)N * szT));
    CK(dev_alloc(ctx, &ctx->U, 3 * (size_t)N * szT));
    if (p->store_macros) {
        CK(dev_alloc(ctx, &ctx->Ui, 3 * (size_t)N * szT));
        CK(dev_alloc(ctx, &ctx->rho, (size_t)N * szT));
        CK(dev_alloc(ctx, &ctx->rhom, (size_t)N * szT));
    }
    if (p->wall_motion) {
        CK(dev_alloc(ctx, &ctx->mw, 2 * (size_t)p->lx * p->lz * szT));
        CK(lattice_fill(ctx, ctx->mw, 2 * (size_t)p->lx * p->lz, 0.0, false));
    }
    CK(dev_alloc(ctx, (void **)&ctx->step, sizeof(DevStep)));
    CK(dev_alloc(ctx, (void **)&ctx->status, sizeof(DevStatus)));

    if (cells && cells->nc > 0) {
        const int nc = cells->nc;
        if (!cells->voff || !cells->eoff || !cells->x || !cells->y || !cells->z || !cells->ele) return fail(LBMIBM_ERR_ARG);
        const int nv = cells->voff[nc], nt = cells->eoff[nc];
        if (cells->voff[0] != 0 || cells->eoff[0] != 0 || nv <= 0 || nt <= 0) return fail(LBMIBM_ERR_ARG);
        ctx->nc = nc; ctx->nv = nv; ctx->nt = nt;
        std::vector<double> xr(3 * (size_t)nv);
        for (int j = 0; j < nv; j++) { xr[3 * j] = cells->x[j]; xr[3 * j + 1] = cells->y[j]; xr[3 * j + 2] = cells->z[j]; }
        std::vector<int> nln(nv), ln(7 * (size_t)nv), lo(7 * (size_t)nv), nle(nv), le(6 * (size_t)nv), ec(3 * (size_t)nt);
        std::vector<double> nref(4 * (size_t)nt), c0(3 * (size_t)nt), s0(3 * (size_t)nt);
        lbmibm_status ts = lbmibm_build_topology(nc, cells->voff, cells->eoff, xr.data(), cells->ele, nln.data(),
                                                 ln.data(), lo.data(), nle.data(), le.data(), nref.data(),
                                                 ec.data(), c0.data(), s0.data());
        if (ts != LBMIBM_OK) {
            ctx->err = "mesh topology check failed";
            lbmibm_status e = ts;
            free_all(ctx);
            delete ctx;
            return e;
        }
        std::vector<int> lo8(8 * (size_t)nv), vcell(nv);
        for (int j = 0; j < nv; j++)
            for (int k = 0; k < 8; k++) lo8[8 * (size_t)j + k] = k <= nln[j] && k < 7 ? lo[7 * (size_t)j + k] : j;
        for (int c = 0; c < nc; c++) {
            ctx->max_cell_v = std::max(ctx->max_cell_v, cells->voff[c + 1] - cells->voff[c]);
            for (int j = cells->voff[c]; j < cells->voff[c + 1]; j++) vcell[j] = c;
        }
        auto up_i = [&](int **dst, const int *src, size_t n) -> lbmibm_status {
            lbmibm_status s = dev_alloc(ctx, (void **)dst, n * sizeof(int));
            if (s != LBMIBM_OK) return s;
            LBMIBM_HIP(ctx, hipMemcpy(*dst, src, n * sizeof(int), hipMemcpyHostToDevice));
            return LBMIBM_OK;
        };
        CK(up_i(&ctx->voff, cells->voff, nc + 1));
        CK(up_i(&ctx->eoff, cells->eoff, nc + 1));
        CK(up_i(&ctx->vcell, vcell.data(), nv));
        CK(up_i(&ctx->ele, cells->ele, 3 * (size_t)nt));
        CK(up_i(&ctx->ec, ec.data(), 3 * (size_t)nt));
        CK(up_i(&ctx->nln, nln.data(), nv));
        CK(up_i(&ctx->ln, ln.data(), 7 * (size_t)nv));
        CK(up_i(&ctx->lo8, lo8.data(), 8 * (size_t)nv));
        CK(up_i(&ctx->nle, nle.data(), nv));
        CK(up_i(&ctx->le, le.data(), 6 * (size_t)nv));
        auto up_T = [&](void **dst, const double *src, size_t n) -> lbmibm_status {
            lbmibm_status s = dev_alloc(ctx, dst, n * szT);
            if (s != LBMIBM_OK) return s;
            return src ? put_T(ctx, *dst, src, n) : lattice_fill(ctx, *dst, n, 0.0, false);
        };
        std::vector<double> xs(3 * (size_t)nv);
        std::copy(cells->x, cells->x + nv, xs.begin());
        std::copy(cells->y, cells->y + nv, xs.begin() + nv);
        std::copy(cells->z, cells->z + nv, xs.begin() + 2 * (size_t)nv);
        // geometry in FP64 in every precision mode
        auto up_D = [&](void **dst, const double *src, size_t n) -> lbmibm_status {
            lbmibm_status s = dev_alloc(ctx, dst, n * sizeof(double));
            if (s != LBMIBM_OK) return s;
            LBMIBM_HIP(ctx, hipMemcpy(*dst, src, n * sizeof(double), hipMemcpyHostToDevice));
            return LBMIBM_OK;
        };
        CK(up_D(&ctx->xs, xs.data(), 3 * (size_t)nv));
        CK(up_D(&ctx->xd, xr.data(), 3 * (size_t)nv));
        CK(up_D(&ctx->c0, c0.data(), 3 * (size_t)nt));
        CK(up_D(&ctx->s0, s0.data(), 3 * (size_t)nt));
        CK(up_D(&ctx->nref, nref.data(), 4 * (size_t)nt));
        CK(up_T(&ctx->uvw, nullptr, 3 * (size_t)nv));
        CK(up_T(&ctx->cf, nullptr, 3 * (size_t)nv));
        CK(up_T(&ctx->tt, nullptr, 4 * (size_t)nt));
        CK(up_T(&ctx->rg, nullptr, (size_t)nt));
        CK(up_T(&ctx->tp, nullptr, 2 * (size_t)nt));
        for (void **v : {&ctx->qv, &ctx->qi, &ctx->bv, &ctx->rv, &ctx->rh, &ctx->pv, &ctx->vv, &ctx->tv})
            CK(up_T(v, nullptr, 3 * (size_t)nv));
        if (p->bicg_warm_start == 2) CK(up_T(&ctx->qp, nullptr, 3 * (size_t)nv));
        CK(up_T(&ctx->coef, nullptr, 6 * (size_t)nv));
        CK(up_T(&ctx->dsv, nullptr, (size_t)nv));
        // static system structure: cell-local column ids of the 6 off-diagonal
        // entries (lo order without the vertex itself, init2.cu:169-203), the
        // corner of each vertex in its incident elements and, per column, the
        // two incident elements that share the edge
        {
            std::vector<unsigned short> nbr16(6 * (size_t)nv);
            std::vector<int> cslot(3 * (size_t)nt, 0);
            std::vector<unsigned char> epair(6 * (size_t)nv, 0xff), eslot(3 * (size_t)nt, 0);
            for (int j = 0; j < nv; j++) {
                const int n = nln[j], base = cells->voff[vcell[j]];
                for (int sl = 0; sl < n; sl++) {
                    const int e = le[6 * (size_t)j + sl];
                    const int *ev = cells->ele + 3 * (size_t)e;
                    const size_t k = 3 * (size_t)e + (ev[0] == j ? 0 : (ev[1] == j ? 1 : 2));
                    cslot[k] = sl * nv + j;
                    eslot[k] = (unsigned char)sl;
                }
                int m = 0;
                for (int k = 0; k <= n && k < 7; k++) {
                    const int l = lo[7 * (size_t)j + k];
                    if (l == j) continue;
                    int sa = -1, sb = -1;
                    for (int sl = 0; sl < n; sl++) {
                        const int *ev = cells->ele + 3 * (size_t)le[6 * (size_t)j + sl];
                        if (ev[0] == l || ev[1] == l || ev[2] == l) {
                            if (sa < 0) sa = sl;
                            else if (sb < 0) sb = sl;
                        }
                    }
                    if (m >= 6 || sa < 0 || sb < 0) return fail(LBMIBM_ERR_TOPOLOGY);
                    nbr16[(size_t)m * nv + j] = (unsigned short)(l - base);
                    epair[(size_t)m * nv + j] = (unsigned char)(sa | (sb << 4));
                    m++;
                }
                for (; m < 6; m++) nbr16[(size_t)m * nv + j] = (unsigned short)(j - base);
            }
            CK(dev_alloc(ctx, (void **)&ctx->nbr, 6 * (size_t)nv * sizeof(unsigned short)));
            LBMIBM_HIP(ctx, hipMemcpy(ctx->nbr, nbr16.data(), nbr16.size() * sizeof(unsigned short),
                                      hipMemcpyHostToDevice));
            CK(up_i(&ctx->cslot, cslot.data(), cslot.size()));
            CK(dev_alloc(ctx, (void **)&ctx->epair, epair.size()));
            LBMIBM_HIP(ctx, hipMemcpy(ctx->epair, epair.data(), epair.size(), hipMemcpyHostToDevice));
            CK(dev_alloc(ctx, &ctx->corner, 6 * (size_t)nv * 4 * szT));
            CK(dev_alloc(ctx, (void **)&ctx->eslot, eslot.size()));
            LBMIBM_HIP(ctx, hipMemcpy(ctx->eslot, eslot.data(), eslot.size(), hipMemcpyHostToDevice));
        }
        // cells solved on chip and the rest
        std::vector<int> cls[3];
        for (int c = 0; c < nc; c++) {
            const int V = cells->voff[c + 1] - cells->voff[c];
            if (V > 65535) return fail(LBMIBM_ERR_ARG);  // 16-bit local column ids
            const int k = bicg_class(ctx->prec, V);
            cls[k].push_back(c);
            ctx->bicg_maxv[k] = std::max(ctx->bicg_maxv[k], V);
        }
        for (int k = 0; k < 3; k++) {
            ctx->bicg_n[k] = (int)cls[k].size();
            if (!cls[k].empty()) CK(up_i(&ctx->bicg_cells[k], cls[k].data(), cls[k].size()));
        }
        {
            // on-chip classes together (assembled after the large cells)
            std::vector<int> onchip(cls[0]);
            onchip.insert(onchip.end(), cls[1].begin(), cls[1].end());
            ctx->n_onchip = (int)onchip.size();
            if (!onchip.empty()) CK(up_i(&ctx->onchip_cells, onchip.data(), onchip.size()));
            for (int c = 0; c < nc; c++) {
                const int big = bicg_class(ctx->prec, cells->voff[c + 1] - cells->voff[c]) == 2;
                ctx->cls_maxe[big] = std::max(ctx->cls_maxe[big], cells->eoff[c + 1] - cells->eoff[c]);
                ctx->cls_maxv[big] = std::max(ctx->cls_maxv[big], cells->voff[c + 1] - cells->voff[c]);
            }
        }
        {
            // fused assembly: each cell's vertices in chunks of ASM_CH with the
            // elements incident to them (ascending), and the chunk lists of
            // all cells, the large cells (class 2) and the on-chip cells
            std::vector<int4> chunks;
            std::vector<int> elems, lists[3], mark(nt, -1);
            std::vector<char> big(nc, 0);
            for (int c : cls[2]) big[c] = 1;
            for (int c = 0; c < nc; c++)
                for (int v0 = cells->voff[c]; v0 < cells->voff[c + 1]; v0 += ASM_CH) {
                    const int cnv = std::min(ASM_CH, cells->voff[c + 1] - v0), k = (int)chunks.size();
                    const size_t e0 = elems.size();
                    for (int j = v0; j < v0 + cnv; j++)
                        for (int sl = 0; sl < nln[j]; sl++) {
                            const int e = le[6 * (size_t)j + sl];
                            if (mark[e] != k) { mark[e] = k; elems.push_back(e); }
                        }
                    std::sort(elems.begin() + e0, elems.end());
                    chunks.push_back(make_int4(v0, cnv, (int)e0, (int)(elems.size() - e0)));
                    lists[0].push_back(k);
                    lists[big[c] ? 1 : 2].push_back(k);
                }
            if (elems.size() > (size_t)INT32_MAX) return fail(LBMIBM_ERR_ARG);
            CK(dev_alloc(ctx, (void **)&ctx->asm_chunks, chunks.size() * sizeof(int4)));
            LBMIBM_HIP(ctx, hipMemcpy(ctx->asm_chunks, chunks.data(), chunks.size() * sizeof(int4),
                                      hipMemcpyHostToDevice));
            {
                std::vector<int4> er(elems.size());
                for (size_t i = 0; i < elems.size(); i++) {
                    const int e = elems[i];
                    er[i] = make_int4(e, cells->ele[3 * (size_t)e], cells->ele[3 * (size_t)e + 1],
                                      cells->ele[3 * (size_t)e + 2]);
                }
                CK(dev_alloc(ctx, (void **)&ctx->asm_elems, er.size() * sizeof(int4)));
                LBMIBM_HIP(ctx, hipMemcpy(ctx->asm_elems, er.data(), er.size() * sizeof(int4), hipMemcpyHostToDevice));
            }
            for (int k = 0; k < 3; k++) {
                ctx->asm_n[k] = (int)lists[k].size();
                if (!lists[k].empty()) CK(up_i(&ctx->asm_list[k], lists[k].data(), lists[k].size()));
            }
        }
        std::vector<int> vcls[2];
        for (int c = 0; c < nc; c++) vcls[cells->voff[c + 1] - cells->voff[c] > VOL_SMALL].push_back(c);
        for (int k = 0; k < 2; k++) {
            ctx->vol_n[k] = (int)vcls[k].size();
            if (!vcls[k].empty()) CK(up_i(&ctx->vol_cells[k], vcls[k].data(), vcls[k].size()));
        }
        {
            // volume-correction chunks of VOL_CH elements / vertices, each
            // within one cell
            std::vector<int4> ech, vch, range(nc);
            for (int c = 0; c < nc; c++) {
                range[c] = make_int4((int)ech.size(), 0, (int)vch.size(), 0);
                for (int e = cells->eoff[c]; e < cells->eoff[c + 1]; e += VOL_CH)
                    ech.push_back(make_int4(c, e, std::min(VOL_CH, cells->eoff[c + 1] - e), 0));
                for (int v = cells->voff[c]; v < cells->voff[c + 1]; v += VOL_CH)
                    vch.push_back(make_int4(c, v, std::min(VOL_CH, cells->voff[c + 1] - v), 0));
                range[c].y = (int)ech.size() - range[c].x;
                range[c].w = (int)vch.size() - range[c].z;
            }
            ctx->vol_ne = (int)ech.size();
            ctx->vol_nvc = (int)vch.size();
            for (auto *t : {&ech, &vch, &range}) {
                int4 **dst = t == &ech ? &ctx->vol_echunks : t == &vch ? &ctx->vol_vchunks : &ctx->vol_crange;
                CK(dev_alloc(ctx, (void **)dst, t->size() * sizeof(int4)));
                LBMIBM_HIP(ctx, hipMemcpy(*dst, t->data(), t->size() * sizeof(int4), hipMemcpyHostToDevice));
            }
            CK(dev_alloc(ctx, (void **)&ctx->vol_epart, ech.size() * sizeof(double2)));
            CK(dev_alloc(ctx, (void **)&ctx->vol_vpart, vch.size() * sizeof(double4)));
            CK(dev_alloc(ctx, (void **)&ctx->vol_need, nc * sizeof(int)));
        }
        CK(membrane_setup(ctx));
        std::vector<double> ones(nc, 1.0);
        CK(up_T(&ctx->gscale, ones.data(), nc));
        // lattice tiles and vertex bins of the immersed-boundary kernels
        ctx->ntx = (ctx->nx + TILE_X - 1) / TILE_X;
        ctx->nty = (ctx->ny + TILE_Y - 1) / TILE_Y;
        ctx->ntz = (ctx->nz + TILE_Z - 1) / TILE_Z;
        ctx->ntiles = ctx->ntx * ctx->nty * ctx->ntz;
        ctx->nblk_bin = (nv + 255) / 256;
        if (ctx->nx > 30000 || ctx->ny > 29990 || ctx->nz > 30000) return fail(LBMIBM_ERR_ARG);  // short4 bins
        CK(dev_alloc(ctx, (void **)&ctx->node0, (size_t)nv * sizeof(short4)));
        CK(dev_alloc(ctx, &ctx->srec, (size_t)nv * 6 * szT));
        CK(dev_alloc(ctx, (void **)&ctx->home, (size_t)nv * sizeof(int)));
        CK(dev_alloc(ctx, (void **)&ctx->home_list, (size_t)nv * sizeof(int)));
        CK(dev_alloc(ctx, (void **)&ctx->tile_count, (size_t)ctx->ntiles * sizeof(int)));
        CK(dev_alloc(ctx, (void **)&ctx->tile_off, (size_t)ctx->ntiles * sizeof(int)));
        CK(dev_alloc(ctx, (void **)&ctx->tile_cur, (size_t)ctx->ntiles * sizeof(int)));
        CK(dev_alloc(ctx, (void **)&ctx->tile_cand, (size_t)ctx->ntiles * sizeof(int)));
        {
            // tile orders start as the identity (any permutation is valid)
            std::vector<int> ident(ctx->ntiles);
            for (int t = 0; t < ctx->ntiles; t++) ident[t] = t;
            CK(up_i(&ctx->order_spread, ident.data(), ident.size()));
            CK(up_i(&ctx->order_interp, ident.data(), ident.size()));
        }
        {
            const int zero = 0;
            CK(up_i(&ctx->order_calls, &zero, 1));
        }
        CK(dev_alloc(ctx, (void **)&ctx->blk_fmax, (size_t)ctx->nblk_bin * sizeof(float)));
        CK(dev_alloc(ctx, (void **)&ctx->absmax, sizeof(unsigned)));
        CK(ibm_setup(ctx));
        CK(dev_alloc(ctx, (void **)&ctx->normal, 3 * (size_t)nt * sizeof(double)));
        CK(dev_alloc(ctx, (void **)&ctx->vol0, nc * sizeof(double)));
        CK(dev_alloc(ctx, (void **)&ctx->vol, nc * sizeof(double)));
        CK(dev_alloc(ctx, (void **)&ctx->area, nc * sizeof(double)));
        CK(dev_alloc(ctx, (void **)&ctx->gx, 3 * nc * sizeof(double)));
        CK(dev_alloc(ctx, (void **)&ctx->bicg_iters, nc * sizeof(int)));
        CK(dev_alloc(ctx, (void **)&ctx->vol_passes, nc * sizeof(int)));
        if (hipMemset(ctx->bicg_iters, 0, nc * sizeof(int)) != hipSuccess) return fail(LBMIBM_ERR_HIP);
        if (hipMemset(ctx->vol_passes, 0, nc * sizeof(int)) != hipSuccess) return fail(LBMIBM_ERR_HIP);
        if (hipMemset(ctx->gx, 0, 3 * nc * sizeof(double)) != hipSuccess) return fail(LBMIBM_ERR_HIP);
    }
    // fluid at rest: f = w (deviation 0), macros rho = 1, u = 0 (init1.cu:413-449)
    CK(lattice_fill(ctx, ctx->f[0], (size_t)Q * N, 0.0, true));
    CK(lattice_fill(ctx, ctx->f[1], (size_t)Q * N, 0.0, true));
    CK(lattice_fill(ctx, ctx->F, 3 * (size_t)N, 0.0, false));
    CK(lattice_fill(ctx, ctx->U, 3 * (size_t)N, 0.0, false));
    if (p->store_macros) {
        CK(lattice_fill(ctx, ctx->Ui, 3 * (size_t)N, 0.0, false));
        CK(lattice_fill(ctx, ctx->rho, (size_t)N, 1.0, false));
        CK(lattice_fill(ctx, ctx->rhom, (size_t)N, 1.0, false));
    }
    CK(set_step(ctx, 0, false));
    CK(membrane_autotune(ctx));
    CK(ibm_autotune(ctx));
    if (p->vof_period > 0) CK(vof_alloc(ctx));
    if (p->tube_radius > 0.0) {
        // per-column tube words: bit 0 inside, bit q the pull source of q
        // outside; and the wall fractions of the cut links, by q
        const size_t ncol = (size_t)ctx->nx * ctx->ny;
        std::vector<unsigned> col(ncol, 0u);
        ctx->tube_rh.assign(ncol * Q, -1.0);
        for (int iy = 0; iy < ctx->ny; iy++)
            for (int ix = 0; ix < ctx->nx; ix++) {
                const size_t c = ix + (size_t)ctx->nx * iy;
                if (!tube_inside(ctx, ix, iy)) continue;
                unsigned w = 1u;
                for (int q = 1; q < Q; q++)
                    if (!tube_inside(ctx, ix - CXh[q], iy - CYh[q])) {
                        w |= 1u << q;
                        if (!p->tube_staircase) ctx->tube_rh[c * Q + q] = tube_fraction(ctx, ix, iy, opph(q));
                    }
                col[c] = w;
            }
        CK(dev_alloc(ctx, &ctx->tube_col, ncol * sizeof(unsigned)));
        LBMIBM_HIP(ctx, hipMemcpy(ctx->tube_col, col.data(), ncol * sizeof(unsigned), hipMemcpyHostToDevice));
        if (!p->tube_staircase) {  // the staircase bounces every cut link half-way
            // the lattice kernel's link coefficients: 2r (r < 1/2), -1/(2r)
            // (1/2 <= r <= 1), -2 (no interpolation)
            std::vector<double> sv(ncol * Q);
            for (size_t k = 0; k < sv.size(); k++) {
                const double r = ctx->tube_rh[k];
                sv[k] = r < 0.0 || r > 1.0 ? -2.0 : (r < 0.5 ? 2.0 * r : -(1.0 / (2.0 * r)));
            }
            CK(dev_alloc(ctx, &ctx->tube_r, ncol * Q * sizeof(double)));
            LBMIBM_HIP(ctx, hipMemcpy(ctx->tube_r, sv.data(), ncol * Q * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    CK(reset_status(ctx));
    CK(sync_check(ctx));
#undef CK
    *out = ctx;
    return LBMIBM_OK;
}

void lbmibm_destroy(lbmibm_ctx *ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    delete ctx;
}

// ---- stage entry points -------------------------------------------------
lbmibm_status lbmibm_lbm_d3q19(lbmibm_ctx *ctx, int iter) {
    if (!ctx) return LBMIBM_ERR_ARG;
    lbmibm_status s = set_step(ctx, iter, false);
    if (s == LBMIBM_OK) s = lattice_step(ctx);
    if (s == LBMIBM_OK) s = sync_check(ctx);  // the reference syncs (lbm.cu:1573)
    return s;
}

lbmibm_status lbmibm_ibm_st1(lbmibm_ctx *ctx) {
    if (!ctx) return LBMIBM_ERR_ARG;
    lbmibm_status s = reset_status(ctx);
    if (s == LBMIBM_OK) s = ibm_spread(ctx);
    return s == LBMIBM_OK ? check_flags(ctx, nullptr) : s;
}

lbmibm_status lbmibm_ibm_st2(lbmibm_ctx *ctx) {
    if (!ctx) return LBMIBM_ERR_ARG;
    lbmibm_status s = ibm_interp(ctx);
    return s == LBMIBM_OK ? sync_check(ctx) : s;
}

lbmibm_status lbmibm_stress(lbmibm_ctx *ctx, int iter) {
    if (!ctx) return LBMIBM_ERR_ARG;
    lbmibm_status s = set_step(ctx, iter, false);
    if (s == LBMIBM_OK) s = reset_status(ctx);
    if (s == LBMIBM_OK) s = membrane_forces(ctx);
    return s == LBMIBM_OK ? check_flags(ctx, nullptr) : s;
}

lbmibm_status lbmibm_volume_init(lbmibm_ctx *ctx) {
    if (!ctx) return LBMIBM_ERR_ARG;
    lbmibm_status s = volume_run(ctx, 0);
    return s == LBMIBM_OK ? sync_check(ctx) : s;
}

lbmibm_status lbmibm_volume_correction(lbmibm_ctx *ctx, int *iterations) {
    if (!ctx) return LBMIBM_ERR_ARG;
    lbmibm_status s = reset_status(ctx);
    if (s == LBMIBM_OK) s = volume_run(ctx, 1);
    int k = 0;
    if (s == LBMIBM_OK) s = check_flags(ctx, &k);
    if (iterations) *iterations = k;
    return s;
}

lbmibm_status lbmibm_step(lbmibm_ctx *ctx, int iter) {
    if (!ctx) return LBMIBM_ERR_ARG;
    lbmibm_status s = reset_status(ctx);
    if (s == LBMIBM_OK) s = enqueue_step(ctx, iter, false);
    return s == LBMIBM_OK ? check_flags(ctx, nullptr) : s;
}

lbmibm_status lbmibm_run(lbmibm_ctx *ctx, int iter0, int nsteps, int use_graph) {
    if (!ctx || nsteps < 0) return LBMIBM_ERR_ARG;
    lbmibm_status s = reset_status(ctx);
    if (s != LBMIBM_OK) return s;
    const int M = ctx->p.m;
    if (ctx->p.vof_period > 0) use_graph = 0;  // front tracking synchronises with the host
    // a replayable block is whole M-cycles with an even number of lattice
    // steps, so that the population buffers are back in place, repeated to at
    // least 8 steps: a step's volume correction overlaps the next step's
    // membrane force, except across block ends
    const int L0 = (M % 2 == 0) ? M : 2 * M;
    const int L = L0 * ((8 + L0 - 1) / L0);
    // LBMIBM_OVERLAP_VOLUME=0 keeps the volume correction on the stream
    const char *ov = getenv("LBMIBM_OVERLAP_VOLUME");
    const bool overlap = !(ov && ov[0] == '0');
    int it = iter0;
    const int end = iter0 + nsteps;
    s = set_step(ctx, iter0, false);
    // eager steps until the next step starts an M-cycle
    while (s == LBMIBM_OK && it < end && (use_graph == 0 || (it % M) != 0)) {
        s = enqueue_step(ctx, it + 1, true, overlap);
        it++;
    }
    if (s == LBMIBM_OK) s = join_volume(ctx);  // nothing pending across a capture boundary
    if (s == LBMIBM_OK && use_graph && end - it >= L) {
        if (ctx->graph_ready && (ctx->graph_cur != ctx->cur || ctx->graph_force_zero != ctx->force_zero))
            ctx->graph_ready = false;
        if (!ctx->graph_ready && ctx->graph) {  // recorded before a state change: released once it has run
            LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
            (void)hipGraphExecDestroy(ctx->graph);
            ctx->graph = nullptr;
        }
        if (!ctx->graph_ready) {
            ctx->graph_cur = ctx->cur;
            ctx->graph_force_zero = ctx->force_zero;
            // the block's first spreading must bin: on replay the positions
            // are the previous block's, whatever the host flag says now
            invalidate_bins(ctx);
            hipGraph_t g = nullptr;
            LBMIBM_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            lbmibm_status cs = LBMIBM_OK;
            for (int k = 0; k < L && cs == LBMIBM_OK; k++) cs = enqueue_step(ctx, it + k + 1, true, overlap);
            if (cs == LBMIBM_OK) cs = join_volume(ctx);
            hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            if (cs != LBMIBM_OK) return cs;
            if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
            e = hipGraphInstantiate(&ctx->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return hip_fail(ctx, e, "hipGraphInstantiate");
            ctx->graph_ready = true;
        }
        while (end - it >= L) {
            LBMIBM_HIP(ctx, hipGraphLaunch(ctx->graph, ctx->stream));
            it += L;
        }
    }
    while (s == LBMIBM_OK && it < end) {
        s = enqueue_step(ctx, it + 1, true, overlap);
        it++;
    }
    if (s == LBMIBM_OK) {
        s = join_volume(ctx);
    } else {  // an enqueue failed: let anything left on the third stream finish
        ctx->vol_pending = false;
        (void)hipStreamSynchronize(ctx->vstream);
    }
    return s == LBMIBM_OK ? check_flags(ctx, nullptr) : s;
}

lbmibm_status lbmibm_front_tracking(lbmibm_ctx *ctx, int *iterations, int *attempts) {
    if (!ctx) return LBMIBM_ERR_ARG;
    if (ctx->p.vof_period <= 0 || ctx->nc < 1) {
        ctx->err = "front tracking needs vof_period > 0";
        return LBMIBM_ERR_STATE;
    }
    lbmibm_status s = join_volume(ctx);
    if (s == LBMIBM_OK) s = vof_front_tracking(ctx, iterations, attempts);
    return s == LBMIBM_OK ? sync_check(ctx) : s;
}

lbmibm_status lbmibm_thinc_wlic(lbmibm_ctx *ctx) {
    if (!ctx) return LBMIBM_ERR_ARG;
    if (ctx->p.vof_period <= 0) {
        ctx->err = "THINC/WLIC needs vof_period > 0";
        return LBMIBM_ERR_STATE;
    }
    return vof_thinc(ctx);
}

lbmibm_status lbmibm_download_vf(lbmibm_ctx *ctx, double *vf) {
    if (!ctx || !vf) return LBMIBM_ERR_ARG;
    const long long N = ctx->N;
    std::vector<double> v((size_t)N, 0.0);
    if (ctx->vof[0]) {
        LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        LBMIBM_HIP(ctx, hipMemcpy(v.data(), ctx->vof[ctx->vof_cur], N * sizeof(double), hipMemcpyDeviceToHost));
    }
    for (int pz = 0; pz < ctx->nzp; pz++)
        for (int py = 0; py < ctx->nyp; py++)
            for (int px = 0; px < ctx->nxp; px++) {
                bool comp;
                const long long i = padded_image(ctx, px, py, pz, comp);
                vf[px + (long long)ctx->nxp * (py + (long long)ctx->nyp * pz)] = i < 0 ? 0.0 : v[i];
            }
    return LBMIBM_OK;
}

lbmibm_status lbmibm_upload_vf(lbmibm_ctx *ctx, const double *vf) {
    if (!ctx || !vf) return LBMIBM_ERR_ARG;
    if (!ctx->vof[0]) {
        ctx->err = "vf needs vof_period > 0";
        return LBMIBM_ERR_STATE;
    }
    std::vector<double> v((size_t)ctx->N);
    for (int iz = 0; iz < ctx->nz; iz++)
        for (int iy = 0; iy < ctx->ny; iy++)
            for (int ix = 0; ix < ctx->nx; ix++)
                v[ix + (size_t)ctx->nx * (iy + (size_t)ctx->ny * iz)] = vf[padded_index(ctx, ix, iy, iz)];
    LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LBMIBM_HIP(ctx, hipMemcpy(ctx->vof[ctx->vof_cur], v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
    return LBMIBM_OK;
}

lbmibm_status lbmibm_adhesion_setup(lbmibm_ctx *ctx, const lbmibm_adhesion_params *a, const unsigned char *is_knob) {
    if (!ctx || !a) return LBMIBM_ERR_ARG;
    if (ctx->nc < 1 || ctx->p.periodic_y || !(a->receptor_density > 0.0) ||
        !(a->rc > 0.0) || !(a->shear_rate > 0.0) || !(a->kbt > 0.0) || a->gs == 0.0 || a->limit < 1 ||
        a->max_neighwall < 1 || a->springs_per_knob < 1) {
        ctx->err = "adhesion: needs cells, walls (the slab or the tube) and positive constants";
        return LBMIBM_ERR_ARG;
    }
    lbmibm_status s = join_volume(ctx);
    if (s == LBMIBM_OK) s = adhesion_setup(ctx, a, is_knob);
    ctx->graph_ready = false;
    return s;
}

lbmibm_status lbmibm_adhesion(lbmibm_ctx *ctx) {
    if (!ctx) return LBMIBM_ERR_ARG;
    if (!ctx->adh_on) {
        ctx->err = "adhesion: call lbmibm_adhesion_setup first";
        return LBMIBM_ERR_STATE;
    }
    lbmibm_status s = join_volume(ctx);
    return s == LBMIBM_OK ? adhesion_step(ctx) : s;
}

lbmibm_status lbmibm_adhesion_state(lbmibm_ctx *ctx, unsigned char *is_knob, int *id_wall, double *r_wall,
                                    int *attach, double *r_ij, int *is_attached, double *f_adh) {
    if (!ctx) return LBMIBM_ERR_ARG;
    if (!ctx->adh_on) {
        ctx->err = "adhesion: call lbmibm_adhesion_setup first";
        return LBMIBM_ERR_STATE;
    }
    LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const size_t nv = ctx->nv, nn = nv * ctx->adh.max_neighwall, ns = nv * ctx->adh.springs_per_knob;
    const size_t nrec = (size_t)ctx->adh_nxw * ctx->adh_nzw;
    if (is_knob) LBMIBM_HIP(ctx, hipMemcpy(is_knob, ctx->adh_knob, nv, hipMemcpyDeviceToHost));
    if (id_wall) LBMIBM_HIP(ctx, hipMemcpy(id_wall, ctx->adh_idw, nn * sizeof(int), hipMemcpyDeviceToHost));
    if (r_wall) LBMIBM_HIP(ctx, hipMemcpy(r_wall, ctx->adh_rw, nn * sizeof(double), hipMemcpyDeviceToHost));
    if (attach) LBMIBM_HIP(ctx, hipMemcpy(attach, ctx->adh_attach, ns * sizeof(int), hipMemcpyDeviceToHost));
    if (r_ij) LBMIBM_HIP(ctx, hipMemcpy(r_ij, ctx->adh_rij, ns * sizeof(double), hipMemcpyDeviceToHost));
    if (is_attached) LBMIBM_HIP(ctx, hipMemcpy(is_attached, ctx->adh_isatt, nrec * sizeof(int), hipMemcpyDeviceToHost));
    if (f_adh) LBMIBM_HIP(ctx, hipMemcpy(f_adh, ctx->adh_fadh, 3 * nv * sizeof(double), hipMemcpyDeviceToHost));
    return LBMIBM_OK;
}

lbmibm_status lbmibm_time_stages(lbmibm_ctx *ctx, int iter0, int nsteps, double *stage_ms, int *stage_calls) {
    if (!ctx || nsteps < 1 || !stage_ms) return LBMIBM_ERR_ARG;
    hipEvent_t ev[6];
    for (auto &e : ev) LBMIBM_HIP(ctx, hipEventCreate(&e));
    double acc[6] = {0, 0, 0, 0, 0, 0};
    int calls[6] = {0, 0, 0, 0, 0, 0};
    lbmibm_status s = reset_status(ctx);
    const int M = ctx->p.m;
    for (int k = 0; k < nsteps && s == LBMIBM_OK; k++) {
        const int it = iter0 + k + 1;
        const bool mem = (it - 1) % M == 0 && ctx->nc > 0, post = it % M == 0 && ctx->nc > 0;
        s = set_step(ctx, it, false);
        if (s != LBMIBM_OK) break;
        (void)hipEventRecord(ev[0], ctx->stream);
        if (mem) s = membrane_forces(ctx);
        (void)hipEventRecord(ev[1], ctx->stream);
        if (s == LBMIBM_OK && mem) s = ibm_spread(ctx);
        (void)hipEventRecord(ev[2], ctx->stream);
        if (s == LBMIBM_OK) s = lattice_step(ctx);
        (void)hipEventRecord(ev[3], ctx->stream);
        if (s == LBMIBM_OK && post) s = ibm_interp(ctx);
        (void)hipEventRecord(ev[4], ctx->stream);
        if (s == LBMIBM_OK && post && ctx->p.volume_correction) s = volume_run(ctx, 1);
        (void)hipEventRecord(ev[5], ctx->stream);
        if (hipEventSynchronize(ev[5]) != hipSuccess) { s = LBMIBM_ERR_HIP; break; }
        const bool ran[5] = {mem, mem, true, post, post && ctx->p.volume_correction != 0};
        for (int j = 0; j < 5; j++) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev[j], ev[j + 1]);
            if (ran[j]) { acc[j] += ms; calls[j]++; }
        }
        float tot = 0.f;
        (void)hipEventElapsedTime(&tot, ev[0], ev[5]);
        acc[5] += tot;
        calls[5]++;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    for (int j = 0; j < 6; j++) {
        stage_ms[j] = calls[j] ? acc[j] / calls[j] : 0.0;
        if (stage_calls) stage_calls[j] = calls[j];
    }
    return s == LBMIBM_OK ? check_flags(ctx, nullptr) : s;
}

lbmibm_status lbmibm_time_kernels(lbmibm_ctx *ctx, int iter0, int nsteps, char *report, int report_len) {
    if (!ctx || nsteps < 1 || !report || report_len < 1) return LBMIBM_ERR_ARG;
    report[0] = 0;
    lbmibm_status s = reset_status(ctx);
    if (s == LBMIBM_OK) s = join_volume(ctx);
    ctx->kt_recs.clear();
    ctx->kt_on = true;
    for (int k = 0; k < nsteps && s == LBMIBM_OK; k++) s = enqueue_step(ctx, iter0 + k + 1, false);
    ctx->kt_on = false;
    if (s == LBMIBM_OK) s = join_volume(ctx);
    if (hipDeviceSynchronize() != hipSuccess && s == LBMIBM_OK) s = LBMIBM_ERR_HIP;
    // per name: launches and total time, in order of first appearance
    std::vector<std::pair<std::string, std::pair<int, double>>> acc;
    for (const auto &r : ctx->kt_recs) {
        float ms = 0.0f;
        if (s == LBMIBM_OK) (void)hipEventElapsedTime(&ms, r.a, r.b);
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
        auto it = std::find_if(acc.begin(), acc.end(), [&](const auto &e) { return e.first == r.name; });
        if (it == acc.end()) acc.push_back({r.name, {1, (double)ms}});
        else { it->second.first++; it->second.second += ms; }
    }
    ctx->kt_recs.clear();
    std::string out;
    char line[160];
    for (const auto &e : acc) {
        snprintf(line, sizeof line, "%s\t%d\t%.6f\n", e.first.c_str(), e.second.first,
                 e.second.second / e.second.first);
        out += line;
    }
    snprintf(report, (size_t)report_len, "%s", out.c_str());
    return s == LBMIBM_OK ? check_flags(ctx, nullptr) : s;
}

lbmibm_status lbmibm_synchronize(lbmibm_ctx *ctx) {
    if (!ctx) return LBMIBM_ERR_ARG;
    return sync_check(ctx);
}

// ---- exchange in reference layouts -------------------------------------
lbmibm_status lbmibm_upload_populations(lbmibm_ctx *ctx, const double *fn) {
    if (!ctx || !fn) return LBMIBM_ERR_ARG;
    // f*_q(x) = fn_q(x + c_q), or fn_opp(q)(x) when x + c_q is a wall: the
    // inverse of the push-with-bounce-back of lbm.cu:1273-1310.
    const long long N = ctx->N;
    std::vector<double> fs((size_t)Q * N);
    for (int iz = 0; iz < ctx->nz; iz++)
        for (int iy = 0; iy < ctx->ny; iy++)
            for (int ix = 0; ix < ctx->nx; ix++) {
                const long long i = ix + (long long)ctx->nx * (iy + (long long)ctx->ny * iz);
                const long long pi = padded_index(ctx, ix, iy, iz);
                for (int q = 0; q < Q; q++) {
                    const long long t = push_target(ctx, ix, iy, iz, q);
                    double v;
                    if (t < 0) v = fn[pi * Q + opph(q)];
                    else {
                        const int tx = (int)(t % ctx->nx), ty = (int)((t / ctx->nx) % ctx->ny), tz = (int)(t / ((long long)ctx->nx * ctx->ny));
                        v = fn[padded_index(ctx, tx, ty, tz) * Q + q];
                    }
                    fs[(size_t)q * N + i] = v;
                }
            }
    ctx->graph_ready = false;
    return lattice_set_state_from_fstar(ctx, fs.data());
}

lbmibm_status lbmibm_download_populations(lbmibm_ctx *ctx, double *fn) {
    if (!ctx || !fn) return LBMIBM_ERR_ARG;
    const long long N = ctx->N;
    std::vector<double> fs((size_t)Q * N);
    lbmibm_status s = lattice_get_fstar(ctx, fs.data());
    if (s != LBMIBM_OK) return s;
    // sliding walls: the term the next pull adds to the bounced links
    std::vector<double> mw;
    if (ctx->mw) {
        mw.resize(2 * (size_t)ctx->nx * ctx->nz);
        s = get_T(ctx, mw.data(), ctx->mw, mw.size());
        if (s != LBMIBM_OK) return s;
    }
    for (int iz = 0; iz < ctx->nz; iz++)
        for (int iy = 0; iy < ctx->ny; iy++)
            for (int ix = 0; ix < ctx->nx; ix++) {
                const long long i = ix + (long long)ctx->nx * (iy + (long long)ctx->ny * iz);
                const long long pi = padded_index(ctx, ix, iy, iz);
                for (int q = 0; q < Q; q++) {
                    const long long src = pull_source(ctx, ix, iy, iz, q);
                    double f = src < 0 ? fs[(size_t)opph(q) * N + i] : fs[(size_t)q * N + src];
                    if (src < 0 && !mw.empty() && CZh[q] != 0) {
                        const size_t side = CYh[q] == 1 ? 0 : (size_t)ctx->nx * ctx->nz;
                        f += (double)CZh[q] / WDh[q] * mw[side + ix + (size_t)ctx->nx * iz];
                    }
                    if (src < 0 && ctx->p.tube_radius > 0.0 && tube_inside(ctx, ix, iy)) {
                        // the interpolated bounce-back of the kernel (lbm_cylinder.cu:61-85)
                        const double r = ctx->tube_rh[((size_t)ix + (size_t)ctx->nx * iy) * Q + q];
                        const long long s2 = pull_source(ctx, ix, iy, iz, opph(q));
                        const double own = fs[(size_t)q * N + i];
                        const double f2 = s2 < 0 ? own : fs[(size_t)opph(q) * N + s2];
                        if (r >= 0.0 && r < 0.5) f = 2.0 * r * f + (1.0 - 2.0 * r) * f2;
                        else if (r >= 0.5 && r <= 1.0) f = 1.0 / (2.0 * r) * f + (2.0 * r - 1.0) / (2.0 * r) * own;
                    }
                    fn[pi * Q + q] = f;
                }
            }
    return LBMIBM_OK;
}

lbmibm_status lbmibm_set_equilibrium(lbmibm_ctx *ctx, const double *rho, const double *u, const double *v,
                                     const double *w) {
    if (!ctx) return LBMIBM_ERR_ARG;
    // flu_init (init1.cu:413-449) on the padded grid, then uploaded with the
    // reference's post-stream meaning.
    const long long np = ctx->np;
    const double Cr = ctx->p.cr, Cr2 = Cr * Cr;
    std::vector<double> fn((size_t)np * Q);
    for (long long i = 0; i < np; i++) {
        const double td = rho ? rho[i] : 1.0, tu = u ? u[i] : 0.0, tv = v ? v[i] : 0.0, tw = w ? w[i] : 0.0;
        const double uvw = -1.5 * (tu * tu + tv * tv + tw * tw) * Cr2;
        for (int q = 0; q < Q; q++) {
            const double cu = CXh[q] * tu + CYh[q] * tv + CZh[q] * tw;
            fn[i * Q + q] = td / WDh[q] * (1.0 + 3.0 * cu * Cr + 4.5 * cu * cu * Cr2 + uvw);
        }
    }
    lbmibm_status s = lbmibm_upload_populations(ctx, fn.data());
    if (s != LBMIBM_OK) return s;
    // macros take their initial values on the fluid nodes
    const long long N = ctx->N;
    std::vector<double> uu(3 * (size_t)N), rr((size_t)N);
    for (int iz = 0; iz < ctx->nz; iz++)
        for (int iy = 0; iy < ctx->ny; iy++)
            for (int ix = 0; ix < ctx->nx; ix++) {
                const long long i = ix + (long long)ctx->nx * (iy + (long long)ctx->ny * iz);
                const long long pi = padded_index(ctx, ix, iy, iz);
                uu[i] = u ? u[pi] : 0.0;
                uu[N + i] = v ? v[pi] : 0.0;
                uu[2 * N + i] = w ? w[pi] : 0.0;
                rr[i] = rho ? rho[pi] : 1.0;
            }
    s = put_T(ctx, ctx->U, uu.data(), 3 * (size_t)N);
    if (s == LBMIBM_OK && ctx->p.store_macros) {
        s = put_T(ctx, ctx->Ui, uu.data(), 3 * (size_t)N);
        if (s == LBMIBM_OK) s = put_T(ctx, ctx->rho, rr.data(), (size_t)N);
        if (s == LBMIBM_OK) s = put_T(ctx, ctx->rhom, rr.data(), (size_t)N);
    }
    if (s == LBMIBM_OK) s = lattice_fill(ctx, ctx->F, 3 * (size_t)N, 0.0, false);
    ctx->force_zero = true;
    return s == LBMIBM_OK ? sync_check(ctx) : s;
}

lbmibm_status lbmibm_download_macros(lbmibm_ctx *ctx, double *dn, double *un, double *vn, double *wn, double *dm,
                                     double *um, double *vm, double *wm) {
    if (!ctx) return LBMIBM_ERR_ARG;
    if ((dn || un || vn || wn || dm) && !ctx->p.store_macros) {
        ctx->err = "dn/un/vn/wn/dm need store_macros=1";
        return LBMIBM_ERR_STATE;
    }
    const long long N = ctx->N;
    std::vector<double> U(3 * (size_t)N), Ui, R, RM;
    lbmibm_status s = get_T(ctx, U.data(), ctx->U, 3 * (size_t)N);
    if (s == LBMIBM_OK && ctx->p.store_macros) {
        Ui.resize(3 * (size_t)N); R.resize(N); RM.resize(N);
        s = get_T(ctx, Ui.data(), ctx->Ui, 3 * (size_t)N);
        if (s == LBMIBM_OK) s = get_T(ctx, R.data(), ctx->rho, N);
        if (s == LBMIBM_OK) s = get_T(ctx, RM.data(), ctx->rhom, N);
    }
    if (s != LBMIBM_OK) return s;
    for (int pz = 0; pz < ctx->nzp; pz++)
        for (int py = 0; py < ctx->nyp; py++)
            for (int px = 0; px < ctx->nxp; px++) {
                bool comp;
                const long long i = padded_image(ctx, px, py, pz, comp);
                if (i < 0) continue;
                const long long pi = px + (long long)ctx->nxp * (py + (long long)ctx->nyp * pz);
                if (um) um[pi] = U[i];
                if (vm) vm[pi] = U[N + i];
                if (wm) wm[pi] = U[2 * N + i];
                if (dn) dn[pi] = R[i];
                if (un) un[pi] = Ui[i];
                if (vn) vn[pi] = Ui[N + i];
                if (wn) wn[pi] = Ui[2 * N + i];
                if (dm) dm[pi] = RM[i];
            }
    return LBMIBM_OK;
}

lbmibm_status lbmibm_upload_macros(lbmibm_ctx *ctx, const double *um, const double *vm, const double *wm) {
    if (!ctx || !um || !vm || !wm) return LBMIBM_ERR_ARG;
    const long long N = ctx->N;
    std::vector<double> U(3 * (size_t)N);
    for (int iz = 0; iz < ctx->nz; iz++)
        for (int iy = 0; iy < ctx->ny; iy++)
            for (int ix = 0; ix < ctx->nx; ix++) {
                const long long i = ix + (long long)ctx->nx * (iy + (long long)ctx->ny * iz);
                const long long pi = padded_index(ctx, ix, iy, iz);
                U[i] = um[pi]; U[N + i] = vm[pi]; U[2 * N + i] = wm[pi];
            }
    lbmibm_status s = put_T(ctx, ctx->U, U.data(), 3 * (size_t)N);
    return s == LBMIBM_OK ? sync_check(ctx) : s;
}

lbmibm_status lbmibm_download_lattice_force(lbmibm_ctx *ctx, double *fx, double *fy, double *fz) {
    if (!ctx) return LBMIBM_ERR_ARG;
    const long long N = ctx->N;
    std::vector<double> F(3 * (size_t)N);
    lbmibm_status s = get_T(ctx, F.data(), ctx->F, 3 * (size_t)N);
    if (s != LBMIBM_OK) return s;
    for (int iz = 0; iz < ctx->nz; iz++)
        for (int iy = 0; iy < ctx->ny; iy++)
            for (int ix = 0; ix < ctx->nx; ix++) {
                const long long i = ix + (long long)ctx->nx * (iy + (long long)ctx->ny * iz);
                const long long pi = padded_index(ctx, ix, iy, iz);
                if (fx) fx[pi] = F[i];
                if (fy) fy[pi] = F[N + i];
                if (fz) fz[pi] = F[2 * N + i];
            }
    return LBMIBM_OK;
}

lbmibm_status lbmibm_upload_lattice_force(lbmibm_ctx *ctx, const double *fx, const double *fy, const double *fz) {
    if (!ctx || !fx || !fy || !fz) return LBMIBM_ERR_ARG;
    const long long N = ctx->N;
    std::vector<double> F(3 * (size_t)N);
    bool zero = true;
    for (int iz = 0; iz < ctx->nz; iz++)
        for (int iy = 0; iy < ctx->ny; iy++)
            for (int ix = 0; ix < ctx->nx; ix++) {
                const long long i = ix + (long long)ctx->nx * (iy + (long long)ctx->ny * iz);
                const long long pi = padded_index(ctx, ix, iy, iz);
                F[i] = fx[pi]; F[N + i] = fy[pi]; F[2 * N + i] = fz[pi];
                zero = zero && F[i] == 0.0 && F[N + i] == 0.0 && F[2 * N + i] == 0.0;
            }
    lbmibm_status s = put_T(ctx, ctx->F, F.data(), 3 * (size_t)N);
    ctx->force_zero = zero;
    ctx->graph_ready = false;
    return s == LBMIBM_OK ? sync_check(ctx) : s;
}

lbmibm_status lbmibm_download_cells(lbmibm_ctx *ctx, double *x, double *y, double *z, double *u, double *v,
                                    double *w, double *fx, double *fy, double *fz, double *xd) {
    if (!ctx) return LBMIBM_ERR_ARG;
    const size_t nv = ctx->nv;
    if (nv == 0) return LBMIBM_OK;
    std::vector<double> a(3 * nv);
    struct { void *src; bool dbl; double *d[3]; } sets[3] = {
        {ctx->xs, true, {x, y, z}}, {ctx->uvw, false, {u, v, w}}, {ctx->cf, false, {fx, fy, fz}}};
    LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto &st : sets) {
        if (!st.d[0] && !st.d[1] && !st.d[2]) continue;
        if (st.dbl) {
            LBMIBM_HIP(ctx, hipMemcpy(a.data(), st.src, 3 * nv * sizeof(double), hipMemcpyDeviceToHost));
        } else {
            lbmibm_status s = get_T(ctx, a.data(), st.src, 3 * nv);
            if (s != LBMIBM_OK) return s;
        }
        for (int d = 0; d < 3; d++)
            if (st.d[d]) std::copy(a.begin() + d * nv, a.begin() + (d + 1) * nv, st.d[d]);
    }
    if (xd) LBMIBM_HIP(ctx, hipMemcpy(xd, ctx->xd, 3 * nv * sizeof(double), hipMemcpyDeviceToHost));
    return LBMIBM_OK;
}

lbmibm_status lbmibm_upload_cells(lbmibm_ctx *ctx, const double *x, const double *y, const double *z,
                                  const double *xd, const double *fx, const double *fy, const double *fz) {
    if (!ctx) return LBMIBM_ERR_ARG;
    const size_t nv = ctx->nv;
    if (nv == 0) return LBMIBM_OK;
    std::vector<double> a(3 * nv);
    lbmibm_status s = LBMIBM_OK;
    if (x && y && z) {
        std::copy(x, x + nv, a.begin());
        std::copy(y, y + nv, a.begin() + nv);
        std::copy(z, z + nv, a.begin() + 2 * nv);
        LBMIBM_HIP(ctx, hipMemcpy(ctx->xs, a.data(), 3 * nv * sizeof(double), hipMemcpyHostToDevice));
        invalidate_bins(ctx);
    }
    if (xd) LBMIBM_HIP(ctx, hipMemcpy(ctx->xd, xd, 3 * nv * sizeof(double), hipMemcpyHostToDevice));
    if (s == LBMIBM_OK && fx && fy && fz) {
        std::copy(fx, fx + nv, a.begin());
        std::copy(fy, fy + nv, a.begin() + nv);
        std::copy(fz, fz + nv, a.begin() + 2 * nv);
        s = put_T(ctx, ctx->cf, a.data(), 3 * nv);
    }
    return s == LBMIBM_OK ? sync_check(ctx) : s;
}

lbmibm_status lbmibm_download_membrane(lbmibm_ctx *ctx, double *q, double *t, double *rg, double *tp, double *qi,
                                       int *bicg_iters) {
    if (!ctx) return LBMIBM_ERR_ARG;
    if (ctx->nc == 0) return LBMIBM_OK;
    lbmibm_status s = LBMIBM_OK;
    if (q) s = get_T(ctx, q, ctx->qv, 3 * (size_t)ctx->nv);
    if (s == LBMIBM_OK && t) s = get_T(ctx, t, ctx->tt, 4 * (size_t)ctx->nt);
    if (s == LBMIBM_OK && rg) s = get_T(ctx, rg, ctx->rg, (size_t)ctx->nt);
    if (s == LBMIBM_OK && tp) s = get_T(ctx, tp, ctx->tp, 2 * (size_t)ctx->nt);
    if (s == LBMIBM_OK && qi) s = get_T(ctx, qi, ctx->qi, 3 * (size_t)ctx->nv);
    if (s == LBMIBM_OK && bicg_iters) {
        LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        LBMIBM_HIP(ctx, hipMemcpy(bicg_iters, ctx->bicg_iters, ctx->nc * sizeof(int), hipMemcpyDeviceToHost));
    }
    return s;
}

lbmibm_status lbmibm_download_volume(lbmibm_ctx *ctx, double *volume0, double *volume, double *area, double *gx) {
    if (!ctx) return LBMIBM_ERR_ARG;
    if (ctx->nc == 0) return LBMIBM_OK;
    LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const size_t nc = ctx->nc;
    if (volume0) LBMIBM_HIP(ctx, hipMemcpy(volume0, ctx->vol0, nc * 8, hipMemcpyDeviceToHost));
    if (volume) LBMIBM_HIP(ctx, hipMemcpy(volume, ctx->vol, nc * 8, hipMemcpyDeviceToHost));
    if (area) LBMIBM_HIP(ctx, hipMemcpy(area, ctx->area, nc * 8, hipMemcpyDeviceToHost));
    if (gx) LBMIBM_HIP(ctx, hipMemcpy(gx, ctx->gx, 3 * nc * 8, hipMemcpyDeviceToHost));
    return LBMIBM_OK;
}

lbmibm_status lbmibm_set_cell_force_scale(lbmibm_ctx *ctx, const double *scale) {
    if (!ctx) return LBMIBM_ERR_ARG;
    if (ctx->nc == 0) return LBMIBM_OK;
    std::vector<double> s(ctx->nc, 1.0);
    if (scale) std::copy(scale, scale + ctx->nc, s.begin());
    lbmibm_status r = put_T(ctx, ctx->gscale, s.data(), ctx->nc);
    return r == LBMIBM_OK ? sync_check(ctx) : r;
}

}  // extern "C"
