// Reverse mode of the coupled timestep: vector-Jacobian products of the
// stages (lattice collide-stream, spreading, interpolation + advection,
// Skalak membrane force) as C entry points, the wall-shear objective and its
// gradient with respect to a per-node body-force field and the per-cell force
// scales, with checkpoint-and-recompute every K steps.
//
// No reference counterpart: the reference simulator has no derivatives.  The
// forward it differentiates is the main.cu loop body with M = 1 and no volume
// correction (stress -> ibm_st1 -> lbm_d3q19 -> ibm_st2, main.cu:80-101).
//
// Adjoint state lives in the context: the adjoint of the populations (the
// library's post-collision state f*, exchanged in the reference's
// post-stream layout, which is a permutation of it), of the velocity field
// the interpolation reads, of the lattice force, of the vertex positions and
// of the nodal forces, and the accumulated gradients.  Every stage VJP is
// taken at the current forward state, i.e. it is the derivative of the stage
// as it would run next.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "common.hpp"

using namespace lbmibm;

namespace {

template <typename T>
__global__ void k_wall_shear(const T *__restrict__ U, double *__restrict__ acc, int nx, int ny, int nz, double ex,
                             double ey, double ez, double coef) {
    // one thread per (x, z) column: both wall-adjacent nodes
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (k < nx * nz) {
        const int ix = k % nx, iz = k / nx;
        const size_t N = (size_t)nx * ny * nz;
        for (int iy : {0, ny - 1}) {
            const size_t n = (size_t)ix + (size_t)nx * ((size_t)iy + (size_t)ny * iz);
            v += ex * (double)U[n] + ey * (double)U[N + n] + ez * (double)U[2 * N + n];
        }
    }
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0 && v != 0.0) atomicAdd(acc, coef * v);
}

template <typename T>
__global__ void k_wall_seed(T *__restrict__ ua, int nx, int ny, int nz, double ex, double ey, double ez,
                            double coef) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nx * nz) return;
    const int ix = k % nx, iz = k / nx;
    const size_t N = (size_t)nx * ny * nz;
    for (int iy : {0, ny - 1}) {
        const size_t n = (size_t)ix + (size_t)nx * ((size_t)iy + (size_t)ny * iz);
        ua[n] += (T)(coef * ex);
        ua[N + n] += (T)(coef * ey);
        ua[2 * N + n] += (T)(coef * ez);
    }
}

__global__ void k_soa_to_aos(const double *__restrict__ xs, double *__restrict__ xd, int nv) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nv) return;
    xd[3 * (size_t)j] = xs[j];
    xd[3 * (size_t)j + 1] = xs[nv + j];
    xd[3 * (size_t)j + 2] = xs[2 * (size_t)nv + j];
}

lbmibm_status zero(lbmibm_ctx *ctx, void *p, size_t bytes) {
    if (p && bytes) LBMIBM_HIP(ctx, hipMemsetAsync(p, 0, bytes, ctx->stream));
    return LBMIBM_OK;
}

// Allocates the adjoint buffers on first use.
lbmibm_status adj_alloc(lbmibm_ctx *ctx) {
    if (ctx->adj_u) return LBMIBM_OK;
    const size_t N = ctx->N, nv = ctx->nv, T = ctx->szT;
    struct { void **p; size_t bytes; } a[] = {
        {&ctx->adj_f[0], Q * N * T}, {&ctx->adj_f[1], Q * N * T}, {&ctx->adj_u, 3 * N * T},
        {&ctx->adj_F, 3 * N * T}, {(void **)&ctx->adj_x, 3 * nv * 8}, {&ctx->adj_cf, 3 * nv * T},
        {(void **)&ctx->grad_H, 3 * N * 8}, {(void **)&ctx->grad_gs, (size_t)ctx->nc * 8},
        {&ctx->adj_lam, 3 * nv * T}, {(void **)&ctx->adj_dsv, nv * 8}, {(void **)&ctx->adj_corner, 24 * nv * 8},
        {&ctx->adj_scratch, 3 * nv * T}, {(void **)&ctx->adj_cscale, (size_t)ctx->nc * 8},
        {(void **)&ctx->adj_gpart, nv * 8}, {(void **)&ctx->adj_xmax, 256},
        {&ctx->adj_rec, 4 * nv * 48}};
    for (auto &e : a) {
        lbmibm_status s = dev_alloc(ctx, e.p, e.bytes);
        if (s != LBMIBM_OK) return s;
        s = zero(ctx, *e.p, e.bytes);
        if (s != LBMIBM_OK) return s;
    }
    ctx->adj_cur = 0;
    return LBMIBM_OK;
}

lbmibm_status adj_reset(lbmibm_ctx *ctx) {
    lbmibm_status s = adj_alloc(ctx);
    if (s != LBMIBM_OK) return s;
    const size_t N = ctx->N, nv = ctx->nv, T = ctx->szT;
    s = zero(ctx, ctx->adj_f[ctx->adj_cur], Q * N * T);
    if (s == LBMIBM_OK) s = zero(ctx, ctx->adj_u, 3 * N * T);
    if (s == LBMIBM_OK) s = zero(ctx, ctx->adj_F, 3 * N * T);
    if (s == LBMIBM_OK) s = zero(ctx, ctx->adj_x, 3 * nv * 8);
    if (s == LBMIBM_OK) s = zero(ctx, ctx->adj_cf, 3 * nv * T);
    if (s == LBMIBM_OK) s = zero(ctx, ctx->grad_H, 3 * N * 8);
    if (s == LBMIBM_OK) s = zero(ctx, ctx->grad_gs, (size_t)ctx->nc * 8);
    return s;
}

lbmibm_status need(lbmibm_ctx *ctx, bool ok, const char *what) {
    if (ok) return LBMIBM_OK;
    ctx->err = std::string("adjoint: ") + what;
    return LBMIBM_ERR_STATE;
}

// Conditions shared by every stage adjoint.
lbmibm_status adj_common(lbmibm_ctx *ctx) {
    lbmibm_status s = need(ctx, ctx->prec != LBMIBM_F16, "FP16 population storage has no adjoint");
    if (s == LBMIBM_OK) s = need(ctx, ctx->p.m == 1, "membrane sub-cycling M must be 1");
    if (s == LBMIBM_OK) s = need(ctx, ctx->p.vof_period == 0, "the VOF indicator has no adjoint");
    if (s == LBMIBM_OK) s = need(ctx, ctx->p.tube_radius == 0.0, "the tube geometry has no adjoint");
    if (s == LBMIBM_OK) s = need(ctx, ctx->p.wall_motion == 0, "sliding walls have no adjoint");
    if (s == LBMIBM_OK) s = need(ctx, !ctx->adh_on, "cytoadhesion has no adjoint");
    if (s == LBMIBM_OK) s = adj_alloc(ctx);
    return s;
}

lbmibm_status wall_coef(lbmibm_ctx *ctx, int window, double &coef) {
    lbmibm_status s = need(ctx, ctx->p.periodic_y == 0, "the wall-shear objective needs the y walls");
    if (s != LBMIBM_OK) return s;
    // tau_w = 2 nu (e . u)/dx at the first fluid node (dx/2 from the
    // half-way wall), nu = Df dx^2/dt; mean over both walls and the window
    const lbmibm_params &p = ctx->p;
    coef = 2.0 * p.df * p.dx / p.dt / (2.0 * ctx->nx * (double)ctx->nz) / (double)window;
    return LBMIBM_OK;
}

double e_norm(const lbmibm_params &p, double e[3]) {
    const double n = std::sqrt(p.ex * p.ex + p.ey * p.ey + p.ez * p.ez);
    e[0] = n > 0 ? p.ex / n : 1.0;
    e[1] = n > 0 ? p.ey / n : 0.0;
    e[2] = n > 0 ? p.ez / n : 0.0;
    return n;
}

lbmibm_status wall_shear_add(lbmibm_ctx *ctx, double *acc, double coef) {
    double e[3];
    e_norm(ctx->p, e);
    const int cols = ctx->nx * ctx->nz;
    if (ctx->prec == LBMIBM_F64)
        hipLaunchKernelGGL(k_wall_shear<double>, dim3((cols + 255) / 256), dim3(256), 0, ctx->stream,
                           (const double *)ctx->U, acc, ctx->nx, ctx->ny, ctx->nz, e[0], e[1], e[2], coef);
    else
        hipLaunchKernelGGL(k_wall_shear<float>, dim3((cols + 255) / 256), dim3(256), 0, ctx->stream,
                           (const float *)ctx->U, acc, ctx->nx, ctx->ny, ctx->nz, e[0], e[1], e[2], coef);
    LBMIBM_LAUNCHED(ctx, "k_wall_shear");
    return LBMIBM_OK;
}

lbmibm_status wall_seed(lbmibm_ctx *ctx, double coef) {
    double e[3];
    e_norm(ctx->p, e);
    const int cols = ctx->nx * ctx->nz;
    if (ctx->prec == LBMIBM_F64)
        hipLaunchKernelGGL(k_wall_seed<double>, dim3((cols + 255) / 256), dim3(256), 0, ctx->stream,
                           (double *)ctx->adj_u, ctx->nx, ctx->ny, ctx->nz, e[0], e[1], e[2], coef);
    else
        hipLaunchKernelGGL(k_wall_seed<float>, dim3((cols + 255) / 256), dim3(256), 0, ctx->stream,
                           (float *)ctx->adj_u, ctx->nx, ctx->ny, ctx->nz, e[0], e[1], e[2], coef);
    LBMIBM_LAUNCHED(ctx, "k_wall_seed");
    return LBMIBM_OK;
}

lbmibm_status positions_from_soa(lbmibm_ctx *ctx) {
    if (ctx->nv == 0) return LBMIBM_OK;
    hipLaunchKernelGGL(k_soa_to_aos, dim3((ctx->nv + 255) / 256), dim3(256), 0, ctx->stream,
                       (const double *)ctx->xs, (double *)ctx->xd, ctx->nv);
    LBMIBM_LAUNCHED(ctx, "k_soa_to_aos");
    invalidate_bins(ctx);
    return LBMIBM_OK;
}

lbmibm_status d2d(lbmibm_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (bytes) LBMIBM_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return LBMIBM_OK;
}

// Stage adjoints on the stream (no synchronisation).
lbmibm_status adj_lattice(lbmibm_ctx *ctx, int iter) {
    lbmibm_status s = set_step(ctx, iter, false);
    if (s == LBMIBM_OK) s = lattice_adjoint(ctx, ctx->adj_f[ctx->adj_cur], ctx->adj_f[ctx->adj_cur ^ 1]);
    if (s == LBMIBM_OK) {
        ctx->adj_cur ^= 1;
        s = zero(ctx, ctx->adj_u, 3 * (size_t)ctx->N * ctx->szT);  // consumed
    }
    return s;
}

// Padded [n] field <-> [3][N] device array of a given element size.
lbmibm_status field_get(lbmibm_ctx *ctx, const void *src, bool dbl, double *out) {
    const long long N = ctx->N;
    std::vector<double> a(3 * (size_t)N);
    if (dbl) {
        LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        LBMIBM_HIP(ctx, hipMemcpy(a.data(), src, 3 * (size_t)N * 8, hipMemcpyDeviceToHost));
    } else {
        lbmibm_status s = get_T(ctx, a.data(), src, 3 * (size_t)N);
        if (s != LBMIBM_OK) return s;
    }
    std::fill(out, out + 3 * ctx->np, 0.0);
    for (int iz = 0; iz < ctx->nz; iz++)
        for (int iy = 0; iy < ctx->ny; iy++)
            for (int ix = 0; ix < ctx->nx; ix++) {
                const long long i = ix + (long long)ctx->nx * (iy + (long long)ctx->ny * iz);
                const long long pi = padded_index(ctx, ix, iy, iz);
                for (int d = 0; d < 3; d++) out[d * ctx->np + pi] = a[d * N + i];
            }
    return LBMIBM_OK;
}

lbmibm_status field_set(lbmibm_ctx *ctx, void *dst, bool dbl, const double *in) {
    const long long N = ctx->N;
    std::vector<double> a(3 * (size_t)N);
    for (int iz = 0; iz < ctx->nz; iz++)
        for (int iy = 0; iy < ctx->ny; iy++)
            for (int ix = 0; ix < ctx->nx; ix++) {
                const long long i = ix + (long long)ctx->nx * (iy + (long long)ctx->ny * iz);
                const long long pi = padded_index(ctx, ix, iy, iz);
                for (int d = 0; d < 3; d++) a[d * N + i] = in[d * ctx->np + pi];
            }
    if (dbl) {
        LBMIBM_HIP(ctx, hipMemcpy(dst, a.data(), 3 * (size_t)N * 8, hipMemcpyHostToDevice));
        return LBMIBM_OK;
    }
    return put_T(ctx, dst, a.data(), 3 * (size_t)N);
}

}  // namespace

namespace {
// The gradient's forward runs and recomputes start every BiCGSTAB at x0 = 0
// (bicg_warm_start is set aside while they run), so that a recomputed step
// repeats its recorded solve and the finite-difference runs use the same
// map as the adjoint.
struct ColdSolves {
    lbmibm_ctx *c;
    bool prev;
    explicit ColdSolves(lbmibm_ctx *ctx) : c(ctx), prev(ctx->solve_cold) { c->solve_cold = true; }
    ~ColdSolves() { c->solve_cold = prev; }
};

// ---- wall shear without allocation -------------------------------------
// Partial sums of the wall-shear objective, one slot per 64-column wave of
// k_wall_shear_part, and the total in the slot after them.  The array is a
// per-device static: every call on a device shares it (calls that overlap on
// one device race), and it keeps the last call's values, so each call clears
// the slots it uses first.
constexpr int WS_MAX_PARTS = 65536;  // 64 columns each: nx*nz <= 4,194,304
__device__ __align__(256) double scratch[WS_MAX_PARTS + 1];

constexpr int WS_MAX_DEVICES = 64;
std::mutex ws_mutex;
double *ws_ptr[WS_MAX_DEVICES] = {};

// The device address of scratch on the current device, looked up once per
// device (the host-side symbol is not a device address).
lbmibm_status ws_scratch(lbmibm_ctx *ctx, double **out) {
    int dev = 0;
    LBMIBM_HIP(ctx, hipGetDevice(&dev));
    if (dev < 0 || dev >= WS_MAX_DEVICES) {
        ctx->err = "wall_shear_noalloc: device index out of range";
        return LBMIBM_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(ws_mutex);
    if (!ws_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            ctx->err = std::string("wall_shear_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e);
            return LBMIBM_ERR_HIP;
        }
        ws_ptr[dev] = (double *)p;
    }
    *out = ws_ptr[dev];
    return LBMIBM_OK;
}

__global__ void k_ws_clear(double *__restrict__ part, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) part[k] = 0.0;
}

// As k_wall_shear, but each wave adds its sum to its own slot (no atomics).
template <typename T>
__global__ void k_wall_shear_part(const T *__restrict__ U, double *__restrict__ part, int nx, int ny, int nz,
                                  double ex, double ey, double ez, double coef) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (k < nx * nz) {
        const int ix = k % nx, iz = k / nx;
        const size_t N = (size_t)nx * ny * nz;
        for (int iy : {0, ny - 1}) {
            const size_t n = (size_t)ix + (size_t)nx * ((size_t)iy + (size_t)ny * iz);
            v += ex * (double)U[n] + ey * (double)U[N + n] + ez * (double)U[2 * N + n];
        }
    }
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0 && k - (int)(threadIdx.x & 63) < nx * nz) part[k >> 6] += coef * v;
}

// One 256-thread work group: part[nparts] = the sum of part[0..nparts) in a
// fixed order (strided per thread, then the waves, then the four wave sums).
__global__ __launch_bounds__(256) void k_ws_total(double *__restrict__ part, int nparts) {
    __shared__ double w[4];
    double v = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 256) v += part[i];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[nparts] = (w[0] + w[1]) + (w[2] + w[3]);
}
}  // namespace

extern "C" {

lbmibm_status lbmibm_set_body_force_field(lbmibm_ctx *ctx, const double *gx, const double *gy, const double *gz) {
    if (!ctx) return LBMIBM_ERR_ARG;
    ctx->graph_ready = false;  // captured steps hold the old pointer
    if (!gx && !gy && !gz) {
        if (ctx->Hb) {
            LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
            (void)hipFree(ctx->Hb);
        }
        ctx->Hb = nullptr;
        return LBMIBM_OK;
    }
    if (!gx || !gy || !gz) return LBMIBM_ERR_ARG;
    if (!ctx->Hb) {
        lbmibm_status s = dev_alloc(ctx, &ctx->Hb, 3 * (size_t)ctx->N * ctx->szT);
        if (s != LBMIBM_OK) return s;
    }
    std::vector<double> h(3 * (size_t)ctx->np);
    std::copy(gx, gx + ctx->np, h.begin());
    std::copy(gy, gy + ctx->np, h.begin() + ctx->np);
    std::copy(gz, gz + ctx->np, h.begin() + 2 * ctx->np);
    lbmibm_status s = field_set(ctx, ctx->Hb, false, h.data());
    return s == LBMIBM_OK ? sync_check(ctx) : s;
}

lbmibm_status lbmibm_adjoint_reset(lbmibm_ctx *ctx) {
    if (!ctx) return LBMIBM_ERR_ARG;
    lbmibm_status s = adj_reset(ctx);
    return s == LBMIBM_OK ? sync_check(ctx) : s;
}

lbmibm_status lbmibm_adjoint_get(lbmibm_ctx *ctx, int which, double *out) {
    if (!ctx || !out) return LBMIBM_ERR_ARG;
    lbmibm_status s = adj_alloc(ctx);
    if (s != LBMIBM_OK) return s;
    const long long N = ctx->N;
    const size_t nv = ctx->nv;
    switch (which) {
        case LBMIBM_ADJ_POPULATIONS: {
            std::vector<double> a((size_t)Q * N);
            s = get_T(ctx, a.data(), ctx->adj_f[ctx->adj_cur], (size_t)Q * N);
            if (s != LBMIBM_OK) return s;
            std::fill(out, out + (size_t)Q * ctx->np, 0.0);
            for (int iz = 0; iz < ctx->nz; iz++)
                for (int iy = 0; iy < ctx->ny; iy++)
                    for (int ix = 0; ix < ctx->nx; ix++) {
                        const long long i = ix + (long long)ctx->nx * (iy + (long long)ctx->ny * iz);
                        const long long pi = padded_index(ctx, ix, iy, iz);
                        for (int q = 0; q < Q; q++) {
                            const long long src = pull_source(ctx, ix, iy, iz, q);
                            const int qo = q == 0 ? 0 : q + ((q % 2) << 1) - 1;
                            out[pi * Q + q] = src < 0 ? a[(size_t)qo * N + i] : a[(size_t)q * N + src];
                        }
                    }
            return LBMIBM_OK;
        }
        case LBMIBM_ADJ_VELOCITY: return field_get(ctx, ctx->adj_u, false, out);
        case LBMIBM_ADJ_LATTICE_FORCE: return field_get(ctx, ctx->adj_F, false, out);
        case LBMIBM_ADJ_GRAD_BODY_FORCE: return field_get(ctx, ctx->grad_H, true, out);
        case LBMIBM_ADJ_POSITIONS:
            if (nv == 0) return LBMIBM_OK;
            LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
            LBMIBM_HIP(ctx, hipMemcpy(out, ctx->adj_x, 3 * nv * 8, hipMemcpyDeviceToHost));
            return LBMIBM_OK;
        case LBMIBM_ADJ_NODAL_FORCE: return nv ? get_T(ctx, out, ctx->adj_cf, 3 * nv) : LBMIBM_OK;
        case LBMIBM_ADJ_GRAD_FORCE_SCALE:
            if (ctx->nc == 0) return LBMIBM_OK;
            LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
            LBMIBM_HIP(ctx, hipMemcpy(out, ctx->grad_gs, (size_t)ctx->nc * 8, hipMemcpyDeviceToHost));
            return LBMIBM_OK;
    }
    return LBMIBM_ERR_ARG;
}

lbmibm_status lbmibm_adjoint_set(lbmibm_ctx *ctx, int which, const double *in) {
    if (!ctx || !in) return LBMIBM_ERR_ARG;
    lbmibm_status s = adj_alloc(ctx);
    if (s != LBMIBM_OK) return s;
    const long long N = ctx->N;
    const size_t nv = ctx->nv;
    switch (which) {
        case LBMIBM_ADJ_POPULATIONS: {
            std::vector<double> a((size_t)Q * N);
            for (int iz = 0; iz < ctx->nz; iz++)
                for (int iy = 0; iy < ctx->ny; iy++)
                    for (int ix = 0; ix < ctx->nx; ix++) {
                        const long long i = ix + (long long)ctx->nx * (iy + (long long)ctx->ny * iz);
                        const long long pi = padded_index(ctx, ix, iy, iz);
                        for (int q = 0; q < Q; q++) {
                            const long long t = push_target(ctx, ix, iy, iz, q);
                            const int qo = q == 0 ? 0 : q + ((q % 2) << 1) - 1;
                            double v;
                            if (t < 0) v = in[pi * Q + qo];
                            else {
                                const int tx = (int)(t % ctx->nx), ty = (int)((t / ctx->nx) % ctx->ny),
                                          tz = (int)(t / ((long long)ctx->nx * ctx->ny));
                                v = in[padded_index(ctx, tx, ty, tz) * Q + q];
                            }
                            a[(size_t)q * N + i] = v;
                        }
                    }
            s = put_T(ctx, ctx->adj_f[ctx->adj_cur], a.data(), (size_t)Q * N);
            break;
        }
        case LBMIBM_ADJ_VELOCITY: s = field_set(ctx, ctx->adj_u, false, in); break;
        case LBMIBM_ADJ_LATTICE_FORCE: s = field_set(ctx, ctx->adj_F, false, in); break;
        case LBMIBM_ADJ_GRAD_BODY_FORCE: s = field_set(ctx, ctx->grad_H, true, in); break;
        case LBMIBM_ADJ_POSITIONS:
            if (nv) LBMIBM_HIP(ctx, hipMemcpy(ctx->adj_x, in, 3 * nv * 8, hipMemcpyHostToDevice));
            break;
        case LBMIBM_ADJ_NODAL_FORCE: s = nv ? put_T(ctx, ctx->adj_cf, in, 3 * nv) : LBMIBM_OK; break;
        case LBMIBM_ADJ_GRAD_FORCE_SCALE:
            if (ctx->nc) LBMIBM_HIP(ctx, hipMemcpy(ctx->grad_gs, in, (size_t)ctx->nc * 8, hipMemcpyHostToDevice));
            break;
        default: return LBMIBM_ERR_ARG;
    }
    return s == LBMIBM_OK ? sync_check(ctx) : s;
}

lbmibm_status lbmibm_adj_lbm_d3q19(lbmibm_ctx *ctx, int iter) {
    if (!ctx) return LBMIBM_ERR_ARG;
    lbmibm_status s = adj_common(ctx);
    if (s == LBMIBM_OK) s = adj_lattice(ctx, iter);
    return s == LBMIBM_OK ? sync_check(ctx) : s;
}

lbmibm_status lbmibm_adj_ibm_st1(lbmibm_ctx *ctx) {
    if (!ctx) return LBMIBM_ERR_ARG;
    lbmibm_status s = adj_common(ctx);
    if (s == LBMIBM_OK) s = ibm_spread_adjoint(ctx);
    return s == LBMIBM_OK ? sync_check(ctx) : s;
}

lbmibm_status lbmibm_adj_ibm_st2(lbmibm_ctx *ctx) {
    if (!ctx) return LBMIBM_ERR_ARG;
    lbmibm_status s = adj_common(ctx);
    if (s == LBMIBM_OK) s = ibm_interp_adjoint(ctx);
    return s == LBMIBM_OK ? sync_check(ctx) : s;
}

lbmibm_status lbmibm_adj_stress(lbmibm_ctx *ctx, int iter) {
    if (!ctx) return LBMIBM_ERR_ARG;
    lbmibm_status s = adj_common(ctx);
    if (s == LBMIBM_OK) s = need(ctx, ctx->p.fileoutput <= 0, "FILEOUTPUT accumulators must be off");
    if (s == LBMIBM_OK) s = set_step(ctx, iter, false);
    if (s == LBMIBM_OK) s = reset_status(ctx);
    if (s == LBMIBM_OK) s = membrane_forces_adjoint(ctx, false);
    return s == LBMIBM_OK ? check_flags(ctx, nullptr) : s;
}

lbmibm_status lbmibm_wall_shear(lbmibm_ctx *ctx, int iter0, int nsteps, int window, double *J) {
    if (!ctx || nsteps < 1 || window < 1 || window > nsteps) return LBMIBM_ERR_ARG;
    double coef = 0.0;
    lbmibm_status s = wall_coef(ctx, window, coef);
    if (s != LBMIBM_OK) return s;
    s = need(ctx, ctx->p.m == 1 && !ctx->p.volume_correction, "needs M = 1 and no volume correction");
    if (s != LBMIBM_OK) return s;
    ColdSolves cold(ctx);
    double *acc = nullptr;
    LBMIBM_HIP(ctx, hipMalloc(&acc, sizeof(double)));
    s = zero(ctx, acc, sizeof(double));
    if (s == LBMIBM_OK) s = reset_status(ctx);
    for (int k = 0; k < nsteps && s == LBMIBM_OK; k++) {
        s = enqueue_step(ctx, iter0 + k + 1, false);
        if (s == LBMIBM_OK && k >= nsteps - window) s = wall_shear_add(ctx, acc, coef);
    }
    if (s == LBMIBM_OK) s = check_flags(ctx, nullptr);
    if (s == LBMIBM_OK && hipMemcpy(J, acc, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) s = LBMIBM_ERR_HIP;
    (void)hipFree(acc);
    return s;
}

lbmibm_status lbmibm_wall_shear_noalloc(lbmibm_ctx *ctx, int iter0, int nsteps, int window, double *J) {
    if (!ctx || !J || nsteps < 1 || window < 1 || window > nsteps) return LBMIBM_ERR_ARG;
    const long long cols = (long long)ctx->nx * ctx->nz;
    const int nparts = (int)((cols + 63) / 64);
    if (cols > 64LL * WS_MAX_PARTS) {
        ctx->err = "wall_shear_noalloc: nx*nz = " + std::to_string(cols) + " columns exceeds the scratch limit of " +
                   std::to_string(64LL * WS_MAX_PARTS);
        return LBMIBM_ERR_ARG;
    }
    double coef = 0.0;
    lbmibm_status s = wall_coef(ctx, window, coef);
    if (s != LBMIBM_OK) return s;
    s = need(ctx, ctx->p.m == 1 && !ctx->p.volume_correction, "needs M = 1 and no volume correction");
    if (s != LBMIBM_OK) return s;
    double *part = nullptr;
    s = ws_scratch(ctx, &part);
    if (s != LBMIBM_OK) return s;
    ColdSolves cold(ctx);
    double e[3];
    e_norm(ctx->p, e);
    const dim3 grid((unsigned)((cols + 255) / 256));
    hipLaunchKernelGGL(k_ws_clear, dim3((nparts + 256) / 256), dim3(256), 0, ctx->stream, part, nparts + 1);
    LBMIBM_LAUNCHED(ctx, "k_ws_clear");
    s = reset_status(ctx);
    for (int k = 0; k < nsteps && s == LBMIBM_OK; k++) {
        s = enqueue_step(ctx, iter0 + k + 1, false);
        if (s != LBMIBM_OK || k < nsteps - window) continue;
        if (ctx->prec == LBMIBM_F64)
            hipLaunchKernelGGL(k_wall_shear_part<double>, grid, dim3(256), 0, ctx->stream, (const double *)ctx->U,
                               part, ctx->nx, ctx->ny, ctx->nz, e[0], e[1], e[2], coef);
        else
            hipLaunchKernelGGL(k_wall_shear_part<float>, grid, dim3(256), 0, ctx->stream, (const float *)ctx->U,
                               part, ctx->nx, ctx->ny, ctx->nz, e[0], e[1], e[2], coef);
        LBMIBM_LAUNCHED(ctx, "k_wall_shear_part");
    }
    if (s == LBMIBM_OK) {
        hipLaunchKernelGGL(k_ws_total, dim3(1), dim3(256), 0, ctx->stream, part, nparts);
        LBMIBM_LAUNCHED(ctx, "k_ws_total");
    }
    if (s == LBMIBM_OK) s = check_flags(ctx, nullptr);
    if (s == LBMIBM_OK &&
        hipMemcpyAsync(J, part + nparts, sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        s = LBMIBM_ERR_HIP;
    if (s == LBMIBM_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) s = LBMIBM_ERR_HIP;
    return s;
}

lbmibm_status lbmibm_adjoint_wall_shear(lbmibm_ctx *ctx, int iter0, int nsteps, int window, int checkpoint_every,
                                        double *J, double *gx, double *gy, double *gz, double *dscale) {
    if (!ctx || nsteps < 1 || window < 1 || window > nsteps || checkpoint_every < 1) return LBMIBM_ERR_ARG;
    double coef = 0.0;
    lbmibm_status s = wall_coef(ctx, window, coef);
    if (s == LBMIBM_OK) s = adj_common(ctx);
    if (s == LBMIBM_OK) s = need(ctx, !ctx->p.volume_correction, "volume correction must be off");
    if (s == LBMIBM_OK) s = need(ctx, ctx->p.fileoutput <= 0, "FILEOUTPUT accumulators must be off");
    if (s == LBMIBM_OK && ctx->nc > 0) s = need(ctx, ctx->p.runge_kutta == 4 && ctx->p.br == 0.0, "needs RK4, BR = 0");
    if (s != LBMIBM_OK) return s;
    ColdSolves cold(ctx);
    const int K = std::min(checkpoint_every, nsteps);
    const int nseg = (nsteps + K - 1) / K;
    const size_t N = ctx->N, nv = ctx->nv, szS = ctx->szS, szT = ctx->szT;
    const size_t popB = Q * N * szS, fldB = 3 * N * szT, xB = 3 * nv * 8, vB = 3 * nv * szT;
    // the tape keeps each step's populations, positions, nodal forces, load
    // q, lattice force and velocity
    // checkpoints (populations, positions) at segment starts plus the end
    // state, and the tape of one segment
    void *ck = nullptr, *tape = nullptr;
    double *acc = nullptr;
    const size_t ckB = popB + xB, tapeB = popB + xB + 2 * vB + 2 * fldB;
    LBMIBM_HIP(ctx, hipMalloc(&ck, ckB * (nseg + 1)));
    if (hipMalloc(&tape, tapeB * K) != hipSuccess || hipMalloc(&acc, sizeof(double)) != hipSuccess) {
        (void)hipFree(ck);
        (void)hipFree(tape);
        ctx->err = "adjoint: out of device memory for the checkpoints";
        return LBMIBM_ERR_HIP;
    }
    auto ck_pop = [&](int k) { return (char *)ck + ckB * k; };
    auto ck_x = [&](int k) { return (char *)ck + ckB * k + popB; };
    auto tp = [&](int k, int part) {
        const size_t off[6] = {0, popB, popB + xB, popB + xB + vB, popB + xB + 2 * vB, popB + xB + 2 * vB + fldB};
        return (char *)tape + tapeB * k + off[part];
    };
    enum { T_POP, T_X, T_CF, T_Q, T_F, T_U };
    s = zero(ctx, acc, sizeof(double));
    if (s == LBMIBM_OK) s = reset_status(ctx);
    // forward sweep: objective and checkpoints
    for (int k = 0; k < nsteps && s == LBMIBM_OK; k++) {
        if (k % K == 0) {
            s = d2d(ctx, ck_pop(k / K), ctx->f[ctx->cur], popB);
            if (s == LBMIBM_OK) s = d2d(ctx, ck_x(k / K), ctx->xs, xB);
        }
        if (s == LBMIBM_OK) s = enqueue_step(ctx, iter0 + k + 1, false);
        if (s == LBMIBM_OK && k >= nsteps - window) s = wall_shear_add(ctx, acc, coef);
    }
    if (s == LBMIBM_OK) s = d2d(ctx, ck_pop(nseg), ctx->f[ctx->cur], popB);
    if (s == LBMIBM_OK) s = d2d(ctx, ck_x(nseg), ctx->xs, xB);
    const int cur_end = ctx->cur;
    if (s == LBMIBM_OK) s = check_flags(ctx, nullptr);
    if (s == LBMIBM_OK && J && hipMemcpy(J, acc, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        s = LBMIBM_ERR_HIP;
    // reverse sweep.  The recompute writes each step's state straight into
    // the tape (the buffers the steps write are pointed at the tape slots)
    // and the reverse steps point at the slots again: no copies.  A step's
    // populations and positions come from the previous step's outputs, the
    // first step's from the checkpoint.
    struct Owned {  // restores the context's own buffers on every way out
        lbmibm_ctx *c;
        void *f0, *f1, *U, *F, *cf, *qv, *xs;
        bool active;
        void restore() {
            if (!active) return;
            c->f[0] = f0; c->f[1] = f1; c->U = U; c->F = F; c->cf = cf; c->qv = qv; c->xs = xs;
            c->xs_out = nullptr;
            invalidate_bins(c);
            active = false;
        }
        ~Owned() { restore(); }
    } owned{ctx, ctx->f[0], ctx->f[1], ctx->U, ctx->F, ctx->cf, ctx->qv, ctx->xs, true};
    if (s == LBMIBM_OK) s = adj_reset(ctx);
    for (int seg = nseg - 1; seg >= 0 && s == LBMIBM_OK; seg--) {
        const int k0 = seg * K, len = std::min(K, nsteps - k0);
        auto popk = [&](int k) -> void * { return k == 0 ? (void *)ck_pop(seg) : (void *)tp(k, T_POP); };
        auto xk = [&](int k) -> double * { return k == 0 ? (double *)ck_x(seg) : (double *)tp(k, T_X); };
        ctx->xs = xk(0);
        s = positions_from_soa(ctx);
        // recompute the segment
        for (int k = 0; k < len && s == LBMIBM_OK; k++) {
            const int iter = iter0 + k0 + k + 1;
            ctx->f[ctx->cur] = popk(k);
            ctx->f[1 - ctx->cur] = k + 1 < len ? popk(k + 1) : owned.f0;  // the last output is not needed
            ctx->xs = xk(k);
            ctx->xs_out = k + 1 < len ? xk(k + 1) : (double *)owned.xs;
            invalidate_bins(ctx);
            ctx->cf = tp(k, T_CF); ctx->qv = tp(k, T_Q); ctx->F = tp(k, T_F); ctx->U = tp(k, T_U);
            s = set_step(ctx, iter, false);
            if (s == LBMIBM_OK && ctx->nc > 0) s = membrane_forces(ctx);
            if (s == LBMIBM_OK && ctx->nc > 0) s = ibm_spread(ctx);
            if (s == LBMIBM_OK) s = lattice_step(ctx);
            if (s == LBMIBM_OK && ctx->nc > 0) s = ibm_interp(ctx);
        }
        ctx->xs_out = nullptr;
        for (int k = len - 1; k >= 0 && s == LBMIBM_OK; k--) {
            const int iter = iter0 + k0 + k + 1;
            // advection: positions before it and the velocity it read
            ctx->xs = xk(k);
            s = positions_from_soa(ctx);
            ctx->U = tp(k, T_U);
            if (s == LBMIBM_OK && ctx->nc > 0) s = ibm_interp_adjoint(ctx);
            if (s == LBMIBM_OK && k0 + k >= nsteps - window) s = wall_seed(ctx, coef);
            // lattice step from its pre-step populations and force
            ctx->f[ctx->cur] = popk(k);
            ctx->F = tp(k, T_F);
            if (s == LBMIBM_OK) {
                ctx->force_zero = ctx->nc == 0;
                s = adj_lattice(ctx, iter);
            }
            if (s == LBMIBM_OK && ctx->nc > 0) {
                ctx->cf = tp(k, T_CF);
                s = ibm_spread_adjoint(ctx);
                ctx->qv = tp(k, T_Q);
                if (s == LBMIBM_OK) s = membrane_forces_adjoint(ctx, true);
            }
        }
        if (s == LBMIBM_OK) s = check_flags(ctx, nullptr);
    }
    // the context's own buffers back (the tape is freed below)
    owned.restore();
    // gradients out, forward end state back
    if (s == LBMIBM_OK && (gx || gy || gz)) {
        std::vector<double> g(3 * (size_t)ctx->np);
        s = field_get(ctx, ctx->grad_H, true, g.data());
        if (s == LBMIBM_OK) {
            if (gx) std::copy(g.begin(), g.begin() + ctx->np, gx);
            if (gy) std::copy(g.begin() + ctx->np, g.begin() + 2 * ctx->np, gy);
            if (gz) std::copy(g.begin() + 2 * ctx->np, g.end(), gz);
        }
    }
    if (s == LBMIBM_OK && dscale && ctx->nc > 0) {
        LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        LBMIBM_HIP(ctx, hipMemcpy(dscale, ctx->grad_gs, (size_t)ctx->nc * 8, hipMemcpyDeviceToHost));
    }
    if (s == LBMIBM_OK) {
        ctx->cur = cur_end;
        s = d2d(ctx, ctx->f[ctx->cur], ck_pop(nseg), popB);
        if (s == LBMIBM_OK) s = d2d(ctx, ctx->xs, ck_x(nseg), xB);
        if (s == LBMIBM_OK) s = positions_from_soa(ctx);
        if (s == LBMIBM_OK) s = sync_check(ctx);
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ck);
    (void)hipFree(tape);
    (void)hipFree(acc);
    ctx->graph_ready = false;
    return s;
}

}  // extern "C"
