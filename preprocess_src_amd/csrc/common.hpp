// Internal declarations shared by the HIP translation units of liblbmibm.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "lbmibm.h"

namespace lbmibm {

constexpr int Q = 19;
constexpr int WG = 256;  // work-group size of the streaming kernels (4 wave64)

// Per-step values that kernels read from device memory, so that a captured
// graph replays unchanged: the iteration counter, the body-force amplitude
// e(t) of lbm.cu:1032-1035 and the sliding-wall speed sra U, 0 before
// FLOWSTART (lbm.cu:755).
struct DevStep {
    int iter;
    int pad;
    double amp;
    double wall;
};

// Device status words, checked by the host after a synchronisation.
struct DevStatus {
    int bicg_flag;    // max BiCGSTAB flag over cells (1 breakdown, 2 ITRMAX)
    int vol_passes;   // max volume-correction passes over cells
    int pad0, pad1;
};

// Lattice tiles of the immersed-boundary kernels.
constexpr int VOL_SMALL = 3072;
constexpr int VOL_CH = 1024;  // elements / vertices per volume-correction chunk
constexpr int TILE_X = 16, TILE_Y = 8, TILE_Z = 8;
constexpr int ASM_CH = 256;  // vertices per chunk of the fused membrane assembly (k_assemble)

// Lattice geometry of the vertex binning: unpadded and padded sizes, walls
// beyond the y range or periodic y, spacing and tiles per axis.
struct BinGeom {
    int nx, ny, nz, nxp, nyp, nzp, walls, ntx, nty;
    double dx;
};

// First stencil node of one coordinate (ibm.cu:103-121: floor(xtr/dx) - 1 in
// the padded frame, xtr = X + (n_p - 1) dx / 2), unpadded.
__device__ __forceinline__ int bin_first(double x, int np, double dx) {
    const double xtr = x + (double)(np - 1) * dx / 2.0;
    const double q = dx == 1.0 ? xtr : xtr / dx;
    return (int)floor(q) - 1 - 2;
}

// The tile of a vertex's first stencil node: x and z wrapped once, y wrapped
// (periodic) or clamped (walls); positions far outside the box (or NaN) land
// in a valid tile, where their stencils meet no node and are dropped.
__device__ __forceinline__ int home_tile(const BinGeom &g, double x, double y, double z) {
    int xw = bin_first(x, g.nxp, g.dx), yh = bin_first(y, g.nyp, g.dx), zw = bin_first(z, g.nzp, g.dx);
    if (xw < 0) xw += g.nx; else if (xw >= g.nx) xw -= g.nx;
    if (zw < 0) zw += g.nz; else if (zw >= g.nz) zw -= g.nz;
    if (g.walls) yh = min(max(yh, 0), g.ny - 1);
    else if (yh < 0) yh += g.ny; else if (yh >= g.ny) yh -= g.ny;
    xw = min(max(xw, 0), g.nx - 1);
    zw = min(max(zw, 0), g.nz - 1);
    yh = min(max(yh, 0), g.ny - 1);
    return xw / TILE_X + g.ntx * (yh / TILE_Y + g.nty * (zw / TILE_Z));
}

// Slot for a lane in counters[key], lanes with equal keys aggregated into one
// atomic per wave.  All lanes of the wave must call it.
__device__ __forceinline__ int wave_aggregated_add(int *counters, int key, bool active) {
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(active);
    int result = 0;
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int lkey = __shfl(key, leader);
        const unsigned long long grp = __ballot(active && key == lkey) & pending;
        int base = 0;
        if (lane == leader) base = atomicAdd(&counters[lkey], __popcll(grp));
        base = __shfl(base, leader);
        if ((grp >> lane) & 1ull) result = base + __popcll(grp & ((1ull << lane) - 1ull));
        pending &= ~grp;
    }
    return result;
}

}  // namespace lbmibm

struct lbmibm_ctx {
    lbmibm_params p;
    int nx, ny, nz;            // fluid nodes (unpadded)
    long long N;               // nx*ny*nz
    int nxp, nyp, nzp;         // reference padded sizes (+4)
    long long np;
    int nc = 0, nv = 0, nt = 0;
    int prec;                  // lbmibm_precision
    size_t szS, szT;           // storage / arithmetic type sizes
    hipStream_t stream = nullptr;
    // second stream (default priority) for the large-cell solves, forked from
    // and joined back into `stream` with events; captured graphs keep the fork
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;    // membrane solve
    hipEvent_t ev_afork = nullptr;                       // on-chip assembly done (class 0 on `side`)
    hipEvent_t ev_vfork = nullptr, ev_vjoin = nullptr;  // volume correction classes
    // third stream (default priority): inside lbmibm_run, the volume
    // correction of one step and the binning of its positions overlap the
    // membrane force of the next, which reads fem->xd and not the corrected
    // cel->x (when BR = 0); joined before the next spreading or interpolation
    hipStream_t vstream = nullptr;
    hipEvent_t ev_pfork = nullptr, ev_pjoin = nullptr;
    bool vol_pending = false;
    std::string err;

    // lattice (device)
    void *f[2] = {nullptr, nullptr};  // SoA [19][N] post-collision populations, type S
    int cur = 0;                      // f[cur] holds the current state
    void *F = nullptr;                // [3][N] lattice IBM force (T)
    void *U = nullptr;                // [3][N] M-averaged velocity um,vm,wm (T)
    void *Ui = nullptr;               // [3][N] instantaneous velocity (store_macros)
    void *rho = nullptr;              // [N] instantaneous density (store_macros)
    void *rhom = nullptr;             // [N] accumulated density dm (store_macros)
    bool force_zero = true;           // lattice force known to be identically zero
    void *Hb = nullptr;               // [3][N] per-node body-force field (T), or none

    // adjoint state (allocated on first use): adjoints of the post-collision
    // populations, the velocity field, the lattice force, the vertex
    // positions and nodal forces; accumulated gradients with respect to the
    // body-force field and the per-cell force scales
    void *adj_f[2] = {nullptr, nullptr};  // [19][N] (T); adj_f[adj_cur] is current
    int adj_cur = 0;
    void *adj_u = nullptr;            // [3][N] (T)
    void *adj_F = nullptr;            // [3][N] (T)
    double *adj_x = nullptr;          // [3][Nv]
    void *adj_cf = nullptr;           // [3][Nv] (T)
    double *grad_H = nullptr;         // [3][N]
    double *grad_gs = nullptr;        // [nc]
    void *adj_lam = nullptr;          // [3Nv] (T) adjoint load of the membrane solve
    double *adj_dsv = nullptr;        // [Nv] adjoint of ds x force scale
    double *adj_corner = nullptr;     // [6Nv][4] per-(vertex, slot) position adjoints
    void *adj_scratch = nullptr;      // [3][Nv] (T) discarded solver epilogue output
    double *adj_cscale = nullptr;     // [nc] power-of-two scale of each cell's adjoint right-hand side
    double *adj_gpart = nullptr;      // [Nv] per-vertex terms of the force-scale gradient
    double *xs_out = nullptr;         // when set, the interpolation writes the new positions here (adjoint tape)
    unsigned *adj_xmax = nullptr;     // [1] max over vertices of sum_s |kb_s| (float bits), advection adjoint scale
    void *adj_rec = nullptr;          // [4 Nv] stage records of the advection adjoint

    // cells (device): positions, metric and bending reference in FP64; the
    // other membrane arrays in the arithmetic type T
    int *voff = nullptr, *eoff = nullptr, *vcell = nullptr;
    int *ele = nullptr, *ec = nullptr;
    int *nln = nullptr, *ln = nullptr, *lo8 = nullptr, *nle = nullptr, *le = nullptr;
    void *xs = nullptr;    // [3][Nv] SoA positions (cel->x,y,z), FP64
    void *xd = nullptr;    // [3Nv] AoS positions (fem->xd), FP64
    void *uvw = nullptr;   // [3][Nv] interpolated velocity (cel->u,v,w)
    void *cf = nullptr;    // [3][Nv] nodal force (cel->fx,fy,fz)
    void *c0 = nullptr, *s0 = nullptr, *nref = nullptr;  // FP64
    void *tt = nullptr, *rg = nullptr, *tp = nullptr;
    void *qv = nullptr, *qi = nullptr;                 // [3Nv]
    void *qp = nullptr;                                // [3Nv] the loads before qv (bicg_warm_start 2)
    // fused assembly: chunks (first vertex, vertices, first entry of
    // asm_elems, elements) and the chunk lists of all / large / on-chip cells
    int4 *asm_chunks = nullptr;
    int4 *asm_elems = nullptr;  // (element, its three vertices)
    int *asm_list[3] = {nullptr, nullptr, nullptr};
    int asm_n[3] = {0, 0, 0};
    unsigned char *eslot = nullptr;                    // [3Nt] slot of each corner in its vertex's list
    void *coef = nullptr;                              // [6][Nv] off-diagonal Jacobi-scaled coefficients
    unsigned short *nbr = nullptr;                     // [6][Nv] cell-local neighbour ids (self for unused)
    // BiCGSTAB cell classes: 0 on chip with one vertex per thread, 1 on chip
    // with three vertices per thread, 2 vectors in global memory
    int *bicg_cells[3] = {nullptr, nullptr, nullptr};
    int bicg_n[3] = {0, 0, 0}, bicg_maxv[3] = {0, 0, 0};
    int bicg_var[3] = {0, 0, 0};  // shape variant of the on-chip classes, form of class 2 (autotuned)
    int interp_form = 0;          // 0: tiles staged in LDS, 1: one thread per vertex, rows from memory (autotuned)
    int spread_form = 0;          // 0: source bins per tile, 1: home boxes + gather (autotuned)
    void *sbox = nullptr;         // [ntiles][3][SB_N] staging boxes of spread form 1
    int *sb_cand = nullptr;       // [(nx+ny+nz)*4] box entries per coordinate (spread form 1)
    int big_lds = 0;              // dynamic LDS of the class-2 forms that use LDS (the whole CU)
    int *onchip_cells = nullptr;  // cells of classes 0 and 1
    int n_onchip = 0;
    int cls_maxe[2] = {0, 0}, cls_maxv[2] = {0, 0};  // largest element / vertex count: on-chip, large
    std::string tune_report;
    // volume-correction work-group classes: up to VOL_SMALL vertices, larger
    int *vol_cells[2] = {nullptr, nullptr};
    int vol_n[2] = {0, 0};
    // volume-correction chunks: elements (first pass sums) and vertices
    // (centroid, rewrap) of one cell each, VOL_CH per chunk; per cell the
    // first element chunk, their count, the first vertex chunk, their count
    int4 *vol_echunks = nullptr, *vol_vchunks = nullptr, *vol_crange = nullptr;
    int vol_ne = 0, vol_nvc = 0;
    double2 *vol_epart = nullptr;  // [vol_ne] (V, A) of each element chunk
    double4 *vol_vpart = nullptr;  // [vol_nvc] centroid sums of each vertex chunk
    int *vol_need = nullptr;       // [nc] 1: the cell ran correction passes
    void *bv = nullptr, *rv = nullptr, *rh = nullptr, *pv = nullptr, *vv = nullptr, *tv = nullptr;
    void *gscale = nullptr;                            // [nc] (T)
    void *dsv = nullptr;                               // [Nv] ring area x force scale (T)
    void *corner = nullptr;                            // [6Nv] per-(vertex, slot) rhs terms and rg (T4)
    int *cslot = nullptr;                              // [3Nt] slot 6j+s of each element corner
    unsigned char *epair = nullptr;                    // [6Nv] slots sharing the edge to column m
    double *normal = nullptr;                          // [3Nt] FP64 element normals
    double *vol0 = nullptr, *vol = nullptr, *area = nullptr, *gx = nullptr;  // [nc], [3nc]
    int *bicg_iters = nullptr, *vol_passes = nullptr;  // [nc]
    int max_cell_v = 0;

    // immersed boundary: vertices binned by the lattice tile of their first
    // stencil node (rebuilt whenever the positions change)
    int ntx = 0, nty = 0, ntz = 0, ntiles = 0, nblk_bin = 0;
    short4 *node0 = nullptr;                 // [Nv] first stencil node (x,z wrapped, y+3), bin order
    void *srec = nullptr;                    // [Nv] stencil fractions and force (6 T), bin order
    int *home = nullptr, *home_list = nullptr;  // [Nv]
    int *tile_count = nullptr, *tile_off = nullptr, *tile_cur = nullptr;  // [ntiles]
    int *tile_cand = nullptr;                          // [ntiles] spreading candidates
    int *order_spread = nullptr, *order_interp = nullptr;  // [ntiles] heavy-first tile orders
    int *order_calls = nullptr;              // device counter of binnings (tile-order refresh)
    float *blk_fmax = nullptr;               // [nblk_bin] max |f| per binning block
    unsigned *absmax = nullptr;
    bool bins_valid = false;                 // bins match the current positions
    // tile counts, offsets, orders and home tiles match the current positions
    // (counted by the volume correction's centroid pass); the spreading then
    // only scatters the records
    bool counts_ready = false;

    // VOF indicator (vof_period > 0, allocated on first use): vf [N] FP64 in
    // vof[vof_cur] and its partner for the THINC/WLIC sweeps; front-tracking
    // work arrays on the padded lattice: fixed-point G [3][np], SOR iterate
    // and right-hand side [np], per-block partial sums, the SOR status word
    double *vof[2] = {nullptr, nullptr};
    int vof_cur = 0;
    unsigned long long *ft_g = nullptr;
    double *ft_x = nullptr, *ft_b = nullptr, *ft_part = nullptr;
    void *ft_ctl = nullptr;
    int ft_nparts = 0;
    // tube geometry (tube_radius > 0): per-column words (bit 0 inside the
    // disc, bit q the pull source of q is a wall) and the interpolation
    // coefficients of the cut links [col][19] (2r, -1/(2r) or -2 from the
    // wall fractions r)
    void *tube_col = nullptr, *tube_r = nullptr;  // tube_r stays null for the staircase
    std::vector<double> tube_rh;  // host copy of the fractions (-1 everywhere for the staircase)
    // sliding walls (wall_motion > 0): per wall-row node, 6 rho Cr u_w(iter) of
    // the step that last collided it, [2][nx*nz] (T), bottom row then top row;
    // the next pull adds w_q c_qz times it to the bounced links (lbm.cu:951-970)
    void *mw = nullptr;
    // cytoadhesion (after lbmibm_adhesion_setup): constants, receptor grid,
    // per-vertex knob flags [Nv], receptor lists [Nv*MAXN], springs
    // [Nv*NS] (receptor or -1, r_ij, round state), receptor flags and
    // claims [nrec], f_adh [3Nv] AoS FP64 and the spreading force cf + f_adh
    // [3][Nv] (T)
    bool adh_on = false;
    lbmibm_adhesion_params adh{};
    int adh_nxw = 0, adh_nzw = 0;
    double adh_dxw = 0.0, adh_dzw = 0.0;
    void *adh_knob = nullptr;
    int *adh_idw = nullptr, *adh_attach = nullptr, *adh_pos = nullptr, *adh_active = nullptr, *adh_prop = nullptr;
    int *adh_isatt = nullptr, *adh_nprop = nullptr;  // nprop: proposals per round
    unsigned long long *adh_claim = nullptr;          // per receptor: the highest bid key
    unsigned long long *adh_calls = nullptr;          // [1] calls of the Bell update so far
    int *adh_klist = nullptr, adh_nk = 0;  // the knob vertices
    double *adh_rw = nullptr, *adh_rij = nullptr, *adh_fadh = nullptr;
    double *adh_ring = nullptr;  // [nxw][2] ring receptor coordinates (tube)
    bool solve_cold = false;     // the checkpointed adjoint's forward and recompute start every solve at 0
    // per-kernel HIP-event timing of eager steps (lbmibm_time_kernels only)
    struct KtRec { const char *name; hipEvent_t a, b; };
    bool kt_on = false;
    std::vector<KtRec> kt_recs;
    void *adh_fsum = nullptr;
    lbmibm::DevStep *step = nullptr;
    lbmibm::DevStatus *status = nullptr;

    // whole-M-cycle graph
    hipGraphExec_t graph = nullptr;
    bool graph_ready = false;
    int graph_cur = 0;              // population buffer current when captured
    bool graph_force_zero = true;   // lattice-force variant when captured
};

namespace lbmibm {

// Launchers implemented per translation unit; each dispatches on ctx->prec.
lbmibm_status lattice_step(lbmibm_ctx *ctx);                 // lattice.hip
lbmibm_status lattice_set_state_from_fstar(lbmibm_ctx *ctx, const double *fstar_soa);  // lattice.hip
lbmibm_status lattice_get_fstar(lbmibm_ctx *ctx, double *fstar_soa);  // lattice.hip
lbmibm_status lattice_fill(lbmibm_ctx *ctx, void *ptr, size_t count, double value, bool storage_type);
lbmibm_status lattice_adjoint(lbmibm_ctx *ctx, const void *ga_out, void *ga_in);  // lattice.hip
lbmibm_status ibm_spread(lbmibm_ctx *ctx);                   // ibm.hip
lbmibm_status ibm_interp(lbmibm_ctx *ctx);                   // ibm.hip
lbmibm_status ibm_setup(lbmibm_ctx *ctx);                    // ibm.hip
lbmibm_status ibm_spread_adjoint(lbmibm_ctx *ctx);           // ibm.hip
lbmibm_status ibm_interp_adjoint(lbmibm_ctx *ctx);           // ibm.hip
lbmibm_status membrane_forces(lbmibm_ctx *ctx);              // membrane.hip
// have_q: q (qv) already holds the load at the current positions
lbmibm_status membrane_forces_adjoint(lbmibm_ctx *ctx, bool have_q);  // membrane.hip
int bicg_class(int prec, int nvert);                         // membrane.hip
lbmibm_status membrane_setup(lbmibm_ctx *ctx);               // membrane.hip
lbmibm_status membrane_autotune(lbmibm_ctx *ctx);            // membrane.hip
lbmibm_status ibm_autotune(lbmibm_ctx *ctx);                 // ibm.hip
lbmibm_status volume_run(lbmibm_ctx *ctx, int mode);         // volume.hip (0 init, 1 correct)
lbmibm_status vof_alloc(lbmibm_ctx *ctx);                    // vof.hip
lbmibm_status vof_front_tracking(lbmibm_ctx *ctx, int *iterations, int *attempts);  // vof.hip (host sync)
lbmibm_status vof_thinc(lbmibm_ctx *ctx);                    // vof.hip
lbmibm_status adhesion_setup(lbmibm_ctx *ctx, const lbmibm_adhesion_params *a, const unsigned char *is_knob);  // adhesion.hip
lbmibm_status adhesion_step(lbmibm_ctx *ctx);                // adhesion.hip
// the nodal force the spreading reads: cf, or cf + f_adh with cytoadhesion
inline const void *spread_force(const lbmibm_ctx *ctx) { return ctx->adh_on ? ctx->adh_fsum : ctx->cf; }
// both classes on the given stream (not the context's main stream)
lbmibm_status volume_run_on(lbmibm_ctx *ctx, int mode, hipStream_t stream);  // volume.hip
lbmibm_status ibm_count_begin(lbmibm_ctx *ctx, hipStream_t stream);  // ibm.hip: zero the tile counts
lbmibm_status ibm_count_end(lbmibm_ctx *ctx, hipStream_t stream);    // ibm.hip: tile offsets and orders
// the positions changed: the bins and the tile counts are stale
inline void invalidate_bins(lbmibm_ctx *ctx) {
    ctx->bins_valid = false;
    ctx->counts_ready = false;
}
lbmibm_status set_step(lbmibm_ctx *ctx, int iter, bool advance);  // api.hip
// api.hip helpers shared with adjoint.hip
lbmibm_status dev_alloc(lbmibm_ctx *ctx, void **p, size_t bytes);
lbmibm_status put_T(lbmibm_ctx *ctx, void *dst, const double *src, size_t n);  // host double -> device T
lbmibm_status get_T(lbmibm_ctx *ctx, double *dst, const void *src, size_t n);  // device T -> host double
long long padded_index(const lbmibm_ctx *ctx, int ix, int iy, int iz);
long long pull_source(const lbmibm_ctx *ctx, int ix, int iy, int iz, int q);  // -1 beyond a wall
bool tube_inside(const lbmibm_ctx *ctx, int ix, int iy);  // tube mask at fluid-frame column (ix, iy)
long long push_target(const lbmibm_ctx *ctx, int ix, int iy, int iz, int q);  // -1 beyond a wall
lbmibm_status sync_check(lbmibm_ctx *ctx);
lbmibm_status check_flags(lbmibm_ctx *ctx, int *vol_iters);
lbmibm_status reset_status(lbmibm_ctx *ctx);
lbmibm_status enqueue_step(lbmibm_ctx *ctx, int iter, bool device_advance, bool overlap_volume = false);
lbmibm_status join_volume(lbmibm_ctx *ctx);  // joins a volume correction left on vstream

inline lbmibm_status hip_fail(lbmibm_ctx *ctx, hipError_t e, const char *what) {
    ctx->err = std::string(what) + ": " + hipGetErrorString(e);
    return LBMIBM_ERR_HIP;
}

#define LBMIBM_HIP(ctx, call)                                           \
    do {                                                                \
        hipError_t e_ = (call);                                         \
        if (e_ != hipSuccess) return ::lbmibm::hip_fail(ctx, e_, #call); \
    } while (0)

#define LBMIBM_LAUNCHED(ctx, name)                                      \
    do {                                                                \
        hipError_t e_ = hipGetLastError();                              \
        if (e_ != hipSuccess) return ::lbmibm::hip_fail(ctx, e_, name); \
    } while (0)

// Timed launch: with the kernel timer on (eager steps of lbmibm_time_kernels,
// never under capture) an event pair brackets the launch on its stream.
inline hipEvent_t kt_begin(lbmibm_ctx *ctx, hipStream_t st) {
    if (!ctx->kt_on) return nullptr;
    hipEvent_t a = nullptr;
    if (hipEventCreate(&a) != hipSuccess) return nullptr;
    (void)hipEventRecord(a, st);
    return a;
}
inline void kt_end(lbmibm_ctx *ctx, hipStream_t st, const char *name, hipEvent_t a) {
    if (!a) return;
    hipEvent_t b = nullptr;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return;
    }
    (void)hipEventRecord(b, st);
    ctx->kt_recs.push_back({name, a, b});
}
#define LBMIBM_KT(ctx, st, name, ...)                                   \
    do {                                                                \
        hipEvent_t kt_a_ = ::lbmibm::kt_begin(ctx, st);                 \
        __VA_ARGS__;                                                    \
        ::lbmibm::kt_end(ctx, st, name, kt_a_);                         \
    } while (0)

// Work-group remapping so that each of the 8 XCDs (which receive work groups
// round-robin) walks one contiguous eighth of the index space: neighbouring
// rows read by the stencil then hit that XCD's own L2.
__device__ __forceinline__ int xcd_block(int nblocks) {
    const int per = (nblocks + 7) >> 3;
    const int b = blockIdx.x;
    return (b & 7) * per + (b >> 3);
}
inline int xcd_grid(int nblocks) { return ((nblocks + 7) >> 3) << 3; }

template <typename T> __device__ __forceinline__ T dev_min_image(T d, T L) {
    if (d > L * T(0.5)) d -= L;
    else if (d < -L * T(0.5)) d += L;
    return d;
}

// ---- FP64 reductions ------------------------------------------------------------
//
// Wave sums by DPP row shifts and row broadcasts (a fixed pattern, so the
// result does not depend on timing), then one value per wave through LDS and
// a sequential sum of the wave totals by every thread.  Buffers alternate
// between consecutive reductions, so each needs a single barrier: a buffer
// is rewritten only after the barrier of the reduction in between.

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    int lo = (int)(unsigned)(u & 0xffffffffull), hi = (int)(unsigned)(u >> 32);
    lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, 0xf, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over the 64 lanes, returned to every lane.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_move<0x111, 0xf>(v);  // row_shr:1
    v += dpp_move<0x112, 0xf>(v);  // row_shr:2
    v += dpp_move<0x114, 0xf>(v);  // row_shr:4
    v += dpp_move<0x118, 0xf>(v);  // row_shr:8  -> lane 15 of each row holds the row sum
    v += dpp_move<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_move<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3 -> lane 63 holds the sum
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)(u & 0xffffffffull), 63);
    const int hi = __builtin_amdgcn_readlane((int)(unsigned)(u >> 32), 63);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

struct WgReduce {
    double *sh;  // [2][4][16]
    int parity;
    template <int NV> __device__ __forceinline__ void sum(double (&v)[NV]) {
        double *buf = sh + parity * 64;
        parity ^= 1;
        const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
        for (int i = 0; i < NV; i++) v[i] = wave_sum(v[i]);
        if ((threadIdx.x & 63) == 0)
#pragma unroll
            for (int i = 0; i < NV; i++) buf[i * 16 + w] = v[i];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NV; i++) {
            double s = 0.0;
            for (int k = 0; k < nw; k++) s += buf[i * 16 + k];
            v[i] = s;
        }
    }
};


}  // namespace lbmibm
