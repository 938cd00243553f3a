// Skalak membrane forces on gfx950: stress() = calc_tension + calc_load
// (per-cell consistent-mass system, BiCGSTAB) + q2f + fourpoint_bend
// (stress.cu:29-659, matrix_solver.cu:225-377, dot_product.cu:192-223).
//
// The reference assembles, per cell and one cell at a time, an ELL matrix of
// 3V rows and runs a host-driven BiCGSTAB with a blocking device-to-host copy
// per inner product.  Here:
//   - k_elem: one thread per element (all cells in one launch) computes the
//     tension and each corner's right-hand-side terms and area;
//   - k_vert: one thread per vertex gathers its corners and writes the 6
//     Jacobi-scaled off-diagonal coefficients once (x, y and z rows share
//     them, stress.cu:164-171), the three scaled right-hand sides and the
//     ring area; the column ids are static and built at set-up;
//   - k_bicg_onchip / k_bicg_global: one 1024-thread work group per cell runs
//     the reference's BiCGSTAB iteration to convergence on the device (FP64
//     inner products reduced in a fixed order, device-side convergence and
//     breakdown flags); cells of up to 3072 vertices keep every vector and
//     their matrix rows in registers and LDS, larger ones use global memory;
//   - q2f is fused: k_vert also forms the ring area ds of each vertex
//     and the solver epilogue writes f = -q ds with the load q;
//   - k_bend as the reference.
// Vertex positions, the reference metric and all geometric arithmetic
// (edges, metric invariants, tension, matrix entries, ring areas, hinge
// angles) are FP64 in every precision mode: the strains are small
// differences of positions, and with C = 100 the area term turns FP32
// rounding of |x| ~ 10-100 positions into percent-level force errors.  The
// stored tension, matrix rows, right-hand sides, BiCGSTAB vectors and nodal
// forces use the arithmetic type T; inner products are FP64.
#include <algorithm>

#include "common.hpp"

namespace lbmibm {
namespace {

struct PeriodicBox {
    double lx, ly, lz;
    int py;
};

template <typename T> __device__ __forceinline__ void min_img3(T v[3], const PeriodicBox &b) {
    v[0] = dev_min_image<T>(v[0], (T)b.lx);
    v[2] = dev_min_image<T>(v[2], (T)b.lz);
    if (b.py) v[1] = dev_min_image<T>(v[1], (T)b.ly);
}

struct TensionArgs {
    double c, rgs, wtp;  // wtp = M / FILEOUTPUT
    int neo, fileoutput, m;
};

// Per-element part of stress(): calc_tension (stress.cu:29-121) and, for each
// of the element's three corners, its terms of the consistent-mass system of
// calc_load (stress.cu:182-272): the right-hand side
//   sum_{alpha,beta} rg t^{alpha beta} (e_alpha dN_beta + e_beta dN_alpha) / 4
// per component and the element's rg, from which the per-vertex pass builds
// the mass-matrix row (rg/12 diagonal, rg/24 per shared edge).
template <typename T> struct Corner;
template <> struct Corner<float> { using type = float4; };
template <> struct Corner<double> { using type = double4; };

// One element's calc_tension quantities in FP64 (stress.cu:63-99): edges with
// minimum image, the tension t, gab = rg rounded to T, and the invariants
// for tp.  The corner terms use t unrounded: near the reference state they
// are small sums of O(1) products, so a tension rounded to FP32 would cost
// up to 1e-3 of the nodal force.
template <typename T> struct ElemTension {
    double v1[3], v2[3];
    double t[4];
    double rg, gab, i1, i2;
    __device__ __forceinline__ ElemTension(const double *__restrict__ xd, int ea, int eb, int ec,
                                           const double *__restrict__ nr, const PeriodicBox &box,
                                           const TensionArgs &a) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            v1[d] = xd[3 * eb + d] - xd[3 * ea + d];
            v2[d] = xd[3 * ec + d] - xd[3 * ea + d];
        }
        min_img3<double>(v1, box);
        min_img3<double>(v2, box);
        const double n0 = nr[0], n1 = nr[1], n2 = nr[2], n3 = nr[3];
        // the metric g = (|v1|^2, v1.v2, |v2|^2) and its inverse; one
        // division each for 1/det g and 1/sqrt(det G / det g)
        const double g0 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
        const double g3 = v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2];
        const double g1 = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
        const double dv = g0 * g3 - g1 * g1;
        const double dn = n0 * n3 - n1 * n2;
        rg = sqrt(dv);
        const double idv = 1.0 / dv;
        const double nd0 = g3 * idv, nd1 = -g1 * idv, nd3 = g0 * idv;
        i1 = n0 * g0 + n1 * g1 + n2 * g1 + n3 * g3 - 2.0;
        i2 = dv * dn - 1.0;
        const double js = sqrt(i2 + 1.0), ijs = 1.0 / js;
        const double nn[4] = {n0, n1, n2, n3}, nd[4] = {nd0, nd1, nd1, nd3};
        const double ta = a.rgs * (i1 + 1.0) * ijs, tb = (a.c * i2 - 1.0) * js;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!a.neo) t[k] = ta * nn[k] + tb * nd[k];
            else t[k] = a.rgs * (nn[k] * ijs - nd[k] * js * (ijs * ijs) * (ijs * ijs));
        }
        gab = (double)(T)rg;
    }
    // calc_load terms of corner en: sum_{alpha,beta} rg t^{alpha beta}
    // (e_alpha dN_beta + e_beta dN_alpha) / 4, dN/dxi of the linear triangle
    // (-1,-1), (1,0), (0,1)
    __device__ __forceinline__ void corner(int en, double tb[3]) const {
        const double F[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double vd[2] = {v1[d], v2[d]};
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int al = k >> 1, be = k & 1;
                sum += gab * t[k] * (vd[al] * F[en][be] + vd[be] * F[en][al]) / 4.0;
            }
            tb[d] = sum;
        }
    }
    // principal-tension output accumulator (stress.cu:101-118)
    __device__ __forceinline__ void accumulate_tp(T *__restrict__ tp, int j, int iter, const TensionArgs &a) const {
        double t0 = 0.0, t1 = 0.0;
        if ((iter - 1) % a.fileoutput != 0) { t0 = (double)tp[2 * j]; t1 = (double)tp[2 * j + 1]; }
        const double judge = fabs((i1 + 2.0) * (i1 + 2.0) - 4.0 * (i2 + 1.0));
        const double root = sqrt(judge);
        const double l12 = 0.5 * (i1 + 2.0 + root), l22 = 0.5 * (i1 + 2.0 - root);
        const double l11 = sqrt(l12), l21 = sqrt(l22);
        tp[2 * j] = (T)(t0 + a.wtp * l11 / l21 * (l12 - 1.0 + a.c * l22 * i2));
        tp[2 * j + 1] = (T)(t1 + a.wtp * l21 / l11 * (l22 - 1.0 + a.c * l12 * i2));
    }
};

// The corner records go to the slot of the corner's vertex: cslot[3e+en] =
// s nv + j for vertex j = ele[3e+en] with e its s-th incident element
// (slot-major, so that the per-vertex pass reads slot s of a wave's vertices
// contiguously).
template <typename T>
__global__ __launch_bounds__(256) void k_elem(const double *__restrict__ xd, const int *__restrict__ ele,
                                              const double *__restrict__ nref, T *__restrict__ tt,
                                              T *__restrict__ rg, T *__restrict__ tp, const int *__restrict__ cslot,
                                              typename Corner<T>::type *__restrict__ corner, int ne,
                                              PeriodicBox box, TensionArgs a, const DevStep *__restrict__ st,
                                              const int *__restrict__ cells, const int *__restrict__ eoff) {
    // all elements (cells == nullptr, 1-D grid) or the elements of cell
    // cells[blockIdx.y] (2-D grid)
    int j = blockIdx.x * blockDim.x + threadIdx.x, jend = ne;
    if (cells) {
        const int c = cells[blockIdx.y];
        j += eoff[c];
        jend = eoff[c + 1];
    }
    if (j >= jend) return;
    const ElemTension<T> el(xd, ele[3 * j], ele[3 * j + 1], ele[3 * j + 2], nref + 4 * (size_t)j, box, a);
    if (tt) {  // outputs (td, rgd), not read by the step
        rg[j] = (T)el.rg;
#pragma unroll
        for (int k = 0; k < 4; k++) tt[4 * j + k] = (T)el.t[k];
    }
#pragma unroll
    for (int en = 0; en < 3; en++) {
        double tb[3];
        el.corner(en, tb);
        typename Corner<T>::type cv;
        cv.x = (T)tb[0]; cv.y = (T)tb[1]; cv.z = (T)tb[2]; cv.w = (T)el.gab;
        corner[cslot[3 * (size_t)j + en]] = cv;
    }
    if (tp) el.accumulate_tp(tp, j, st->iter, a);
}

// Per-vertex part of calc_load: the Jacobi-scaled row (stress.cu:263-268) and
// right-hand side of each vertex from the corner terms of its incident
// elements, summed in the reference's order (le list), plus the q2f ring area
// (stress.cu:415-438: a third of the incident triangle areas, rg/6 each)
// times the cell's force scale.  corner[s nv + j] holds the record of vertex
// j in its s-th incident element; epair[m nv + j] holds the two slots
// s_A < s_B whose elements share the edge to the m-th off-diagonal column (4
// bits each, 0xff for an unused column).  rg/12 and rg/6 are formed as 2 and
// 4 times rg/24 (exact: scaling by a power of two commutes with rounding),
// one division per slot.
template <typename T>
__global__ __launch_bounds__(256) void k_vert(const int *__restrict__ nln, const unsigned char *__restrict__ epair,
                                              const typename Corner<T>::type *__restrict__ corner,
                                              const int *__restrict__ vcell, const T *__restrict__ gscale,
                                              T *__restrict__ coef, T *__restrict__ bv, T *__restrict__ dsv,
                                              int nv, const int *__restrict__ cells, const int *__restrict__ voff) {
    // all vertices or those of cell cells[blockIdx.y], as k_elem
    int jj = blockIdx.x * blockDim.x + threadIdx.x, jend = nv;
    if (cells) {
        const int c = cells[blockIdx.y];
        jj += voff[c];
        jend = voff[c + 1];
    }
    if (jj >= jend) return;
    const int n = nln[jj];
    double r24[6], tb[3] = {0.0, 0.0, 0.0}, jac = 0.0, ds = 0.0;
#pragma unroll
    for (int s = 0; s < 6; s++) {
        r24[s] = 0.0;
        if (s < n) {
            const typename Corner<T>::type cv = corner[(size_t)s * nv + jj];
            r24[s] = (double)cv.w / 24.0;
            jac += 2.0 * r24[s];  // rg / 12
            ds += 4.0 * r24[s];   // rg / 2 / 3
            tb[0] += (double)cv.x;
            tb[1] += (double)cv.y;
            tb[2] += (double)cv.z;
        }
    }
    const double ijac = 1.0 / jac;
#pragma unroll
    for (int m = 0; m < 6; m++) {
        const unsigned ep = epair[(size_t)m * nv + jj];
        double ra = 0.0, rb = 0.0;
#pragma unroll
        for (int s = 0; s < 6; s++) {
            if ((int)(ep & 15u) == s) ra = r24[s];
            if ((int)(ep >> 4) == s) rb = r24[s];
        }
        coef[(size_t)m * nv + jj] = ep == 0xffu ? T(0) : (T)((ra + rb) * ijac);
    }
#pragma unroll
    for (int d = 0; d < 3; d++) bv[3 * (size_t)jj + d] = (T)(tb[d] * ijac);
    dsv[jj] = (T)(ds * (double)gscale[vcell[jj]]);
}

// Threads of a k_assemble work group.  Its LDS (ASM_CH x 6 corner slots,
// 48 KB) admits three work groups per CU; with twice as many threads as chunk
// vertices those hold 24 waves where 256 threads held 12, and a chunk's 550-
// 630 elements take two passes of the element loop instead of three.
constexpr int ASM_NT = 512;

// Fused assembly (forward stress): one work group per chunk of up to ASM_CH
// consecutive vertices of a cell computes, for every element incident to
// them (a static list built at set-up; elements on a chunk boundary are
// computed by each chunk they touch), the tension and the corner terms of
// the chunk's vertices into LDS, in FP64, and then writes the rows of
// k_vert from LDS.  The corner records never reach memory, and the
// right-hand sides are summed from unrounded corner terms: k_elem + k_vert
// round each term to T first, which in FP32 loses up to 1e-4 of a small
// nodal force to cancellation.  The tension and rg outputs and the tp
// accumulator are written by the chunk holding the element's first vertex.
template <typename T>
__global__ __launch_bounds__(ASM_NT) void k_assemble(const int *__restrict__ list, const int4 *__restrict__ chunks,
                                                  const int4 *__restrict__ elems, const double *__restrict__ xd,
                                                  const unsigned char *__restrict__ eslot,
                                                  const double *__restrict__ nref, T *__restrict__ tt,
                                                  T *__restrict__ rg, T *__restrict__ tp,
                                                  const int *__restrict__ nln, const unsigned char *__restrict__ epair,
                                                  const int *__restrict__ vcell, const T *__restrict__ gscale,
                                                  T *__restrict__ coef, T *__restrict__ bv, T *__restrict__ dsv,
                                                  int nv, PeriodicBox box, TensionArgs a,
                                                  const DevStep *__restrict__ st, int nchunks) {
    __shared__ double4 slot[ASM_CH * 6];  // [local vertex][incident element]
    // each XCD takes one contiguous eighth of the chunk list: neighbouring
    // chunks of a cell share vertices and boundary elements, whose lines
    // then go into one L2
    const int lb = xcd_block(nchunks);
    if (lb >= nchunks) return;
    const int4 ch = chunks[list[lb]];
    const int cv0 = ch.x, cnv = ch.y, e0 = ch.z, ne = ch.w;
    for (int i = threadIdx.x; i < ne; i += blockDim.x) {
        const int4 er = elems[e0 + i];
        const int e = er.x, ev[3] = {er.y, er.z, er.w};
        const ElemTension<T> el(xd, ev[0], ev[1], ev[2], nref + 4 * (size_t)e, box, a);
        if ((unsigned)(ev[0] - cv0) < (unsigned)cnv) {  // the element's owner chunk
            if (tt) {
                rg[e] = (T)el.rg;
#pragma unroll
                for (int k = 0; k < 4; k++) tt[4 * (size_t)e + k] = (T)el.t[k];
            }
            if (tp) el.accumulate_tp(tp, e, st->iter, a);
        }
#pragma unroll
        for (int en = 0; en < 3; en++) {
            const int jl = ev[en] - cv0;
            if ((unsigned)jl >= (unsigned)cnv) continue;
            double tb[3];
            el.corner(en, tb);
            slot[6 * jl + eslot[3 * (size_t)e + en]] = make_double4(tb[0], tb[1], tb[2], el.gab);
        }
    }
    __syncthreads();
    for (int jl = threadIdx.x; jl < cnv; jl += blockDim.x) {
        const int jj = cv0 + jl;
        const int n = nln[jj];
        double r24[6], tb[3] = {0.0, 0.0, 0.0}, jac = 0.0, ds = 0.0;
#pragma unroll
        for (int s = 0; s < 6; s++) {  // as k_vert
            r24[s] = 0.0;
            if (s < n) {
                const double4 cv = slot[6 * jl + s];
                r24[s] = cv.w / 24.0;
                jac += 2.0 * r24[s];
                ds += 4.0 * r24[s];
                tb[0] += cv.x;
                tb[1] += cv.y;
                tb[2] += cv.z;
            }
        }
        const double ijac = 1.0 / jac;
#pragma unroll
        for (int m = 0; m < 6; m++) {
            const unsigned ep = epair[(size_t)m * nv + jj];
            double ra = 0.0, rb = 0.0;
#pragma unroll
            for (int s = 0; s < 6; s++) {
                if ((int)(ep & 15u) == s) ra = r24[s];
                if ((int)(ep >> 4) == s) rb = r24[s];
            }
            coef[(size_t)m * nv + jj] = ep == 0xffu ? T(0) : (T)((ra + rb) * ijac);
        }
#pragma unroll
        for (int d = 0; d < 3; d++) bv[3 * (size_t)jj + d] = (T)(tb[d] * ijac);
        dsv[jj] = (T)(ds * (double)gscale[vcell[jj]]);
    }
}

// Solver epilogue per vertex: the load q (fem->q, substitute_load
// stress.cu:275-303), its output accumulator qi, and the nodal force of q2f
// f = -q ds (stress.cu:440-442) with ds the ring area times the cell's force
// scale.
template <typename T> struct SolveOut {
    T *q, *cf, *qi;
    T *qp;        // the loads of the solve before q (warm start 2), else null
    const T *ds;
    const DevStep *st;
    int nv, fo;   // fo = FILEOUTPUT period of qi (0: no accumulator)
    double wq;    // M / FILEOUTPUT
    // Start value of entry i (= 3 g + d) of the solve: 0 (cold), the last
    // load q (warm 1) or the linear extrapolation 2 q - qp from the last two
    // loads (warm 2), which then moves q into qp.  Called once per entry, by
    // the entry's owner thread, before the epilogue writes q.
    __device__ __forceinline__ T start(size_t i, int warm) const {
        if (warm == 0) return T(0);
        const T last = q[i];
        if (warm == 1) return last;
        const T prev = qp[i];
        qp[i] = last;
        return (T)(2.0 * (double)last - (double)prev);
    }
    __device__ __forceinline__ void put(size_t g, const T x[3]) const {
        const double d = (double)ds[g];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            q[3 * g + k] = x[k];
            cf[(size_t)k * nv + g] = (T)(-1.0 * (double)x[k] * d);
        }
        if (fo > 0) {
            const bool reset = (st->iter - 1) % fo == 0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double prev = reset ? 0.0 : (double)qi[3 * g + k];
                qi[3 * g + k] = (T)(prev + wq * (double)x[k]);
            }
        }
    }
};

// Three components of vertex j from an LDS vector stored as [V][4].
__device__ __forceinline__ void lds_load3(const float *a, int j, float o[3]) {
    const float4 q = *reinterpret_cast<const float4 *>(a + 4 * j);
    o[0] = q.x; o[1] = q.y; o[2] = q.z;
}
__device__ __forceinline__ void lds_load3(const double *a, int j, double o[3]) {
    const double2 q = *reinterpret_cast<const double2 *>(a + 4 * j);
    o[0] = q.x; o[1] = q.y; o[2] = a[4 * j + 2];
}
template <typename T> __device__ __forceinline__ void lds_store3(T *a, int j, const T v[3]) {
    a[4 * j] = v[0]; a[4 * j + 1] = v[1]; a[4 * j + 2] = v[2];
}

// Work-group shapes (vertices per thread K, threads NT) of the two on-chip
// classes.  Each class has candidate shapes of the same capacity K*NT; the
// one used is picked by timing them on the cells at set-up (autotune).
constexpr int BICG0_CAP = 1024, BICG1_CAP = 3072;
constexpr int BICG_NVAR = 3;
constexpr int BICG0_SHAPES[BICG_NVAR][2] = {{4, 256}, {2, 512}, {1, 1024}};
constexpr int BICG1_SHAPES[BICG_NVAR][2] = {{6, 512}, {3, 1024}, {4, 768}};

// On-chip BiCGSTAB for one cell of at most NT*KMAX vertices: every thread
// keeps the rows of its vertices (r, p, v, t and the 6 off-diagonal
// coefficients) in registers; p (then r) and b are mirrored in LDS ([V][4])
// for the neighbour gathers of the matrix-vector products, and x is kept in
// LDS ([V][4]).  matrix_solver.cu:
// 279-377 in the same update order; (r, r^) of the next iteration is fused
// with the residual reduction.  x starts at 0 (stress.cu:139), so r^ = b; with
// warm != 0 it starts at the previous load out.q and r^ = r0 = b - A x0, the
// stop test staying |r| / |b| < residual (a start that already passes it
// returns x0 after no iteration).
template <typename T, int KMAX, int NT>
__global__ __launch_bounds__(NT) void k_bicg_onchip(const int *__restrict__ cells, const int *__restrict__ voff,
                                                      const T *__restrict__ coef,
                                                      const unsigned short *__restrict__ nbr,
                                                      const T *__restrict__ bv, SolveOut<T> out, int nv,
                                                      int itrmax, double residual, int warm,
                                                      int *__restrict__ iters, DevStatus *__restrict__ status) {
    extern __shared__ double dyn[];
    __shared__ double sh[128];
    WgReduce red{sh, 0};
    const int c = cells[blockIdx.x];
    const int v0 = voff[c], V = voff[c + 1] - v0;
    // p for v = A p, then r for t = A r in the same array (the reduction
    // after v = A p ends the reads of p; the one after t = A r those of r);
    // x lives in LDS too, each vertex touched by its own thread only, which
    // keeps the largest shape within its registers
    T *Ps = (T *)dyn;              // [V][4] p, then r
    T *Bs = Ps + 4 * (size_t)V;    // [V][4] right-hand side = r^
    T *Xs = Bs + 4 * (size_t)V;    // [V][4] x
    const int tid = threadIdx.x;
    T r[KMAX][3], p[KMAX][3], v[KMAX][3], t[KMAX][3], cf[KMAX][6];
    unsigned nb[KMAX][3];  // two 16-bit column ids per register
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        const int j = tid + k * NT;
        const bool on = j < V;
        const size_t g = (size_t)v0 + (on ? j : 0);
#pragma unroll
        for (int m = 0; m < 6; m++) cf[k][m] = on ? coef[(size_t)m * nv + g] : T(0);
#pragma unroll
        for (int m = 0; m < 3; m++)
            nb[k][m] = on ? ((unsigned)nbr[(size_t)(2 * m) * nv + g] | ((unsigned)nbr[(size_t)(2 * m + 1) * nv + g] << 16))
                          : 0u;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const T bb = on ? bv[3 * g + d] : T(0);
            r[k][d] = bb; p[k][d] = bb; t[k][d] = T(0);
            v[k][d] = on ? out.start(3 * (size_t)g + d, warm) : T(0);  // x0 until the loop
            acc[0] += (double)bb * (double)bb;
        }
        if (on) {
            lds_store3(Ps, j, warm ? v[k] : p[k]);
            lds_store3(Bs, j, p[k]);
            lds_store3(Xs, j, v[k]);
        }
    }
    red.sum<1>(*(double(*)[1])acc);  // its barrier also orders the LDS writes above
    const double bb = acc[0];
    if (bb == 0.0) {  // exactly zero load: x = 0 solves the system
        const T zero[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            const int j = tid + k * NT;
            if (j < V) out.put((size_t)v0 + j, zero);
        }
        if (tid == 0) iters[c] = 0;
        return;
    }
    auto spmv = [&](const T *src, int k, const T own[3], T y[3]) {
        y[0] = own[0]; y[1] = own[1]; y[2] = own[2];
#pragma unroll
        for (int m = 0; m < 6; m++) {
            T o[3];
            lds_load3(src, (int)((nb[k][m >> 1] >> (16 * (m & 1))) & 0xffffu), o);
            y[0] += cf[k][m] * o[0]; y[1] += cf[k][m] * o[1]; y[2] += cf[k][m] * o[2];
        }
    };
    double rr = bb, rr_old = 0.0, alfa = 1.0, omeg = 1.0;
    int it = 0, flag = 0;
    bool conv = false;
    if (warm) {  // r = p = r^ = b - A x0
        acc[0] = 0.0;
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            const int j = tid + k * NT;
            if (j < V) {
                spmv(Ps, k, v[k], t[k]);
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    r[k][d] = r[k][d] - t[k][d];
                    p[k][d] = r[k][d];
                    acc[0] += (double)r[k][d] * (double)r[k][d];
                }
            }
#pragma unroll
            for (int d = 0; d < 3; d++) { v[k][d] = T(0); t[k][d] = T(0); }
        }
        red.sum<1>(*(double(*)[1])acc);  // its barrier ends the gathers of x0
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            const int j = tid + k * NT;
            if (j < V) { lds_store3(Ps, j, r[k]); lds_store3(Bs, j, r[k]); }
        }
        __syncthreads();
        rr = acc[0];
        if (sqrt(rr / bb) < residual) { conv = true; it = -1; }  // x0 passes: no iteration
    }
    if (!conv)
    for (it = 0; it < itrmax; it++) {
        if (rr == 0.0) { flag = 1; break; }
        if (it > 0) {
            const T beta = (T)((rr / rr_old) * (alfa / omeg)), om = (T)omeg;
#pragma unroll
            for (int k = 0; k < KMAX; k++) {
                const int j = tid + k * NT;
#pragma unroll
                for (int d = 0; d < 3; d++) p[k][d] = r[k][d] + beta * (p[k][d] - om * v[k][d]);
                if (j < V) lds_store3(Ps, j, p[k]);
            }
            __syncthreads();
        }
        // v = A p ; alfa = (r, r^) / (v, r^)
        acc[0] = 0.0;
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            const int j = tid + k * NT;
            if (j < V) {
                spmv(Ps, k, p[k], v[k]);
                T b3[3];
                lds_load3(Bs, j, b3);
#pragma unroll
                for (int d = 0; d < 3; d++) acc[0] += (double)v[k][d] * (double)b3[d];
            }
        }
        red.sum<1>(*(double(*)[1])acc);
        alfa = rr / acc[0];
        const T al = (T)alfa;
        // r = r - alfa v
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            const int j = tid + k * NT;
#pragma unroll
            for (int d = 0; d < 3; d++) r[k][d] = r[k][d] - al * v[k][d];
            if (j < V) lds_store3(Ps, j, r[k]);
        }
        __syncthreads();
        // t = A r ; omeg = (r, t) / (t, t)
        acc[0] = 0.0; acc[1] = 0.0;
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            const int j = tid + k * NT;
            if (j < V) {
                spmv(Ps, k, r[k], t[k]);
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    acc[0] += (double)r[k][d] * (double)t[k][d];
                    acc[1] += (double)t[k][d] * (double)t[k][d];
                }
            }
        }
        red.sum<2>(acc);
        omeg = acc[0] / acc[1];
        const T om = (T)omeg;
        // x = x + alfa p + omeg r ; r = r - omeg t
        acc[0] = 0.0; acc[1] = 0.0;
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            const int j = tid + k * NT;
            T b3[3] = {T(0), T(0), T(0)}, xv[3] = {T(0), T(0), T(0)};
            if (j < V) { lds_load3(Bs, j, b3); lds_load3(Xs, j, xv); }
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const T ri = r[k][d];
                xv[d] = xv[d] + al * p[k][d] + om * ri;
                const T rn = ri - om * t[k][d];
                r[k][d] = rn;
                if (j < V) {
                    acc[0] += (double)rn * (double)rn;
                    acc[1] += (double)rn * (double)b3[d];
                }
            }
            if (j < V) lds_store3(Xs, j, xv);
        }
        red.sum<2>(acc);
        rr_old = rr;
        rr = acc[1];
        if (sqrt(acc[0] / bb) < residual) { conv = true; break; }
    }
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        const int j = tid + k * NT;
        if (j < V) {
            T xv[3];
            lds_load3(Xs, j, xv);
            out.put((size_t)v0 + j, xv);
        }
    }
    if (tid == 0) {
        iters[c] = conv ? it + 1 : it;
        if (!conv && flag == 0) flag = 2;
        if (flag) atomicMax(&status->bicg_flag, flag);
    }
}

// Large cells, second form: p lives only in LDS (SoA [3][V]), so the first
// matrix-vector product gathers from LDS and p never travels to memory; r, v,
// t, x in global memory.  Every phase walks the thread's vertices in batches
// of GB with all loads of a batch issued before use.  For cells whose p fits
// LDS (V <= BIG_LDS_V); the autotuner times it against k_bicg_global.  With
// warm != 0 the solve starts at x0 = out.q and r^ = r0 = b - A x0 replaces b
// in bv (each vertex's entries are read by their own thread only).
constexpr int BIG_LDS_V = 13568;  // 3 * 4 B * V within 160 KB, FP32

template <typename T, int GB>
__global__ __launch_bounds__(1024) void k_bicg_big(const int *__restrict__ cells, const int *__restrict__ voff,
                                                   const T *__restrict__ coef,
                                                   const unsigned short *__restrict__ nbr,
                                                   T *__restrict__ bv, T *__restrict__ x, T *__restrict__ r,
                                                   T *__restrict__ v, T *__restrict__ t, SolveOut<T> out, int nv,
                                                   int itrmax, double residual, int warm, int *__restrict__ iters,
                                                   DevStatus *__restrict__ status) {
    extern __shared__ double dyn[];
    __shared__ double sh[128];
    WgReduce red{sh, 0};
    const int c = cells[blockIdx.x];
    const int v0 = voff[c], V = voff[c + 1] - v0;
    T *P = (T *)dyn;  // [3][V]
    const int tid = threadIdx.x;
    constexpr int NT = 1024;
    // y = A in, with in read by get(local vertex, component)
    auto spmv_batch = [&](int base, auto get, T y[GB][3]) {
        T cm[GB][6];
        int col[GB][6];
#pragma unroll
        for (int i = 0; i < GB; i++) {
            const int j = min(base + i * NT, V - 1);
            const size_t g = (size_t)v0 + j;
#pragma unroll
            for (int m = 0; m < 6; m++) {
                cm[i][m] = coef[(size_t)m * nv + g];
                col[i][m] = nbr[(size_t)m * nv + g];
            }
#pragma unroll
            for (int d = 0; d < 3; d++) y[i][d] = get(j, d);
        }
#pragma unroll
        for (int i = 0; i < GB; i++)
#pragma unroll
            for (int m = 0; m < 6; m++)
#pragma unroll
                for (int d = 0; d < 3; d++) y[i][d] += cm[i][m] * get(col[i][m], d);
    };
    auto getP = [&](int j, int d) { return P[(size_t)d * V + j]; };
    auto getR = [&](int j, int d) { return r[3 * ((size_t)v0 + j) + d]; };
    double acc[2] = {0.0, 0.0};
    for (int j = tid; j < V; j += NT) {
        const size_t g = (size_t)v0 + j;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const T b = bv[3 * g + d], x0 = out.start(3 * (size_t)g + d, warm);
            x[3 * g + d] = x0; r[3 * g + d] = b; P[(size_t)d * V + j] = warm ? x0 : b;
            acc[0] += (double)b * (double)b;
        }
    }
    red.sum<1>(*(double(*)[1])acc);  // its barrier also orders the writes above
    const double bb = acc[0];
    if (bb == 0.0) {
        const T zero[3] = {T(0), T(0), T(0)};
        for (int j = tid; j < V; j += NT) out.put((size_t)v0 + j, zero);
        if (tid == 0) iters[c] = 0;
        return;
    }
    double rr = bb, rr_old = 0.0, alfa = 1.0, omeg = 1.0;
    int it = 0, flag = 0;
    bool conv = false;
    if (warm) {  // r = r^ = b - A x0 (x0 gathered from P), then p = r in P
        acc[0] = 0.0;
        for (int base = tid; base < V; base += GB * NT) {
            T y[GB][3];
            spmv_batch(base, getP, y);
#pragma unroll
            for (int i = 0; i < GB; i++) {
                const int j = base + i * NT;
                if (j < V) {
                    const size_t g = (size_t)v0 + j;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const T rn = r[3 * g + d] - y[i][d];
                        r[3 * g + d] = rn;
                        bv[3 * g + d] = rn;
                        acc[0] += (double)rn * (double)rn;
                    }
                }
            }
        }
        red.sum<1>(*(double(*)[1])acc);  // its barrier ends the gathers of x0
        for (int j = tid; j < V; j += NT)
#pragma unroll
            for (int d = 0; d < 3; d++) P[(size_t)d * V + j] = r[3 * ((size_t)v0 + j) + d];
        __syncthreads();
        rr = acc[0];
        if (sqrt(rr / bb) < residual) { conv = true; it = -1; }  // x0 passes: no iteration
    }
    if (!conv)
    for (it = 0; it < itrmax; it++) {
        if (rr == 0.0) { flag = 1; break; }
        if (it > 0) {  // p = r + beta (p - omeg v)
            const double beta = (rr / rr_old) * (alfa / omeg);
            for (int base = tid; base < V; base += GB * NT) {
                T rv_[GB][3], vv_[GB][3];
#pragma unroll
                for (int i = 0; i < GB; i++) {
                    const size_t g = (size_t)v0 + min(base + i * NT, V - 1);
#pragma unroll
                    for (int d = 0; d < 3; d++) { rv_[i][d] = r[3 * g + d]; vv_[i][d] = v[3 * g + d]; }
                }
#pragma unroll
                for (int i = 0; i < GB; i++) {
                    const int j = base + i * NT;
                    if (j < V)
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            T &pp = P[(size_t)d * V + j];
                            pp = (T)((double)rv_[i][d] + beta * ((double)pp - omeg * (double)vv_[i][d]));
                        }
                }
            }
            __syncthreads();
        }
        // v = A p ; alfa = (r, r^) / (v, r^)
        acc[0] = 0.0;
        for (int base = tid; base < V; base += GB * NT) {
            T y[GB][3];
            spmv_batch(base, getP, y);
#pragma unroll
            for (int i = 0; i < GB; i++) {
                const int j = base + i * NT;
                if (j < V) {
                    const size_t g = (size_t)v0 + j;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        v[3 * g + d] = y[i][d];
                        acc[0] += (double)y[i][d] * (double)bv[3 * g + d];
                    }
                }
            }
        }
        red.sum<1>(*(double(*)[1])acc);
        alfa = rr / acc[0];
        for (int j = tid; j < V; j += NT) {
            const size_t g = (size_t)v0 + j;
#pragma unroll
            for (int d = 0; d < 3; d++) r[3 * g + d] = (T)((double)r[3 * g + d] - alfa * (double)v[3 * g + d]);
        }
        __syncthreads();
        // t = A r ; omeg = (r, t) / (t, t)
        acc[0] = 0.0; acc[1] = 0.0;
        for (int base = tid; base < V; base += GB * NT) {
            T y[GB][3];
            spmv_batch(base, getR, y);
#pragma unroll
            for (int i = 0; i < GB; i++) {
                const int j = base + i * NT;
                if (j < V) {
                    const size_t g = (size_t)v0 + j;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        t[3 * g + d] = y[i][d];
                        acc[0] += (double)r[3 * g + d] * (double)y[i][d];
                        acc[1] += (double)y[i][d] * (double)y[i][d];
                    }
                }
            }
        }
        red.sum<2>(acc);
        omeg = acc[0] / acc[1];
        // x = x + alfa p + omeg r ; r = r - omeg t
        acc[0] = 0.0; acc[1] = 0.0;
        for (int j = tid; j < V; j += NT) {
            const size_t g = (size_t)v0 + j;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double ri = (double)r[3 * g + d];
                x[3 * g + d] = (T)((double)x[3 * g + d] + alfa * (double)P[(size_t)d * V + j] + omeg * ri);
                const T rn = (T)(ri - omeg * (double)t[3 * g + d]);
                r[3 * g + d] = rn;
                acc[0] += (double)rn * (double)rn;
                acc[1] += (double)rn * (double)bv[3 * g + d];
            }
        }
        red.sum<2>(acc);
        rr_old = rr;
        rr = acc[1];
        if (sqrt(acc[0] / bb) < residual) { conv = true; break; }
    }
    __syncthreads();
    for (int j = tid; j < V; j += NT) {
        const size_t g = (size_t)v0 + j;
        const T xv[3] = {x[3 * g], x[3 * g + 1], x[3 * g + 2]};
        out.put(g, xv);
    }
    if (tid == 0) {
        iters[c] = conv ? it + 1 : it;
        if (!conv && flag == 0) flag = 2;
        if (flag) atomicMax(&status->bicg_flag, flag);
    }
}

// Large cells, third form (FP32): every vertex belongs to one thread for the
// whole solve (j = tid + k*1024, k < K), so r (later t, which reuses its
// registers) and v stay in registers.  One LDS vector [V][3] holds p for
// v = A p and then s for t = A s, so both products gather from LDS.  p and x
// are the only vectors in memory, each touched by its owner thread only, so
// they need no barriers.  Per vertex and iteration that is 156 B of memory
// traffic (two matrix reads, b twice, p three times, x twice) against about
// 270 B for the p-in-LDS form.  Same arithmetic and update order as the
// other forms (matrix_solver.cu:279-377).  With warm != 0 the solve starts at
// x0 = out.q and r^ = r0 = b - A x0 replaces b in bv (read by its owner
// thread only).
// A copy of v that the compiler cannot see through: addresses derived from it
// are recomputed where they are used instead of being hoisted out of the
// iteration loop and held in registers for the whole solve.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

template <int K>
__global__ __launch_bounds__(1024) void k_bicg_res(const int *__restrict__ cells, const int *__restrict__ voff,
                                                   const float *__restrict__ coef,
                                                   const unsigned short *__restrict__ nbr,
                                                   float *__restrict__ bv, float *__restrict__ x,
                                                   float *__restrict__ pg, SolveOut<float> out, int nv, int itrmax,
                                                   double residual, int warm, int *__restrict__ iters,
                                                   DevStatus *__restrict__ status) {
    extern __shared__ double dyn[];
    __shared__ double sh[128];
    WgReduce red{sh, 0};
    constexpr int NT = 1024;
    const int c = cells[blockIdx.x];
    const int v0 = voff[c], V = voff[c + 1] - v0;
    float *L = (float *)dyn;  // [V][3]
    float rt[K][3], vr[K][3];  // r (then t), v
    // y = A L at local vertex j (global index g); 32-bit offsets from the
    // uniform bases
    auto spmv = [&](int j, unsigned g, float y[3]) {
        float cm[6];
        int col[6];
#pragma unroll
        for (int m = 0; m < 6; m++) {
            cm[m] = coef[(unsigned)m * (unsigned)nv + g];
            col[m] = nbr[(unsigned)m * (unsigned)nv + g];
        }
#pragma unroll
        for (int d = 0; d < 3; d++) y[d] = L[3 * j + d];
#pragma unroll
        for (int m = 0; m < 6; m++)
#pragma unroll
            for (int d = 0; d < 3; d++) y[d] += cm[m] * L[3 * col[m] + d];
    };
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = threadIdx.x + k * NT;
#pragma unroll
        for (int d = 0; d < 3; d++) { rt[k][d] = 0.0f; vr[k][d] = 0.0f; }
        if (j < V) {
            const unsigned g = (unsigned)(v0 + j);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const float b = bv[3 * g + d], x0 = out.start(3 * (size_t)g + d, warm);
                x[3 * g + d] = x0;
                if (warm) L[3 * j + d] = x0;
                rt[k][d] = b;
                acc[0] += (double)b * (double)b;
            }
        }
    }
    red.sum<1>(*(double(*)[1])acc);  // its barrier also orders the writes of x0 to L
    const double bb = acc[0];
    if (bb == 0.0) {
        const float zero[3] = {0.0f, 0.0f, 0.0f};
        for (int j = threadIdx.x; j < V; j += NT) out.put((size_t)v0 + j, zero);
        if (threadIdx.x == 0) iters[c] = 0;
        return;
    }
    double rr = bb, rr_old = 0.0, alfa = 1.0, omeg = 1.0;
    int it = 0, flag = 0;
    bool conv = false;
    if (warm) {  // r = r^ = b - A x0
        acc[0] = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int j = threadIdx.x + k * NT;
            if (j < V) {
                const unsigned g = (unsigned)(v0 + j);
                spmv(j, g, vr[k]);
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    rt[k][d] = rt[k][d] - vr[k][d];
                    bv[3 * g + d] = rt[k][d];
                    acc[0] += (double)rt[k][d] * (double)rt[k][d];
                    vr[k][d] = 0.0f;
                }
            }
        }
        red.sum<1>(*(double(*)[1])acc);  // its barrier ends the gathers of x0 before p is written
        rr = acc[0];
        if (sqrt(rr / bb) < residual) { conv = true; it = -1; }  // x0 passes: no iteration
    }
    if (!conv)
    for (it = 0; it < itrmax; it++) {
        if (rr == 0.0) { flag = 1; break; }
        // p = r (first pass) or r + beta (p - omeg v), into LDS and memory
        const double beta = it > 0 ? (rr / rr_old) * (alfa / omeg) : 0.0;
        {
            const int t0 = opaque(threadIdx.x), b0 = opaque(v0);
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int j = t0 + k * NT;
                if (j < V) {
                    const unsigned g = (unsigned)(b0 + j);
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        float pn = rt[k][d];
                        if (it > 0)
                            pn = (float)((double)rt[k][d] + beta * ((double)pg[3 * g + d] - omeg * (double)vr[k][d]));
                        pg[3 * g + d] = pn;
                        L[3 * j + d] = pn;
                    }
                }
            }
        }
        __syncthreads();
        // v = A p ; alfa = (r, r^) / (v, r^)
        acc[0] = 0.0;
        {
            const int t0 = opaque(threadIdx.x), b0 = opaque(v0);
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int j = t0 + k * NT;
                if (j < V) {
                    const unsigned g = (unsigned)(b0 + j);
                    spmv(j, g, vr[k]);
#pragma unroll
                    for (int d = 0; d < 3; d++) acc[0] += (double)vr[k][d] * (double)bv[3 * g + d];
                }
            }
        }
        red.sum<1>(*(double(*)[1])acc);  // its barrier ends the reads of p
        alfa = rr / acc[0];
        // s = r - alfa v into LDS
        {
            const int t0 = opaque(threadIdx.x);
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int j = t0 + k * NT;
                if (j < V)
#pragma unroll
                    for (int d = 0; d < 3; d++) L[3 * j + d] = (float)((double)rt[k][d] - alfa * (double)vr[k][d]);
            }
        }
        __syncthreads();
        // t = A s ; omeg = (s, t) / (t, t)
        acc[0] = 0.0; acc[1] = 0.0;
        {
            const int t0 = opaque(threadIdx.x), b0 = opaque(v0);
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int j = t0 + k * NT;
                if (j < V) {
                    spmv(j, (unsigned)(b0 + j), rt[k]);
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        acc[0] += (double)L[3 * j + d] * (double)rt[k][d];
                        acc[1] += (double)rt[k][d] * (double)rt[k][d];
                    }
                }
            }
        }
        red.sum<2>(acc);
        omeg = acc[0] / acc[1];
        // x = x + alfa p + omeg s ; r = s - omeg t
        acc[0] = 0.0; acc[1] = 0.0;
        {
            const int t0 = opaque(threadIdx.x), b0 = opaque(v0);
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int j = t0 + k * NT;
                if (j < V) {
                    const unsigned g = (unsigned)(b0 + j);
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const double si = (double)L[3 * j + d];
                        x[3 * g + d] = (float)((double)x[3 * g + d] + alfa * (double)pg[3 * g + d] + omeg * si);
                        const float rn = (float)(si - omeg * (double)rt[k][d]);
                        rt[k][d] = rn;
                        acc[0] += (double)rn * (double)rn;
                        acc[1] += (double)rn * (double)bv[3 * g + d];
                    }
                }
            }
        }
        red.sum<2>(acc);  // its barrier ends the reads of s before p is written
        rr_old = rr;
        rr = acc[1];
        if (sqrt(acc[0] / bb) < residual) { conv = true; break; }
    }
    for (int j = threadIdx.x; j < V; j += NT) {
        const size_t g = (size_t)v0 + j;
        const float xv[3] = {x[3 * g], x[3 * g + 1], x[3 * g + 2]};
        out.put(g, xv);
    }
    if (threadIdx.x == 0) {
        iters[c] = conv ? it + 1 : it;
        if (!conv && flag == 0) flag = 2;
        if (flag) atomicMax(&status->bicg_flag, flag);
    }
}

// Vertices per thread of the register-resident form (V <= K * 1024).
constexpr int BICG_RES_K = 11;

// Same iteration for cells of any size with the vectors in global memory.
// Each thread walks its vertices in batches of GB so that the loads of a
// batch are in flight together.  With warm != 0 the solve starts at
// x0 = out.q and r^ = r0 = b - A x0 replaces b in bv.
template <typename T>
__global__ __launch_bounds__(1024) void k_bicg_global(const int *__restrict__ cells, const int *__restrict__ voff,
                                                      const T *__restrict__ coef,
                                                      const unsigned short *__restrict__ nbr,
                                                      T *__restrict__ bv, T *__restrict__ x,
                                                      T *__restrict__ r, T *__restrict__ p, T *__restrict__ v,
                                                      T *__restrict__ t, SolveOut<T> out, int nv, int itrmax,
                                                      double residual, int warm, int *__restrict__ iters,
                                                      DevStatus *__restrict__ status) {
    __shared__ double sh[128];
    WgReduce red{sh, 0};
    const int c = cells[blockIdx.x];
    const int v0 = voff[c], v1 = voff[c + 1];
    const int tid = threadIdx.x, nt = blockDim.x;
    constexpr int GB = sizeof(T) == 4 ? 4 : 2;
    // y = A in for the vertices base + i*nt, i < GB (those below v1)
    auto apply = [&](const T *__restrict__ in, int base, T y[GB][3]) {
        int col[GB][6];
        T cm[GB][6];
#pragma unroll
        for (int i = 0; i < GB; i++) {
            const int vv = min(base + i * nt, v1 - 1);
#pragma unroll
            for (int m = 0; m < 6; m++) {
                cm[i][m] = coef[(size_t)m * nv + vv];
                col[i][m] = v0 + nbr[(size_t)m * nv + vv];
            }
#pragma unroll
            for (int d = 0; d < 3; d++) y[i][d] = in[3 * (size_t)vv + d];
        }
#pragma unroll
        for (int i = 0; i < GB; i++)
#pragma unroll
            for (int m = 0; m < 6; m++) {
                const T *o = in + 3 * (size_t)col[i][m];
                y[i][0] += cm[i][m] * o[0]; y[i][1] += cm[i][m] * o[1]; y[i][2] += cm[i][m] * o[2];
            }
    };
    double acc[2] = {0.0, 0.0};
    for (int vv = v0 + tid; vv < v1; vv += nt)
        for (int d = 0; d < 3; d++) {
            const size_t k = 3 * (size_t)vv + d;
            const T b = bv[k];
            x[k] = out.start(k, warm); r[k] = b; p[k] = b;
            acc[0] += (double)b * (double)b;
        }
    red.sum<1>(*(double(*)[1])acc);  // its barrier also orders the writes above
    const double bb = acc[0];
    if (bb == 0.0) {
        for (int vv = v0 + tid; vv < v1; vv += nt) {
            const T zero[3] = {T(0), T(0), T(0)};
            out.put(vv, zero);
        }
        if (tid == 0) iters[c] = 0;
        return;
    }
    double rr = bb, rr_old = 0.0, alfa = 1.0, omeg = 1.0;
    int it = 0, flag = 0;
    bool conv = false;
    if (warm) {  // r = p = r^ = b - A x0
        acc[0] = 0.0;
        for (int base = v0 + tid; base < v1; base += GB * nt) {
            T y[GB][3];
            apply(x, base, y);
#pragma unroll
            for (int i = 0; i < GB; i++) {
                const int vv = base + i * nt;
                if (vv < v1)
                    for (int d = 0; d < 3; d++) {
                        const size_t k = 3 * (size_t)vv + d;
                        const T rn = r[k] - y[i][d];
                        r[k] = rn; p[k] = rn; bv[k] = rn;
                        acc[0] += (double)rn * (double)rn;
                    }
            }
        }
        red.sum<1>(*(double(*)[1])acc);  // its barrier orders the writes of p before its gathers
        rr = acc[0];
        if (sqrt(rr / bb) < residual) { conv = true; it = -1; }  // x0 passes: no iteration
    }
    if (!conv)
    for (it = 0; it < itrmax; it++) {
        if (rr == 0.0) { flag = 1; break; }
        if (it > 0) {
            const double beta = (rr / rr_old) * (alfa / omeg);
#pragma unroll 4
            for (int vv = v0 + tid; vv < v1; vv += nt)
                for (int d = 0; d < 3; d++) {
                    const size_t k = 3 * (size_t)vv + d;
                    p[k] = (T)((double)r[k] + beta * ((double)p[k] - omeg * (double)v[k]));
                }
            __syncthreads();
        }
        acc[0] = 0.0;
        for (int base = v0 + tid; base < v1; base += GB * nt) {
            T y[GB][3];
            apply(p, base, y);
#pragma unroll
            for (int i = 0; i < GB; i++) {
                const int vv = base + i * nt;
                if (vv < v1)
                    for (int d = 0; d < 3; d++) {
                        const size_t k = 3 * (size_t)vv + d;
                        v[k] = y[i][d];
                        acc[0] += (double)y[i][d] * (double)bv[k];
                    }
            }
        }
        red.sum<1>(*(double(*)[1])acc);
        alfa = rr / acc[0];
#pragma unroll 4
        for (int vv = v0 + tid; vv < v1; vv += nt)
            for (int d = 0; d < 3; d++) {
                const size_t k = 3 * (size_t)vv + d;
                r[k] = (T)((double)r[k] - alfa * (double)v[k]);
            }
        __syncthreads();
        acc[0] = 0.0; acc[1] = 0.0;
        for (int base = v0 + tid; base < v1; base += GB * nt) {
            T y[GB][3];
            apply(r, base, y);
#pragma unroll
            for (int i = 0; i < GB; i++) {
                const int vv = base + i * nt;
                if (vv < v1)
                    for (int d = 0; d < 3; d++) {
                        const size_t k = 3 * (size_t)vv + d;
                        t[k] = y[i][d];
                        acc[0] += (double)r[k] * (double)y[i][d];
                        acc[1] += (double)y[i][d] * (double)y[i][d];
                    }
            }
        }
        red.sum<2>(acc);
        omeg = acc[0] / acc[1];
        acc[0] = 0.0; acc[1] = 0.0;
#pragma unroll 4
        for (int vv = v0 + tid; vv < v1; vv += nt)
            for (int d = 0; d < 3; d++) {
                const size_t k = 3 * (size_t)vv + d;
                const double ri = (double)r[k];
                x[k] = (T)((double)x[k] + alfa * (double)p[k] + omeg * ri);
                const T rn = (T)(ri - omeg * (double)t[k]);
                r[k] = rn;
                acc[0] += (double)rn * (double)rn;
                acc[1] += (double)rn * (double)bv[k];
            }
        red.sum<2>(acc);
        rr_old = rr;
        rr = acc[1];
        if (sqrt(acc[0] / bb) < residual) { conv = true; break; }
    }
    __syncthreads();  // every thread's last x update is in memory
    for (int vv = v0 + tid; vv < v1; vv += nt) {
        const T xv[3] = {x[3 * (size_t)vv], x[3 * (size_t)vv + 1], x[3 * (size_t)vv + 2]};
        out.put(vv, xv);
    }
    if (tid == 0) {
        iters[c] = conv ? it + 1 : it;
        if (!conv && flag == 0) flag = 2;
        if (flag) atomicMax(&status->bicg_flag, flag);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bend(const double *__restrict__ xs, const int *__restrict__ ele,
                                              const int *__restrict__ ec, const double *__restrict__ c0,
                                              const double *__restrict__ s0, T *__restrict__ cf, int nv, int ne,
                                              double br, PeriodicBox box) {
    const int e1 = blockIdx.x * blockDim.x + threadIdx.x;
    if (e1 >= ne) return;
    const int i1 = ele[3 * e1], i2 = ele[3 * e1 + 1], i3 = ele[3 * e1 + 2];
    for (int k = 0; k < 3; k++) {
        const int e2 = ec[3 * e1 + k];
        const int j1 = ele[3 * e2], j2 = ele[3 * e2 + 1], j3 = ele[3 * e2 + 2];
        int k1, k2, k3, k4;
        // hinge identification (stress.cu:497-505)
        if      ((i1==j1 && i2==j2)||(i1==j2 && i2==j1)) { k1 = i3; k2 = i1; k3 = i2; k4 = j3; }
        else if ((i1==j2 && i2==j3)||(i1==j3 && i2==j2)) { k1 = i3; k2 = i1; k3 = i2; k4 = j1; }
        else if ((i1==j3 && i2==j1)||(i1==j1 && i2==j3)) { k1 = i3; k2 = i1; k3 = i2; k4 = j2; }
        else if ((i2==j1 && i3==j2)||(i2==j2 && i3==j1)) { k1 = i1; k2 = i2; k3 = i3; k4 = j3; }
        else if ((i2==j2 && i3==j3)||(i2==j3 && i3==j2)) { k1 = i1; k2 = i2; k3 = i3; k4 = j1; }
        else if ((i2==j3 && i3==j1)||(i2==j1 && i3==j3)) { k1 = i1; k2 = i2; k3 = i3; k4 = j2; }
        else if ((i3==j1 && i1==j2)||(i3==j2 && i1==j1)) { k1 = i2; k2 = i3; k3 = i1; k4 = j3; }
        else if ((i3==j2 && i1==j3)||(i3==j3 && i1==j2)) { k1 = i2; k2 = i3; k3 = i1; k4 = j1; }
        else if ((i3==j3 && i1==j1)||(i3==j1 && i1==j3)) { k1 = i2; k2 = i3; k3 = i1; k4 = j2; }
        else continue;
        double a1[3], a2[3], a3[3], a4[3];
        const int ks[4] = {k1, k2, k3, k4};
        double *as[4] = {a1, a2, a3, a4};
        for (int m = 0; m < 4; m++)
            for (int d = 0; d < 3; d++) as[m][d] = xs[(size_t)d * nv + ks[m]];
        // wrap the hinge vertices next to a1 (stress.cu:513-518)
        const double Ls[3] = {box.lx, box.ly, box.lz};
        for (int m = 1; m < 4; m++)
            for (int d = 0; d < 3; d++) {
                if (d == 1 && !box.py) continue;
                if (as[m][d] - a1[d] > Ls[d] * 0.5) as[m][d] -= Ls[d];
                else if (as[m][d] - a1[d] < -Ls[d] * 0.5) as[m][d] += Ls[d];
            }
        double a21[3], a31[3], a34[3], a24[3], a23[3], t1[3], t2[3];
        for (int d = 0; d < 3; d++) {
            a21[d] = a2[d] - a1[d]; a31[d] = a3[d] - a1[d]; a34[d] = a3[d] - a4[d];
            a24[d] = a2[d] - a4[d]; a23[d] = a2[d] - a3[d];
            t1[d] = (a1[d] + a2[d] + a3[d]) / 3.0; t2[d] = (a4[d] + a2[d] + a3[d]) / 3.0;
        }
        const double xi[3] = {a21[1] * a31[2] - a21[2] * a31[1], a21[2] * a31[0] - a21[0] * a31[2],
                         a21[0] * a31[1] - a21[1] * a31[0]};
        const double ze[3] = {a34[1] * a24[2] - a34[2] * a24[1], a34[2] * a24[0] - a34[0] * a24[2],
                         a34[0] * a24[1] - a34[1] * a24[0]};
        const double lxi = sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
        const double lze = sqrt(ze[0] * ze[0] + ze[1] * ze[1] + ze[2] * ze[2]);
        const double cost = (xi[0] * ze[0] + xi[1] * ze[1] + xi[2] * ze[2]) / lxi / lze;
        const double cost2 = 1.0 - cost * cost;
        if (cost2 < 1.0e-6) continue;
        const double cost0 = c0[3 * e1 + k], sint0 = s0[3 * e1 + k];
        const double side = (xi[0] - ze[0]) * (t1[0] - t2[0]) + (xi[1] - ze[1]) * (t1[1] - t2[1]) +
                       (xi[2] - ze[2]) * (t1[2] - t2[2]);
        const double kb = 2.0 / sqrt(3.0) * br;
        double sint, beta;
        if (side >= 0.0) { sint = sqrt(cost2); beta = kb * (sint * cost0 - cost * sint0) / sqrt(cost2); }
        else { sint = -sqrt(cost2); beta = -kb * (sint * cost0 - cost * sint0) / sqrt(cost2); }
        const double b11 = -beta * cost / (xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
        const double b12 = beta / lxi / lze;
        const double b22 = -beta * cost / (ze[0] * ze[0] + ze[1] * ze[1] + ze[2] * ze[2]);
        double f[3];
        f[0] = b11 * (xi[1] * -a23[2] - xi[2] * -a23[1]) + b12 * (ze[1] * -a23[2] - ze[2] * -a23[1]);
        f[1] = b11 * (xi[2] * -a23[0] - xi[0] * -a23[2]) + b12 * (ze[2] * -a23[0] - ze[0] * -a23[2]);
        f[2] = b11 * (xi[0] * -a23[1] - xi[1] * -a23[0]) + b12 * (ze[0] * -a23[1] - ze[1] * -a23[0]);
        for (int d = 0; d < 3; d++) unsafeAtomicAdd(cf + (size_t)d * nv + k1, (T)f[d]);
        f[0] = b11 * (xi[1] * -a31[2] - xi[2] * -a31[1]) + b22 * (ze[1] * a34[2] - ze[2] * a34[1])
             + b12 * ((xi[1] * a34[2] - xi[2] * a34[1]) + (ze[1] * -a31[2] - ze[2] * -a31[1]));
        f[1] = b11 * (xi[2] * -a31[0] - xi[0] * -a31[2]) + b22 * (ze[2] * a34[0] - ze[0] * a34[2])
             + b12 * ((xi[2] * a34[0] - xi[0] * a34[2]) + (ze[2] * -a31[0] - ze[0] * -a31[2]));
        f[2] = b11 * (xi[0] * -a31[1] - xi[1] * -a31[0]) + b22 * (ze[0] * a34[1] - ze[1] * a34[0])
             + b12 * ((xi[0] * a34[1] - xi[1] * a34[0]) + (ze[0] * -a31[1] - ze[1] * -a31[0]));
        for (int d = 0; d < 3; d++) unsafeAtomicAdd(cf + (size_t)d * nv + k2, (T)(0.5 * f[d]));
        f[0] = b11 * (xi[1] * a21[2] - xi[2] * a21[1]) + b22 * (ze[1] * -a24[2] - ze[2] * -a24[1])
             + b12 * ((xi[1] * -a24[2] - xi[2] * -a24[1]) + (ze[1] * a21[2] - ze[2] * a21[1]));
        f[1] = b11 * (xi[2] * a21[0] - xi[0] * a21[2]) + b22 * (ze[2] * -a24[0] - ze[0] * -a24[2])
             + b12 * ((xi[2] * -a24[0] - xi[0] * -a24[2]) + (ze[2] * a21[0] - ze[0] * a21[2]));
        f[2] = b11 * (xi[0] * a21[1] - xi[1] * a21[0]) + b22 * (ze[0] * -a24[1] - ze[1] * -a24[0])
             + b12 * ((xi[0] * -a24[1] - xi[1] * -a24[0]) + (ze[0] * a21[1] - ze[1] * a21[0]));
        for (int d = 0; d < 3; d++) unsafeAtomicAdd(cf + (size_t)d * nv + k3, (T)(0.5 * f[d]));
    }
}

// The stop test's relative residual: RESIDUAL, raised to the floor with FP32
// solver arithmetic (below about 1e-7 an FP32 iterate stops changing).
double solver_residual(const lbmibm_ctx *ctx) {
    const lbmibm_params &p = ctx->p;
    return ctx->prec == LBMIBM_F64 ? p.residual : std::max(p.residual, p.residual_floor);
}

// One on-chip class with the given shape variant.
template <typename T>
lbmibm_status launch_onchip(lbmibm_ctx *ctx, int cls, int var, const SolveOut<T> &out, hipStream_t stream,
                            int warm = 0) {
    const lbmibm_params &p = ctx->p;
    const double residual = solver_residual(ctx);
    const size_t lds = 3 * 4 * (size_t)ctx->bicg_maxv[cls] * sizeof(T);
    const int n = ctx->bicg_n[cls];
#define LBMIBM_ONCHIP(TT, K, NT)                                                                                \
    hipLaunchKernelGGL((k_bicg_onchip<TT, K, NT>), dim3(n), dim3(NT), lds, stream, ctx->bicg_cells[cls],          \
                       ctx->voff, (const TT *)ctx->coef, ctx->nbr, (const TT *)ctx->bv,                           \
                       *reinterpret_cast<const SolveOut<TT> *>(&out), ctx->nv, p.itrmax, residual, warm,         \
                       ctx->bicg_iters, ctx->status)
    const hipEvent_t kt = kt_begin(ctx, stream);
    if (cls == 0) {
        switch (var) {
            case 0: LBMIBM_ONCHIP(T, BICG0_SHAPES[0][0], BICG0_SHAPES[0][1]); break;
            case 1: LBMIBM_ONCHIP(T, BICG0_SHAPES[1][0], BICG0_SHAPES[1][1]); break;
            default: LBMIBM_ONCHIP(T, BICG0_SHAPES[2][0], BICG0_SHAPES[2][1]); break;
        }
    } else {
        if constexpr (sizeof(T) == 4) {  // class 1 is FP32 only (bicg_class)
            switch (var) {
                case 0: LBMIBM_ONCHIP(float, BICG1_SHAPES[0][0], BICG1_SHAPES[0][1]); break;
                case 1: LBMIBM_ONCHIP(float, BICG1_SHAPES[1][0], BICG1_SHAPES[1][1]); break;
                default: LBMIBM_ONCHIP(float, BICG1_SHAPES[2][0], BICG1_SHAPES[2][1]); break;
            }
        } else {
            return LBMIBM_ERR_STATE;
        }
    }
#undef LBMIBM_ONCHIP
    kt_end(ctx, stream, cls == 0 ? "k_bicg_onchip class 0" : "k_bicg_onchip class 1", kt);
    LBMIBM_LAUNCHED(ctx, "k_bicg_onchip");
    return LBMIBM_OK;
}

// Forms of the large-cell class: 0 = vectors in global memory, 1 = p in LDS,
// 2 = register-resident (FP32; 1 and 2 only when the largest cell fits).
constexpr int BICG_NFORM = 3;
constexpr const char *BICG_FORM_NAMES[BICG_NFORM] = {"global", "p-in-LDS", "resident"};

// Dynamic LDS of the large-cell forms: what the largest cell needs, raised
// to the whole of the CU's LDS (ctx->big_lds, set up in membrane_setup), so
// that no other work group (the on-chip classes run concurrently) shares the
// CU and its issue slots with a large-cell solve.
size_t big_lds(const lbmibm_ctx *ctx) {
    const size_t need = 3 * (size_t)ctx->bicg_maxv[2] * sizeof(float);
    return std::max(need, (size_t)ctx->big_lds);
}

bool big_form_fits(const lbmibm_ctx *ctx, int form) {
    if (form == 0) return true;
    if (ctx->prec == LBMIBM_F64) return false;
    if (form == 1) return ctx->bicg_maxv[2] <= BIG_LDS_V;
    return ctx->bicg_maxv[2] <= BICG_RES_K * 1024;
}

template <typename T>
lbmibm_status launch_big(lbmibm_ctx *ctx, int var, const SolveOut<T> &out, hipStream_t stream, int warm = 0) {
    const lbmibm_params &p = ctx->p;
    const double residual = solver_residual(ctx);
    if constexpr (sizeof(T) == 4) {
        if (var == 2 && big_form_fits(ctx, 2)) {
            const size_t lds = big_lds(ctx);
            LBMIBM_KT(ctx, stream, "k_bicg_res",
                      hipLaunchKernelGGL(k_bicg_res<BICG_RES_K>, dim3(ctx->bicg_n[2]), dim3(1024), lds, stream,
                                         ctx->bicg_cells[2], ctx->voff, (const float *)ctx->coef, ctx->nbr,
                                         (float *)ctx->bv, (float *)ctx->rh, (float *)ctx->pv, out, ctx->nv,
                                         p.itrmax, residual, warm, ctx->bicg_iters, ctx->status));
            LBMIBM_LAUNCHED(ctx, "k_bicg_res");
            return LBMIBM_OK;
        }
        if (var == 1 && big_form_fits(ctx, 1)) {
            const size_t lds = big_lds(ctx);
            LBMIBM_KT(ctx, stream, "k_bicg_big",
                      hipLaunchKernelGGL((k_bicg_big<T, 4>), dim3(ctx->bicg_n[2]), dim3(1024), lds, stream,
                                         ctx->bicg_cells[2], ctx->voff, (const T *)ctx->coef, ctx->nbr, (T *)ctx->bv,
                                         (T *)ctx->rh, (T *)ctx->rv, (T *)ctx->vv, (T *)ctx->tv, out, ctx->nv,
                                         p.itrmax, residual, warm, ctx->bicg_iters, ctx->status));
            LBMIBM_LAUNCHED(ctx, "k_bicg_big");
            return LBMIBM_OK;
        }
    }
    LBMIBM_KT(ctx, stream, "k_bicg_global",
              hipLaunchKernelGGL(k_bicg_global<T>, dim3(ctx->bicg_n[2]), dim3(1024), 0, stream, ctx->bicg_cells[2],
                                 ctx->voff, (const T *)ctx->coef, ctx->nbr, (T *)ctx->bv, (T *)ctx->rh, (T *)ctx->rv,
                                 (T *)ctx->pv, (T *)ctx->vv, (T *)ctx->tv, out, ctx->nv, p.itrmax, residual, warm,
                                 ctx->bicg_iters, ctx->status));
    LBMIBM_LAUNCHED(ctx, "k_bicg_global");
    return LBMIBM_OK;
}

template <typename T> SolveOut<T> solve_out(lbmibm_ctx *ctx) {
    SolveOut<T> out;
    out.q = (T *)ctx->qv;
    out.cf = (T *)ctx->cf;
    out.qi = (T *)ctx->qi;
    out.qp = (T *)ctx->qp;
    out.ds = (const T *)ctx->dsv;
    out.st = ctx->step;
    out.nv = ctx->nv;
    out.fo = ctx->p.fileoutput > 0 ? ctx->p.fileoutput : 0;
    out.wq = ctx->p.fileoutput > 0 ? (double)ctx->p.m / (double)ctx->p.fileoutput : 0.0;
    return out;
}

// Pseudo-random right-hand side in [-1, 1) for the autotune solves.
template <typename T> __global__ void k_synth_rhs(T *__restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long z = (unsigned long long)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    b[i] = (T)((double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0);
}

// Tension, corner terms and rows of all cells (cells == nullptr) or of the
// listed cells (n of them, at most max_e elements and max_v vertices each).
// The tension and rg outputs are kept with store_macros.
template <typename T>
lbmibm_status assemble_impl(lbmibm_ctx *ctx, const PeriodicBox &box, const TensionArgs &ta,
                            const int *cells = nullptr, int n = 0, int max_e = 0, int max_v = 0) {
    const lbmibm_params &p = ctx->p;
    const int ne = ctx->nt, nv = ctx->nv;
    const dim3 ge = cells ? dim3((max_e + 255) / 256, n) : dim3((ne + 255) / 256);
    const dim3 gv = cells ? dim3((max_v + 255) / 256, n) : dim3((nv + 255) / 256);
    LBMIBM_KT(ctx, ctx->stream, "k_elem",
              hipLaunchKernelGGL(k_elem<T>, ge, dim3(256), 0, ctx->stream, (const double *)ctx->xd, ctx->ele,
                                 (const double *)ctx->nref, p.store_macros ? (T *)ctx->tt : (T *)nullptr,
                                 (T *)ctx->rg, p.fileoutput > 0 ? (T *)ctx->tp : (T *)nullptr, ctx->cslot,
                                 (typename Corner<T>::type *)ctx->corner, ne, box, ta, ctx->step, cells, ctx->eoff));
    LBMIBM_LAUNCHED(ctx, "k_elem");
    LBMIBM_KT(ctx, ctx->stream, "k_vert",
              hipLaunchKernelGGL(k_vert<T>, gv, dim3(256), 0, ctx->stream, ctx->nln, ctx->epair,
                                 (const typename Corner<T>::type *)ctx->corner, ctx->vcell, (const T *)ctx->gscale,
                                 (T *)ctx->coef, (T *)ctx->bv, (T *)ctx->dsv, nv, cells, ctx->voff));
    LBMIBM_LAUNCHED(ctx, "k_vert");
    return LBMIBM_OK;
}

// Forward assembly of the chunks in list `which` (0 all cells, 1 the large
// cells, 2 the on-chip cells) with k_assemble.
template <typename T>
lbmibm_status assemble_fused(lbmibm_ctx *ctx, const PeriodicBox &box, const TensionArgs &ta, int which) {
    const lbmibm_params &p = ctx->p;
    const int n = ctx->asm_n[which];
    if (n == 0) return LBMIBM_OK;
    LBMIBM_KT(ctx, ctx->stream, "k_assemble",
              hipLaunchKernelGGL(k_assemble<T>, dim3(xcd_grid(n)), dim3(ASM_NT), 0, ctx->stream, ctx->asm_list[which],
                                 ctx->asm_chunks, ctx->asm_elems, (const double *)ctx->xd, ctx->eslot,
                                 (const double *)ctx->nref, p.store_macros ? (T *)ctx->tt : (T *)nullptr,
                                 (T *)ctx->rg, p.fileoutput > 0 ? (T *)ctx->tp : (T *)nullptr, ctx->nln,
                                 ctx->epair, ctx->vcell, (const T *)ctx->gscale, (T *)ctx->coef, (T *)ctx->bv,
                                 (T *)ctx->dsv, ctx->nv, box, ta, ctx->step, n));
    LBMIBM_LAUNCHED(ctx, "k_assemble");
    return LBMIBM_OK;
}

PeriodicBox box_of(const lbmibm_ctx *ctx) {
    return PeriodicBox{ctx->nx * ctx->p.dx, ctx->ny * ctx->p.dx, ctx->nz * ctx->p.dx, ctx->p.periodic_y};
}

TensionArgs tension_args(const lbmibm_ctx *ctx) {
    const lbmibm_params &p = ctx->p;
    TensionArgs ta;
    ta.c = p.c; ta.rgs = p.rgs; ta.neo = p.neohookean; ta.m = p.m;
    ta.fileoutput = p.fileoutput > 0 ? p.fileoutput : 1;
    ta.wtp = (double)p.m / (double)ta.fileoutput;
    return ta;
}

// Times every shape variant of the on-chip classes on the cells' own mass
// matrices (reference state) with a pseudo-random right-hand side and keeps
// the fastest; the solves' outputs are discarded.
template <typename T> lbmibm_status autotune_impl(lbmibm_ctx *ctx) {
    lbmibm_status s = assemble_impl<T>(ctx, box_of(ctx), tension_args(ctx));
    if (s != LBMIBM_OK) return s;
    const size_t n3 = 3 * (size_t)ctx->nv;
    hipLaunchKernelGGL(k_synth_rhs<T>, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, ctx->stream, (T *)ctx->bv,
                       n3);
    LBMIBM_LAUNCHED(ctx, "k_synth_rhs");
    const SolveOut<T> out = solve_out<T>(ctx);
    hipEvent_t e0, e1;
    LBMIBM_HIP(ctx, hipEventCreate(&e0));
    LBMIBM_HIP(ctx, hipEventCreate(&e1));
    char line[160];
    for (int cls = 0; cls < 2; cls++) {
        if (ctx->bicg_n[cls] == 0 || (cls == 1 && sizeof(T) != 4)) continue;
        float best = 0.0f;
        for (int v = 0; v < BICG_NVAR; v++) {
            s = launch_onchip<T>(ctx, cls, v, out, ctx->stream);  // warm-up
            if (s != LBMIBM_OK) break;
            LBMIBM_HIP(ctx, hipEventRecord(e0, ctx->stream));
            for (int r = 0; r < 2 && s == LBMIBM_OK; r++) s = launch_onchip<T>(ctx, cls, v, out, ctx->stream);
            LBMIBM_HIP(ctx, hipEventRecord(e1, ctx->stream));
            LBMIBM_HIP(ctx, hipEventSynchronize(e1));
            float ms = 0.0f;
            LBMIBM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
            ms *= 0.5f;
            const int *shape = cls == 0 ? BICG0_SHAPES[v] : BICG1_SHAPES[v];
            snprintf(line, sizeof line, "bicg class %d shape %dx%d: %.4f ms\n", cls, shape[0], shape[1], ms);
            ctx->tune_report += line;
            if (v == 0 || ms < best) { best = ms; ctx->bicg_var[cls] = v; }
        }
        if (s != LBMIBM_OK) break;
        const int *shape = cls == 0 ? BICG0_SHAPES[ctx->bicg_var[cls]] : BICG1_SHAPES[ctx->bicg_var[cls]];
        snprintf(line, sizeof line, "bicg class %d: using %dx%d\n", cls, shape[0], shape[1]);
        ctx->tune_report += line;
    }
    if (s == LBMIBM_OK && ctx->bicg_n[2] > 0 && sizeof(T) == 4) {
        float best = 0.0f;
        for (int v = 0; v < BICG_NFORM && s == LBMIBM_OK; v++) {
            if (!big_form_fits(ctx, v)) continue;
            s = launch_big<T>(ctx, v, out, ctx->stream);  // warm-up
            if (s != LBMIBM_OK) break;
            LBMIBM_HIP(ctx, hipEventRecord(e0, ctx->stream));
            s = launch_big<T>(ctx, v, out, ctx->stream);
            LBMIBM_HIP(ctx, hipEventRecord(e1, ctx->stream));
            LBMIBM_HIP(ctx, hipEventSynchronize(e1));
            float ms = 0.0f;
            LBMIBM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
            snprintf(line, sizeof line, "bicg class 2 form %s: %.4f ms\n", BICG_FORM_NAMES[v], ms);
            ctx->tune_report += line;
            if (v == 0 || ms < best) { best = ms; ctx->bicg_var[2] = v; }
        }
        snprintf(line, sizeof line, "bicg class 2: using %s\n", BICG_FORM_NAMES[ctx->bicg_var[2]]);
        ctx->tune_report += line;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // the timed solves wrote loads of the synthetic systems: the first warm
    // started solve begins at 0, as a cold one
    if (s == LBMIBM_OK) LBMIBM_HIP(ctx, hipMemsetAsync(ctx->qv, 0, 3 * (size_t)ctx->nv * sizeof(T), ctx->stream));
    return s;
}

// warm: the start of the solves (bicg_warm_start: 0 from zero, 1 from the
// previous loads, 2 extrapolated from the two previous loads).
template <typename T> lbmibm_status forces_impl(lbmibm_ctx *ctx, int warm) {
    const lbmibm_params &p = ctx->p;
    const PeriodicBox box = box_of(ctx);
    const TensionArgs ta = tension_args(ctx);
    const int ne = ctx->nt, nv = ctx->nv;
    // with large cells, their systems are assembled first so that their
    // solve (the longest) starts while the others are assembled and solved
    const bool split = ctx->bicg_n[2] > 0;
    lbmibm_status st = assemble_fused<T>(ctx, box, ta, split ? 1 : 0);
    if (st != LBMIBM_OK) return st;
    const SolveOut<T> out = solve_out<T>(ctx);
    if (ctx->bicg_n[2] > 0) {
        // the large cells (one work group each, vectors in global memory) run
        // on the side stream, concurrently with the on-chip classes; x in its
        // own buffer, the epilogue writes q
        LBMIBM_HIP(ctx, hipEventRecord(ctx->ev_fork, ctx->stream));
        LBMIBM_HIP(ctx, hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
        lbmibm_status sb = launch_big<T>(ctx, ctx->bicg_var[2], out, ctx->side, warm);
        if (sb != LBMIBM_OK) return sb;
        if (split && ctx->n_onchip > 0) {
            sb = assemble_fused<T>(ctx, box, ta, 2);
            if (sb != LBMIBM_OK) return sb;
        }
    }
    // With all three classes, class 0 follows the large-cell solve on the
    // side stream and class 1 runs alone here: the large-cell solve holds its
    // CUs for about as long as the on-chip assembly and class 1 take on the
    // others, so the two streams end close together.  Otherwise class 1 runs
    // first (its work groups take a whole CU's LDS, one cell per CU), and
    // class 0's short work groups fill the CUs its last round leaves free.
    // Class 1 is enqueued before class 0's fork, so that in a captured graph
    // it is the assembly's first dependent and keeps the assembly's queue
    // (a dependent on another queue starts about 9 us after its parent).
    const bool c0_side = split && ctx->bicg_n[0] > 0 && ctx->bicg_n[1] > 0;
    if (c0_side) LBMIBM_HIP(ctx, hipEventRecord(ctx->ev_afork, ctx->stream));
    for (int cls = 1; cls >= (c0_side ? 1 : 0); cls--)
        if (ctx->bicg_n[cls] > 0) {
            lbmibm_status s = launch_onchip<T>(ctx, cls, ctx->bicg_var[cls], out, ctx->stream, warm);
            if (s != LBMIBM_OK) return s;
        }
    if (c0_side) {
        LBMIBM_HIP(ctx, hipStreamWaitEvent(ctx->side, ctx->ev_afork, 0));
        lbmibm_status s = launch_onchip<T>(ctx, 0, ctx->bicg_var[0], out, ctx->side, warm);
        if (s != LBMIBM_OK) return s;
    }
    if (ctx->bicg_n[2] > 0) {
        LBMIBM_HIP(ctx, hipEventRecord(ctx->ev_join, ctx->side));
        LBMIBM_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
    }
    if (p.br != 0.0) {
        hipLaunchKernelGGL(k_bend<T>, dim3((ne + 255) / 256), dim3(256), 0, ctx->stream, (const double *)ctx->xs,
                           ctx->ele, ctx->ec, (const double *)ctx->c0, (const double *)ctx->s0, (T *)ctx->cf, nv, ne, p.br,
                           box);
        LBMIBM_LAUNCHED(ctx, "k_bend");
    }
    return LBMIBM_OK;
}

// ---- adjoint of the membrane force ---------------------------------------------
//
// Forward per cell: q = M(X)^-1 b(X) (the Jacobi scaling does not change the
// solution), cf = -q dsv with dsv = ds(X) s_c (ring area times the cell's
// force scale).  Given cfa, the adjoint of cf:
//   qbar = -dsv cfa,  dsvbar = -cfa . q,  s_c-bar += dsvbar ds,
//   lambda = M^-1 qbar (M is symmetric: the same solver, right-hand side
//   qbar / jac), bbar = lambda, Mbar = -lambda q^T,
//   rg_e-bar = -(1/24)(sum_j lambda_j.q_j + (sum_j lambda_j).(sum_k q_k))
//              + s_c sum_j dsvbar_j / 6                (M and ds from rg_e),
//   Xbar += sum_e rg_e-bar grad rg_e + J_b^T lambda.
// b is the gradient of the discrete Skalak (or neo-Hookean) energy, so its
// Jacobian is symmetric and J_b^T lambda is the directional derivative of the
// element right-hand sides along lambda, taken here in forward mode with dual
// numbers through the same element formula as k_elem.

struct Dual {
    double v, d;
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
    const double r = a.v / b.v;
    return {r, (a.d - r * b.d) / b.v};
}
__device__ __forceinline__ Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
    const double r = a / b.v;
    return {r, -r * b.d / b.v};
}
__device__ __forceinline__ Dual dsqrt(Dual a) {
    const double r = sqrt(a.v);
    return {r, a.d / (2.0 * r)};
}
__device__ __forceinline__ double dsqrt(double a) { return sqrt(a); }

// Right-hand-side terms tb[corner][component] of one element (k_elem's
// formula, without the rounding of the stored tension to T).
template <typename R>
__device__ __forceinline__ void elem_rhs(const R v1[3], const R v2[3], const double nn[4], const TensionArgs &a,
                                         R tb[3][3]) {
    const R g0 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
    const R g3 = v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2];
    const R g1 = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
    const R dv = g0 * g3 - g1 * g1;
    const double dn = nn[0] * nn[3] - nn[1] * nn[2];
    const R nd0 = g3 / dv, nd1 = -g1 / dv, nd3 = g0 / dv;
    const R i1 = g0 * nn[0] + g1 * nn[1] + g1 * nn[2] + g3 * nn[3] - 2.0;
    const R i2 = dv * dn - 1.0;
    const R js = dsqrt(i2 + 1.0);
    const R nd[4] = {nd0, nd1, nd1, nd3};
    R tk[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (!a.neo) tk[k] = a.rgs * nn[k] * ((i1 + 1.0) / js) + (a.c * i2 - 1.0) * js * nd[k];
        else tk[k] = a.rgs * (nn[k] / js - nd[k] * js / (i2 + 1.0) / (i2 + 1.0));
    }
    const R gab = dsqrt(dv);
    const double F[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
#pragma unroll
    for (int en = 0; en < 3; en++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const R vd[2] = {v1[d], v2[d]};
            R sum = gab * 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int al = k >> 1, be = k & 1;
                sum = sum + gab * tk[k] * (vd[al] * F[en][be] + vd[be] * F[en][al]) / 4.0;
            }
            tb[en][d] = sum;
        }
}

// Per vertex: the adjoint right-hand side qbar/jac into bv, dsvbar, and the
// per-cell force-scale gradient.
template <typename T>
__global__ __launch_bounds__(256) void k_adj_load(const int *__restrict__ nln,
                                                  const typename Corner<T>::type *__restrict__ corner,
                                                  const T *__restrict__ dsv, const T *__restrict__ q,
                                                  const T *__restrict__ cfa, T *__restrict__ bv,
                                                  double *__restrict__ dsva, double *__restrict__ gpart, int nv) {
    const int jj = blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= nv) return;
    const int n = nln[jj];
    double jac = 0.0, ds = 0.0;
    for (int s = 0; s < n; s++) {
        const double r24 = (double)corner[(size_t)s * nv + jj].w / 24.0;
        jac += 2.0 * r24;  // rg / 12, as k_vert
        ds += 4.0 * r24;   // rg / 2 / 3
    }
    const double d = (double)dsv[jj];
    double sb = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double a = (double)cfa[(size_t)k * nv + jj];
        bv[3 * (size_t)jj + k] = (T)(-d * a / jac);
        sb -= a * (double)q[3 * (size_t)jj + k];
    }
    dsva[jj] = sb;
    gpart[jj] = sb * ds;  // this vertex's term of the cell's force-scale gradient
}

// Per cell (one work group): scale the adjoint right-hand side by a power of
// two so that its largest entry is of order 1, and keep the scale.  The
// adjoint loads of cells far from where the objective is measured can be
// many orders of magnitude below those near it; in FP32 they would reach the
// subnormal range, where the solver cannot reach a relative residual.  Power
// of two scaling is exact, and the solution is scaled back after the solve.
// The same work group also adds the cell's force-scale gradient (its
// vertices' terms gpart, in a fixed order) to gsa, without atomics.
template <typename T>
__global__ __launch_bounds__(256) void k_rhs_scale(const int *__restrict__ voff, T *__restrict__ bv,
                                                   double *__restrict__ cscale, const double *__restrict__ gpart,
                                                   double *__restrict__ gsa) {
    const int c = blockIdx.x;
    const size_t v0 = voff[c], v1 = voff[c + 1];
    const size_t b0 = 3 * v0, b1 = 3 * v1;
    __shared__ float wmax[4];
    __shared__ double wsum[4];
    float m = 0.0f;
    double gs = 0.0;
    for (size_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) m = fmaxf(m, fabsf((float)bv[i]));
    for (size_t i = v0 + threadIdx.x; i < v1; i += blockDim.x) gs += gpart[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    gs = wave_sum(gs);
    if ((threadIdx.x & 63) == 0) {
        wmax[threadIdx.x >> 6] = m;
        wsum[threadIdx.x >> 6] = gs;
    }
    __syncthreads();
    m = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
    const double s = m > 0.0f ? ldexp(1.0, -ilogb((double)m)) : 1.0;
    for (size_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) bv[i] = (T)((double)bv[i] * s);
    if (threadIdx.x == 0) {
        cscale[c] = s;
        if (gsa) gsa[c] += ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
    }
}

// Per vertex: undo the right-hand-side scale on the solution.
template <typename T>
__global__ __launch_bounds__(256) void k_lam_unscale(const int *__restrict__ vcell, const double *__restrict__ cscale,
                                                     T *__restrict__ lam, int nv) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nv) return;
    const double s = 1.0 / cscale[vcell[j]];
#pragma unroll
    for (int k = 0; k < 3; k++) lam[3 * (size_t)j + k] = (T)((double)lam[3 * (size_t)j + k] * s);
}

// Per element: its position adjoints (rg terms and J_b^T lambda) into the
// slots of its corners.
template <typename T>
__global__ __launch_bounds__(256) void k_elem_adj(const double *__restrict__ xd, const int *__restrict__ ele,
                                                  const double *__restrict__ nref, const int *__restrict__ cslot,
                                                  const int *__restrict__ vcell, const T *__restrict__ gscale,
                                                  const T *__restrict__ lam, const T *__restrict__ q,
                                                  const double *__restrict__ dsva, double *__restrict__ cadj,
                                                  int ne, PeriodicBox box, TensionArgs a) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ne) return;
    const int vid[3] = {ele[3 * j], ele[3 * j + 1], ele[3 * j + 2]};
    double v1[3], v2[3], L[3][3], Qv[3][3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        v1[d] = xd[3 * vid[1] + d] - xd[3 * vid[0] + d];
        v2[d] = xd[3 * vid[2] + d] - xd[3 * vid[0] + d];
    }
    min_img3<double>(v1, box);
    min_img3<double>(v2, box);
#pragma unroll
    for (int en = 0; en < 3; en++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            L[en][d] = (double)lam[3 * (size_t)vid[en] + d];
            Qv[en][d] = (double)q[3 * (size_t)vid[en] + d];
        }
    const double nn[4] = {nref[4 * j], nref[4 * j + 1], nref[4 * j + 2], nref[4 * j + 3]};
    // rg_e-bar from the mass matrix and the ring areas
    double lq = 0.0, sl[3] = {0.0, 0.0, 0.0}, sq[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int en = 0; en < 3; en++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            lq += L[en][d] * Qv[en][d];
            sl[d] += L[en][d];
            sq[d] += Qv[en][d];
        }
    const double gsc = (double)gscale[vcell[vid[0]]];
    const double rgb = -(lq + sl[0] * sq[0] + sl[1] * sq[1] + sl[2] * sq[2]) / 24.0 +
                       gsc * (dsva[vid[0]] + dsva[vid[1]] + dsva[vid[2]]) / 6.0;
    // grad rg, rg = |v1 x v2|
    const double g0 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
    const double g3 = v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2];
    const double g1 = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
    const double rg = sqrt(g0 * g3 - g1 * g1);
    double d1[3], d2[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        d1[d] = (g3 * v1[d] - g1 * v2[d]) / rg;
        d2[d] = (g0 * v2[d] - g1 * v1[d]) / rg;
    }
    // J_b^T lambda: tangent of the element terms along lambda
    Dual w1[3], w2[3], tb[3][3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        w1[d] = {v1[d], L[1][d] - L[0][d]};
        w2[d] = {v2[d], L[2][d] - L[0][d]};
    }
    elem_rhs<Dual>(w1, w2, nn, a, tb);
#pragma unroll
    for (int en = 0; en < 3; en++) {
        double xa[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double grg = en == 0 ? -(d1[d] + d2[d]) : (en == 1 ? d1[d] : d2[d]);
            xa[d] = rgb * grg + tb[en][d].d;
        }
        double *o = cadj + 4 * (size_t)cslot[3 * (size_t)j + en];
        o[0] = xa[0]; o[1] = xa[1]; o[2] = xa[2]; o[3] = 0.0;
    }
}

// Per vertex: sum of its corner adjoints into the position adjoint.
__global__ __launch_bounds__(256) void k_vert_adj(const int *__restrict__ nln, const double *__restrict__ cadj,
                                                  double *__restrict__ xa, int nv) {
    const int jj = blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= nv) return;
    const int n = nln[jj];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int s = 0; s < n; s++) {
        const double *c = cadj + 4 * ((size_t)s * nv + jj);
        s0 += c[0]; s1 += c[1]; s2 += c[2];
    }
    xa[jj] += s0;
    xa[nv + jj] += s1;
    xa[2 * (size_t)nv + jj] += s2;
}

template <typename T> lbmibm_status forces_adj_impl(lbmibm_ctx *ctx, bool have_q) {
    const lbmibm_params &p = ctx->p;
    const int ne = ctx->nt, nv = ctx->nv;
    // forward at the current positions: q unless the caller restored it (qv)
    // from a recording of this step, then the rows, ring areas and the corner
    // records (k_elem + k_vert: k_adj_load reads the records)
    lbmibm_status s = have_q ? LBMIBM_OK : forces_impl<T>(ctx, false);
    if (s == LBMIBM_OK) s = assemble_impl<T>(ctx, box_of(ctx), tension_args(ctx));
    if (s != LBMIBM_OK) return s;
    hipLaunchKernelGGL(k_adj_load<T>, dim3((nv + 255) / 256), dim3(256), 0, ctx->stream, ctx->nln,
                       (const typename Corner<T>::type *)ctx->corner, (const T *)ctx->dsv, (const T *)ctx->qv,
                       (const T *)ctx->adj_cf, (T *)ctx->bv, ctx->adj_dsv, ctx->adj_gpart, nv);
    LBMIBM_LAUNCHED(ctx, "k_adj_load");
    hipLaunchKernelGGL(k_rhs_scale<T>, dim3(ctx->nc), dim3(256), 0, ctx->stream, ctx->voff, (T *)ctx->bv,
                       ctx->adj_cscale, ctx->adj_gpart, ctx->grad_gs);
    LBMIBM_LAUNCHED(ctx, "k_rhs_scale");
    SolveOut<T> out = solve_out<T>(ctx);
    out.q = (T *)ctx->adj_lam;
    out.cf = (T *)ctx->adj_scratch;
    out.qi = nullptr;
    out.fo = 0;
    for (int cls = 0; cls < 2; cls++)
        if (ctx->bicg_n[cls] > 0) {
            s = launch_onchip<T>(ctx, cls, ctx->bicg_var[cls], out, ctx->stream);
            if (s != LBMIBM_OK) return s;
        }
    if (ctx->bicg_n[2] > 0) {
        s = launch_big<T>(ctx, ctx->bicg_var[2], out, ctx->stream);
        if (s != LBMIBM_OK) return s;
    }
    hipLaunchKernelGGL(k_lam_unscale<T>, dim3((nv + 255) / 256), dim3(256), 0, ctx->stream, ctx->vcell,
                       ctx->adj_cscale, (T *)ctx->adj_lam, nv);
    LBMIBM_LAUNCHED(ctx, "k_lam_unscale");
    hipLaunchKernelGGL(k_elem_adj<T>, dim3((ne + 255) / 256), dim3(256), 0, ctx->stream, (const double *)ctx->xd,
                       ctx->ele, (const double *)ctx->nref, ctx->cslot, ctx->vcell, (const T *)ctx->gscale,
                       (const T *)ctx->adj_lam, (const T *)ctx->qv, ctx->adj_dsv, ctx->adj_corner, ne, box_of(ctx),
                       tension_args(ctx));
    LBMIBM_LAUNCHED(ctx, "k_elem_adj");
    hipLaunchKernelGGL(k_vert_adj, dim3((nv + 255) / 256), dim3(256), 0, ctx->stream, ctx->nln, ctx->adj_corner,
                       ctx->adj_x, nv);
    LBMIBM_LAUNCHED(ctx, "k_vert_adj");
    (void)p;
    return LBMIBM_OK;
}

}  // namespace

lbmibm_status membrane_forces_adjoint(lbmibm_ctx *ctx, bool have_q) {
    if (ctx->nc == 0) return LBMIBM_OK;
    if (ctx->p.br != 0.0) {
        ctx->err = "adjoint: bending (BR != 0) has no adjoint";
        return LBMIBM_ERR_STATE;
    }
    return ctx->prec == LBMIBM_F64 ? forces_adj_impl<double>(ctx, have_q) : forces_adj_impl<float>(ctx, have_q);
}

int bicg_class(int prec, int nvert) {
    // on chip: up to 1024 vertices (any precision) and up to 3072 (FP32: the
    // FP64 rows do not fit the register budget); else vectors in global memory
    if (nvert <= BICG0_CAP) return 0;
    if (nvert <= BICG1_CAP && prec != LBMIBM_F64) return 1;
    return 2;
}

lbmibm_status membrane_setup(lbmibm_ctx *ctx) {
    // allow every on-chip shape the LDS its class's largest cell needs (p, r, b as [V][4])
    static_assert(BICG0_SHAPES[0][0] * BICG0_SHAPES[0][1] == BICG0_CAP && BICG0_SHAPES[1][0] * BICG0_SHAPES[1][1] ==
                      BICG0_CAP && BICG0_SHAPES[2][0] * BICG0_SHAPES[2][1] == BICG0_CAP, "class 0 shapes");
    static_assert(BICG1_SHAPES[0][0] * BICG1_SHAPES[0][1] == BICG1_CAP && BICG1_SHAPES[1][0] * BICG1_SHAPES[1][1] ==
                      BICG1_CAP && BICG1_SHAPES[2][0] * BICG1_SHAPES[2][1] == BICG1_CAP, "class 1 shapes");
    const hipFuncAttribute A = hipFuncAttributeMaxDynamicSharedMemorySize;
    if (ctx->bicg_n[0] > 0) {
        const int lds = (int)(3 * 4 * (size_t)ctx->bicg_maxv[0] * ctx->szT);
        if (ctx->prec == LBMIBM_F64) {
            LBMIBM_HIP(ctx, hipFuncSetAttribute((const void *)k_bicg_onchip<double, 4, 256>, A, lds));
            LBMIBM_HIP(ctx, hipFuncSetAttribute((const void *)k_bicg_onchip<double, 2, 512>, A, lds));
            LBMIBM_HIP(ctx, hipFuncSetAttribute((const void *)k_bicg_onchip<double, 1, 1024>, A, lds));
        } else {
            LBMIBM_HIP(ctx, hipFuncSetAttribute((const void *)k_bicg_onchip<float, 4, 256>, A, lds));
            LBMIBM_HIP(ctx, hipFuncSetAttribute((const void *)k_bicg_onchip<float, 2, 512>, A, lds));
            LBMIBM_HIP(ctx, hipFuncSetAttribute((const void *)k_bicg_onchip<float, 1, 1024>, A, lds));
        }
    }
    // large-cell forms with LDS: the whole CU (see big_lds), less their static LDS
    ctx->big_lds = 0;
    if (ctx->bicg_n[2] > 0 && (big_form_fits(ctx, 1) || big_form_fits(ctx, 2))) {
        int dev = 0, lmax = 0;
        LBMIBM_HIP(ctx, hipGetDevice(&dev));
        LBMIBM_HIP(ctx, hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        hipFuncAttributes fa1{}, fa2{};
        LBMIBM_HIP(ctx, hipFuncGetAttributes(&fa1, (const void *)k_bicg_big<float, 4>));
        LBMIBM_HIP(ctx, hipFuncGetAttributes(&fa2, (const void *)k_bicg_res<BICG_RES_K>));
        const int stat = (int)std::max(fa1.sharedSizeBytes, fa2.sharedSizeBytes);
        ctx->big_lds = std::max(0, lmax - stat);
    }
    if (ctx->bicg_n[2] > 0 && big_form_fits(ctx, 1)) {
        const int lds = (int)big_lds(ctx);
        LBMIBM_HIP(ctx, hipFuncSetAttribute((const void *)k_bicg_big<float, 4>, A, lds));
    }
    if (ctx->bicg_n[2] > 0 && big_form_fits(ctx, 2)) {
        const int lds = (int)big_lds(ctx);
        LBMIBM_HIP(ctx, hipFuncSetAttribute((const void *)k_bicg_res<BICG_RES_K>, A, lds));
    }
    if (ctx->bicg_n[1] > 0) {
        const int lds = (int)(3 * 4 * (size_t)ctx->bicg_maxv[1] * sizeof(float));
        LBMIBM_HIP(ctx, hipFuncSetAttribute((const void *)k_bicg_onchip<float, 6, 512>, A, lds));
        LBMIBM_HIP(ctx, hipFuncSetAttribute((const void *)k_bicg_onchip<float, 3, 1024>, A, lds));
        LBMIBM_HIP(ctx, hipFuncSetAttribute((const void *)k_bicg_onchip<float, 4, 768>, A, lds));
    }
    return LBMIBM_OK;
}

lbmibm_status membrane_autotune(lbmibm_ctx *ctx) {
    ctx->bicg_var[0] = ctx->bicg_var[1] = ctx->bicg_var[2] = 0;
    ctx->tune_report.clear();
    if (ctx->nc == 0) return LBMIBM_OK;
    const char *env = getenv("LBMIBM_AUTOTUNE");
    if (env && env[0] == '0') {
        ctx->tune_report = "autotune off (LBMIBM_AUTOTUNE=0): first shapes\n";
        return LBMIBM_OK;
    }
    // LBMIBM_BICG_SHAPES="a,b[,c]" fixes the shape variants of classes 0 and 1
    // and the form of class 2
    const char *fixed = getenv("LBMIBM_BICG_SHAPES");
    if (fixed && fixed[0]) {
        int a = 0, b = 0, cc = 0;
        const int nread = sscanf(fixed, "%d,%d,%d", &a, &b, &cc);
        if (nread < 2 || a < 0 || a >= BICG_NVAR || b < 0 || b >= BICG_NVAR || cc < 0 || cc >= BICG_NFORM)
            return LBMIBM_ERR_ARG;
        ctx->bicg_var[0] = a;
        ctx->bicg_var[1] = b;
        ctx->bicg_var[2] = cc;
        ctx->tune_report = "shapes fixed by LBMIBM_BICG_SHAPES\n";
        return LBMIBM_OK;
    }
    lbmibm_status s = ctx->prec == LBMIBM_F64 ? autotune_impl<double>(ctx) : autotune_impl<float>(ctx);
    if (s != LBMIBM_OK) return s;
    // discard what the timing solves left in the output accumulators
    s = lattice_fill(ctx, ctx->qi, 3 * (size_t)ctx->nv, 0.0, false);
    if (s == LBMIBM_OK) s = lattice_fill(ctx, ctx->tp, 2 * (size_t)ctx->nt, 0.0, false);
    return s;
}

lbmibm_status membrane_forces(lbmibm_ctx *ctx) {
    if (ctx->nc == 0) return LBMIBM_OK;
    const int warm = ctx->solve_cold ? 0 : ctx->p.bicg_warm_start;
    return ctx->prec == LBMIBM_F64 ? forces_impl<double>(ctx, warm) : forces_impl<float>(ctx, warm);
}

}  // namespace lbmibm
