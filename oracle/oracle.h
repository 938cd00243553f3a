/*
 * FP64 host ORACLE for the coupled D3Q19-LBM / IBM / Skalak timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch CPU restatement of the
 * reference algorithm, used as the checker by tests/, by
 * __graft_entry__.smoke() and as the cpu_baseline leg of bench.py.  Nothing
 * in the product (preprocess_src_amd/) links, loads or calls it.
 *
 * It deliberately keeps the reference's own data model so that it is an
 * independent statement of the algorithm, not a copy of the GPU design:
 *   - padded lattice with two halo layers per side, nx = LX/dx + 4
 *     (init1.cu:145-147), node index i = ix + nx*iy + nx*ny*iz,
 *   - AoS populations f[i*Q+q] in FP64, push streaming with half-way
 *     bounce-back into the node's own opposite slot (lbm.cu:1273-1310),
 *   - bc mask COMP/BUFF/WALL (init1.cu:164-178),
 *   - per-cell ELL mass matrix solved by BiCGSTAB (stress.cu:146-272,
 *     matrix_solver.cu:279-377),
 *   - per-cell volume/area sums (volume_correction.cu:246-281).
 * The one generalisation is that cells may have different meshes: cell c owns
 * vertices voff[c]..voff[c+1]-1 and elements eoff[c]..eoff[c+1]-1 (the
 * reference assumes one mesh per species, jid = ic*vertex + j, init1.cu:642).
 *
 * Parity pinning (what checks each function beyond this restatement) is listed
 * per function below and in the test files.
 */
#ifndef LBMIBM_ORACLE_H
#define LBMIBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_COMP = 0, ORC_BUFF = 1, ORC_WALL = 2 };

typedef struct orc_params {
    int    nx, ny, nz;      /* padded lattice, LX/dx + 4 per axis (init1.cu:145-147)      */
    int    periodic_y;      /* 0: WALL outside |y|<LY/2 (slab channel), 1: fully periodic  */
    double dx, dt, cr;      /* lattice spacing, time step, Courant number Cr              */
    double df, lambda1;     /* omega = 1/(3(1+(lambda1-1)vf)Df + 0.5)  (lbm.cu:1131)       */
    double p;               /* body force magnitude P (Guo force 2 rho P e, lbm.cu:590)  */
    double ex, ey, ez;      /* body force direction e                                      */
    int    force_modulated; /* 1: e*(1+0.5 sin(pi t/FREQ)) after FLOWSTART (lbm.cu:1032);  */
                            /* 2: e after FLOWSTART, 0 before (lbm_cylinder.cu:292-293)     */
    int    flowstart;
    double freq;
    double b;               /* IBM force coupling B (lbm.cu:1116)                          */
    int    m;               /* membrane sub-cycling M (main.cu:80-101)                     */
    double lx, ly, lz;      /* domain extents LX, LY, LZ                                   */
    double c;               /* Skalak area dilation C (stress.cu:96)                       */
    double rgs;             /* cel->rg, 1.0 for RBCs (init1.cu:524)                        */
    double br;              /* bending rigidity (stress.cu:538)                            */
    int    itrmax;          /* BiCGSTAB ITRMAX (matrix_solver.cu:330)                      */
    double residual;        /* BiCGSTAB RESIDUAL (matrix_solver.cu:370)                    */
    int    runge_kutta;     /* 1, 2 or 4 (ibm.cu:430)                                      */
    int    neohookean;      /* 0: SKALAK, 1: NEOHOOKEAN (stress.cu:95-99)                  */
    int    fileoutput;      /* FILEOUTPUT period of the tp/qi accumulators                 */
    int    volume_correction; /* 1: run volume_correction in orc_step (main.cu:101)        */
    int    vof_period;      /* K: front tracking every K steps, THINC/WLIC every M steps  */
                            /* (main.cu:93-97); 0: no VOF indicator (vf stays 0)          */
    int    sor_itrmax;      /* ITRMAX, RESIDUAL of the red-black SOR (matrix_solver.cu:   */
    double sor_residual;    /* 188-215), the same constants as BiCGSTAB's in the reference */
    int    tube;            /* 1: CapsuleCHANNELFLOW tube: COMP where sqrt(x^2+y^2) <= R,  */
    double tube_radius;     /* periodic in z only, interpolated bounce-back on the links   */
                            /* that cross r = R (init1.cu:190-236, lbm_cylinder.cu)         */
    int    tube_staircase;  /* with tube: 1 half-way bounce-back on the mask only          */
                            /* (lbm_CapsuleCHANNELFLOW, lbm.cu:1050-1094, 1273-1310)       */
    int    wall_motion;     /* slab walls sliding in z (lbm.cu:951-970): 0 none, 1 allwall */
                            /* (+U at y = +LY/2, -U at y = -LY/2), 2 topwall (+U at top)  */
    double wall_velocity;   /* U; sra = 0 while iter < FLOWSTART (lbm.cu:755)              */
} orc_params;

/* Pointer set equivalent to the reference's lattice/cell/fem structs
 * (Appendix of def.h reconstructed from use).  Lattice arrays are padded
 * [n = nx*ny*nz]; membrane arrays use global vertex / element ids. */
typedef struct orc_state {
    /* lattice */
    double *fn, *fm;                       /* [n*19]  (swapped after each lbm step)     */
    double *dn, *un, *vn, *wn;             /* [n]                                        */
    double *dm, *um, *vm, *wm;             /* [n]                                        */
    double *fx, *fy, *fz;                  /* [n]  lattice IBM force                     */
    double *vf;                            /* [n] or NULL (NULL = vf=0 everywhere)       */
    int    *bc;                            /* [n]                                        */
    /* cells */
    int     nc;
    const int *voff, *eoff;                /* [nc+1]                                     */
    double *x, *y, *z;                     /* [Nv] SoA positions (cel->x)                */
    double *u, *v, *w;                     /* [Nv]                                       */
    double *cfx, *cfy, *cfz;               /* [Nv] nodal membrane force (cel->fx)        */
    double *xd;                            /* [3Nv] AoS positions used by FEM (fem->xd)  */
    const int *ele;                        /* [3Nt] global vertex ids                    */
    const int *ec;                         /* [3Nt] edge-adjacent elements               */
    const double *c0, *s0;                 /* [3Nt] bending reference angle              */
    const double *nref;                    /* [4Nt] reference contravariant metric       */
    double *t, *rg, *tp;                   /* [4Nt], [Nt], [2Nt] (tp may be NULL)        */
    const int *nln, *ln, *lo, *nle, *le;   /* [Nv], [7Nv], [7Nv], [Nv], [6Nv]            */
    double *q, *qi;                        /* [3Nv] load; qi may be NULL                 */
    const double *gs_scale;                /* [nc] per-cell force scale or NULL (=1)     */
    double *volume0, *volume, *area, *gx;  /* [nc], [nc], [nc], [3nc]                    */
    int    *bicg_iters;                    /* [nc] iterations of the last solve or NULL  */
    /* per-node body-force field [n] added to P e (Guo force 2 rho (P e + g)),
     * NULL = none; not in the reference, used by the gradient checks */
    const double *gbx, *gby, *gbz;
    /* cytoadhesion (adhesion.cu) or NULL: orc_step runs Adhesion() after
     * stress and spreads cfx + f_adh (ibm_st1bb, ibm.cu:133-212) */
    struct orc_adhesion *adh;
} orc_state;

/* Cytoadhesion of the slab (CapsuleSHEARFLOW) with the linear spring law
 * (adhesion.cu, LINEAR_SPRING) or the worm-like chain (WLC).  Receptors on the wall y = -LY/2 on an
 * nxw x nzw grid (wall_init, adhesion.cu:435-466).  Per vertex: knob flag,
 * MAXN nearest receptors, springs_per_knob springs.  Randomness: the
 * counter-based orc_adh_uniform(seed, iter, spring) in (0, 1], one draw per
 * spring per call (curand_uniform in the reference, adhesion.cu:550-558).
 * A receptor wanted by several springs in the same call goes to the lowest
 * spring id (vertex * springs_per_knob + spring); the others try their
 * next candidate (the reference resolves this with an atomicOr race). */
typedef struct orc_adhesion {
    double receptor_density, rc, k_ad, l0, k_on0, k_off0, chi_beta, kbt, shear_rate, gs;
    int    limit, max_neighwall, springs_per_knob;
    unsigned long long seed;
    int    wlc;                 /* 0: LINEAR_SPRING, 1: WLC (adhesion.cu:564-567, 677-690) */
    double lcont, f_wlc, ra, alpha1, alpha2;  /* WLC: Lcont, fWLC, RA, ALPHA1, ALPHA2   */
    int    nxw, nzw;            /* receptor grid (set by orc_adhesion_init)  */
    double dxw, dzw;
    unsigned char *is_knob;     /* [Nv]            */
    int    *id_wall;            /* [Nv*MAXN]       */
    double *r_wall;             /* [Nv*MAXN]       */
    int    *attach;             /* [Nv*NS], -1 free */
    double *r_ij;               /* [Nv*NS]         */
    int    *is_attached;        /* [nxw*nzw]       */
    double *f_adh;              /* [3Nv] AoS       */
} orc_adhesion;

/* ---- lattice ---------------------------------------------------------- */
/* cdo_init mask, slab (walls outside |y|<LY/2, BUFF halo in x,z) or fully
 * periodic (BUFF halo in x,y,z).  init1.cu:133-188 */
void orc_bc_init(const orc_params *p, int *bc);
/* flu_init: f = w rho (1 + 3cU + 4.5(cU)^2 - 1.5U^2), U = Cr u on every node;
 * rho/u/v/w are padded [n] arrays (NULL rho -> 1, NULL velocity -> 0);
 * fm = 0, macros = initial values, lattice force = 0.  init1.cu:281-449 */
void orc_flu_init(const orc_params *p, orc_state *s, const double *rho,
                  const double *u, const double *v, const double *w);
/* One lbm_d3q19 call: fused collide + push + BB, BUFF macro copy, M-average,
 * then the fn/fm pointer swap.  lbm.cu:706-1312, 1316-1460, 1463-1579 */
void orc_lbm_step(const orc_params *p, orc_state *s, int iter);

/* ---- immersed boundary ----------------------------------------------- */
/* ibm_st1: zero lattice force, spread nodal forces with the 4-point cosine
 * delta (weight phi^3/dx^3), fold BUFF contributions onto their periodic
 * COMP images.  ibm.cu:27-131, 214-369 */
void orc_ibm_st1(const orc_params *p, orc_state *s);
/* ibm_st2: interpolate (um,vm,wm) with weight phi^3, RK1/2/4 on the frozen
 * field, advect by M*dt, wrap x and z; writes x,y,z,u,v,w and xd.
 * ibm.cu:372-694 */
void orc_ibm_st2(const orc_params *p, orc_state *s);
/* ---- VOF indicator ------------------------------------------------------
 * Slab connectivity only (walls in y, CapsuleSHEARFLOW).  Arrays are padded
 * [n]; s->vf must be set. */
/* front_tracking_source1-3: G = sum over elements of delta * (inward unit
 * normal) * area, spread from the element centroid over the floor-1..ceil+1
 * node box, then the x/z halo contributions folded onto the COMP images and
 * copied back.  fronttrack.cu:274-499 */
void orc_ft_gradient(const orc_params *p, const orc_state *s, double *gx, double *gy, double *gz);
/* front_tracking_source4 + jacobi_source: b = div G / (-6/dx^2) on COMP, 0
 * elsewhere.  fronttrack.cu:502-556, 314-333 */
void orc_ft_rhs(const orc_params *p, const int *bc, const double *gx, const double *gy, const double *gz,
                double *b);
/* One redblacksor call on the Jacobi-scaled front-tracking matrix
 * (fronttrack.cu:27-238): x[n] holds the guess on entry and the iterate on
 * exit; the expand2pow2(n) - n identity rows of the padding start at `pad0`.
 * Returns 0 (converged), 2 (ITRMAX).  matrix_solver.cu:112-216 */
int orc_ft_sor(const orc_params *p, const int *bc, const double *b, double *x, double pad0, int *iters);
/* front_tracking host loop: gradient, rhs, SOR from x = cnt for cnt = 0, 1,
 * ... (error on the 11th attempt); vf = x.  Returns 0 or 5.
 * fronttrack.cu:628-724 */
int orc_front_tracking(const orc_params *p, orc_state *s, int *iters, int *attempts);
/* normal3d_st1/st2: the 2x2x2-cube difference gradient, averaged over the 8
 * cubes around the node; written for 1 <= i < n-1 per axis.
 * thincwlic.cu:12-106 */
void orc_vof_normals(const orc_params *p, const double *vf, double *n1x, double *n1y, double *n1z);
/* The THINC part of wlic1d4md_st1 for the upwind cell value C, orientation
 * alpha = +-1, face velocity uc and isgn (1: upwind cell below the face).
 * thincwlic.cu:187-197 */
double orc_thinc_flux(double c, double alpha, double uc, int isgn, int m, double dt, double dx);
/* thinc_wlic: three directional sweeps (normals, fluxes, COMP update, BUFF
 * copy / WALL zero) on s->vf with the M-averaged velocity (um, vm, wm).
 * thincwlic.cu:108-313, 316-372 */
void orc_thinc_wlic(const orc_params *p, orc_state *s);

/* SetDistanceFromWall_GPU (lbm_cylinder.cu:118-190) for the link from padded
 * node i along direction q: the fraction |x_i - x_wall| / |c_q dx| of the
 * tube-wall crossing, or -1 where the reference records no crossing. */
double orc_tube_fraction(const orc_params *p, int i, int q);

/* Cytoadhesion entry points (orc_adhesion above).  The arrays are the
 * caller's.  orc_adhesion_grid fills nxw, nzw, dxw, dzw and returns the
 * receptor count.  orc_adhesion_init: adhesion_init's state, the neighbour
 * search and the initial attachment to the nearest free receptors
 * (adhesion.cu:316-404, 278-314).  orc_adhesion_step: Adhesion()
 * (adhesion.cu:772-790): spring force, then the search and the Bell
 * attach / detach (adhesion.cu:513-615). */
int    orc_adhesion_grid(const orc_params *p, orc_adhesion *a);
void   orc_adhesion_init(const orc_params *p, orc_state *s, orc_adhesion *a);
void   orc_adhesion_step(const orc_params *p, orc_state *s, orc_adhesion *a, int iter);
double orc_adh_uniform(unsigned long long seed, long long iter, long long sid);

/* Single-point helpers used by the delta-function tests. */
double orc_delta_weight(double r_over_dx);                 /* (1+cos(pi r/2))/4     */

/* ---- membrane ---------------------------------------------------------- */
/* stress(): calc_tension, per-cell ELL assembly + BiCGSTAB, q2f,
 * fourpoint_bend.  Returns the first non-zero BiCGSTAB flag (1 breakdown,
 * 2 not converged) or 0.  stress.cu:29-659 */
int  orc_stress(const orc_params *p, orc_state *s, int iter);
/* calc_tension only (stress.cu:29-121). */
void orc_calc_tension(const orc_params *p, orc_state *s, int iter);
/* Assemble the Jacobi-scaled ELL system of cell c (stress.cu:146-272):
 * ptr[mn], index[7*mn], value[7*mn], b[mn] with mn = expand2pow2(3V).
 * Requires orc_calc_tension first.  Returns mn. */
int  orc_assemble_cell(const orc_params *p, const orc_state *s, int c,
                       int *ptr, int *index, double *value, double *b);
/* bicgstab(): x starts at the given value (the reference passes x=0).
 * Returns 0, 1 (breakdown) or 2 (ITRMAX).  iters_out may be NULL.
 * matrix_solver.cu:279-377 */
int  orc_bicgstab(const int *ptr, const int *index, const double *value,
                  double *x, const double *b, int n, int itrmax,
                  double residual, int *iters_out);
int  orc_expand2pow2(int n);                                 /* etc.cu:425-446 */

/* ---- volume ------------------------------------------------------------- */
/* Per-cell volume and area by the divergence theorem on SoA x,y,z
 * (volume_correction.cu:138-200, 246-281). */
void orc_calc_volume(const orc_params *p, const orc_state *s, double *volume,
                     double *area);
void orc_volume_init(const orc_params *p, orc_state *s);   /* :427-450 */
/* Returns iter-1 as the reference does, or -1 when 1000 passes were hit. */
int  orc_volume_correction(const orc_params *p, orc_state *s); /* :453-524 */

/* ---- topology (init1.cu:626-853, init2.cu:12-400) -------------------- */
/* Builds nln/ln (cyclically sorted ring), lo (ring+self ascending),
 * nle/le, reference contravariant metric nref from the reference positions
 * xr (AoS [3Nv]), ec (edge neighbours) and the flat bending reference
 * c0=1, s0=0.  Returns 0 or 3 (the reference's error(3)). */
int  orc_fem_topology(int nc, const int *voff, const int *eoff,
                      const double *xr, const int *ele,
                      int *nln, int *ln, int *lo, int *nle, int *le,
                      double *nref, int *ec, double *c0, double *s0);

/* ---- one main.cu iteration body with the M gating (main.cu:73-110) ---- */
/* With vof_period > 0: thinc_wlic after ibm_st2 (iter%M==0) and
 * front_tracking (iter%K==0) before the volume correction (main.cu:93-97).
 * Returns 0, or the stress flag (1/2), 3 when volume correction failed, 5
 * when front tracking failed.  vol_iters may be NULL. */
int  orc_step(const orc_params *p, orc_state *s, int iter, int *vol_iters);

/* Discrete Skalak energy sum_e A0_e (1/4)(I1^2 + 2 I1 - 2 I2 + C I2^2)
 * (normalised by Gs), used only to pin the tension/RHS by finite
 * differences.  Uses xd and nref. */
double orc_skalak_energy(const orc_params *p, const orc_state *s);
/* Bending energy sum over hinges of (2/sqrt3) br (1 - cos(theta - theta0))
 * with the sign convention of fourpoint_bend, used to pin its forces; like
 * the force, it leaves out the nearly flat hinges (sin^2 < 1e-6). */
double orc_bending_energy(const orc_params *p, const orc_state *s);

/* OpenMP threads the oracle runs with (omp_get_max_threads). */
int  orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
