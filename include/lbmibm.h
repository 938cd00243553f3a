/*
 * lbmibm.h -- C ABI of the MI355X-native coupled D3Q19-LBM / IBM / Skalak
 * membrane timestep (liblbmibm.so, built from preprocess_src_amd/csrc).
 *
 * The five stage entry points replace the reference's host functions one for
 * one, so a driver keeps main.cu's call sequence and M-gating
 * (main.cu:80-106):
 *
 *   lbmibm_stress            <- void stress(process*, domain*, cell*, fem*, lattice*)   stress.cu:580
 *   lbmibm_ibm_st1           <- void ibm_st1(domain*, lattice*, cell*)                 ibm.cu:299
 *   lbmibm_lbm_d3q19         <- void lbm_d3q19(process*, domain*, lattice*)            lbm.cu:1463
 *   lbmibm_ibm_st2           <- void ibm_st2(domain*, lattice*, cell*, fem*)           ibm.cu:642
 *   lbmibm_volume_correction <- int  volume_correction(domain*, cell*, fem*)           volume_correction.cu:453
 *   lbmibm_volume_init       <- void volume_init(domain*, cell*)                       volume_correction.cu:427
 *
 * Host arrays crossing the boundary use the reference layouts:
 *   - padded lattice of n = (LX/dx+4)(LY/dx+4)(LZ/dx+4) nodes, index
 *     i = ix + nx*iy + nx*ny*iz (init1.cu:145-167);
 *   - populations AoS f[i*19+q] in FP64, D3Q19 order of lbm.cu:1-20, with the
 *     reference's post-stream meaning of ltc->fn (lbm.cu:1575);
 *   - macroscopic fields [n] (dn,un,vn,wn,dm,um,vm,wm), velocities in units
 *     of Uc (lattice velocity = Cr*u, lbm.cu:1116);
 *   - vertex positions SoA x,y,z [Nv] (cel->x) and AoS xd [3Nv] (fem->xd),
 *     element vertex ids ele [3Nt] (cel->ele), global vertex ids.
 * Cells may have different meshes: cell c owns vertices voff[c]..voff[c+1]-1
 * and elements eoff[c]..eoff[c+1]-1.
 *
 * Every call returns an lbmibm_status; failures that the reference answers
 * with exit(1) (stress.cu:364-366, volume_correction.cu:490-493) come back as
 * status codes.  lbmibm_last_error() gives a message.
 */
#ifndef LBMIBM_H
#define LBMIBM_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum lbmibm_status {
    LBMIBM_OK = 0,
    LBMIBM_ERR_ARG = 1,            /* bad argument or size                                  */
    LBMIBM_ERR_HIP = 2,            /* HIP runtime error or no device                        */
    LBMIBM_ERR_BICG_BREAKDOWN = 3, /* BiCGSTAB (r,r^)=0 with b!=0, flag 1 (matrix_solver.cu:334) */
    LBMIBM_ERR_BICG_NOCONV = 4,    /* BiCGSTAB hit ITRMAX, flag 2 (matrix_solver.cu:374)     */
    LBMIBM_ERR_VOLUME = 5,         /* volume correction hit 1000 passes (:487-493)          */
    LBMIBM_ERR_TOPOLOGY = 6,       /* mesh topology check failed, error(3) (init2.cu:161)   */
    LBMIBM_ERR_STATE = 7,          /* call not valid in the current state                   */
    LBMIBM_ERR_VOF = 8             /* front-tracking SOR failed from every start, error(5)
                                      (fronttrack.cu:707-720)                               */
} lbmibm_status;

typedef enum lbmibm_precision {
    LBMIBM_F64 = 0, /* FP64 populations and arithmetic (the reference's precision)            */
    LBMIBM_F32 = 1, /* FP32 populations stored as f - w*rho0, FP32 arithmetic; FP64 inner
                       products and FP64 per-cell volume/area reductions                       */
    LBMIBM_F16 = 2  /* populations stored in FP16 as (f - w*rho0)/w, FP32 arithmetic           */
} lbmibm_precision;

/* Runtime replacement of the def.h macros and cdo_init-derived constants. */
typedef struct lbmibm_params {
    int    lx, ly, lz;       /* fluid nodes per axis, LX/dx, LY/dx, LZ/dx                    */
    int    periodic_y;       /* 0: half-way bounce-back walls at y=+-LY/2 (slab channel of
                                lbm.cu:706-979); 1: fully periodic                           */
    double dx;               /* lattice spacing  dx = Lc/KL   (init1.cu:138)                 */
    double dt;               /* time step        dt = Cr dx/Uc (init1.cu:139)                */
    double cr;               /* Courant number Cr                                            */
    double df;               /* diffusion number Df; tau = (6Df+1)/2 (init1.cu:140)          */
    double lambda1;          /* viscosity ratio LAMBDA1 (lbm.cu:1131)                        */
    double p;                /* body force magnitude P (Guo force 2 rho P e)                 */
    double ex, ey, ez;       /* body force direction e                                       */
    int    force_modulated;  /* 1: e (1 + 0.5 sin(pi t/FREQ)) for iter > FLOWSTART, else 0
                                (lbm.cu:1032-1035); 2: e for iter > FLOWSTART, else 0 (the
                                tube, lbm_cylinder.cu:292-293); 0: constant e (lbm.cu:518)   */
    int    flowstart;        /* FLOWSTART                                                    */
    double freq;             /* FREQ                                                         */
    double b;                /* IBM coupling B (lbm.cu:1116-1118, 1135)                      */
    int    m;                /* membrane sub-cycling M (main.cu:80-101)                      */
    double c;                /* Skalak area dilation C (stress.cu:96)                        */
    double rgs;              /* cel->rg (1.0 for RBCs, init1.cu:524)                         */
    double br;               /* bending rigidity BR (stress.cu:538)                          */
    int    itrmax;           /* BiCGSTAB ITRMAX                                              */
    double residual;         /* BiCGSTAB RESIDUAL (relative)                                 */
    int    runge_kutta;      /* 1, 2 or 4 (ibm.cu:430)                                       */
    int    neohookean;       /* 0: SKALAK, 1: NEOHOOKEAN (stress.cu:95-99)                   */
    int    volume_correction;/* 1: lbmibm_step runs volume correction (main.cu:101-106)      */
    int    precision;        /* lbmibm_precision                                             */
    int    store_macros;     /* 1: keep dn,un,vn,wn and the dm accumulation each step, the
                                tension and rg of each element and the vertex velocity (the
                                reference's outputs); 0: only what the next step reads      */
    int    fileoutput;       /* FILEOUTPUT period of the tp / qi accumulators (0: off)       */
    int    vof_period;       /* K: lbmibm_step runs THINC/WLIC every M steps and front
                                tracking every K steps (main.cu:93-97); 0: no VOF indicator
                                (vf = 0).  Needs cells and walls: the slab or the tube (not
                                periodic_y); LAMBDA1 != 1 needs it                           */
    int    sor_itrmax;       /* ITRMAX of the front-tracking red-black SOR                   */
    double sor_residual;     /* its RESIDUAL (relative, matrix_solver.cu:188-215); the
                                reference shares both constants with BiCGSTAB                */
    double tube_radius;      /* > 0: the CapsuleCHANNELFLOW tube of radius Lc/2 (= R of
                                lbm_cylinder.cu:167) along z: nodes with sqrt(x^2+y^2) > R
                                are walls, z is periodic, links crossing r = R take the
                                interpolated bounce-back (lbm_cylinder.cu:61-235); needs
                                R <= min(LX,LY)/2 - dx and periodic_y = 0. 0: off          */
    int    tube_staircase;   /* with tube_radius > 0: 1 takes the half-way bounce-back of
                                lbm_CapsuleCHANNELFLOW on the tube mask (lbm.cu:1050-1094)
                                on every link into a wall node; 0 the interpolated bounce-
                                back of lbm_cylinder.cu:61-235                               */
    int    wall_motion;      /* slab walls sliding along z, with Ladd's momentum term
                                -6 w rho Cr (c . u_w) on the bounced links (lbm.cu:951-970):
                                0 at rest; 1 both walls, y = +LY/2 at +U and y = -LY/2 at -U
                                (allwall, and lbm_SHEARFLOW, lbm.cu:455-468); 2 only
                                y = +LY/2 at +U (topwall).  The walls rest while
                                iter < FLOWSTART (sra = 0, lbm.cu:755).  Needs the slab     */
    double wall_velocity;    /* U in units of Uc (lattice speed Cr U): sr LY/2 for allwall,
                                sr LY for topwall, ub = 1 in lbm_SHEARFLOW                  */
    int    bicg_warm_start;  /* start of each stress() BiCGSTAB: 0 from x0 = 0 as stress.cu:139
                                (r^ = b); 1 from the cell's previous load q; 2 from the linear
                                extrapolation 2 q_n - q_(n-1) of its last two loads (both zero
                                before the first solves).  1 and 2 take r^ = r0 = b - A x0 and
                                stop at the same relative residual |r|/|b| as 0, so their loads
                                agree with it to that residual.  The adjoint's solves always
                                start from 0                                                   */
    double residual_floor;   /* with FP32 solver arithmetic (FP32, FP16 storage) the stop test
                                uses max(RESIDUAL, residual_floor): below about 1e-7 an FP32
                                iterate no longer changes, only the recursively updated
                                residual keeps falling.  0 keeps RESIDUAL.  FP64 ignores it */
    int    fp32_collide;     /* LBMIBM_F32 only: 0 collides the FP32-stored populations in FP64
                                arithmetic (the lattice step stays bound by its FP32 HBM
                                traffic, and the collision's FP32 rounding does not drift over
                                long runs); 1 collides in FP32                              */
} lbmibm_params;

/* Initial cell set: reference SoA positions and element vertex ids. */
typedef struct lbmibm_cells {
    int           nc;
    const int    *voff;      /* [nc+1] */
    const int    *eoff;      /* [nc+1] */
    const double *x, *y, *z; /* [Nv]   */
    const int    *ele;       /* [3Nt]  */
} lbmibm_cells;

typedef struct lbmibm_ctx lbmibm_ctx;

/* Defaults for an lx*ly*lz slab channel: dx=dt=Cr=1, Df=0.1 (tau=0.8),
 * LAMBDA1=1, no body force, B=0, M=1, C=100, rg=1, BR=0, ITRMAX=1000,
 * RESIDUAL=1e-10, RK4, Skalak, volume correction on, FP32, store_macros=0,
 * no VOF indicator (SOR ITRMAX=1000, RESIDUAL=1e-10), BiCGSTAB from x0 = 0,
 * FP32 residual floor 1e-7, FP32 storage collided in FP64. */
void          lbmibm_default_params(lbmibm_params *p, int lx, int ly, int lz);

/* cdo_init + flu_init + cel_init + fem_init + mem_init (init1.cu:13-1092,
 * init2.cu:258-400): allocates every device buffer once, builds the FEM
 * topology from the cells (may be NULL) and sets the fluid to rest (rho=1,
 * u=0).  volume_init is separate, as in main.cu:60. */
lbmibm_status lbmibm_create(const lbmibm_params *p, const lbmibm_cells *cells, lbmibm_ctx **out);
void          lbmibm_destroy(lbmibm_ctx *ctx);
const char   *lbmibm_last_error(const lbmibm_ctx *ctx);
/* No reference counterpart.  lbmibm_create times the candidate work-group
 * shapes of the on-chip membrane solvers and the forms of the large-cell
 * solver (0 vectors in memory, 1 p in LDS, 2 register-resident) on the
 * cells' own matrices, and the two forms of the interpolation (0 tiles staged
 * in LDS, 1 one thread per vertex reading the field's rows) and of the
 * spreading (0 source bins per tile, 1 home boxes summed by a gather) on the
 * initial positions, and keeps the fastest of each.  In the environment,
 * LBMIBM_AUTOTUNE=0 skips it and uses the first shapes and forms 0;
 * LBMIBM_BICG_SHAPES="a,b[,c]" fixes the variants of the two on-chip classes
 * and the large-cell form (a form the largest cell does not fit falls back to
 * form 0); LBMIBM_INTERP_FORM=0|1 and LBMIBM_SPREAD_FORM=0|1 fix those forms
 * (spread form 1 is refused on lattices much smaller than its boxes).  This
 * returns the timings and choices as text, owned by ctx. */
const char   *lbmibm_autotune_report(const lbmibm_ctx *ctx);
/* Sizes: padded lattice nodes n, vertices, elements, cells. */
void          lbmibm_sizes(const lbmibm_ctx *ctx, long long *n_padded, int *nv, int *nt, int *nc);

/* ---- stage entry points (reference host functions) ------------------- */
lbmibm_status lbmibm_lbm_d3q19(lbmibm_ctx *ctx, int iter);                 /* lbm.cu:1463          */
lbmibm_status lbmibm_ibm_st1(lbmibm_ctx *ctx);                             /* ibm.cu:299           */
lbmibm_status lbmibm_ibm_st2(lbmibm_ctx *ctx);                             /* ibm.cu:642           */
lbmibm_status lbmibm_stress(lbmibm_ctx *ctx, int iter);                    /* stress.cu:580        */
lbmibm_status lbmibm_volume_init(lbmibm_ctx *ctx);                         /* volume_correction.cu:427 */
lbmibm_status lbmibm_volume_correction(lbmibm_ctx *ctx, int *iterations);  /* volume_correction.cu:453;
                                                                              *iterations = iter-1 */
/* One main.cu loop body (main.cu:75-106) with the M gating. */
lbmibm_status lbmibm_step(lbmibm_ctx *ctx, int iter);
/* Steps iter0+1 .. iter0+nsteps without host synchronisation; whole
 * M-cycles are replayed from a captured hipGraph when use_graph != 0.
 * Device-side solver/volume flags are checked once at the end. */
lbmibm_status lbmibm_run(lbmibm_ctx *ctx, int iter0, int nsteps, int use_graph);
lbmibm_status lbmibm_synchronize(lbmibm_ctx *ctx);
/* Timers (no reference counterpart; the reference's timer is commented out,
 * main.cu:67-70): runs steps iter0+1 .. iter0+nsteps eagerly with HIP events
 * on the context's stream around every stage and returns the average
 * milliseconds per executed call in stage_ms[6] and the call counts in
 * stage_calls[6], in the order stress, ibm_st1, lbm_d3q19, ibm_st2,
 * volume_correction, whole step. */
lbmibm_status lbmibm_time_stages(lbmibm_ctx *ctx, int iter0, int nsteps, double *stage_ms, int *stage_calls);
/* Per-kernel timer (no reference counterpart): runs steps iter0+1 ..
 * iter0+nsteps eagerly, as lbmibm_step does, with a HIP event pair on its own
 * stream around each launch of the path's main kernels (k_lbm, binning,
 * k_spread_*, k_interp_*, k_elem, k_vert, the BiCGSTAB classes and forms,
 * k_volume).  report receives one line per kernel, "name\tlaunches\tavg_ms",
 * truncated to report_len bytes. */
lbmibm_status lbmibm_time_kernels(lbmibm_ctx *ctx, int iter0, int nsteps, char *report, int report_len);

/* ---- data exchange in the reference layouts --------------------------- */
/* flu_init: f = f_eq(rho, u) on every node; rho,u,v,w padded [n] or NULL
 * (rho=1, u=0).  Also resets macros and the lattice force. init1.cu:281-449 */
lbmibm_status lbmibm_set_equilibrium(lbmibm_ctx *ctx, const double *rho, const double *u,
                                     const double *v, const double *w);
lbmibm_status lbmibm_upload_populations(lbmibm_ctx *ctx, const double *fn);   /* ltc->fnH -> device */
lbmibm_status lbmibm_download_populations(lbmibm_ctx *ctx, double *fn);       /* COMP nodes filled  */
/* Any pointer may be NULL.  dn,un,vn,wn,dm need store_macros=1.  COMP and
 * BUFF nodes are filled (BUFF = periodic image, lbm.cu:1316-1432). */
lbmibm_status lbmibm_download_macros(lbmibm_ctx *ctx, double *dn, double *un, double *vn, double *wn,
                                     double *dm, double *um, double *vm, double *wm);
lbmibm_status lbmibm_upload_macros(lbmibm_ctx *ctx, const double *um, const double *vm, const double *wm);
lbmibm_status lbmibm_download_lattice_force(lbmibm_ctx *ctx, double *fx, double *fy, double *fz);
lbmibm_status lbmibm_upload_lattice_force(lbmibm_ctx *ctx, const double *fx, const double *fy,
                                          const double *fz);
/* Cell state; any pointer may be NULL.  x,y,z,u,v,w,fx,fy,fz [Nv]; xd [3Nv].
 * The interpolated vertex velocity u,v,w (cel->u..) is kept with
 * store_macros=1, like the other outputs that the step itself does not read. */
lbmibm_status lbmibm_download_cells(lbmibm_ctx *ctx, double *x, double *y, double *z, double *u,
                                    double *v, double *w, double *fx, double *fy, double *fz, double *xd);
lbmibm_status lbmibm_upload_cells(lbmibm_ctx *ctx, const double *x, const double *y, const double *z,
                                  const double *xd, const double *fx, const double *fy, const double *fz);
/* Membrane internals; any pointer may be NULL.  q [3Nv] (fem->q without the
 * per-cell power-of-two padding), t [4Nt], rg [Nt], tp [2Nt], qi [3Nv]. */
lbmibm_status lbmibm_download_membrane(lbmibm_ctx *ctx, double *q, double *t, double *rg, double *tp,
                                       double *qi, int *bicg_iters);
/* Per-cell volume0, volume, area [nc] and centroid gx [3nc] (FP64). */
lbmibm_status lbmibm_download_volume(lbmibm_ctx *ctx, double *volume0, double *volume, double *area,
                                     double *gx);
/* Per-cell force scale (per-cell Gs / reference Gs), [nc]; NULL resets to 1. */
lbmibm_status lbmibm_set_cell_force_scale(lbmibm_ctx *ctx, const double *scale);

/* ---- VOF indicator (vof_period > 0) --------------------------------------
 * The indicator vf of the cells' interiors sets the relaxation rate of the
 * collision, omega = 1/(3(1 + (LAMBDA1-1) vf) Df + 0.5) (lbm.cu:1131).  It
 * is kept in FP64 whatever the lattice precision. */
/* front_tracking (fronttrack.cu:628-724), also vof_init (init1.cu:1095-1123,
 * once before the time loop): vf from the current positions by spreading
 * the elements' inward area vectors, the divergence and a red-black SOR
 * Poisson solve (matrix_solver.cu:112-216) started from x = 0, 1, 2, ... after
 * each failure; LBMIBM_ERR_VOF after 10 failed starts.  *iterations: SOR
 * iterations of the start that converged, *attempts: starts used (either
 * may be NULL).  Synchronises the host once per batch of SOR iterations. */
lbmibm_status lbmibm_front_tracking(lbmibm_ctx *ctx, int *iterations, int *attempts);
/* thinc_wlic (thincwlic.cu:316-372): advects vf with the M-averaged
 * velocity over M dt, one directional sweep per axis.  No host sync. */
lbmibm_status lbmibm_thinc_wlic(lbmibm_ctx *ctx);
/* ltc->vfH on the padded lattice [n]: COMP values, BUFF nodes their periodic
 * image, WALL nodes 0.  Upload takes the COMP values. */
lbmibm_status lbmibm_download_vf(lbmibm_ctx *ctx, double *vf);
lbmibm_status lbmibm_upload_vf(lbmibm_ctx *ctx, const double *vf);

/* ---- cytoadhesion (adhesion.cu; slab wall or tube ring, linear or WLC) --
 * Knob vertices bind receptors (on the wall y = -LY/2, or on the ring r = R
 * of the tube) through springs_per_knob
 * springs each, by the Bell model; the spring force joins the membrane force
 * in the spreading (ibm_st1bb, ibm.cu:133-212).  Once set up, lbmibm_step
 * runs Adhesion() after stress (main.cu:84-86).  Random numbers: splitmix64
 * of (seed, iteration, spring), one draw per spring per call; a receptor
 * wanted by several springs in one call goes to the lowest spring id
 * (vertex * springs_per_knob + spring), the others try their next one.
 * The reference draws from curand seeded by clock() and races on the
 * receptor flag. */
typedef struct lbmibm_adhesion_params {
    double receptor_density;  /* RECEPTOR_DENSITY; receptor spacing sqrt(1/density)/RC   */
    double rc;                /* RC: physical length of one unit                           */
    double k_ad, l0;          /* K_AD (spring constant), L0 (rest length in dx)            */
    double k_on0, k_off0;     /* K_ON0, K_OFF0 (rates, divided by the shear rate)          */
    double chi_beta, kbt;     /* ChiBETA (reactive compliance), kBT                        */
    double shear_rate;        /* cdo->sr                                                   */
    double gs;                /* GS: the spring force is K_AD/GS (r - L0 dx) e             */
    int    limit;             /* LIMIT: search radius in dx                                */
    int    max_neighwall;     /* MAX_NEIGHWALL: receptors kept per knob, nearest first     */
    int    springs_per_knob;  /* NUMSPRING_KNOB                                            */
    double knob;              /* KNOB: expected knobs per cell when is_knob is NULL        */
    unsigned long long seed;
    int    wlc;               /* 0: LINEAR_SPRING; 1: WLC (adhesion.cu:564-566, 677-690)   */
    double lcont, f_wlc, ra;  /* WLC: contour length Lcont (in dx), fWLC, RA               */
    double alpha1, alpha2;    /* WLC: wall repulsion ALPHA1, ALPHA2                        */
} lbmibm_adhesion_params;
/* adhesion_init + wall_init (adhesion.cu:316-466): receptor grid, knobs
 * (is_knob [Nv], the reference's ./input/icosa%d.dat flags, or NULL to draw
 * them with probability KNOB / cell vertices), the neighbour search and the
 * attachment of every spring to the nearest free receptors.  Needs cells and
 * walls (the slab or the tube; not periodic_y). */
lbmibm_status lbmibm_adhesion_setup(lbmibm_ctx *ctx, const lbmibm_adhesion_params *a, const unsigned char *is_knob);
/* Adhesion() (adhesion.cu:772-790): spring force, then search + Bell. */
lbmibm_status lbmibm_adhesion(lbmibm_ctx *ctx);
/* Any pointer may be NULL: is_knob [Nv], id_wall / r_wall [Nv*MAXN], attach
 * (receptor or -1) / r_ij [Nv*NS], is_attached [receptors], f_adh [3Nv] AoS. */
lbmibm_status lbmibm_adhesion_state(lbmibm_ctx *ctx, unsigned char *is_knob, int *id_wall, double *r_wall,
                                    int *attach, double *r_ij, int *is_attached, double *f_adh);
/* The file output of Adhesion() (adhesion.cu:785-822): writes
 * dir/adhesion/adhe%05d.bin (fnum_ad, iter, f_adh [3Nv], attachment_point
 * [Nv*NS], is_knob [Nv] as bytes, is_attached [receptors], r_ij [Nv*NS]) and
 * dir/adhesion/restart_adhe.dat, then increments *fnum_ad.  The reference
 * calls it when (iter-1) % FILEOUTPUT_ADHESION == 0. */
lbmibm_status lbmibm_write_adhesion(lbmibm_ctx *ctx, const char *dir, int iter, int *fnum_ad);

/* ---- the reference's binary files (etc.cu) ------------------------------
 * Byte layouts of output_data / restart_file: native-endian int32 and FP64
 * arrays one after another, no headers.  The cell sections assume one mesh
 * for all cells (cel->vertex, cel->element): a context whose cells differ
 * gets LBMIBM_ERR_ARG.  Fields this path does not model are written as the
 * reference initialises them and ignored on reading: vf2, the indicator of a
 * second species (0, init1.cu:396-411), and vf when vof_period = 0.  dm (the density sum) and the vertex
 * velocities u, v, w are kept only with store_macros (0 otherwise); the
 * per-cell rx, ry, rz, which no reference code assigns, are written as the
 * mean of each cell's vertex positions. */
/* bc codes of the padded lattice (cdo_init, init1.cu:164-178); the
 * reference's numeric values were in the absent def.h. */
typedef enum lbmibm_bc { LBMIBM_BC_COMP = 0, LBMIBM_BC_BUFF = 1, LBMIBM_BC_WALL = 2 } lbmibm_bc;
lbmibm_status lbmibm_bc_mask(const lbmibm_ctx *ctx, int *bc);                 /* ltc->bcH [n]    */
/* restart.bin (output_data, etc.cu:738-789): iter, fnum, fn [n*19], dm, um,
 * vm, wm, vf, vf2 [n], then x, y, z [Nv], rx, ry, rz [nc], u, v, w [Nv],
 * qi [nc*mn] (cell c at c*3V, mn = expand2pow2(3V), init2.cu:286), tp [2Nt]. */
lbmibm_status lbmibm_write_restart(lbmibm_ctx *ctx, const char *path, int iter, int fnum);
/* restart_file (etc.cu:164-323): reads restart.bin into the context
 * (populations, um..wm, dm with store_macros, positions with fem->x from them,
 * u, v, w, qi, tp) and returns iter and fnum, both 0 when fnum > 100
 * (etc.cu:205).  The file size must match this lattice and these cells. */
lbmibm_status lbmibm_read_restart(lbmibm_ctx *ctx, const char *path, int *iter, int *fnum);
/* prof%04d.bin (etc.cu:547-560): iter, dm, um, vm, wm, vf, vf2 [n], bc [n]. */
lbmibm_status lbmibm_write_prof(lbmibm_ctx *ctx, const char *path, int iter);
/* cell%04d.bin (etc.cu:565-611): n, vertex, element, x, y, z, u, v, w [Nv],
 * qi [3Nv], tp [2Nt], ele [3Nt] (global vertex ids). */
lbmibm_status lbmibm_write_cell(lbmibm_ctx *ctx, const char *path);
/* output_data's file part (etc.cu:544-789): dir/prof%04d.bin when
 * fnum % 50 == 0, dir/cell/cell%04d.bin (the directory must exist), then
 * dir/restart%04d.bin when fnum % 100 == 0 and dir/restart.bin otherwise;
 * increments *fnum. */
lbmibm_status lbmibm_output_data(lbmibm_ctx *ctx, const char *dir, int iter, int *fnum);

/* ---- reverse mode (no reference counterpart) --------------------------
 * Adjoint (vector-Jacobian product) entry points for the main.cu loop body
 * with M = 1, RUNGE_KUTTA = 4, BR = 0, no volume correction and no FILEOUTPUT
 * accumulators, in FP64 or FP32 (not FP16 storage).  The context holds the
 * adjoint state below; each stage VJP is taken at the current forward state,
 * i.e. it differentiates the stage as it would run next, and leaves the
 * forward state unchanged (except that lbmibm_adj_stress recomputes the
 * membrane force at the current positions). */
enum {
    LBMIBM_ADJ_POPULATIONS = 0,     /* [n*19] AoS, the reference's post-stream layout of ltc->fn  */
    LBMIBM_ADJ_VELOCITY = 1,        /* [3n] padded blocks x|y|z: the velocity field u (um..wm)     */
    LBMIBM_ADJ_LATTICE_FORCE = 2,   /* [3n]: the lattice IBM force (ltc->fx..)                     */
    LBMIBM_ADJ_POSITIONS = 3,       /* [3Nv] SoA blocks x|y|z: vertex positions (cel->x.., fem->xd) */
    LBMIBM_ADJ_NODAL_FORCE = 4,     /* [3Nv]: nodal forces (cel->fx..)                             */
    LBMIBM_ADJ_GRAD_BODY_FORCE = 5, /* [3n]: accumulated gradient w.r.t. the body-force field      */
    LBMIBM_ADJ_GRAD_FORCE_SCALE = 6 /* [nc]: accumulated gradient w.r.t. the per-cell force scale  */
};
/* Per-node body-force field g, added to P e in the Guo force 2 rho (P e + B F
 * + g); three padded [n] arrays (COMP nodes read), or all NULL to remove it. */
lbmibm_status lbmibm_set_body_force_field(lbmibm_ctx *ctx, const double *gx, const double *gy, const double *gz);
/* Zeroes every adjoint buffer and gradient (allocated on first use). */
lbmibm_status lbmibm_adjoint_reset(lbmibm_ctx *ctx);
/* Exchange of one adjoint buffer (which = LBMIBM_ADJ_*) in the layouts above;
 * padded entries that are not COMP nodes read back as 0 and are ignored. */
lbmibm_status lbmibm_adjoint_get(lbmibm_ctx *ctx, int which, double *out);
lbmibm_status lbmibm_adjoint_set(lbmibm_ctx *ctx, int which, const double *in);
/* Collide-stream: POPULATIONS and VELOCITY (adjoints of the step's outputs) ->
 * POPULATIONS (of its input) and LATTICE_FORCE; adds to GRAD_BODY_FORCE and
 * clears VELOCITY. */
lbmibm_status lbmibm_adj_lbm_d3q19(lbmibm_ctx *ctx, int iter);
/* Spreading: LATTICE_FORCE -> NODAL_FORCE; adds to POSITIONS. */
lbmibm_status lbmibm_adj_ibm_st1(lbmibm_ctx *ctx);
/* Interpolation + RK4 advection: POSITIONS (after) -> POSITIONS (before);
 * adds to VELOCITY. */
lbmibm_status lbmibm_adj_ibm_st2(lbmibm_ctx *ctx);
/* Skalak membrane force (tension, consistent-mass solve, q2f): NODAL_FORCE ->
 * adds to POSITIONS and GRAD_FORCE_SCALE (one transposed solve per cell). */
lbmibm_status lbmibm_adj_stress(lbmibm_ctx *ctx, int iter);
/* Wall-shear objective J = mean over the last `window` of steps iter0+1 ..
 * iter0+nsteps of tau_w = 2 Df (dx/dt) (e . u) at the first fluid nodes of
 * both y walls, averaged over the walls (e = unit body-force direction).
 * Runs the steps. */
lbmibm_status lbmibm_wall_shear(lbmibm_ctx *ctx, int iter0, int nsteps, int window, double *J);
/* J as lbmibm_wall_shear, allocating nothing: no counterpart in the
 * reference.  For callers that cannot allocate device memory once set up
 * (harnesses that forbid it, loops that evaluate J repeatedly).  Its partial
 * sums (one per 64 wall columns, summed in a fixed order) live in a static
 * device array of the library, 65,537 doubles (512 KiB) per device, held
 * until the process ends: nx*nz up to 4,194,304 columns, larger lattices get
 * LBMIBM_ERR_ARG.  Every call on a device shares that array (each call
 * clears it first), so calls that overlap on one device race. */
lbmibm_status lbmibm_wall_shear_noalloc(lbmibm_ctx *ctx, int iter0, int nsteps, int window, double *J);
/* J as lbmibm_wall_shear and its gradient with respect to the body-force
 * field (padded [n] arrays, any may be NULL) and the per-cell force scales
 * ([nc] or NULL), by checkpoints every `checkpoint_every` steps and per-segment
 * recomputation (device memory ~ checkpoint_every steps of populations,
 * fields and positions).  The populations and positions end at step
 * iter0+nsteps, as after lbmibm_run. */
lbmibm_status lbmibm_adjoint_wall_shear(lbmibm_ctx *ctx, int iter0, int nsteps, int window, int checkpoint_every,
                                        double *J, double *gx, double *gy, double *gz, double *dscale);

/* ---- host-side preprocessing (no GPU needed) ------------------------- */
/* Icosphere of the given subdivision level on the unit sphere, outward
 * triangles: 10*4^level+2 vertices, 20*4^level triangles, vertices numbered
 * along a Morton curve and triangles sorted by their lowest vertex id (each
 * rotated to start there) for memory locality.  Call with NULL arrays to get
 * the counts.  Replaces ./cell/icosa_div files (init1.cu:490). */
lbmibm_status lbmibm_mesh_icosphere(int level, int *nv, int *nt, double *xyz, int *tri);
/* Evans-Fung biconcave map of unit-sphere points (axis z), in place; replaces
 * ./cell/rbc_div files (init1.cu:506). */
lbmibm_status lbmibm_mesh_biconcave(int nv, double *xyz);
/* Seeded placement of nc cells of radius r (bounding sphere) in the domain:
 * mode 0 = random sequential addition (centre distance >= 2r+1, wall
 * distance >= r+1), mode 1 = jittered grid of gx*gy*gz slots (nc must fit).
 * Writes centres [3nc] and rotation matrices [9nc] (Shoemake-uniform). */
lbmibm_status lbmibm_place_cells(int nc, double r, double lxd, double lyd, double lzd, int periodic_y,
                                 unsigned long long seed, int mode, int gx, int gy, int gz,
                                 double *centres, double *rot);
/* fem_init / cel_init connectivity (init1.cu:626-853, init2.cu:12-400).
 * LBMIBM_ERR_TOPOLOGY where the reference calls error(3): a vertex with more
 * than 6 neighbours or 6 incident elements, an edge count that breaks
 * Euler's formula, an unmatched edge, a degenerate reference triangle, or a
 * cell of fewer than 86 vertices (per-cell system below 256 rows,
 * init2.cu:286-287). */
lbmibm_status lbmibm_build_topology(int nc, const int *voff, const int *eoff, const double *xr,
                                    const int *ele, int *nln, int *ln, int *lo, int *nle, int *le,
                                    double *nref, int *ec, double *c0, double *s0);

#ifdef __cplusplus
}
#endif
#endif
