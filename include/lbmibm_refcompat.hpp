/*
 * lbmibm_refcompat.hpp -- header-only adaptor that gives a reference-style
 * driver (main.cu) the five stage functions it calls, by name, on top of the
 * C ABI of lbmibm.h.
 *
 * The reference declares these in its def.h (not shipped) with C++ linkage
 * and struct-pointer arguments:
 *   void lbm_d3q19(process*, domain*, lattice*)                 lbm.cu:1463
 *   void ibm_st1(domain*, lattice*, cell*)                       ibm.cu:299
 *   void ibm_st2(domain*, lattice*, cell*, fem*)                 ibm.cu:642
 *   void stress(process*, domain*, cell*, fem*, lattice*)        stress.cu:580
 *   int  volume_correction(domain*, cell*, fem*)                 volume_correction.cu:453
 *   void volume_init(domain*, cell*)                             volume_correction.cu:427
 *   void output_data(process*, domain*, lattice*, cell*, fem*, Wall*, Wall*)   etc.cu:449
 *   void restart_file(process*, domain*, lattice*, cell*, fem*, Wall*, Wall*)  etc.cu:164
 * The adaptor's functions are templates over those struct types, so they
 * compile against whatever def.h defines; they read only prc->iter (and
 * prc->fnum, cdo->restart for the two file functions).  The
 * device state lives in one lbmibm_ctx created by lbmibm_ref::attach() after
 * the reference's own host initialisation (cdo_init, flu_init, cel_init,
 * fem_init), which replaces mem_init.  lbmibm_ref::pull_lattice() and
 * pull_cells() fill the reference's host mirrors (fnH, dnH..wmH, cel->x..,
 * fem->x) for output_data / restart_file.
 *
 * Errors: the reference exits the process on a failed solve or volume
 * correction (stress.cu:364-366, volume_correction.cu:490-493); the adaptor
 * does the same after printing lbmibm_last_error(), so main.cu's behaviour is
 * unchanged.
 */
#ifndef LBMIBM_REFCOMPAT_HPP
#define LBMIBM_REFCOMPAT_HPP

#include <cstdio>
#include <cstdlib>

#include "lbmibm.h"

namespace lbmibm_ref {

inline lbmibm_ctx *&context() {
    static lbmibm_ctx *ctx = nullptr;
    return ctx;
}

inline void check(lbmibm_status s, const char *where) {
    if (s == LBMIBM_OK) return;
    std::fprintf(stderr, "%s: %s\n", where, lbmibm_last_error(context()));
    std::exit(1);
}

/* Creates the device context.  The caller fills p from its def.h constants
 * (LX/LY/LZ over dx, Cr, Df, LAMBDA1, P, B, M, K, C, GS, BR, ITRMAX,
 * RESIDUAL, RUNGE_KUTTA) and cells from cel->x/y/z and cel->ele (one mesh per cell
 * species or any mix, as voff/eoff say). */
inline void attach(const lbmibm_params &p, const lbmibm_cells &cells) {
    check(lbmibm_create(&p, cells.nc > 0 ? &cells : nullptr, &context()), "lbmibm_create");
}

inline void detach() {
    lbmibm_destroy(context());
    context() = nullptr;
}

/* Host mirrors in the reference layouts: fn [n*19] post-stream populations,
 * macros [n]; any pointer may be null. */
inline void pull_lattice(double *fn, double *dn, double *un, double *vn, double *wn, double *dm, double *um,
                         double *vm, double *wm) {
    if (fn) check(lbmibm_download_populations(context(), fn), "lbmibm_download_populations");
    check(lbmibm_download_macros(context(), dn, un, vn, wn, dm, um, vm, wm), "lbmibm_download_macros");
}

inline void pull_cells(double *x, double *y, double *z, double *u, double *v, double *w, double *fx, double *fy,
                       double *fz, double *xd) {
    check(lbmibm_download_cells(context(), x, y, z, u, v, w, fx, fy, fz, xd), "lbmibm_download_cells");
}

}  // namespace lbmibm_ref

template <class Process, class Domain, class Lattice>
inline void lbm_d3q19(Process *prc, Domain *, Lattice *) {
    lbmibm_ref::check(lbmibm_lbm_d3q19(lbmibm_ref::context(), prc->iter), "lbm_d3q19");
}

template <class Domain, class Lattice, class Cell>
inline void ibm_st1(Domain *, Lattice *, Cell *) {
    lbmibm_ref::check(lbmibm_ibm_st1(lbmibm_ref::context()), "ibm_st1");
}

template <class Domain, class Lattice, class Cell, class Fem>
inline void ibm_st2(Domain *, Lattice *, Cell *, Fem *) {
    lbmibm_ref::check(lbmibm_ibm_st2(lbmibm_ref::context()), "ibm_st2");
}

template <class Process, class Domain, class Cell, class Fem, class Lattice>
inline void stress(Process *prc, Domain *, Cell *, Fem *, Lattice *) {
    lbmibm_ref::check(lbmibm_stress(lbmibm_ref::context(), prc->iter), "stress");
}

template <class Domain, class Cell>
inline void volume_init(Domain *, Cell *) {
    lbmibm_ref::check(lbmibm_volume_init(lbmibm_ref::context()), "volume_init");
}

template <class Domain, class Cell, class Fem>
inline int volume_correction(Domain *, Cell *, Fem *) {
    int iterations = 0;
    lbmibm_ref::check(lbmibm_volume_correction(lbmibm_ref::context(), &iterations), "volume_correction");
    return iterations;
}

// vof_init (init1.cu:1095-1123) and the two VOF stages of the loop
// (main.cu:93-97).  They need p.vof_period > 0 (K) at attach; lbmibm_step
// calls the same two stages itself.
template <class Domain, class Lattice, class Cell>
inline void vof_init(Domain *, Lattice *, Cell *) {
    std::printf("Initializing volume fraction ---------------------------------------\n");
    lbmibm_ref::check(lbmibm_front_tracking(lbmibm_ref::context(), nullptr, nullptr), "vof_init");
}

template <class Domain, class Lattice, class Cell>
inline void front_tracking(Domain *, Lattice *, Cell *) {
    lbmibm_ref::check(lbmibm_front_tracking(lbmibm_ref::context(), nullptr, nullptr), "front_tracking");
}

template <class Domain, class Lattice>
inline void thinc_wlic(Domain *, Lattice *) {
    lbmibm_ref::check(lbmibm_thinc_wlic(lbmibm_ref::context()), "thinc_wlic");
}

// Cytoadhesion: wall_init + adhesion_init in one call, given the def.h
// constants and the knob flags of ./input/icosa%d.dat (or NULL to draw
// them), then Adhesion() of the loop (main.cu:84-86), which lbmibm_step also
// runs by itself once set up.  With FILEOUTPUT_ADHESION > 0, Adhesion()
// writes ./result/adhesion/adhe%05d.bin and restart_adhe.dat when
// (iter-1) % FILEOUTPUT_ADHESION == 0 and advances prc->fnum_ad
// (adhesion.cu:785-822).
namespace lbmibm_ref {
inline int &adhesion_file_period() {
    static int k = 0;
    return k;
}
inline void attach_adhesion(const lbmibm_adhesion_params &a, const unsigned char *is_knob,
                            int fileoutput_adhesion = 0) {
    check(lbmibm_adhesion_setup(context(), &a, is_knob), "adhesion_init");
    adhesion_file_period() = fileoutput_adhesion;
}
}  // namespace lbmibm_ref

template <class Process, class Domain, class Cell, class Wall>
inline void Adhesion(Process *prc, Domain *, Cell *, Wall *, Wall *) {
    lbmibm_ref::check(lbmibm_adhesion(lbmibm_ref::context()), "Adhesion");
    const int k = lbmibm_ref::adhesion_file_period();
    if (k > 0 && (prc->iter - 1) % k == 0)
        lbmibm_ref::check(lbmibm_write_adhesion(lbmibm_ref::context(), "./result", prc->iter, &prc->fnum_ad),
                          "Adhesion file output");
}

// The binary files of output_data into ./result with the reference's names
// (prof%04d.bin every 50th file, cell/cell%04d.bin, restart.bin or
// restart%04d.bin every 100th); increments prc->fnum.  The text files
// (config.dat, center.dat, surface_volume.dat) are not written.
template <class Process, class Domain, class Lattice, class Cell, class Fem, class Wall>
inline void output_data(Process *prc, Domain *, Lattice *, Cell *, Fem *, Wall *, Wall *) {
    std::printf("Result file %04d writing...\n", prc->fnum);
    lbmibm_ref::check(lbmibm_output_data(lbmibm_ref::context(), "./result", prc->iter, &prc->fnum), "output_data");
}

// With cdo->restart set, ./result/restart.bin back into the context, and
// prc->iter, prc->fnum from it (both 0 when fnum > 100, etc.cu:205).
template <class Process, class Domain, class Lattice, class Cell, class Fem, class Wall>
inline void restart_file(Process *prc, Domain *cdo, Lattice *, Cell *, Fem *, Wall *, Wall *) {
    if (cdo->restart == 0) return;
    std::printf("Initializing restart process\n");
    lbmibm_ref::check(lbmibm_read_restart(lbmibm_ref::context(), "./result/restart.bin", &prc->iter, &prc->fnum),
                      "restart_file");
}

#endif
