// Reference-style driver loop compiled against include/lbmibm_refcompat.hpp
// with stand-in structs of the reference's names (def.h is not shipped); the
// loop body is main.cu:73-110 with the M gating.  Built and run by
// tests/test_refcompat.py: without arguments it only checks that the
// adaptor compiles and links and that default parameters come back; with
// "run" it drives a small channel on the GPU, with "vof" the VOF stages,
// and with "cells <in> <out>" a set of cells read from a case file.
#include <cstdio>
#include <cstring>
#include <vector>

#include "lbmibm_refcompat.hpp"

struct process { int iter, fnum, fnum_ad; };
struct domain { int restart; };
struct lattice {};
struct cell {};
struct fem {};
struct Wall {};

template <class T> static bool get(std::FILE *f, T *v, size_t n) { return std::fread(v, sizeof(T), n, f) == n; }
template <class T> static void put(std::FILE *f, const T *v, size_t n) { std::fwrite(v, sizeof(T), n, f); }

// "cells <in> <out>": the case file <in> holds int32 lx, ly, lz, nc, precision,
// steps, FP64 tau, P, B, C, RESIDUAL, int32 voff[nc+1], eoff[nc+1], FP64
// x, y, z [Nv], int32 ele [3Nt].  The cells go through attach, volume_init
// and main.cu's loop with the volume correction every M steps; <out> gets
// FP64 fn [19n], um, vm, wm [n], x, y, z, fx, fy, fz [Nv], volume, area
// [nc] and int32 the last correction's pass count.
static int run_cells(const char *in, const char *out) {
    std::FILE *f = std::fopen(in, "rb");
    if (!f) return 10;
    int hdr[6];
    double par[5];
    if (!get(f, hdr, 6) || !get(f, par, 5)) return 11;
    const int nc = hdr[3], steps = hdr[5];
    std::vector<int> voff(nc + 1), eoff(nc + 1);
    if (!get(f, voff.data(), nc + 1) || !get(f, eoff.data(), nc + 1)) return 11;
    const int nv = voff[nc], nt = eoff[nc];
    std::vector<double> x(nv), y(nv), z(nv);
    std::vector<int> ele(3 * (size_t)nt);
    if (!get(f, x.data(), nv) || !get(f, y.data(), nv) || !get(f, z.data(), nv) || !get(f, ele.data(), 3 * (size_t)nt))
        return 11;
    std::fclose(f);
    lbmibm_params p;
    lbmibm_default_params(&p, hdr[0], hdr[1], hdr[2]);
    p.precision = hdr[4];
    p.df = (2.0 * par[0] - 1.0) / 6.0;  // tau = (6 Df + 1)/2
    p.p = par[1];
    p.b = par[2];
    p.c = par[3];
    p.residual = par[4];
    p.volume_correction = 1;
    p.store_macros = 1;
    lbmibm_cells cs{nc, voff.data(), eoff.data(), x.data(), y.data(), z.data(), ele.data()};
    lbmibm_ref::attach(p, cs);
    process prc{0, 0, 0};
    domain cdo{0};
    lattice ltc;
    cell cel;
    fem fm;
    const int M = p.m;
    int passes = 0;
    volume_init(&cdo, &cel);
    while (prc.iter < steps) {
        prc.iter++;
        if ((prc.iter - 1) % M == 0) stress(&prc, &cdo, &cel, &fm, &ltc);
        if ((prc.iter - 1) % M == 0) ibm_st1(&cdo, &ltc, &cel);
        lbm_d3q19(&prc, &cdo, &ltc);
        if (prc.iter % M == 0) ibm_st2(&cdo, &ltc, &cel, &fm);
        if (prc.iter % M == 0) passes = volume_correction(&cdo, &cel, &fm);
    }
    long long n = 0;
    lbmibm_sizes(lbmibm_ref::context(), &n, nullptr, nullptr, nullptr);
    std::vector<double> fn(19 * (size_t)n), um(n), vm(n), wm(n), fx(nv), fy(nv), fz(nv), v0(nc), vol(nc), area(nc);
    lbmibm_ref::pull_lattice(fn.data(), nullptr, nullptr, nullptr, nullptr, nullptr, um.data(), vm.data(), wm.data());
    lbmibm_ref::pull_cells(x.data(), y.data(), z.data(), nullptr, nullptr, nullptr, fx.data(), fy.data(), fz.data(),
                           nullptr);
    lbmibm_ref::check(lbmibm_download_volume(lbmibm_ref::context(), v0.data(), vol.data(), area.data(), nullptr),
                      "download_volume");
    lbmibm_ref::detach();
    std::FILE *o = std::fopen(out, "wb");
    if (!o) return 12;
    put(o, fn.data(), fn.size());
    for (auto *a : {&um, &vm, &wm}) put(o, a->data(), a->size());
    for (auto *a : {&x, &y, &z, &fx, &fy, &fz}) put(o, a->data(), a->size());
    put(o, vol.data(), vol.size());
    put(o, area.data(), area.size());
    put(o, &passes, 1);
    std::fclose(o);
    std::printf("cells %d vertices %d steps %d passes %d\n", nc, nv, steps, passes);
    return 0;
}

int main(int argc, char **argv) {
    lbmibm_params p;
    lbmibm_default_params(&p, 16, 12, 10);
    if (p.lx != 16 || p.m != 1) return 2;
    if (argc == 4 && std::strcmp(argv[1], "cells") == 0) return run_cells(argv[2], argv[3]);
    if (argc < 2 || (std::strcmp(argv[1], "run") != 0 && std::strcmp(argv[1], "vof") != 0)) {
        std::printf("adaptor ok\n");
        return 0;
    }
    if (std::strcmp(argv[1], "vof") == 0) {
        // one spherical cell (radius 3) and a viscosity contrast: vof_init,
        // then the loop with thinc_wlic every M and front_tracking every K
        int nv = 0, nt = 0;
        lbmibm_ref::check(lbmibm_mesh_icosphere(2, &nv, &nt, nullptr, nullptr), "icosphere");
        std::vector<double> xyz(3 * (size_t)nv), x(nv), y(nv), z(nv);
        std::vector<int> tri(3 * (size_t)nt);
        lbmibm_ref::check(lbmibm_mesh_icosphere(2, &nv, &nt, xyz.data(), tri.data()), "icosphere");
        for (int j = 0; j < nv; j++) {
            x[j] = 3.0 * xyz[3 * j] + 0.3;
            y[j] = 3.0 * xyz[3 * j + 1] - 0.2;
            z[j] = 3.0 * xyz[3 * j + 2] + 0.1;
        }
        int voff[2] = {0, nv}, eoff[2] = {0, nt};
        lbmibm_cells one{1, voff, eoff, x.data(), y.data(), z.data(), tri.data()};
        p.lambda1 = 4.0;
        p.vof_period = 5;  // K
        p.sor_itrmax = 5000;
        p.sor_residual = 1e-8;
        p.volume_correction = 0;
        lbmibm_ref::attach(p, one);
        process prc{0, 0, 0};
        domain cdo{0};
        lattice ltc;
        cell cel;
        fem fm;
        const int M = p.m, K = p.vof_period;
        vof_init(&cdo, &ltc, &cel);
        while (prc.iter < 10) {
            prc.iter++;
            if ((prc.iter - 1) % M == 0) stress(&prc, &cdo, &cel, &fm, &ltc);
            if ((prc.iter - 1) % M == 0) ibm_st1(&cdo, &ltc, &cel);
            lbm_d3q19(&prc, &cdo, &ltc);
            if (prc.iter % M == 0) ibm_st2(&cdo, &ltc, &cel, &fm);
            if (prc.iter % M == 0) thinc_wlic(&cdo, &ltc);
            if (prc.iter % K == 0) front_tracking(&cdo, &ltc, &cel);
        }
        long long n = 0;
        lbmibm_sizes(lbmibm_ref::context(), &n, nullptr, nullptr, nullptr);
        std::vector<double> vf(n);
        lbmibm_ref::check(lbmibm_download_vf(lbmibm_ref::context(), vf.data()), "download_vf");
        double vmax = 0.0, vsum = 0.0;
        const int nx = p.lx + 4, ny = p.ly + 4, nz = p.lz + 4;
        for (int iz = 2; iz < nz - 2; iz++)
            for (int iy = 2; iy < ny - 2; iy++)
                for (int ix = 2; ix < nx - 2; ix++) {
                    const double v = vf[ix + (size_t)nx * (iy + (size_t)ny * iz)];
                    vmax = v > vmax ? v : vmax;
                    vsum += v;
                }
        std::printf("vf max %.4f sum %.3f\n", vmax, vsum);
        lbmibm_ref::detach();
        // a radius-3 sphere: vf near 1 at its centre, volume about 113
        return vmax > 0.95 && vmax < 1.05 && vsum > 100.0 && vsum < 125.0 ? 0 : 5;
    }
    p.p = 5e-6;
    lbmibm_cells none{};
    lbmibm_ref::attach(p, none);
    process prc{0, 0, 0};
    domain cdo{0};
    lattice ltc;
    cell cel;
    fem fm;
    Wall wall;
    const int M = p.m;
    output_data(&prc, &cdo, &ltc, &cel, &fm, &wall, &wall);  // main.cu:65, into ./result
    while (prc.iter < 20) {
        prc.iter++;
        if ((prc.iter - 1) % M == 0) stress(&prc, &cdo, &cel, &fm, &ltc);
        if ((prc.iter - 1) % M == 0) ibm_st1(&cdo, &ltc, &cel);
        lbm_d3q19(&prc, &cdo, &ltc);
        if (prc.iter % M == 0) ibm_st2(&cdo, &ltc, &cel, &fm);
        if (prc.iter % M == 0) (void)volume_correction(&cdo, &cel, &fm);
        if (prc.iter % 10 == 0) output_data(&prc, &cdo, &ltc, &cel, &fm, &wall, &wall);
    }
    // restart.bin was last written at iter 20 with file number 2
    process back{0, 0, 0};
    cdo.restart = 1;
    restart_file(&back, &cdo, &ltc, &cel, &fm, &wall, &wall);
    if (back.iter != 20 || back.fnum != 2) return 4;
    long long n = 0;
    int nv = 0, nt = 0, nc = 0;
    lbmibm_sizes(lbmibm_ref::context(), &n, &nv, &nt, &nc);
    std::vector<double> un(n);
    lbmibm_ref::pull_lattice(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, un.data(), nullptr, nullptr);
    double umax = 0.0;
    for (double u : un) umax = u > umax ? u : umax;
    std::printf("umax %.6e\n", umax);
    lbmibm_ref::detach();
    return umax > 0.0 ? 0 : 3;
}
