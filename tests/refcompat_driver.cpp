// Reference-style driver loop compiled against include/lbmibm_refcompat.hpp
// with stand-in structs of the reference's names (def.h is not shipped); the
// loop body is main.cu:73-110 with the M gating.  Built and run by
// tests/test_refcompat.py: without arguments it only checks that the
// adaptor compiles and links and that default parameters come back; with
// "run" it drives a small channel on the GPU.
#include <cstdio>
#include <cstring>
#include <vector>

#include "lbmibm_refcompat.hpp"

struct process { int iter, fnum; };
struct domain { int restart; };
struct lattice {};
struct cell {};
struct fem {};
struct Wall {};

int main(int argc, char **argv) {
    lbmibm_params p;
    lbmibm_default_params(&p, 16, 12, 10);
    if (p.lx != 16 || p.m != 1) return 2;
    if (argc < 2 || std::strcmp(argv[1], "run") != 0) {
        std::printf("adaptor ok\n");
        return 0;
    }
    p.p = 5e-6;
    lbmibm_cells none{};
    lbmibm_ref::attach(p, none);
    process prc{0, 0};
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
    process back{0, 0};
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
