// The reference's binary files: restart.bin written and read back, and the
// prof%04d.bin / cell%04d.bin outputs (etc.cu output_data:449-961 and
// restart_file:164-323), in its byte layout: native-endian int32 and FP64
// arrays one after another, no headers.  Host code on top of the exchange
// entry points, plus the few device buffers (accumulators, the density sum)
// that those do not upload.
//
// The reference's cell arrays assume one mesh for all cells (cel->vertex,
// cel->element); a context whose cells differ in mesh size is refused.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

using namespace lbmibm;

namespace {

// expand2pow2 (etc.cu:425-446): the power of two the ELL rows are padded to
int expand_pow2(int n) {
    int cnt = 0;
    while (n > 1) {
        if (n % 2 == 1) n++;
        n /= 2;
        cnt++;
    }
    return 1 << cnt;
}

struct Layout {
    long long n;     // padded lattice nodes
    int nc, V, E;    // cells, vertices and elements per cell
    int mn;          // padded system rows per cell (init2.cu:286)
};

lbmibm_status layout_of(lbmibm_ctx *ctx, Layout &L, bool need_cells) {
    L.n = ctx->np;
    L.nc = ctx->nc;
    L.V = L.E = 0;
    L.mn = 0;
    if (ctx->nc == 0) return need_cells ? LBMIBM_ERR_STATE : LBMIBM_OK;
    std::vector<int> voff(ctx->nc + 1), eoff(ctx->nc + 1);
    LBMIBM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    LBMIBM_HIP(ctx, hipMemcpy(voff.data(), ctx->voff, voff.size() * sizeof(int), hipMemcpyDeviceToHost));
    LBMIBM_HIP(ctx, hipMemcpy(eoff.data(), ctx->eoff, eoff.size() * sizeof(int), hipMemcpyDeviceToHost));
    L.V = voff[1] - voff[0];
    L.E = eoff[1] - eoff[0];
    for (int c = 0; c < ctx->nc; c++)
        if (voff[c + 1] - voff[c] != L.V || eoff[c + 1] - eoff[c] != L.E) {
            ctx->err = "reference file formats need every cell on the same mesh (cel->vertex, cel->element)";
            return LBMIBM_ERR_ARG;
        }
    L.mn = expand_pow2(3 * L.V);
    return LBMIBM_OK;
}

struct Out {
    FILE *f = nullptr;
    bool ok = true;
    ~Out() {
        if (f) fclose(f);
    }
    template <typename V> void put(const V *p, size_t n) {
        if (ok && n && fwrite(p, sizeof(V), n, f) != n) ok = false;
    }
    template <typename V> void put(const std::vector<V> &a) { put(a.data(), a.size()); }
    void zeros(size_t n) {
        std::vector<double> z(std::min<size_t>(n, 1 << 16), 0.0);
        while (ok && n) {
            const size_t k = std::min(n, z.size());
            put(z.data(), k);
            n -= k;
        }
    }
};

lbmibm_status open_out(lbmibm_ctx *ctx, Out &o, const char *path) {
    if (!path) return LBMIBM_ERR_ARG;
    o.f = fopen(path, "wb");
    if (!o.f) {
        ctx->err = std::string("cannot open ") + path + " for writing";
        return LBMIBM_ERR_ARG;
    }
    return LBMIBM_OK;
}

lbmibm_status close_out(lbmibm_ctx *ctx, Out &o, const char *path) {
    const bool ok = o.ok && fclose(o.f) == 0;
    o.f = nullptr;
    if (!ok) {
        ctx->err = std::string("write to ") + path + " failed";
        return LBMIBM_ERR_ARG;
    }
    return LBMIBM_OK;
}

// The fields of a lattice record: dm (the density sum, zero without
// store_macros), um, vm, wm (the M-averaged velocity / Cr), vf (the VOF
// indicator, 0 without vof_period) and vf2 (the indicator of a second cell
// species, not modelled: 0, the reference's initial value, init1.cu:396-411).
lbmibm_status put_lattice_fields(lbmibm_ctx *ctx, Out &o, long long n) {
    std::vector<double> dm(n, 0.0), um(n), vm(n), wm(n);
    lbmibm_status s = lbmibm_download_macros(ctx, nullptr, nullptr, nullptr, nullptr,
                                             ctx->p.store_macros ? dm.data() : nullptr, um.data(), vm.data(),
                                             wm.data());
    if (s != LBMIBM_OK) return s;
    o.put(dm);
    o.put(um);
    o.put(vm);
    o.put(wm);
    if (ctx->p.vof_period > 0) {
        std::vector<double> vf(n);
        s = lbmibm_download_vf(ctx, vf.data());
        if (s != LBMIBM_OK) return s;
        o.put(vf);
    } else {
        o.zeros((size_t)n);
    }
    o.zeros((size_t)n);
    return LBMIBM_OK;
}

// Per-cell centres: the reference writes cel->rx,ry,rz, which no code on the
// path assigns (init1.cu:547, 937); written here as the mean of each cell's
// vertex positions.
void cell_means(const Layout &L, const std::vector<double> &x, std::vector<double> &r) {
    r.assign(L.nc, 0.0);
    for (int c = 0; c < L.nc; c++) {
        double s = 0.0;
        for (int j = 0; j < L.V; j++) s += x[(size_t)c * L.V + j];
        r[c] = s / L.V;
    }
}

}  // namespace

extern "C" {

lbmibm_status lbmibm_bc_mask(const lbmibm_ctx *ctx, int *bc) {
    if (!ctx || !bc) return LBMIBM_ERR_ARG;
    // cdo_init (init1.cu:164-178): COMP inside, the two-node halo BUFF where
    // it is a periodic image (x, z; y when periodic), WALL beyond the walls
    for (int pz = 0; pz < ctx->nzp; pz++)
        for (int py = 0; py < ctx->nyp; py++)
            for (int px = 0; px < ctx->nxp; px++) {
                const bool inx = px >= 2 && px <= ctx->nxp - 3, iny = py >= 2 && py <= ctx->nyp - 3,
                           inz = pz >= 2 && pz <= ctx->nzp - 3;
                const long long i = px + (long long)ctx->nxp * (py + (long long)ctx->nyp * pz);
                if (ctx->p.tube_radius > 0.0) {  // init1.cu:217-229
                    bc[i] = !tube_inside(ctx, px - 2, py - 2) ? LBMIBM_BC_WALL : (inz ? LBMIBM_BC_COMP : LBMIBM_BC_BUFF);
                    continue;
                }
                bc[i] = inx && iny && inz ? LBMIBM_BC_COMP
                                          : (iny || ctx->p.periodic_y ? LBMIBM_BC_BUFF : LBMIBM_BC_WALL);
            }
    return LBMIBM_OK;
}

lbmibm_status lbmibm_write_restart(lbmibm_ctx *ctx, const char *path, int iter, int fnum) {
    if (!ctx) return LBMIBM_ERR_ARG;
    Layout L;
    lbmibm_status s = layout_of(ctx, L, false);
    if (s != LBMIBM_OK) return s;
    std::vector<double> fn((size_t)L.n * Q);
    s = lbmibm_download_populations(ctx, fn.data());
    if (s != LBMIBM_OK) return s;
    Out o;
    s = open_out(ctx, o, path);
    if (s != LBMIBM_OK) return s;
    o.put(&iter, 1);
    o.put(&fnum, 1);
    o.put(fn);
    s = put_lattice_fields(ctx, o, L.n);
    if (s != LBMIBM_OK) return s;
    if (L.nc > 0) {
        const size_t nv = (size_t)L.nc * L.V, ne = (size_t)L.nc * L.E;
        std::vector<double> x(nv), y(nv), z(nv), u(nv), v(nv), w(nv), qi(3 * nv), tp(2 * ne), r;
        s = lbmibm_download_cells(ctx, x.data(), y.data(), z.data(), u.data(), v.data(), w.data(), nullptr,
                                  nullptr, nullptr, nullptr);
        if (s == LBMIBM_OK)
            s = lbmibm_download_membrane(ctx, nullptr, nullptr, nullptr, tp.data(), qi.data(), nullptr);
        if (s != LBMIBM_OK) return s;
        o.put(x);
        o.put(y);
        o.put(z);
        for (const auto *a : {&x, &y, &z}) {
            cell_means(L, *a, r);
            o.put(r);
        }
        o.put(u);
        o.put(v);
        o.put(w);
        // fem->qi: [nc * mn], cell c's loads at c*3V.. (substitute_load,
        // stress.cu:292-300), the rest of the allocation unused
        o.put(qi);
        o.zeros((size_t)L.nc * L.mn - 3 * nv);
        o.put(tp);
    }
    return close_out(ctx, o, path);
}

lbmibm_status lbmibm_read_restart(lbmibm_ctx *ctx, const char *path, int *iter, int *fnum) {
    if (!ctx || !path) return LBMIBM_ERR_ARG;
    Layout L;
    lbmibm_status s = layout_of(ctx, L, false);
    if (s != LBMIBM_OK) return s;
    const size_t nv = (size_t)L.nc * L.V, ne = (size_t)L.nc * L.E;
    const size_t want = 2 * sizeof(int) + sizeof(double) * ((size_t)L.n * Q + 6 * (size_t)L.n) +
                        (L.nc > 0 ? sizeof(double) * (6 * nv + 3 * (size_t)L.nc + (size_t)L.nc * L.mn + 2 * ne) : 0);
    FILE *f = fopen(path, "rb");
    if (!f) {
        ctx->err = std::string("cannot open ") + path;
        return LBMIBM_ERR_ARG;
    }
    std::vector<char> buf;
    fseek(f, 0, SEEK_END);
    const long size = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (size < 0 || (size_t)size != want) {
        fclose(f);
        ctx->err = std::string(path) + ": size does not match this lattice and these cells";
        return LBMIBM_ERR_ARG;
    }
    buf.resize(want);
    const bool ok = fread(buf.data(), 1, want, f) == want;
    fclose(f);
    if (!ok) {
        ctx->err = std::string("read of ") + path + " failed";
        return LBMIBM_ERR_ARG;
    }
    size_t off = 0;
    auto take_i = [&]() { int v; std::memcpy(&v, buf.data() + off, sizeof v); off += sizeof v; return v; };
    auto take_d = [&](size_t n) {
        std::vector<double> a(n);
        std::memcpy(a.data(), buf.data() + off, n * sizeof(double));
        off += n * sizeof(double);
        return a;
    };
    int it = take_i(), fno = take_i();
    if (fno > 100) { it = 0; fno = 0; }  // as restart_file (etc.cu:205)
    const std::vector<double> fn = take_d((size_t)L.n * Q);
    const std::vector<double> dm = take_d(L.n), um = take_d(L.n), vm = take_d(L.n), wm = take_d(L.n);
    const std::vector<double> vf = take_d(L.n);
    off += (size_t)L.n * sizeof(double);  // vf2: one cell species here
    s = lbmibm_upload_populations(ctx, fn.data());
    if (s == LBMIBM_OK && ctx->p.vof_period > 0) s = lbmibm_upload_vf(ctx, vf.data());
    if (s == LBMIBM_OK) s = lbmibm_upload_macros(ctx, um.data(), vm.data(), wm.data());
    if (s == LBMIBM_OK && ctx->p.store_macros) {
        std::vector<double> a(ctx->N);
        for (int iz = 0; iz < ctx->nz; iz++)
            for (int iy = 0; iy < ctx->ny; iy++)
                for (int ix = 0; ix < ctx->nx; ix++)
                    a[ix + (size_t)ctx->nx * (iy + (size_t)ctx->ny * iz)] = dm[padded_index(ctx, ix, iy, iz)];
        s = put_T(ctx, ctx->rhom, a.data(), (size_t)ctx->N);
    }
    if (s == LBMIBM_OK && L.nc > 0) {
        const std::vector<double> x = take_d(nv), y = take_d(nv), z = take_d(nv);
        off += 3 * (size_t)L.nc * sizeof(double);  // rx, ry, rz
        const std::vector<double> u = take_d(nv), v = take_d(nv), w = take_d(nv);
        const std::vector<double> qi = take_d((size_t)L.nc * L.mn), tp = take_d(2 * ne);
        // fem->x from cel->x (etc.cu:254-262)
        std::vector<double> xd(3 * nv);
        for (size_t j = 0; j < nv; j++) { xd[3 * j] = x[j]; xd[3 * j + 1] = y[j]; xd[3 * j + 2] = z[j]; }
        s = lbmibm_upload_cells(ctx, x.data(), y.data(), z.data(), xd.data(), nullptr, nullptr, nullptr);
        std::vector<double> uvw(3 * nv);
        std::copy(u.begin(), u.end(), uvw.begin());
        std::copy(v.begin(), v.end(), uvw.begin() + nv);
        std::copy(w.begin(), w.end(), uvw.begin() + 2 * nv);
        if (s == LBMIBM_OK) s = put_T(ctx, ctx->uvw, uvw.data(), 3 * nv);
        if (s == LBMIBM_OK) s = put_T(ctx, ctx->qi, qi.data(), 3 * nv);
        if (s == LBMIBM_OK) s = put_T(ctx, ctx->tp, tp.data(), 2 * ne);
    }
    if (s != LBMIBM_OK) return s;
    ctx->graph_ready = false;
    if (iter) *iter = it;
    if (fnum) *fnum = fno;
    return sync_check(ctx);
}

lbmibm_status lbmibm_write_prof(lbmibm_ctx *ctx, const char *path, int iter) {
    if (!ctx) return LBMIBM_ERR_ARG;
    Out o;
    lbmibm_status s = open_out(ctx, o, path);
    if (s != LBMIBM_OK) return s;
    o.put(&iter, 1);
    s = put_lattice_fields(ctx, o, ctx->np);
    if (s != LBMIBM_OK) return s;
    std::vector<int> bc(ctx->np);
    s = lbmibm_bc_mask(ctx, bc.data());
    if (s != LBMIBM_OK) return s;
    o.put(bc);
    return close_out(ctx, o, path);
}

lbmibm_status lbmibm_write_cell(lbmibm_ctx *ctx, const char *path) {
    if (!ctx) return LBMIBM_ERR_ARG;
    Layout L;
    lbmibm_status s = layout_of(ctx, L, true);
    if (s != LBMIBM_OK) return s;
    const size_t nv = (size_t)L.nc * L.V, ne = (size_t)L.nc * L.E;
    std::vector<double> x(nv), y(nv), z(nv), u(nv), v(nv), w(nv), qi(3 * nv), tp(2 * ne);
    std::vector<int> ele(3 * ne);
    s = lbmibm_download_cells(ctx, x.data(), y.data(), z.data(), u.data(), v.data(), w.data(), nullptr, nullptr,
                              nullptr, nullptr);
    if (s == LBMIBM_OK) s = lbmibm_download_membrane(ctx, nullptr, nullptr, nullptr, tp.data(), qi.data(), nullptr);
    if (s != LBMIBM_OK) return s;
    LBMIBM_HIP(ctx, hipMemcpy(ele.data(), ctx->ele, ele.size() * sizeof(int), hipMemcpyDeviceToHost));
    Out o;
    s = open_out(ctx, o, path);
    if (s != LBMIBM_OK) return s;
    o.put(&L.nc, 1);
    o.put(&L.V, 1);
    o.put(&L.E, 1);
    for (const auto *a : {&x, &y, &z, &u, &v, &w}) o.put(*a);
    o.put(qi);  // the first nc*vertex*3 entries of fem->qi (etc.cu:577)
    o.put(tp);
    o.put(ele);  // global vertex ids, as cel->ele (init1.cu:672-688)
    return close_out(ctx, o, path);
}

lbmibm_status lbmibm_write_adhesion(lbmibm_ctx *ctx, const char *dir, int iter, int *fnum_ad) {
    // Adhesion() file output (adhesion.cu:785-822): dir/adhesion/adhe%05d.bin
    // with fnum_ad, iter, f_adh [3 Nv] (AoS), attachment_point [Nv NS],
    // is_knob [Nv] (one byte each), the receptors' is_attached [num_wall]
    // and r_ij [Nv NS], then dir/adhesion/restart_adhe.dat ("%d\n%d\n",
    // fnum_ad, iter); fnum_ad is incremented.
    if (!ctx || !dir || !fnum_ad) return LBMIBM_ERR_ARG;
    if (!ctx->adh_on) {
        ctx->err = "adhesion: call lbmibm_adhesion_setup first";
        return LBMIBM_ERR_STATE;
    }
    const size_t nv = ctx->nv, ns = nv * ctx->adh.springs_per_knob;
    const size_t nrec = (size_t)ctx->adh_nxw * ctx->adh_nzw;
    std::vector<double> fadh(3 * nv), rij(ns);
    std::vector<int> attach(ns), isatt(nrec);
    std::vector<unsigned char> knob(nv);
    lbmibm_status s = lbmibm_adhesion_state(ctx, knob.data(), nullptr, nullptr, attach.data(), rij.data(),
                                            isatt.data(), fadh.data());
    if (s != LBMIBM_OK) return s;
    char name[4096];
    snprintf(name, sizeof name, "%s/adhesion/adhe%05d.bin", dir, *fnum_ad);
    Out o;
    s = open_out(ctx, o, name);
    if (s != LBMIBM_OK) return s;
    o.put(fnum_ad, 1);
    o.put(&iter, 1);
    o.put(fadh);
    o.put(attach);
    o.put(knob);  // sizeof(bool) = 1
    o.put(isatt);
    o.put(rij);
    s = close_out(ctx, o, name);
    if (s != LBMIBM_OK) return s;
    snprintf(name, sizeof name, "%s/adhesion/restart_adhe.dat", dir);
    FILE *f = fopen(name, "w");
    if (!f) {
        ctx->err = std::string("cannot open ") + name + " for writing";
        return LBMIBM_ERR_ARG;
    }
    const bool ok = fprintf(f, "%d\n%d\n", *fnum_ad, iter) > 0;
    if (fclose(f) != 0 || !ok) {
        ctx->err = std::string("write to ") + name + " failed";
        return LBMIBM_ERR_ARG;
    }
    (*fnum_ad)++;
    return LBMIBM_OK;
}

lbmibm_status lbmibm_output_data(lbmibm_ctx *ctx, const char *dir, int iter, int *fnum) {
    if (!ctx || !dir || !fnum) return LBMIBM_ERR_ARG;
    char name[4096];
    lbmibm_status s = LBMIBM_OK;
    if (*fnum % 50 == 0) {
        snprintf(name, sizeof name, "%s/prof%04d.bin", dir, *fnum);
        s = lbmibm_write_prof(ctx, name, iter);
    }
    if (s == LBMIBM_OK && ctx->nc > 0) {
        snprintf(name, sizeof name, "%s/cell/cell%04d.bin", dir, *fnum);
        s = lbmibm_write_cell(ctx, name);
    }
    if (s == LBMIBM_OK) {
        if (*fnum % 100 == 0) snprintf(name, sizeof name, "%s/restart%04d.bin", dir, *fnum);
        else snprintf(name, sizeof name, "%s/restart.bin", dir);
        s = lbmibm_write_restart(ctx, name, iter, *fnum);
    }
    if (s == LBMIBM_OK) (*fnum)++;
    return s;
}

}  // extern "C"
