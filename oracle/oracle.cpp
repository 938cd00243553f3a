// FP64 host ORACLE (test infrastructure only; see oracle.h).
//
// Every function restates the reference's algorithm from the text of the
// cited file:line ranges, in the reference's own padded/AoS data model.  It
// is compiled with -ffp-contract=off so that it is a plain IEEE FP64 program.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

namespace {

constexpr int Q = 19;
// D3Q19 directions in the reference order (lbm.cu:1-20).
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 0, 0, 0, 0};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
// Weights 1/3, 1/18, 1/36 written as the reference's divisors (lbm.cu:1136-1211).
constexpr double WDIV[Q] = {3.0, 18.0, 18.0, 18.0, 18.0, 18.0, 18.0, 36.0, 36.0, 36.0,
                            36.0, 36.0, 36.0, 36.0, 36.0, 36.0, 36.0, 36.0, 36.0};
// Opposite direction q + ((q%2)<<1) - 1 (lbm_cylinder.cu:219).
inline int opp(int q) { return q == 0 ? 0 : q + ((q % 2) << 1) - 1; }

struct Grid {
    int nx, ny, nz;
    bool py;
    int idx(int ix, int iy, int iz) const { return ix + nx * iy + nx * ny * iz; }
    void coords(int i, int &ix, int &iy, int &iz) const {
        iz = i / (nx * ny);
        iy = i / nx - ny * iz;
        ix = i - nx * iy - nx * ny * iz;
    }
    // Periodic image of a halo coordinate: the COMP range is [2, n-3], the
    // period is n-4 (lbm.cu:1350-1353, 1410-1411).
    static int wrap(int a, int n) {
        if (a < 2) return a + (n - 4);
        if (a > n - 3) return a - (n - 4);
        return a;
    }
    int image(int i) const {
        int ix, iy, iz;
        coords(i, ix, iy, iz);
        ix = wrap(ix, nx);
        iz = wrap(iz, nz);
        if (py) iy = wrap(iy, ny);
        return idx(ix, iy, iz);
    }
};

inline Grid grid_of(const orc_params *p) { return Grid{p->nx, p->ny, p->nz, p->periodic_y != 0}; }

inline double min_image(double d, double L) {
    if (d > L * 0.5) d -= L;
    else if (d < -L * 0.5) d += L;
    return d;
}

}  // namespace

extern "C" {

int orc_num_threads(void) { return omp_get_max_threads(); }

double orc_delta_weight(double r) { return (1.0 + std::cos(M_PI * r / 2.0)) / 4.0; }

double orc_tube_fraction(const orc_params *p, int i, int q) {
    // lbm_cylinder.cu:147-186, with R = the tube radius
    const Grid g = grid_of(p);
    int ix, iy, iz;
    g.coords(i, ix, iy, iz);
    const double dxl = p->dx, R = p->tube_radius;
    const double x_i = -p->lx / 2.0 - 1.5 * dxl + (double)ix * dxl;
    const double y_i = -p->ly / 2.0 - 1.5 * dxl + (double)iy * dxl;
    const double z_i = -p->lz / 2.0 - 1.5 * dxl + (double)iz * dxl;
    const double x_j = -p->lx / 2.0 - 1.5 * dxl + (double)(ix + CX[q]) * dxl;
    const double y_j = -p->ly / 2.0 - 1.5 * dxl + (double)(iy + CY[q]) * dxl;
    const double z_j = -p->lz / 2.0 - 1.5 * dxl + (double)(iz + CZ[q]) * dxl;
    const double dx = x_j - x_i, dy = y_j - y_i, dz = z_j - z_i;
    const double r0 = std::sqrt(dx * dx + dy * dy + dz * dz);
    const double Det = (x_i * dx + y_i * dy) * (x_i * dx + y_i * dy) - (dx * dx + dy * dy) * (x_i * x_i + y_i * y_i - R * R);
    if (Det < 0.0) return -1.0;
    const double t1 = (-x_i * dx - y_i * dy + std::sqrt(Det)) / (dx * dx + dy * dy);
    const double t2 = (-x_i * dx - y_i * dy - std::sqrt(Det)) / (dx * dx + dy * dy);
    double t;
    if (t1 > 0.0 && t1 < 1.0) t = t1;
    else if (t2 > 0.0 && t2 < 1.0) t = t2;
    else return -1.0;
    const double xw = x_i + t * dx, yw = y_i + t * dy, zw = z_i + t * dz;
    const double r = std::sqrt((x_i - xw) * (x_i - xw) + (y_i - yw) * (y_i - yw) + (z_i - zw) * (z_i - zw));
    return r / r0;
}

int orc_expand2pow2(int n) {
    // etc.cu:425-446
    int cnt = 0;
    while (n > 1) {
        if (n % 2 == 1) n++;
        n = n / 2;
        cnt++;
    }
    return (int)std::pow(2.0, cnt);
}

// ------------------------------------------------------------------------
// Lattice
// ------------------------------------------------------------------------
void orc_bc_init(const orc_params *p, int *bc) {
    // init1.cu:164-178 (slab: COMP inside, BUFF in the x/z halo, WALL beyond
    // |y| >= LY/2).  The node coordinate x = -LX/2 - 1.5dx + ix*dx lies
    // strictly inside (-LX/2, LX/2) exactly for ix in [2, nx-3].
    const Grid g = grid_of(p);
    const int n = g.nx * g.ny * g.nz;
    if (p->tube) {
        // init1.cu:217-229: the tube r <= Lc/2, BUFF in the z halo
        for (int i = 0; i < n; i++) {
            int ix, iy, iz;
            g.coords(i, ix, iy, iz);
            const double x = -p->lx / 2.0 - 1.5 * p->dx + (double)ix * p->dx;
            const double y = -p->ly / 2.0 - 1.5 * p->dx + (double)iy * p->dx;
            const double r = std::sqrt(x * x + y * y);
            const bool inz = iz >= 2 && iz <= g.nz - 3;
            if (r <= p->tube_radius && inz) bc[i] = ORC_COMP;
            else if (r <= p->tube_radius) bc[i] = ORC_BUFF;
            else bc[i] = ORC_WALL;
        }
        return;
    }
    for (int i = 0; i < n; i++) {
        int ix, iy, iz;
        g.coords(i, ix, iy, iz);
        bool inx = ix >= 2 && ix <= g.nx - 3;
        bool iny = iy >= 2 && iy <= g.ny - 3;
        bool inz = iz >= 2 && iz <= g.nz - 3;
        if (inx && iny && inz) bc[i] = ORC_COMP;
        else if (iny || g.py) bc[i] = ORC_BUFF;
        else bc[i] = ORC_WALL;
    }
}

void orc_flu_init(const orc_params *p, orc_state *s, const double *rho, const double *u,
                  const double *v, const double *w) {
    // init1.cu:281-449
    const Grid g = grid_of(p);
    const int n = g.nx * g.ny * g.nz;
    const double Cr = p->cr, Cr2 = Cr * Cr;
    for (int i = 0; i < n; i++) {
        double td = rho ? rho[i] : 1.0;
        double tu = u ? u[i] : 0.0, tv = v ? v[i] : 0.0, tw = w ? w[i] : 0.0;
        s->dn[i] = td; s->un[i] = tu; s->vn[i] = tv; s->wn[i] = tw;
        s->dm[i] = td; s->um[i] = tu; s->vm[i] = tv; s->wm[i] = tw;
        s->fx[i] = 0.0; s->fy[i] = 0.0; s->fz[i] = 0.0;
        if (s->vf) s->vf[i] = 0.0;
        double uvw = -1.5 * (tu * tu + tv * tv + tw * tw) * Cr2;
        for (int q = 0; q < Q; q++) {
            double cu = CX[q] * tu + CY[q] * tv + CZ[q] * tw;
            s->fn[i * Q + q] = td / WDIV[q] * (1.0 + 3.0 * cu * Cr + 4.5 * cu * cu * Cr2 + uvw);
            s->fm[i * Q + q] = 0.0;
        }
    }
}

void orc_lbm_step(const orc_params *p, orc_state *s, int iter) {
    // lbm_CapsuleSHEARFLOW / lbm_CapsuleCHANNELFLOW (lbm.cu:706-1312) with the
    // slab connectivity (walls in y, BUFF -> periodic image) and the channel
    // body force of lbm_CHANNELFLOW / CapsuleCHANNELFLOW (lbm.cu:518, 1032-1035).
    const Grid g = grid_of(p);
    const int n = g.nx * g.ny * g.nz;
    const double Cr = p->cr, Cr2 = Cr * Cr;
    double ui = p->ex, vi = p->ey, wi = p->ez;
    if (p->force_modulated == 1) {
        double t = p->dt * (double)iter;
        double amp = (iter > p->flowstart) ? 1.0 + 0.5 * std::sin(M_PI * t / p->freq) : 0.0;
        ui *= amp; vi *= amp; wi *= amp;
    } else if (p->force_modulated == 2) {  // lbm_cylinder.cu:292-293
        const double amp = iter > p->flowstart ? 1.0 : 0.0;
        ui *= amp; vi *= amp; wi *= amp;
    }
    const bool reset = ((iter - 1) % p->m) == 0;
    const double *fn = s->fn;
    double *fm = s->fm;
    // sliding walls (lbm_CapsuleSHEARFLOW, lbm.cu:755, 951-970; lbm_SHEARFLOW,
    // lbm.cu:455-468): z velocity of the wall above (+y) and below (-y)
    const double sra = iter < p->flowstart ? 0.0 : p->wall_velocity;
    const double uw_top = p->wall_motion ? sra : 0.0;
    const double uw_bot = p->wall_motion == 1 ? -sra : 0.0;

#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        if (s->bc[i] != ORC_COMP) continue;
        int ix, iy, iz;
        g.coords(i, ix, iy, iz);
        double f[Q];
        for (int q = 0; q < Q; q++) f[q] = fn[i * Q + q];

        if (reset) { s->dm[i] = 0.0; s->um[i] = 0.0; s->vm[i] = 0.0; s->wm[i] = 0.0; }
        const double tfx = s->fx[i], tfy = s->fy[i], tfz = s->fz[i];
        const double tvf = s->vf ? s->vf[i] : 0.0;
        // optional per-node body-force field, same convention as P e
        const double gbx = s->gbx ? s->gbx[i] : 0.0, gby = s->gby ? s->gby[i] : 0.0, gbz = s->gbz ? s->gbz[i] : 0.0;

        double td = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
        for (int q = 0; q < Q; q++) td += f[q];
        for (int q = 0; q < Q; q++) { jx += CX[q] * f[q]; jy += CY[q] * f[q]; jz += CZ[q] * f[q]; }
        const double tu = (jx + td * p->p * ui + td * p->b * tfx + td * gbx) / (td * Cr);
        const double tv = (jy + td * p->p * vi + td * p->b * tfy + td * gby) / (td * Cr);
        const double tw = (jz + td * p->p * wi + td * p->b * tfz + td * gbz) / (td * Cr);

        s->dn[i] = td; s->un[i] = tu; s->vn[i] = tv; s->wn[i] = tw;
        s->dm[i] += td; s->um[i] += tu; s->vm[i] += tv; s->wm[i] += tw;

        // Collision (lbm.cu:1129-1213): BGK + two Guo source terms.
        const double om = 1.0 / (3.0 * (1.0 + (p->lambda1 - 1.0) * tvf) * p->df + 0.5);
        const double om1 = 1.0 - om;
        const double om2 = (1.0 - om / 2.0) * 2.0 * td * p->p;
        const double om3 = (1.0 - om / 2.0) * 2.0 * td * p->b;
        const double om4 = (1.0 - om / 2.0) * 2.0 * td;
        const double uvw = -1.5 * (tu * tu + tv * tv + tw * tw) * Cr2;
        const double Ux = tu * Cr, Uy = tv * Cr, Uz = tw * Cr;
        for (int q = 0; q < Q; q++) {
            const double cu = CX[q] * Ux + CY[q] * Uy + CZ[q] * Uz;
            const double gx = 3.0 * (CX[q] - Ux) + 9.0 * cu * CX[q];
            const double gy = 3.0 * (CY[q] - Uy) + 9.0 * cu * CY[q];
            const double gz = 3.0 * (CZ[q] - Uz) + 9.0 * cu * CZ[q];
            const double fs = om1 * f[q] + om * td / WDIV[q] * (1.0 + 3.0 * cu + 4.5 * cu * cu + uvw)
                            + om2 / WDIV[q] * (gx * ui + gy * vi + gz * wi)
                            + om3 / WDIV[q] * (gx * tfx + gy * tfy + gz * tfz)
                            + om4 / WDIV[q] * (gx * gbx + gy * gby + gz * gbz);
            // Push with half-way bounce-back (lbm.cu:1273-1310).
            const int t = g.idx(ix + CX[q], iy + CY[q], iz + CZ[q]);
            int j;
            double fw = fs;
            if (s->bc[t] == ORC_WALL) {
                j = i * Q + opp(q);
                // Ladd's term of a wall moving along z: f - 6 w td Cr (c . u_w)
                if (p->wall_motion && !p->tube && CZ[q] != 0 && CY[q] != 0) {
                    const double uw = CY[q] > 0 ? uw_top : uw_bot;
                    fw = fs - 6.0 * td * Cr / WDIV[q] * (CZ[q] * uw);
                }
            }
            else if (s->bc[t] == ORC_BUFF) j = g.image(t) * Q + q;
            else j = t * Q + q;
            fm[j] = fw;
        }
    }

    // The tube: linear interpolation on every link that crosses the wall
    // (lbm_BounceBack_GPU, lbm_cylinder.cu:193-235).  All links read the
    // values the push wrote and then write: where a node has wall nodes on
    // both sides of a link the reference's threads race on those slots, and
    // its f3 then reads a wall node's unused slot; here f2 and f3 are the
    // node's own half-way-bounced value.
    if (p->tube && !p->tube_staircase) {
        std::vector<std::pair<size_t, double>> upd;
        for (int i = 0; i < n; i++) {
            if (s->bc[i] != ORC_COMP) continue;
            int ix, iy, iz;
            g.coords(i, ix, iy, iz);
            for (int d = 1; d < Q; d++) {
                const int t = g.idx(ix + CX[d], iy + CY[d], iz + CZ[d]);
                if (s->bc[t] != ORC_WALL) continue;
                const double r = orc_tube_fraction(p, i, d);
                if (r < 0.0) continue;
                const int q = opp(d);
                int jn = g.idx(ix - CX[d], iy - CY[d], iz - CZ[d]);
                const bool jwall = s->bc[jn] == ORC_WALL;
                if (s->bc[jn] == ORC_BUFF) jn = g.image(jn);
                const double f1 = fm[(size_t)i * Q + q];
                const double f2 = fm[(size_t)i * Q + d];
                const double f3 = jwall ? fm[(size_t)i * Q + d] : fm[(size_t)jn * Q + q];
                double f;
                if (r < 0.5) f = 2.0 * r * f1 + (1.0 - 2.0 * r) * f2;
                else if (r <= 1.0) f = 1.0 / (2.0 * r) * f1 + (2.0 * r - 1.0) / (2.0 * r) * f3;
                else f = f1;
                upd.emplace_back((size_t)i * Q + q, f);
            }
        }
        for (const auto &u : upd) fm[u.first] = u.second;
    }

    // BUFF nodes take the macros of their periodic image (lbm.cu:1316-1432).
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        if (s->bc[i] != ORC_BUFF) continue;
        const int i2 = g.image(i);
        s->dn[i] = s->dn[i2]; s->un[i] = s->un[i2]; s->vn[i] = s->vn[i2]; s->wn[i] = s->wn[i2];
        s->dm[i] = s->dm[i2]; s->um[i] = s->um[i2]; s->vm[i] = s->vm[i2]; s->wm[i] = s->wm[i2];
    }
    // lbm_AVERAGE on every node, dm not divided (lbm.cu:1435-1460, 1568-1571).
    if (iter % p->m == 0) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) {
            s->um[i] /= (double)p->m;
            s->vm[i] /= (double)p->m;
            s->wm[i] /= (double)p->m;
        }
    }
    std::swap(s->fn, s->fm);  // lbm.cu:1575
}

// ------------------------------------------------------------------------
// Immersed boundary
// ------------------------------------------------------------------------
void orc_ibm_st1(const orc_params *p, orc_state *s) {
    const Grid g = grid_of(p);
    const int n = g.nx * g.ny * g.nz;
    const double dx = p->dx;
    // ibm_st1a (ibm.cu:27-52)
    for (int i = 0; i < n; i++) { s->fx[i] = 0.0; s->fy[i] = 0.0; s->fz[i] = 0.0; }
    if (s->nc == 0) return;
    // ibm_st1b (ibm.cu:55-131), summed in vertex order.
    const int nv = s->voff[s->nc];
    for (int jid = 0; jid < nv; jid++) {
        const double xtr = s->x[jid] + (double)(g.nx - 1) * dx / 2.0;
        const double ytr = s->y[jid] + (double)(g.ny - 1) * dx / 2.0;
        const double ztr = s->z[jid] + (double)(g.nz - 1) * dx / 2.0;
        const int x1 = (int)std::floor(xtr / dx) - 1, x2 = (int)std::ceil(xtr / dx) + 1;
        const int y1 = (int)std::floor(ytr / dx) - 1, y2 = (int)std::ceil(ytr / dx) + 1;
        const int z1 = (int)std::floor(ztr / dx) - 1, z2 = (int)std::ceil(ztr / dx) + 1;
        for (int iz = z1; iz <= z2; iz++)
            for (int iy = y1; iy <= y2; iy++)
                for (int ix = x1; ix <= x2; ix++) {
                    if (ix < 0 || ix >= g.nx || iy < 0 || iy >= g.ny || iz < 0 || iz >= g.nz) continue;
                    const int i = g.idx(ix, iy, iz);
                    const double xx = std::fabs((double)ix * dx - xtr);
                    const double yy = std::fabs((double)iy * dx - ytr);
                    const double zz = std::fabs((double)iz * dx - ztr);
                    const double xcos = 1.0 + std::cos(M_PI * xx / (2.0 * dx));
                    const double ycos = 1.0 + std::cos(M_PI * yy / (2.0 * dx));
                    const double zcos = 1.0 + std::cos(M_PI * zz / (2.0 * dx));
                    const double dlt = xcos * ycos * zcos / (64.0 * dx * dx * dx);
                    const double *fa = s->adh ? s->adh->f_adh + 3 * (size_t)jid : nullptr;
                    s->fx[i] += dlt * (fa ? s->cfx[jid] + fa[0] : s->cfx[jid]);
                    s->fy[i] += dlt * (fa ? s->cfy[jid] + fa[1] : s->cfy[jid]);
                    s->fz[i] += dlt * (fa ? s->cfz[jid] + fa[2] : s->cfz[jid]);
                }
    }
    // ibm_st1c (ibm.cu:214-296): every COMP node near a periodic face adds
    // the halo nodes that are its images (x and z in the slab, plus y when
    // the domain is fully periodic).
    std::vector<double> ox(n), oy(n), oz(n);
    std::memcpy(ox.data(), s->fx, sizeof(double) * n);
    std::memcpy(oy.data(), s->fy, sizeof(double) * n);
    std::memcpy(oz.data(), s->fz, sizeof(double) * n);
    for (int i = 0; i < n; i++) {
        if (s->bc[i] != ORC_COMP) continue;
        int ix, iy, iz;
        g.coords(i, ix, iy, iz);
        int sx[2] = {0, 0}, sy[2] = {0, 0}, sz[2] = {0, 0};
        int nsx = 1, nsy = 1, nsz = 1;
        if (ix <= 3) sx[nsx++] = g.nx - 4; else if (ix >= g.nx - 4) sx[nsx++] = -(g.nx - 4);
        if (iz <= 3) sz[nsz++] = g.nz - 4; else if (iz >= g.nz - 4) sz[nsz++] = -(g.nz - 4);
        if (g.py) { if (iy <= 3) sy[nsy++] = g.ny - 4; else if (iy >= g.ny - 4) sy[nsy++] = -(g.ny - 4); }
        for (int a = 0; a < nsx; a++)
            for (int b = 0; b < nsy; b++)
                for (int c = 0; c < nsz; c++) {
                    if (a == 0 && b == 0 && c == 0) continue;
                    const int i2 = g.idx(ix + sx[a], iy + sy[b], iz + sz[c]);
                    s->fx[i] += ox[i2];
                    s->fy[i] += oy[i2];
                    s->fz[i] += oz[i2];
                }
    }
}

void orc_ibm_st2(const orc_params *p, orc_state *s) {
    // ibm_st2a (ibm.cu:372-639)
    if (s->nc == 0) return;
    const Grid g = grid_of(p);
    const double dx = p->dx, dt = p->dt, M = (double)p->m;
    const double xmin = -p->lx / 2.0, xmax = p->lx / 2.0;
    const double ymin = -p->ly / 2.0, ymax = p->ly / 2.0;
    const double zmin = -p->lz / 2.0, zmax = p->lz / 2.0;
    const int nv = s->voff[s->nc];
    const int rkn = p->runge_kutta;

#pragma omp parallel for schedule(static)
    for (int jid = 0; jid < nv; jid++) {
        const double xn = s->x[jid], yn = s->y[jid], zn = s->z[jid];
        double xr = xn, yr = yn, zr = zn;
        double ku[4] = {0, 0, 0, 0}, kv[4] = {0, 0, 0, 0}, kw[4] = {0, 0, 0, 0};
        for (int rk = 0; rk < rkn; rk++) {
            double ur = 0.0, vr = 0.0, wr = 0.0;
            int x1, x2, y1, y2, z1, z2;
            double xtr, ytr, ztr, scale;
            if (rkn == 4) {
                xtr = xr + (double)(g.nx - 1) * dx / 2.0;
                ytr = yr + (double)(g.ny - 1) * dx / 2.0;
                ztr = zr + (double)(g.nz - 1) * dx / 2.0;
                x1 = (int)std::floor(xtr / dx) - 1; x2 = (int)std::ceil(xtr / dx) + 1;
                y1 = (int)std::floor(ytr / dx) - 1; y2 = (int)std::ceil(ytr / dx) + 1;
                z1 = (int)std::floor(ztr / dx) - 1; z2 = (int)std::ceil(ztr / dx) + 1;
                scale = dx;
            } else {
                // RK2 / Euler branches hard-code dx = 1 (ibm.cu:524-551, 581-608).
                xtr = xr + (double)(g.nx - 1) / 2.0;
                ytr = yr + (double)(g.ny - 1) / 2.0;
                ztr = zr + (double)(g.nz - 1) / 2.0;
                x1 = (int)std::floor(xtr); x2 = (int)std::ceil(xtr);
                y1 = (int)std::floor(ytr); y2 = (int)std::ceil(ytr);
                z1 = (int)std::floor(ztr); z2 = (int)std::ceil(ztr);
                if (x1 == x2) { x1 -= 1; x2 += 1; }
                if (y1 == y2) { y1 -= 1; y2 += 1; }
                if (z1 == z2) { z1 -= 1; z2 += 1; }
                x1 -= 1; x2 += 1; y1 -= 1; y2 += 1; z1 -= 1; z2 += 1;
                scale = 1.0;
            }
            for (int iz = z1; iz <= z2; iz++)
                for (int iy = y1; iy <= y2; iy++)
                    for (int ix = x1; ix <= x2; ix++) {
                        if (ix < 0 || ix >= g.nx || iy < 0 || iy >= g.ny || iz < 0 || iz >= g.nz) continue;
                        const int i = g.idx(ix, iy, iz);
                        const double xx = std::fabs((double)ix * scale - xtr);
                        const double yy = std::fabs((double)iy * scale - ytr);
                        const double zz = std::fabs((double)iz * scale - ztr);
                        const double xcos = 1.0 + std::cos(M_PI * xx / (2.0 * scale));
                        const double ycos = 1.0 + std::cos(M_PI * yy / (2.0 * scale));
                        const double zcos = 1.0 + std::cos(M_PI * zz / (2.0 * scale));
                        const double dlt = xcos * ycos * zcos / 64.0;
                        ur += dlt * s->um[i];
                        vr += dlt * s->vm[i];
                        wr += dlt * s->wm[i];
                    }
            ku[rk] = ur; kv[rk] = vr; kw[rk] = wr;
            if (rkn == 4 && rk < 2) {
                xr = xn + ur * M * dt / 2.0; yr = yn + vr * M * dt / 2.0; zr = zn + wr * M * dt / 2.0;
            } else if (rkn == 4 && rk == 2) {
                xr = xn + ur * M * dt; yr = yn + vr * M * dt; zr = zn + wr * M * dt;
            } else if (rkn == 2 && rk == 0) {
                xr = xn + ur * M; yr = yn + vr * M; zr = zn + wr * M;
            }
        }
        double um, vm, wm;
        if (rkn == 4) {
            um = (ku[0] + 2.0 * ku[1] + 2.0 * ku[2] + ku[3]) / 6.0;
            vm = (kv[0] + 2.0 * kv[1] + 2.0 * kv[2] + kv[3]) / 6.0;
            wm = (kw[0] + 2.0 * kw[1] + 2.0 * kw[2] + kw[3]) / 6.0;
        } else if (rkn == 2) {
            um = (ku[0] + ku[1]) / 2.0; vm = (kv[0] + kv[1]) / 2.0; wm = (kw[0] + kw[1]) / 2.0;
        } else {
            um = ku[0]; vm = kv[0]; wm = kw[0];
        }
        xr = xn + um * M * dt;
        yr = yn + vm * M * dt;
        zr = zn + wm * M * dt;
        if (xr < xmin) xr += p->lx;
        if (xr > xmax) xr -= p->lx;
        if (zr < zmin) zr += p->lz;
        if (zr > zmax) zr -= p->lz;
        if (p->periodic_y) {
            if (yr < ymin) yr += p->ly;
            if (yr > ymax) yr -= p->ly;
        }
        s->u[jid] = um; s->x[jid] = xr; s->xd[jid * 3 + 0] = xr;
        s->v[jid] = vm; s->y[jid] = yr; s->xd[jid * 3 + 1] = yr;
        s->w[jid] = wm; s->z[jid] = zr; s->xd[jid * 3 + 2] = zr;
    }
}

// ------------------------------------------------------------------------
// Membrane
// ------------------------------------------------------------------------
static inline void edge_vec(const orc_params *p, const double *x, int a, int b, double v[3]) {
    for (int d = 0; d < 3; d++) v[d] = x[3 * b + d] - x[3 * a + d];
    v[0] = min_image(v[0], p->lx);
    v[2] = min_image(v[2], p->lz);
    if (p->periodic_y) v[1] = min_image(v[1], p->ly);
}

void orc_calc_tension(const orc_params *p, orc_state *s, int iter) {
    // calc_tension (stress.cu:29-121).  All elements are computed: the
    // reference's truncating grid Ne/256 (stress.cu:604-605) drops none at
    // the 1280/5120/20480-element meshes used here.
    if (s->nc == 0) return;
    const int ne = s->eoff[s->nc];
    const int M = p->m, FO = p->fileoutput > 0 ? p->fileoutput : 1;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ne; j++) {
        const int ea = s->ele[3 * j + 0], eb = s->ele[3 * j + 1], ec = s->ele[3 * j + 2];
        double v1d[3], v2d[3];
        edge_vec(p, s->xd, ea, eb, v1d);
        edge_vec(p, s->xd, ea, ec, v2d);
        const double *n = s->nref + 4 * j;
        const double lv1d = std::sqrt(v1d[0] * v1d[0] + v1d[1] * v1d[1] + v1d[2] * v1d[2]);
        const double lv2d = std::sqrt(v2d[0] * v2d[0] + v2d[1] * v2d[1] + v2d[2] * v2d[2]);
        const double inv12d = v1d[0] * v2d[0] + v1d[1] * v2d[1] + v1d[2] * v2d[2];
        const double dv12d = (lv1d * lv1d) * (lv2d * lv2d) - inv12d * inv12d;
        const double dn12 = n[0] * n[3] - n[1] * n[2];
        s->rg[j] = std::sqrt(dv12d);
        const double vd[4] = {lv1d * lv1d, inv12d, inv12d, lv2d * lv2d};
        const double nd[4] = {vd[3] / dv12d, -vd[1] / dv12d, -vd[2] / dv12d, vd[0] / dv12d};
        const double i1 = n[0] * vd[0] + n[1] * vd[1] + n[2] * vd[2] + n[3] * vd[3] - 2.0;
        const double i2 = dv12d * dn12 - 1.0;
        const double js = std::sqrt(i2 + 1.0);
        for (int k = 0; k < 4; k++) {
            if (!p->neohookean)
                s->t[4 * j + k] = p->rgs * (i1 + 1.0) / js * n[k] + (p->c * i2 - 1.0) * js * nd[k];
            else
                s->t[4 * j + k] = p->rgs * (n[k] / js - nd[k] * js / (i2 + 1.0) / (i2 + 1.0));
        }
        if (s->tp) {
            if ((iter - 1) % FO == 0) { s->tp[2 * j] = 0.0; s->tp[2 * j + 1] = 0.0; }
            const double judge = std::fabs((i1 + 2.0) * (i1 + 2.0) - 4.0 * (i2 + 1.0));
            const double root = std::sqrt(judge);
            const double l12 = 0.5 * (i1 + 2.0 + root), l22 = 0.5 * (i1 + 2.0 - root);
            const double l11 = std::sqrt(l12), l21 = std::sqrt(l22);
            s->tp[2 * j + 0] += (double)M / (double)FO * l11 / l21 * (l12 - 1.0 + p->c * l22 * i2);
            s->tp[2 * j + 1] += (double)M / (double)FO * l21 / l11 * (l22 - 1.0 + p->c * l12 * i2);
        }
    }
}

static int assemble_cell(const orc_params *p, const orc_state *s, int c, int *ptr, int *index,
                         double *value, double *b, double *diag) {
    const int base = s->voff[c], V = s->voff[c + 1] - base;
    const int mn = orc_expand2pow2(3 * V);
    // initialize + set_matrix_component (stress.cu:124-179)
    for (int j = 0; j < mn; j++) {
        b[j] = 0.0;
        for (int k = 0; k < 7; k++) { index[k * mn + j] = 0; value[k * mn + j] = 0.0; }
        const int node = j / 3, d = j % 3;
        if (node < V) {
            const int jj = base + node;
            ptr[j] = s->nln[jj] + 1;
            for (int i = 0; i < s->nln[jj] + 1; i++) index[i * mn + j] = 3 * (s->lo[7 * jj + i] - base) + d;
        } else {
            ptr[j] = 1;
            index[j] = j;
        }
    }
    // set_matrix_coefficient (stress.cu:182-272)
    static const double F[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
    for (int j = 0; j < 3 * V; j++) {
        const int node = j / 3, d = j % 3, jj = base + node;
        const int nn = s->nln[jj];
        double tvalue[7] = {0, 0, 0, 0, 0, 0, 0}, tb = 0.0;
        int order = 0;
        for (int i = 0; i < nn; i++) {
            const int element = s->le[6 * jj + i];
            const double gab = s->rg[element];
            const int ea = s->ele[3 * element], eb = s->ele[3 * element + 1], ec = s->ele[3 * element + 2];
            int oa = 0, ob = 0, oc = 0, en = 0;
            for (int k = 0; k < nn + 1; k++) {
                const int l = s->lo[7 * jj + k];
                if (ea == l) oa = k;
                if (eb == l) ob = k;
                if (ec == l) oc = k;
                if (jj == l) order = k;
            }
            if (ea == jj) { tvalue[oa] += gab / 12.0; en = 0; } else tvalue[oa] += gab / 24.0;
            if (eb == jj) { tvalue[ob] += gab / 12.0; en = 1; } else tvalue[ob] += gab / 24.0;
            if (ec == jj) { tvalue[oc] += gab / 12.0; en = 2; } else tvalue[oc] += gab / 24.0;
            double vd[2];
            vd[0] = s->xd[3 * eb + d] - s->xd[3 * ea + d];
            vd[1] = s->xd[3 * ec + d] - s->xd[3 * ea + d];
            if (d == 0) { vd[0] = min_image(vd[0], p->lx); vd[1] = min_image(vd[1], p->lx); }
            else if (d == 2) { vd[0] = min_image(vd[0], p->lz); vd[1] = min_image(vd[1], p->lz); }
            else if (p->periodic_y) { vd[0] = min_image(vd[0], p->ly); vd[1] = min_image(vd[1], p->ly); }
            for (int k = 0; k < 4; k++) {
                const int al = k / 2, be = k % 2;
                tb += gab * s->t[4 * element + k] * (vd[al] * F[en][be] + vd[be] * F[en][al]) / 4.0;
            }
        }
        const double jacobi = tvalue[order];
        for (int k = 0; k < 7; k++) value[k * mn + j] = tvalue[k] / jacobi;
        b[j] = tb / jacobi;
        if (diag) diag[j] = jacobi;
    }
    return mn;
}

int orc_assemble_cell(const orc_params *p, const orc_state *s, int c, int *ptr, int *index,
                      double *value, double *b) {
    return assemble_cell(p, s, c, ptr, index, value, b, nullptr);
}

static void spmv_ell(const int *ptr, const int *index, const double *value, const double *xi,
                     double *xo, int n) {
    // matrix_solver.cu:265-277
    for (int tid = 0; tid < n; tid++) {
        double sum = 0.0;
        for (int i = 0; i < ptr[tid]; i++) sum += value[i * n + tid] * xi[index[i * n + tid]];
        xo[tid] = sum;
    }
}

static double dot(const double *a, const double *b, int n) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

int orc_bicgstab(const int *ptr, const int *index, const double *value, double *x, const double *b,
                 int n, int itrmax, double residual, int *iters_out) {
    // matrix_solver.cu:279-377 (same update order; inner products summed
    // sequentially in FP64 instead of the reference's tree, dot_product.cu).
    std::vector<double> ra(n, 0.0), r(n, 0.0), pp(n, 0.0), m(n, 0.0), nv(n, 0.0);
    double alfa = 1.0, beta = 0.0, omeg = 1.0, r_dot_ra = 0.0, r_dot_ra_old = 0.0;
    spmv_ell(ptr, index, value, x, r.data(), n);
    for (int i = 0; i < n; i++) { const double t = b[i] - r[i]; r[i] = t; ra[i] = t; pp[i] = t; }
    const double b_dot_b = dot(b, b, n);
    int iter;
    for (iter = 0; iter < itrmax; iter++) {
        r_dot_ra_old = r_dot_ra;
        r_dot_ra = dot(r.data(), ra.data(), n);
        if (r_dot_ra == 0) { if (iters_out) *iters_out = iter; return 1; }
        if (iter > 0) {
            beta = (r_dot_ra / r_dot_ra_old) * (alfa / omeg);
            for (int i = 0; i < n; i++) pp[i] = r[i] + beta * (pp[i] - omeg * m[i]);
        }
        spmv_ell(ptr, index, value, pp.data(), m.data(), n);
        alfa = r_dot_ra / dot(m.data(), ra.data(), n);
        for (int i = 0; i < n; i++) r[i] = r[i] - alfa * m[i];
        spmv_ell(ptr, index, value, r.data(), nv.data(), n);
        omeg = dot(r.data(), nv.data(), n) / dot(nv.data(), nv.data(), n);
        for (int i = 0; i < n; i++) {
            const double ri = r[i];
            x[i] = x[i] + alfa * pp[i] + omeg * ri;
            r[i] = ri - omeg * nv[i];
        }
        const double err = std::sqrt(dot(r.data(), r.data(), n) / b_dot_b);
        if (err < residual) break;
    }
    if (iters_out) *iters_out = iter + (iter < itrmax ? 1 : 0);
    if (iter >= itrmax) return 2;
    return 0;
}

static void q2f(const orc_params *p, orc_state *s) {
    // q2f (stress.cu:376-445); f = -q ds, times the optional per-cell scale.
    const int nv = s->voff[s->nc];
#pragma omp parallel for schedule(static)
    for (int j1 = 0; j1 < nv; j1++) {
        double ds = 0.0;
        const double x1 = s->xd[3 * j1], y1 = s->xd[3 * j1 + 1], z1 = s->xd[3 * j1 + 2];
        const int nl = s->nln[j1];
        for (int k = 0; k < nl; k++) {
            const int j2 = s->ln[7 * j1 + k], j3 = s->ln[7 * j1 + (k + 1) % nl];
            double xx1 = s->xd[3 * j2] - x1, yy1 = s->xd[3 * j2 + 1] - y1, zz1 = s->xd[3 * j2 + 2] - z1;
            double xx2 = s->xd[3 * j3] - x1, yy2 = s->xd[3 * j3 + 1] - y1, zz2 = s->xd[3 * j3 + 2] - z1;
            xx1 = min_image(xx1, p->lx); zz1 = min_image(zz1, p->lz);
            xx2 = min_image(xx2, p->lx); zz2 = min_image(zz2, p->lz);
            if (p->periodic_y) { yy1 = min_image(yy1, p->ly); yy2 = min_image(yy2, p->ly); }
            const double nvx = -(yy1 * zz2 - zz1 * yy2);
            const double nvy = -(zz1 * xx2 - xx1 * zz2);
            const double nvz = -(xx1 * yy2 - yy1 * xx2);
            const double nn = std::sqrt(nvx * nvx + nvy * nvy + nvz * nvz);
            ds += nn / 2.0 / 3.0;
        }
        s->cfx[j1] = -1.0 * s->q[j1 * 3 + 0] * ds;
        s->cfy[j1] = -1.0 * s->q[j1 * 3 + 1] * ds;
        s->cfz[j1] = -1.0 * s->q[j1 * 3 + 2] * ds;
    }
    if (s->gs_scale) {
        for (int c = 0; c < s->nc; c++)
            for (int j = s->voff[c]; j < s->voff[c + 1]; j++) {
                s->cfx[j] *= s->gs_scale[c]; s->cfy[j] *= s->gs_scale[c]; s->cfz[j] *= s->gs_scale[c];
            }
    }
}

// Hinge identification of fourpoint_bend / cel_init (stress.cu:497-505).
static void hinge(const int *ele, int e1, int e2, int &k1, int &k2, int &k3, int &k4) {
    const int i1 = ele[3 * e1], i2 = ele[3 * e1 + 1], i3 = ele[3 * e1 + 2];
    const int j1 = ele[3 * e2], j2 = ele[3 * e2 + 1], j3 = ele[3 * e2 + 2];
    k1 = k2 = k3 = k4 = -1;
    if      ((i1==j1 && i2==j2)||(i1==j2 && i2==j1)) { k1 = i3; k2 = i1; k3 = i2; k4 = j3; }
    else if ((i1==j2 && i2==j3)||(i1==j3 && i2==j2)) { k1 = i3; k2 = i1; k3 = i2; k4 = j1; }
    else if ((i1==j3 && i2==j1)||(i1==j1 && i2==j3)) { k1 = i3; k2 = i1; k3 = i2; k4 = j2; }
    else if ((i2==j1 && i3==j2)||(i2==j2 && i3==j1)) { k1 = i1; k2 = i2; k3 = i3; k4 = j3; }
    else if ((i2==j2 && i3==j3)||(i2==j3 && i3==j2)) { k1 = i1; k2 = i2; k3 = i3; k4 = j1; }
    else if ((i2==j3 && i3==j1)||(i2==j1 && i3==j3)) { k1 = i1; k2 = i2; k3 = i3; k4 = j2; }
    else if ((i3==j1 && i1==j2)||(i3==j2 && i1==j1)) { k1 = i2; k2 = i3; k3 = i1; k4 = j3; }
    else if ((i3==j2 && i1==j3)||(i3==j3 && i1==j2)) { k1 = i2; k2 = i3; k3 = i1; k4 = j1; }
    else if ((i3==j3 && i1==j1)||(i3==j1 && i1==j3)) { k1 = i2; k2 = i3; k3 = i1; k4 = j2; }
}

struct HingeGeom {
    double a21[3], a31[3], a34[3], a24[3], a23[3], xi[3], ze[3], t1[3], t2[3];
};

static void hinge_geom(const orc_params *p, const orc_state *s, int k1, int k2, int k3, int k4,
                       HingeGeom &h) {
    double a1[3] = {s->x[k1], s->y[k1], s->z[k1]};
    double a2[3] = {s->x[k2], s->y[k2], s->z[k2]};
    double a3[3] = {s->x[k3], s->y[k3], s->z[k3]};
    double a4[3] = {s->x[k4], s->y[k4], s->z[k4]};
    auto wrapto = [&](double *a, int d, double L) {
        if (a[d] - a1[d] > L * 0.5) a[d] -= L;
        else if (a[d] - a1[d] < -L * 0.5) a[d] += L;
    };
    for (double *a : {a2, a3, a4}) {
        wrapto(a, 0, p->lx);
        wrapto(a, 2, p->lz);
        if (p->periodic_y) wrapto(a, 1, p->ly);
    }
    for (int d = 0; d < 3; d++) {
        h.a21[d] = a2[d] - a1[d]; h.a31[d] = a3[d] - a1[d]; h.a34[d] = a3[d] - a4[d];
        h.a24[d] = a2[d] - a4[d]; h.a23[d] = a2[d] - a3[d];
        h.t1[d] = (a1[d] + a2[d] + a3[d]) / 3.0; h.t2[d] = (a4[d] + a2[d] + a3[d]) / 3.0;
    }
    h.xi[0] = h.a21[1] * h.a31[2] - h.a21[2] * h.a31[1];
    h.xi[1] = h.a21[2] * h.a31[0] - h.a21[0] * h.a31[2];
    h.xi[2] = h.a21[0] * h.a31[1] - h.a21[1] * h.a31[0];
    h.ze[0] = h.a34[1] * h.a24[2] - h.a34[2] * h.a24[1];
    h.ze[1] = h.a34[2] * h.a24[0] - h.a34[0] * h.a24[2];
    h.ze[2] = h.a34[0] * h.a24[1] - h.a34[1] * h.a24[0];
}

static void fourpoint_bend(const orc_params *p, orc_state *s) {
    // fourpoint_bend (stress.cu:448-578), accumulated sequentially.
    const int ne = s->eoff[s->nc];
    const double eps = 1.0e-6;
    for (int e1 = 0; e1 < ne; e1++) {
        for (int k = 0; k < 3; k++) {
            const int e2 = s->ec[3 * e1 + k];
            int k1, k2, k3, k4;
            hinge(s->ele, e1, e2, k1, k2, k3, k4);
            if (k1 < 0) continue;
            HingeGeom h;
            hinge_geom(p, s, k1, k2, k3, k4, h);
            const double *xi = h.xi, *ze = h.ze;
            const double cost0 = s->c0[3 * e1 + k], sint0 = s->s0[3 * e1 + k];
            const double lxi = std::sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
            const double lze = std::sqrt(ze[0] * ze[0] + ze[1] * ze[1] + ze[2] * ze[2]);
            const double cost = (xi[0] * ze[0] + xi[1] * ze[1] + xi[2] * ze[2]) / lxi / lze;
            const double cost2 = 1.0 - cost * cost;
            if (cost2 < eps) continue;
            double sint, beta;
            const double side = (xi[0] - ze[0]) * (h.t1[0] - h.t2[0]) + (xi[1] - ze[1]) * (h.t1[1] - h.t2[1])
                              + (xi[2] - ze[2]) * (h.t1[2] - h.t2[2]);
            if (side >= 0.0) { sint = std::sqrt(cost2); beta = 2.0 / std::sqrt(3.0) * p->br * (sint * cost0 - cost * sint0) / std::sqrt(cost2); }
            else { sint = -std::sqrt(cost2); beta = -2.0 / std::sqrt(3.0) * p->br * (sint * cost0 - cost * sint0) / std::sqrt(cost2); }
            const double b11 = -beta * cost / (xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
            const double b12 = beta / lxi / lze;
            const double b22 = -beta * cost / (ze[0] * ze[0] + ze[1] * ze[1] + ze[2] * ze[2]);
            const double *a21 = h.a21, *a31 = h.a31, *a34 = h.a34, *a24 = h.a24, *a23 = h.a23;
            double fb[3];
            // k1
            fb[0] = b11 * (xi[1] * -a23[2] - xi[2] * -a23[1]) + b12 * (ze[1] * -a23[2] - ze[2] * -a23[1]);
            fb[1] = b11 * (xi[2] * -a23[0] - xi[0] * -a23[2]) + b12 * (ze[2] * -a23[0] - ze[0] * -a23[2]);
            fb[2] = b11 * (xi[0] * -a23[1] - xi[1] * -a23[0]) + b12 * (ze[0] * -a23[1] - ze[1] * -a23[0]);
            s->cfx[k1] += fb[0]; s->cfy[k1] += fb[1]; s->cfz[k1] += fb[2];
            // k2
            fb[0] = b11 * (xi[1] * -a31[2] - xi[2] * -a31[1]) + b22 * (ze[1] * a34[2] - ze[2] * a34[1])
                  + b12 * ((xi[1] * a34[2] - xi[2] * a34[1]) + (ze[1] * -a31[2] - ze[2] * -a31[1]));
            fb[1] = b11 * (xi[2] * -a31[0] - xi[0] * -a31[2]) + b22 * (ze[2] * a34[0] - ze[0] * a34[2])
                  + b12 * ((xi[2] * a34[0] - xi[0] * a34[2]) + (ze[2] * -a31[0] - ze[0] * -a31[2]));
            fb[2] = b11 * (xi[0] * -a31[1] - xi[1] * -a31[0]) + b22 * (ze[0] * a34[1] - ze[1] * a34[0])
                  + b12 * ((xi[0] * a34[1] - xi[1] * a34[0]) + (ze[0] * -a31[1] - ze[1] * -a31[0]));
            s->cfx[k2] += 0.5 * fb[0]; s->cfy[k2] += 0.5 * fb[1]; s->cfz[k2] += 0.5 * fb[2];
            // k3
            fb[0] = b11 * (xi[1] * a21[2] - xi[2] * a21[1]) + b22 * (ze[1] * -a24[2] - ze[2] * -a24[1])
                  + b12 * ((xi[1] * -a24[2] - xi[2] * -a24[1]) + (ze[1] * a21[2] - ze[2] * a21[1]));
            fb[1] = b11 * (xi[2] * a21[0] - xi[0] * a21[2]) + b22 * (ze[2] * -a24[0] - ze[0] * -a24[2])
                  + b12 * ((xi[2] * -a24[0] - xi[0] * -a24[2]) + (ze[2] * a21[0] - ze[0] * a21[2]));
            fb[2] = b11 * (xi[0] * a21[1] - xi[1] * a21[0]) + b22 * (ze[0] * -a24[1] - ze[1] * -a24[0])
                  + b12 * ((xi[0] * -a24[1] - xi[1] * -a24[0]) + (ze[0] * a21[1] - ze[1] * a21[0]));
            s->cfx[k3] += 0.5 * fb[0]; s->cfy[k3] += 0.5 * fb[1]; s->cfz[k3] += 0.5 * fb[2];
        }
    }
}

int orc_stress(const orc_params *p, orc_state *s, int iter) {
    // stress() host launcher (stress.cu:580-659) with calc_load (:306-373).
    if (s->nc == 0) return 0;
    orc_calc_tension(p, s, iter);
    int flag = 0;
    const int FO = p->fileoutput > 0 ? p->fileoutput : 1;
#pragma omp parallel for schedule(dynamic, 1) reduction(max : flag)
    for (int c = 0; c < s->nc; c++) {
        const int base = s->voff[c], V = s->voff[c + 1] - base;
        const int mn = orc_expand2pow2(3 * V);
        std::vector<int> ptr(mn), index(7 * mn);
        std::vector<double> value(7 * mn), b(mn), x(mn, 0.0);
        assemble_cell(p, s, c, ptr.data(), index.data(), value.data(), b.data(), nullptr);
        int its = 0, f = 0;
        // An exactly zero right-hand side is solved by x = 0; the reference
        // would report breakdown and exit (matrix_solver.cu:334, stress.cu:366).
        if (dot(b.data(), b.data(), mn) != 0.0)
            f = orc_bicgstab(ptr.data(), index.data(), value.data(), x.data(), b.data(), mn, p->itrmax,
                             p->residual, &its);
        if (s->bicg_iters) s->bicg_iters[c] = its;
        if (f > flag) flag = f;
        // substitute_load (stress.cu:275-303)
        for (int j = 0; j < 3 * V; j++) {
            s->q[3 * base + j] = x[j];
            if (s->qi) {
                if ((iter - 1) % FO == 0) s->qi[3 * base + j] = 0.0;
                s->qi[3 * base + j] += (double)p->m / (double)FO * x[j];
            }
        }
    }
    if (flag) return flag;
    q2f(p, s);
    if (p->br != 0.0) fourpoint_bend(p, s);
    return 0;
}

// ------------------------------------------------------------------------
// Volume
// ------------------------------------------------------------------------
static void ele_normals(const orc_state *s, std::vector<double> &normal, std::vector<double> &ds,
                        std::vector<double> &dv) {
    // CalcEleNormal_GPU (volume_correction.cu:138-200)
    const int ne = s->eoff[s->nc];
    normal.resize(3 * ne); ds.resize(ne); dv.resize(ne);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; e++) {
        const int na = s->ele[3 * e], nb = s->ele[3 * e + 1], nc = s->ele[3 * e + 2];
        const double v1[3] = {s->x[nb] - s->x[na], s->y[nb] - s->y[na], s->z[nb] - s->z[na]};
        const double v2[3] = {s->x[nc] - s->x[na], s->y[nc] - s->y[na], s->z[nc] - s->z[na]};
        const double cr[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                              v1[0] * v2[1] - v1[1] * v2[0]};
        const double nrm = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        for (int d = 0; d < 3; d++) normal[3 * e + d] = cr[d] / nrm;
        ds[e] = nrm / 2.0;
        const double r[3] = {(s->x[na] + s->x[nb] + s->x[nc]) / 3.0, (s->y[na] + s->y[nb] + s->y[nc]) / 3.0,
                             (s->z[na] + s->z[nb] + s->z[nc]) / 3.0};
        dv[e] = 1.0 / 3.0 * (r[0] * normal[3 * e] + r[1] * normal[3 * e + 1] + r[2] * normal[3 * e + 2]) * ds[e];
    }
}

void orc_calc_volume(const orc_params *, const orc_state *s, double *volume, double *area) {
    // CalcVolume (volume_correction.cu:246-281)
    std::vector<double> normal, ds, dv;
    ele_normals(s, normal, ds, dv);
    for (int c = 0; c < s->nc; c++) {
        double V = 0.0, A = 0.0;
        for (int e = s->eoff[c]; e < s->eoff[c + 1]; e++) { V += dv[e]; A += ds[e]; }
        volume[c] = V;
        area[c] = A;
    }
}

static void calc_center(const orc_params *p, orc_state *s) {
    // CalcCenter (volume_correction.cu:310-352)
    for (int c = 0; c < s->nc; c++) {
        double gx = 0.0, gy = 0.0, gz = 0.0;
        const int V = s->voff[c + 1] - s->voff[c];
        for (int j = s->voff[c]; j < s->voff[c + 1]; j++) { gx += s->x[j]; gy += s->y[j]; gz += s->z[j]; }
        gx /= (double)V; gy /= (double)V; gz /= (double)V;
        if (gx < -p->lx / 2.0) gx += p->lx; else if (gx > p->lx / 2.0) gx -= p->lx;
        if (gz < -p->lz / 2.0) gz += p->lz; else if (gz > p->lz / 2.0) gz -= p->lz;
        if (p->periodic_y) { if (gy < -p->ly / 2.0) gy += p->ly; else if (gy > p->ly / 2.0) gy -= p->ly; }
        s->gx[3 * c] = gx; s->gx[3 * c + 1] = gy; s->gx[3 * c + 2] = gz;
    }
}

void orc_volume_init(const orc_params *p, orc_state *s) {
    // volume_init (volume_correction.cu:427-450)
    if (s->nc == 0) return;
    orc_calc_volume(p, s, s->volume0, s->area);
    calc_center(p, s);
}

int orc_volume_correction(const orc_params *p, orc_state *s) {
    // volume_correction (volume_correction.cu:453-524)
    if (s->nc == 0) return 0;
    const double COEFF = 0.1, EPS = 1.0e-3;
    // CapsulePeriodicBefore (:28-67, :284-308)
    for (int c = 0; c < s->nc; c++)
        for (int j = s->voff[c]; j < s->voff[c + 1]; j++) {
            double d = s->x[j] - s->gx[3 * c];
            if (d > p->lx * 0.5) s->x[j] -= p->lx; else if (d < -p->lx * 0.5) s->x[j] += p->lx;
            d = s->z[j] - s->gx[3 * c + 2];
            if (d > p->lz * 0.5) s->z[j] -= p->lz; else if (d < -p->lz * 0.5) s->z[j] += p->lz;
            if (p->periodic_y) {
                d = s->y[j] - s->gx[3 * c + 1];
                if (d > p->ly * 0.5) s->y[j] -= p->ly; else if (d < -p->ly * 0.5) s->y[j] += p->ly;
            }
        }
    std::vector<char> conv(s->nc, 0);
    std::vector<double> lambda(s->nc, 0.0);
    std::vector<int> owner(s->voff[s->nc]);
    for (int c = 0; c < s->nc; c++)
        for (int j = s->voff[c]; j < s->voff[c + 1]; j++) owner[j] = c;
    std::vector<double> normal, ds, dv;
    int iter = 0;
    bool all;
    do {
        ele_normals(s, normal, ds, dv);
        all = true;
        for (int c = 0; c < s->nc; c++) {  // CalcVolume + CalcLambda1 (:378-399)
            double V = 0.0, A = 0.0;
            for (int e = s->eoff[c]; e < s->eoff[c + 1]; e++) { V += dv[e]; A += ds[e]; }
            s->volume[c] = V; s->area[c] = A;
            const double dV = std::fabs(V - s->volume0[c]) / s->volume0[c] * 100.0;
            if (dV < EPS) { lambda[c] = 0.0; conv[c] = 1; }
            else lambda[c] = 18.0 * (V - s->volume0[c]) / A;
            all = all && conv[c];
        }
        const int nv = s->voff[s->nc];
        // VolumeCorrection_GPU (:97-135)
#pragma omp parallel for schedule(static)
        for (int j = 0; j < nv; j++) {
            double nn[3] = {0.0, 0.0, 0.0};
            const int cnt = s->nle[j];
            for (int k = 0; k < cnt; k++) {
                const int e = s->le[6 * j + k];
                nn[0] += normal[3 * e] / cnt; nn[1] += normal[3 * e + 1] / cnt; nn[2] += normal[3 * e + 2] / cnt;
            }
            const double l = lambda[owner[j]];
            s->x[j] += -l / 6.0 * nn[0] * COEFF;
            s->y[j] += -l / 6.0 * nn[1] * COEFF;
            s->z[j] += -l / 6.0 * nn[2] * COEFF;
        }
        iter++;
    } while (!all && iter < 1000);
    calc_center(p, s);
    // CapsulePeriodicAfter (:70-94)
    for (int j = 0; j < s->voff[s->nc]; j++) {
        if (s->x[j] < -p->lx / 2.0) s->x[j] += p->lx; else if (s->x[j] > p->lx / 2.0) s->x[j] -= p->lx;
        if (s->z[j] < -p->lz / 2.0) s->z[j] += p->lz; else if (s->z[j] > p->lz / 2.0) s->z[j] -= p->lz;
        if (p->periodic_y) { if (s->y[j] < -p->ly / 2.0) s->y[j] += p->ly; else if (s->y[j] > p->ly / 2.0) s->y[j] -= p->ly; }
    }
    if (iter >= 1000) return -1;
    return iter - 1;
}

// ------------------------------------------------------------------------
// Topology
// ------------------------------------------------------------------------
int orc_fem_topology(int nc, const int *voff, const int *eoff, const double *xr, const int *ele,
                     int *nln, int *ln, int *lo, int *nle, int *le, double *nref, int *ec, double *c0,
                     double *s0) {
    const int nv = voff[nc], ne = eoff[nc];
    for (int c = 0; c < nc; c++)  // fem_init rejects mn < 256 (init2.cu:286-287)
        if (orc_expand2pow2(3 * (voff[c + 1] - voff[c])) < 256) return 3;
    for (int j = 0; j < nv; j++) {
        nln[j] = 0; nle[j] = 0;
        for (int k = 0; k < 7; k++) { ln[7 * j + k] = 0; lo[7 * j + k] = 0; }
        for (int k = 0; k < 6; k++) le[6 * j + k] = 0;
    }
    auto add_nlink = [&](int na, int nb) -> int {  // init2.cu:12-42
        for (int i = 0; i < nln[na]; i++)
            if (nb == ln[7 * na + i]) return 0;
        if (nln[na] >= 7) return 3;
        ln[7 * na + nln[na]] = nb;
        nln[na]++;
        return 0;
    };
    auto checklink = [&](int na, int nb) {  // init2.cu:66-89
        for (int i = 0; i < nln[na]; i++)
            if (nb == ln[7 * na + i]) return 1;
        return 0;
    };
    auto topology_cell = [&](int c) -> int {
        const int V = voff[c + 1] - voff[c], T = eoff[c + 1] - eoff[c];
        for (int eid = eoff[c]; eid < eoff[c + 1]; eid++) {  // fem_init (init2.cu:317-341)
            const int ea = ele[3 * eid], eb = ele[3 * eid + 1], ecc = ele[3 * eid + 2];
            if (add_nlink(ea, eb) || add_nlink(ea, ecc) || add_nlink(eb, ecc) || add_nlink(eb, ea) ||
                add_nlink(ecc, ea) || add_nlink(ecc, eb))
                return 3;
            for (int v : {ea, eb, ecc}) {  // add_elink (init2.cu:45-63)
                if (nle[v] >= 6) return 3;
                le[6 * v + nle[v]] = eid;
                nle[v]++;
            }
        }
        int checker = 0;
        for (int j = voff[c]; j < voff[c + 1]; j++) checker += nln[j];
        if (checker != 2 * (V + T - 2)) return 3;
        for (int j = voff[c]; j < voff[c + 1]; j++)
            for (int k = 0; k < nln[j]; k++) lo[7 * j + k] = ln[7 * j + k];
        // sort_link_spring (init2.cu:92-166)
        double gx = 0.0, gy = 0.0, gz = 0.0;
        for (int j = voff[c]; j < voff[c + 1]; j++) { gx += xr[3 * j]; gy += xr[3 * j + 1]; gz += xr[3 * j + 2]; }
        gx /= V; gy /= V; gz /= V;
        for (int jid = voff[c]; jid < voff[c + 1]; jid++) {
            int lns[7];
            int nfound = 0;
            lns[0] = ln[7 * jid];
            for (int k = 1; k < nln[jid]; k++) {
                const int na = lns[k - 1];
                const double va[3] = {xr[3 * na] - xr[3 * jid], xr[3 * na + 1] - xr[3 * jid + 1],
                                      xr[3 * na + 2] - xr[3 * jid + 2]};
                for (int l = 0; l < nln[jid]; l++) {
                    const int nb = ln[7 * jid + l];
                    if (na == nb) continue;
                    if (checklink(na, nb) == 1) {
                        const double vb[3] = {xr[3 * nb] - xr[3 * jid], xr[3 * nb + 1] - xr[3 * jid + 1],
                                              xr[3 * nb + 2] - xr[3 * jid + 2]};
                        const double cc[3] = {va[1] * vb[2] - va[2] * vb[1], va[2] * vb[0] - va[0] * vb[2],
                                              va[0] * vb[1] - va[1] * vb[0]};
                        const double inpro = cc[0] * (xr[3 * jid] - gx) + cc[1] * (xr[3 * jid + 1] - gy) +
                                             cc[2] * (xr[3 * jid + 2] - gz);
                        if (inpro > 0) { lns[k] = nb; nfound++; }
                    }
                }
            }
            if (nfound + 1 != nln[jid]) return 3;
            for (int k = 0; k < nln[jid]; k++) ln[7 * jid + k] = lns[k];
        }
        // sort_link_order (init2.cu:169-203)
        for (int jid = voff[c]; jid < voff[c + 1]; jid++) {
            int *l = lo + 7 * jid;
            const int cnt = nln[jid];
            l[cnt] = jid;
            for (int k = 0; k < cnt + 1; k++)
                for (int m = cnt; m > k; m--)
                    if (l[m - 1] > l[m]) std::swap(l[m - 1], l[m]);
        }
        // Element connectivity (init1.cu:741-795): the edge-adjacent
        // elements of e in ascending element order.  An edge map gives the
        // same list as the reference's all-pairs search.
        std::map<std::pair<int, int>, std::vector<int>> edges;
        for (int eid = eoff[c]; eid < eoff[c + 1]; eid++)
            for (int k = 0; k < 3; k++) {
                int a = ele[3 * eid + k], b = ele[3 * eid + (k + 1) % 3];
                if (a > b) std::swap(a, b);
                edges[{a, b}].push_back(eid);
            }
        for (int eid = eoff[c]; eid < eoff[c + 1]; eid++) {
            std::vector<int> nb;
            for (int k = 0; k < 3; k++) {
                int a = ele[3 * eid + k], b = ele[3 * eid + (k + 1) % 3];
                if (a > b) std::swap(a, b);
                for (int o : edges[{a, b}])
                    if (o != eid) nb.push_back(o);
            }
            std::sort(nb.begin(), nb.end());
            if (nb.size() != 3) return 3;
            for (int k = 0; k < 3; k++) {
                ec[3 * eid + k] = nb[k];
                c0[3 * eid + k] = 1.0;  // reference angle forced flat (init1.cu:847-850)
                s0[3 * eid + k] = 0.0;
            }
        }
            return 0;
    };
    // cells are independent (disjoint vertex and element ranges): one cell per
    // OpenMP task, the first error code of any cell is returned
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(max : err)
    for (int c = 0; c < nc; c++) err = std::max(err, topology_cell(c));
    if (err) return err;
    // set_contravariant (init2.cu:206-255)
    for (int eid = 0; eid < ne; eid++) {
        const int ea = ele[3 * eid], eb = ele[3 * eid + 1], ecc = ele[3 * eid + 2];
        double v1[3], v2[3];
        for (int j = 0; j < 3; j++) { v1[j] = xr[3 * eb + j] - xr[3 * ea + j]; v2[j] = xr[3 * ecc + j] - xr[3 * ea + j]; }
        const double lv1 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
        const double lv2 = v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2];
        const double inv12 = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
        const double dv12 = lv1 * lv2 - inv12 * inv12;
        nref[4 * eid + 0] = lv2 / dv12; nref[4 * eid + 1] = -inv12 / dv12;
        nref[4 * eid + 2] = -inv12 / dv12; nref[4 * eid + 3] = lv1 / dv12;
    }
    return 0;
}

// ------------------------------------------------------------------------
// One time-loop body (main.cu:73-110)
// ------------------------------------------------------------------------
int orc_step(const orc_params *p, orc_state *s, int iter, int *vol_iters) {
    if (vol_iters) *vol_iters = 0;
    if ((iter - 1) % p->m == 0 && s->nc > 0) {
        const int f = orc_stress(p, s, iter);
        if (f) return f;
        if (s->adh) orc_adhesion_step(p, s, s->adh, iter);  // main.cu:84-86
        orc_ibm_st1(p, s);
    }
    orc_lbm_step(p, s, iter);
    if (iter % p->m == 0 && s->nc > 0) {
        orc_ibm_st2(p, s);
        if (p->vof_period > 0) {
            orc_thinc_wlic(p, s);
            if (iter % p->vof_period == 0 && orc_front_tracking(p, s, nullptr, nullptr) != 0) return 5;
        }
        if (p->volume_correction) {
            const int k = orc_volume_correction(p, s);
            if (k < 0) return 3;
            if (vol_iters) *vol_iters = k;
        }
    }
    return 0;
}

// ------------------------------------------------------------------------
// Energies used only by the finite-difference pins
// ------------------------------------------------------------------------
double orc_skalak_energy(const orc_params *p, const orc_state *s) {
    const int ne = s->eoff[s->nc];
    double W = 0.0;
    for (int j = 0; j < ne; j++) {
        const int ea = s->ele[3 * j], eb = s->ele[3 * j + 1], ec = s->ele[3 * j + 2];
        double v1[3], v2[3];
        edge_vec(p, s->xd, ea, eb, v1);
        edge_vec(p, s->xd, ea, ec, v2);
        const double *n = s->nref + 4 * j;
        const double g11 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
        const double g22 = v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2];
        const double g12 = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
        const double detg = g11 * g22 - g12 * g12;
        const double detG_inv = n[0] * n[3] - n[1] * n[2];
        const double A0 = 0.5 / std::sqrt(detG_inv);
        const double I1 = n[0] * g11 + n[1] * g12 + n[2] * g12 + n[3] * g22 - 2.0;
        const double I2 = detg * detG_inv - 1.0;
        W += A0 * 0.25 * (I1 * I1 + 2.0 * I1 - 2.0 * I2 + p->c * I2 * I2);
    }
    return W;
}

double orc_bending_energy(const orc_params *p, const orc_state *s) {
    const int ne = s->eoff[s->nc];
    double E = 0.0;
    for (int e1 = 0; e1 < ne; e1++)
        for (int k = 0; k < 3; k++) {
            int k1, k2, k3, k4;
            hinge(s->ele, e1, s->ec[3 * e1 + k], k1, k2, k3, k4);
            if (k1 < 0) continue;
            HingeGeom h;
            hinge_geom(p, s, k1, k2, k3, k4, h);
            const double *xi = h.xi, *ze = h.ze;
            const double lxi = std::sqrt(xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2]);
            const double lze = std::sqrt(ze[0] * ze[0] + ze[1] * ze[1] + ze[2] * ze[2]);
            const double cost = (xi[0] * ze[0] + xi[1] * ze[1] + xi[2] * ze[2]) / lxi / lze;
            const double side = (xi[0] - ze[0]) * (h.t1[0] - h.t2[0]) + (xi[1] - ze[1]) * (h.t1[1] - h.t2[1])
                              + (xi[2] - ze[2]) * (h.t1[2] - h.t2[2]);
            // hinges that fourpoint_bend skips (sin^2 < 1e-6, stress.cu:536)
            // carry no force: leave them out of the energy as well
            if (1.0 - cost * cost < 1.0e-6) continue;
            double sint = std::sqrt(1.0 - cost * cost);
            if (side < 0.0) sint = -sint;
            const double cost0 = s->c0[3 * e1 + k], sint0 = s->s0[3 * e1 + k];
            // each hinge is visited from both elements: half the energy each
            E += 0.5 * 2.0 / std::sqrt(3.0) * p->br * (1.0 - (cost * cost0 + sint * sint0));
        }
    return E;
}

}  // extern "C"
