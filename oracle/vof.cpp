// FP64 host ORACLE, VOF indicator part (test infrastructure only; see
// oracle.h).  A plain restatement of front tracking (fronttrack.cu), its
// red-black SOR solve (matrix_solver.cu) and the THINC/WLIC advection
// (thincwlic.cu) on the reference's padded lattice, for the slab of
// CapsuleSHEARFLOW (walls in y, periodic x and z) and the tube of
// CapsuleCHANNELFLOW (walls outside the disc, periodic z): everything else
// follows the bc mask.
//
// The one choice the reference leaves open: its red-black SOR updates a BUFF
// row (x_i - x_image = 0) in the same kernel as the COMP rows of its colour,
// so when the image has the same colour the BUFF row reads it while it is
// being written.  Here, within each colour, the COMP / WALL / padding rows
// are updated first and the BUFF rows then read the updated image: one of
// the orders the reference can produce, and a deterministic one.
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int COMP = ORC_COMP, BUFF = ORC_BUFF, WALL = ORC_WALL;

struct PGrid {
    int nx, ny, nz;
    void coords(int i, int &ix, int &iy, int &iz) const {
        iz = i / (nx * ny);
        iy = i / nx - ny * iz;
        ix = i - nx * iy - nx * ny * iz;
    }
    // periodic image of a BUFF node, x and z (fronttrack.cu:117-125,
    // thincwlic.cu:264-270)
    int image(int i) const {
        int ix, iy, iz;
        coords(i, ix, iy, iz);
        int j = i;
        if (ix <= 1) j += nx - 4;
        if (ix >= nx - 2) j -= nx - 4;
        if (iz <= 1) j += nx * ny * (nz - 4);
        if (iz >= nz - 2) j -= nx * ny * (nz - 4);
        return j;
    }
};

inline PGrid pgrid(const orc_params *p) { return PGrid{p->nx, p->ny, p->nz}; }

}  // namespace

extern "C" {

void orc_ft_gradient(const orc_params *p, const orc_state *s, double *gx, double *gy, double *gz) {
    const PGrid g = pgrid(p);
    const int nx = g.nx, ny = g.ny, nz = g.nz, n = nx * ny * nz;
    const double dx = p->dx, LX = p->lx, LZ = p->lz;
    // front_tracking_source1 (fronttrack.cu:314-333)
    for (int i = 0; i < n; i++) gx[i] = gy[i] = gz[i] = 0.0;
    // front_tracking_source2 (fronttrack.cu:336-441), elements in order
    const int ne = s->eoff[s->nc];
    for (int eid = 0; eid < ne; eid++) {
        const int j1 = s->ele[3 * eid], j2 = s->ele[3 * eid + 1], j3 = s->ele[3 * eid + 2];
        double tx1 = s->x[j1], tx2 = s->x[j2], tx3 = s->x[j3];
        double ty1 = s->y[j1], ty2 = s->y[j2], ty3 = s->y[j3];
        double tz1 = s->z[j1], tz2 = s->z[j2], tz3 = s->z[j3];
        double xx1 = tx2 - tx1, yy1 = ty2 - ty1, zz1 = tz2 - tz1;
        if (xx1 > LX * 0.5) { xx1 -= LX; tx1 += LX; }
        else if (xx1 < -LX * 0.5) { xx1 += LX; tx2 += LX; }
        if (zz1 > LZ * 0.5) { zz1 -= LZ; tz1 += LZ; }
        else if (zz1 < -LZ * 0.5) { zz1 += LZ; tz2 += LZ; }
        double xx2 = tx3 - tx1, yy2 = ty3 - ty1, zz2 = tz3 - tz1;
        if (xx2 > LX * 0.5) { xx2 -= LX; tx1 += LX; }
        else if (xx2 < -LX * 0.5) { xx2 += LX; tx3 += LX; }
        if (zz2 > LZ * 0.5) { zz2 -= LZ; tz1 += LZ; }
        else if (zz2 < -LZ * 0.5) { zz2 += LZ; tz3 += LZ; }
        const double xg = (tx1 + tx2 + tx3) / 3.0, yg = (ty1 + ty2 + ty3) / 3.0, zg = (tz1 + tz2 + tz3) / 3.0;
        const double xtr = xg + (double)(nx - 1) * dx / 2.0;
        const double ytr = yg + (double)(ny - 1) * dx / 2.0;
        const double ztr = zg + (double)(nz - 1) * dx / 2.0;
        const int x1 = (int)std::floor(xtr / dx) - 1, x2 = (int)std::ceil(xtr / dx) + 1;
        const int y1 = (int)std::floor(ytr / dx) - 1, y2 = (int)std::ceil(ytr / dx) + 1;
        const int z1 = (int)std::floor(ztr / dx) - 1, z2 = (int)std::ceil(ztr / dx) + 1;
        double nvx = -(yy1 * zz2 - zz1 * yy2);
        double nvy = -(zz1 * xx2 - xx1 * zz2);
        double nvz = -(xx1 * yy2 - yy1 * xx2);
        const double nn = std::sqrt(nvx * nvx + nvy * nvy + nvz * nvz);
        const double ds = nn / 2.0;
        nvx /= nn; nvy /= nn; nvz /= nn;
        for (int iz = z1; iz <= z2; iz++)
            for (int iy = y1; iy <= y2; iy++)
                for (int ix = x1; ix <= x2; ix++) {
                    if (ix < 0 || ix >= nx || iy < 0 || iy >= ny || iz < 0 || iz >= nz) continue;
                    const int i = ix + nx * iy + nx * ny * iz;
                    const double xx = (double)ix * dx - xtr, yy = (double)iy * dx - ytr, zz = (double)iz * dx - ztr;
                    const double xcos = 1.0 + std::cos(M_PI * xx / (2.0 * dx));
                    const double ycos = 1.0 + std::cos(M_PI * yy / (2.0 * dx));
                    const double zcos = 1.0 + std::cos(M_PI * zz / (2.0 * dx));
                    const double dlt = xcos * ycos * zcos / (64.0 * dx * dx * dx);
                    gx[i] += dlt * nvx * ds;
                    gy[i] += dlt * nvy * ds;
                    gz[i] += dlt * nvz * ds;
                }
    }
    // front_tracking_source3 (fronttrack.cu:444-499): each COMP node within
    // two nodes of an x/z end takes its halo images' sums and writes the
    // total back to them (every halo node belongs to one COMP node)
    const int sx = nx - 4, sz = nx * ny * (nz - 4);
    for (int i = 0; i < n; i++) {
        if (s->bc[i] != COMP) continue;
        int ix, iy, iz;
        g.coords(i, ix, iy, iz);
        int i1 = -1, i2 = -1, i3 = -1;
        // the tube (CapsuleCHANNELFLOW, fronttrack.cu:466-482) folds z only
        const bool xl = !p->tube && ix <= 3, xh = !p->tube && ix >= nx - 4, zl = iz <= 3, zh = iz >= nz - 4;
        const int ox = xl ? sx : -sx, oz = zl ? sz : -sz;
        if ((xl || xh) && (zl || zh)) { i1 = i + ox; i2 = i + oz; i3 = i + ox + oz; }
        else if (xl || xh) i1 = i + ox;
        else if (zl || zh) i1 = i + oz;
        for (int k : {i1, i2, i3})
            if (k >= 0) { gx[i] += gx[k]; gy[i] += gy[k]; gz[i] += gz[k]; }
        for (int k : {i1, i2, i3})
            if (k >= 0) { gx[k] = gx[i]; gy[k] = gy[i]; gz[k] = gz[i]; }
    }
}

void orc_ft_rhs(const orc_params *p, const int *bc, const double *gx, const double *gy, const double *gz,
                double *b) {
    const int nx = p->nx, ny = p->ny, n = p->nx * p->ny * p->nz, nxy = nx * ny;
    const double dx = p->dx, diag = -6.0 / (dx * dx);
    for (int i = 0; i < n; i++) {
        double dg = 0.0;
        if (bc[i] == COMP)
            dg = (gx[i + 1] - gx[i - 1] + gy[i + nx] - gy[i - nx] + gz[i + nxy] - gz[i - nxy]) / (2.0 * dx);
        b[i] = bc[i] == COMP ? dg / diag : dg / 1.0;  // jacobi_source: the diagonal is 1 off COMP
    }
}

int orc_ft_sor(const orc_params *p, const int *bc, const double *b, double *x, double pad0, int *iters) {
    const PGrid g = pgrid(p);
    const int nx = g.nx, ny = g.ny, n = nx * ny * g.nz, nxy = nx * ny;
    const int mn = orc_expand2pow2(n);
    const double dx = p->dx, alpha = 1.3;
    // Jacobi-scaled off-diagonal of a COMP row (fronttrack.cu:97-110, 294-311)
    const double a = (1.0 / (dx * dx)) / (-6.0 / (dx * dx));
    std::vector<double> X(mn, pad0), r(mn, 0.0);
    for (int i = 0; i < n; i++) X[i] = x[i];
    double bb = 0.0;
    for (int i = 0; i < n; i++) bb += b[i] * b[i];
    std::vector<int> parity(mn);
    for (int i = 0; i < mn; i++) {
        if (i < n) {
            int ix, iy, iz;
            g.coords(i, ix, iy, iz);
            parity[i] = (ix + iy + iz) % 2;
        } else {
            parity[i] = i % 2;
        }
    }
    auto row = [&](int i) {
        double sum;
        if (i >= n || bc[i] == WALL) {
            sum = 1.0 * X[i];
        } else if (bc[i] == COMP) {
            sum = 0.0;
            sum += a * X[i - nxy];
            sum += a * X[i - nx];
            sum += a * X[i - 1];
            sum += 1.0 * X[i];
            sum += a * X[i + 1];
            sum += a * X[i + nx];
            sum += a * X[i + nxy];
        } else {
            sum = X[i] - X[g.image(i)];
        }
        const double bi = i < n ? b[i] : 0.0;
        r[i] = bi - sum;
        X[i] += alpha * r[i];
    };
    int it;
    bool conv = false;
    for (it = 0; it < p->sor_itrmax; it++) {
        for (int c = 0; c < 2; c++) {
            for (int i = 0; i < mn; i++)
                if (parity[i] == c && !(i < n && bc[i] == BUFF)) row(i);
            for (int i = 0; i < n; i++)
                if (parity[i] == c && bc[i] == BUFF) row(i);
        }
        double rr = 0.0;
        for (int i = 0; i < mn; i++) rr += r[i] * r[i];
        const double err = std::sqrt(rr / bb);
        if (err < p->sor_residual) { conv = true; break; }
    }
    for (int i = 0; i < n; i++) x[i] = X[i];
    if (iters) *iters = conv ? it + 1 : it;
    return conv ? 0 : 2;
}

int orc_front_tracking(const orc_params *p, orc_state *s, int *iters, int *attempts) {
    const int n = p->nx * p->ny * p->nz;
    std::vector<double> gx(n), gy(n), gz(n), b(n), x(n);
    orc_ft_gradient(p, s, gx.data(), gy.data(), gz.data());
    orc_ft_rhs(p, s->bc, gx.data(), gy.data(), gz.data(), b.data());
    int cnt = 0, flag, its = 0;
    do {  // fronttrack.cu:707-720: a fresh start from x = cnt after each failure
        for (int i = 0; i < n; i++) x[i] = (double)cnt;
        flag = orc_ft_sor(p, s->bc, b.data(), x.data(), (double)cnt, &its);
        cnt++;
        if (cnt > 10) {
            if (attempts) *attempts = cnt;
            return 5;
        }
    } while (flag != 0);
    std::memcpy(s->vf, x.data(), sizeof(double) * n);
    if (iters) *iters = its;
    if (attempts) *attempts = cnt;
    return 0;
}

void orc_vof_normals(const orc_params *p, const double *vf, double *n1x, double *n1y, double *n1z) {
    const PGrid g = pgrid(p);
    const int nx = g.nx, ny = g.ny, nz = g.nz, n = nx * ny * nz, nxy = nx * ny;
    const double dx = p->dx;
    std::vector<double> ax(n, 0.0), ay(n, 0.0), az(n, 0.0);
    for (int i = 0; i < n; i++) {  // normal3d_st1 (thincwlic.cu:12-56)
        int ix, iy, iz;
        g.coords(i, ix, iy, iz);
        if (ix == 0 || iy == 0 || iz == 0) continue;
        ax[i] = 0.25 / dx * (vf[i] - vf[i - 1] + vf[i - nx] - vf[i - 1 - nx] + vf[i - nxy] - vf[i - 1 - nxy]
                             + vf[i - nx - nxy] - vf[i - 1 - nx - nxy]);
        ay[i] = 0.25 / dx * (vf[i] - vf[i - nx] + vf[i - 1] - vf[i - 1 - nx] + vf[i - nxy] - vf[i - nx - nxy]
                             + vf[i - 1 - nxy] - vf[i - 1 - nx - nxy]);
        az[i] = 0.25 / dx * (vf[i] - vf[i - nxy] + vf[i - 1] - vf[i - 1 - nxy] + vf[i - nx] - vf[i - nx - nxy]
                             + vf[i - 1 - nx] - vf[i - 1 - nx - nxy]);
    }
    for (int i = 0; i < n; i++) {  // normal3d_st2 (thincwlic.cu:59-106)
        int ix, iy, iz;
        g.coords(i, ix, iy, iz);
        if (ix == 0 || iy == 0 || iz == 0 || ix == nx - 1 || iy == ny - 1 || iz == nz - 1) {
            n1x[i] = n1y[i] = n1z[i] = 0.0;
            continue;
        }
        const double *A[3] = {ax.data(), ay.data(), az.data()};
        double *out[3] = {n1x, n1y, n1z};
        for (int d = 0; d < 3; d++) {
            const double *v = A[d];
            out[d][i] = 0.125 * (v[i] + v[i + nxy] + v[i + 1] + v[i + 1 + nxy] + v[i + nx] + v[i + nx + nxy]
                                 + v[i + 1 + nx] + v[i + 1 + nx + nxy]);
        }
    }
}

double orc_thinc_flux(double c, double alpha, double uc, int isgn, int m, double dt, double dx) {
    // thincwlic.cu:187-197 (expf/coshf on float arguments, as written there)
    const double beta = 2.0;
    const double a1 = expf((float)(beta * (2.0 * c - 1.0) / alpha));
    const double a3 = expf((float)beta);
    const double xc = 0.5 / beta * std::log((a3 * a3 - a1 * a3) / (a1 * a3 - 1.0));
    const double a4 = coshf((float)(beta * ((double)isgn - uc * (double)m * dt / dx - xc)));
    const double a5 = coshf((float)(beta * ((double)isgn - xc)));
    return 0.5 * (uc * (double)m * dt - alpha * dx / beta * std::log(a4 / a5));
}

void orc_thinc_wlic(const orc_params *p, orc_state *s) {
    const PGrid g = pgrid(p);
    const int nx = g.nx, ny = g.ny, nz = g.nz, n = nx * ny * nz;
    const double dt = p->dt, dx = p->dx;
    double *vf = s->vf;
    std::vector<double> n1x(n), n1y(n), n1z(n), flux(n);
    const double *vel[3] = {s->um, s->vm, s->wm};
    const int acs_of[3] = {1, nx, nx * ny};
    for (int idr = 0; idr < 3; idr++) {
        orc_vof_normals(p, vf, n1x.data(), n1y.data(), n1z.data());
        const int acs = acs_of[idr];
        const double *u = vel[idr];
        const double *n1 = idr == 0 ? n1x.data() : (idr == 1 ? n1y.data() : n1z.data());
        for (int i = 0; i < n; i++) {  // wlic1d4md_st1 (thincwlic.cu:108-214)
            flux[i] = 0.0;
            int ix, iy, iz;
            g.coords(i, ix, iy, iz);
            if (ix == 0 || iy == 0 || iz == 0) continue;
            if (ix == 1 || iy == 1 || iz == 1) continue;
            if (ix == nx - 1 || iy == ny - 1 || iz == nz - 1) continue;
            const double up = u[i], um = u[i - acs];
            const double uc = (up + um) / 2.0;
            const int isgn = uc <= 0.0 ? 0 : 1;
            const int ii = uc <= 0.0 ? i : i - acs;
            const double nn = std::fabs(n1x[ii]) + std::fabs(n1y[ii]) + std::fabs(n1z[ii]);
            const double upwind = vf[ii] * uc * (double)p->m * dt;
            if (vf[ii] >= 1.0 - 1.0e-6 || vf[ii] <= 1.0e-6) {
                flux[i] = upwind;
            } else if (nn <= 1.0e-6) {
                flux[i] = upwind;
            } else {
                const int ib = ii - acs;  // the three ib assignments leave ii - acs
                const double alpha = vf[ib] <= vf[ii + acs] ? 1.0 : -1.0;
                const double f = orc_thinc_flux(vf[ii], alpha, uc, isgn, p->m, dt, dx);
                const double weight = std::fabs(n1[ii]) / nn;
                flux[i] = f * weight + vf[ii] * uc * (double)p->m * dt * (1.0 - weight);
            }
            if (s->bc[i] == WALL || s->bc[i - acs] == WALL) flux[i] = 0.0;
        }
        for (int i = 0; i < n; i++)  // wlic1d4md_st2 (thincwlic.cu:217-250)
            if (s->bc[i] == COMP) vf[i] = vf[i] - (flux[i + acs] - flux[i]) / dx;
        for (int i = 0; i < n; i++) {  // wlic1d4md_st3 (thincwlic.cu:253-313)
            if (s->bc[i] == BUFF) vf[i] = vf[g.image(i)];
            else if (s->bc[i] == WALL) vf[i] = 0.0;
        }
    }
}

}  // extern "C"
