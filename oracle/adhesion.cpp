// FP64 host ORACLE, cytoadhesion part (test infrastructure only; see
// oracle.h).  A plain restatement of adhesion.cu for the slab
// (CapsuleSHEARFLOW) with the linear spring or WLC law, in the reference's
// per-vertex arrays.  The two choices the reference leaves to the hardware
// are fixed here and in the library alike: the random numbers come from a
// counter-based generator, and a receptor wanted by several springs in one
// call goes to the lowest spring id, the others moving to their next
// candidate (rounds until no spring proposes).  Detachments are applied
// before attachments, and a spring that detaches does not re-attach in the
// same call (in the reference one draw decides either branch).
#include "oracle.h"

#include <climits>
#include <cmath>
#include <map>
#include <vector>

namespace {

inline double min_image_signed(double d, double L, double xi, double xj) {
    // adhesion.cu:656-668: |d| > L/2 moves d by L towards the image
    if (std::fabs(d) > L / 2.0) {
        if (xi > xj) d += L;
        else if (xi < xj) d -= L;
    }
    return d;
}

struct Rec {
    double x, y, z;
};

// Receptor id = iz * nxw + ix: the slab's wall y = -LY/2 (adhesion.cu:452-
// 457) or, in the tube, the ring of radius R at angle ix * dxw (nxw = ntheta,
// dxw = dtheta, adhesion.cu:475-490).
inline Rec receptor(const orc_params *p, const orc_adhesion *a, int id) {
    const int iz = id / a->nxw, ix = id - iz * a->nxw;
    if (p->tube)
        return Rec{p->tube_radius * std::cos(ix * a->dxw), p->tube_radius * std::sin(ix * a->dxw),
                   (iz - a->nzw / 2) * a->dzw};
    return Rec{(ix - a->nxw / 2) * a->dxw, -0.5 * p->ly, (iz - a->nzw / 2) * a->dzw};
}

// StoreNeighWall_CUBE_GPU (adhesion.cu:109-206): receptors within LIMIT dx,
// nearest first; ties keep the scan order.  At most MAXN are kept (the
// reference writes past its MAXN slots when more are in range).
void search(const orc_params *p, const orc_state *s, orc_adhesion *a, int v) {
    const int MAXN = a->max_neighwall;
    int *ids = a->id_wall + (size_t)v * MAXN;
    double *rs = a->r_wall + (size_t)v * MAXN;
    for (int k = 0; k < MAXN; k++) { ids[k] = -1; rs[k] = 1.0e15; }  // InitNeighWall_GPU
    if (!a->is_knob[v]) return;
    const double dx = p->dx, xi = s->x[v], yi = s->y[v], zi = s->z[v];
    if (p->tube ? std::sqrt(xi * xi + yi * yi) < p->tube_radius - a->limit * dx  // adhesion.cu:248-249
                : std::sqrt(yi * yi) < 0.5 * p->ly - a->limit * dx)
        return;
    const int iz = (int)((zi - (-p->lz / 2.0)) / a->dzw);
    const int span = a->limit * (int)(dx / a->dzw);
    int cnt = 0;
    for (int i = 0; i < 2 * span; i++) {
        int jz = iz + i - (span - 1);
        double zj = (jz - a->nzw / 2) * a->dzw;
        if (jz < 0) { jz += a->nzw; zj += p->lz; }
        else if (jz >= a->nzw) { jz -= a->nzw; zj -= p->lz; }
        for (int jx = 0; jx < a->nxw; jx++) {
            const int id = jz * a->nxw + jx;
            const Rec rc = receptor(p, a, id);
            const double r = std::sqrt((rc.x - xi) * (rc.x - xi) + (rc.y - yi) * (rc.y - yi) + (zj - zi) * (zj - zi));
            if (r > a->limit * dx) continue;
            // stable insertion into the sorted list of at most MAXN
            int pos = cnt < MAXN ? cnt : MAXN;
            while (pos > 0 && rs[pos - 1] > r) pos--;
            if (pos >= MAXN) continue;
            for (int k = (cnt < MAXN ? cnt : MAXN - 1); k > pos; k--) { ids[k] = ids[k - 1]; rs[k] = rs[k - 1]; }
            ids[pos] = id;
            rs[pos] = r;
            if (cnt < MAXN) cnt++;
        }
    }
}

// The extension measure of the Bell rates: LINEAR_SPRING (adhesion.cu:562-
// 563) or WLC with contour length LL = Lcont dx (adhesion.cu:564-566).
double bell_extension(const orc_params *p, const orc_adhesion *a, double r) {
    if (a->wlc) {
        const double LL = a->lcont * p->dx;
        return std::fmax(0.0, 0.25 / ((1.0 - r / LL) * (1.0 - r / LL)) + r / LL - 0.25);
    }
    return std::fmax(0.0, (r - a->l0 * p->dx) * a->rc);
}

double bell_force(const orc_adhesion *a, double dr) {
    return a->wlc ? std::fabs(a->f_wlc * dr) : std::fabs(a->k_ad * dr);
}

double on_probability(const orc_params *p, const orc_adhesion *a, double r, double *dr_out) {
    // adhesion.cu:562-573
    const double dr = bell_extension(p, a, r);
    const double f_ad = bell_force(a, dr);
    const double k_on = a->k_on0 / a->shear_rate * std::exp(f_ad * (a->chi_beta - 0.5 * dr) / a->kbt);
    *dr_out = dr;
    return 1.0 - std::exp(-k_on * p->dt * (double)p->m);
}

// Attachment rounds over the springs in `active` (sorted spring ids):
// each proposes its next acceptable free candidate; the lowest id wins.
void attach_rounds(const orc_params *p, const orc_state *s, orc_adhesion *a, const std::vector<int> &active,
                   bool always, int iter) {
    const int NS = a->springs_per_knob, MAXN = a->max_neighwall;
    std::vector<int> pos(active.size(), 0);
    std::vector<char> done(active.size(), 0);
    for (;;) {
        std::map<int, int> best;  // receptor -> lowest proposing spring (index into active)
        std::vector<int> prop(active.size(), -1);
        for (size_t k = 0; k < active.size(); k++) {
            if (done[k]) continue;
            const int sid = active[k], v = sid / NS;
            const double u = always ? 0.0 : orc_adh_uniform(a->seed, iter, sid);
            int c = pos[k];
            for (; c < MAXN; c++) {
                const int id = a->id_wall[(size_t)v * MAXN + c];
                if (id == -1) continue;
                if (!always) {
                    double dr;
                    if (on_probability(p, a, a->r_wall[(size_t)v * MAXN + c], &dr) <= u) continue;
                }
                if (a->is_attached[id]) continue;
                break;
            }
            if (c >= MAXN) { done[k] = 1; continue; }
            pos[k] = c;
            prop[k] = a->id_wall[(size_t)v * MAXN + c];
            auto it = best.find(prop[k]);
            if (it == best.end() || active[it->second] > sid) best[prop[k]] = (int)k;
        }
        if (best.empty()) break;
        for (size_t k = 0; k < active.size(); k++) {
            if (prop[k] < 0) continue;
            if (best[prop[k]] == (int)k) {
                const int sid = active[k], v = sid / NS;
                a->attach[sid] = prop[k];
                a->is_attached[prop[k]] = 1;
                if (!always) {
                    double dr;
                    on_probability(p, a, a->r_wall[(size_t)v * MAXN + pos[k]], &dr);
                    a->r_ij[sid] = dr;
                }
                done[k] = 1;
            } else {
                pos[k]++;
            }
        }
    }
    (void)s;
}

}  // namespace

extern "C" {

double orc_adh_uniform(unsigned long long seed, long long iter, long long sid) {
    // splitmix64 (Steele, Lea, Flood 2014) of the key, mapped to (0, 1]
    unsigned long long z = seed ^ ((unsigned long long)iter * 0x9E3779B97F4A7C15ull) ^
                           ((unsigned long long)sid * 0xD1B54A32D192ED03ull);
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)((z >> 11) + 1) * (1.0 / 9007199254740992.0);
}

int orc_adhesion_grid(const orc_params *p, orc_adhesion *a) {
    double dz = std::sqrt(1.0 / a->receptor_density) / a->rc;
    if (p->tube) {  // wall_init, CapsuleCHANNELFLOW (adhesion.cu:467-480)
        a->nzw = (int)(p->lz / dz + 1.0e-03);
        a->dzw = p->lz / (double)a->nzw;
        a->dxw = a->dzw / p->tube_radius;
        a->nxw = (int)(2.0 * M_PI / a->dxw);
        return a->nxw * a->nzw;
    }
    // wall_init, CapsuleSHEARFLOW (adhesion.cu:441-458)
    a->nzw = (int)(p->lz / dz + 1.0e-03);
    a->nxw = (int)(p->lx / dz + 1.0e-03);
    a->dzw = p->lz / (double)a->nzw;
    a->dxw = p->lx / (double)a->nxw;
    return a->nxw * a->nzw;
}

void orc_adhesion_init(const orc_params *p, orc_state *s, orc_adhesion *a) {
    const int nv = s->voff[s->nc], NS = a->springs_per_knob;
    const int nrec = orc_adhesion_grid(p, a);
    for (int i = 0; i < 3 * nv; i++) a->f_adh[i] = 0.0;
    for (int i = 0; i < nv * NS; i++) { a->attach[i] = -1; a->r_ij[i] = -1.0e10; }
    for (int i = 0; i < nrec; i++) a->is_attached[i] = 0;
    for (int v = 0; v < nv; v++) search(p, s, a, v);
    std::vector<int> active;
    for (int v = 0; v < nv; v++)
        if (a->is_knob[v])
            for (int k = 0; k < NS; k++) active.push_back(v * NS + k);
    attach_rounds(p, s, a, active, true, 0);  // AttachmentPointAtWall_GPU: nearest free
}

void orc_adhesion_step(const orc_params *p, orc_state *s, orc_adhesion *a, int iter) {
    const int nv = s->voff[s->nc], NS = a->springs_per_knob;
    const double dx = p->dx;
    // AdhesiveSpringForce_GPU (adhesion.cu:617-706).  With WLC the
    // repulsive term is set only where |dy| < L0 dx and otherwise keeps the
    // value of the knob's previous spring, as the reference's variable does.
    for (int v = 0; v < nv; v++) {
        double f[3] = {0.0, 0.0, 0.0};
        if (a->is_knob[v]) {
            const double xi = s->x[v], yi = s->y[v], zi = s->z[v];
            double repulf = 0.0;
            for (int k = 0; k < NS; k++) {
                const int id = a->attach[v * NS + k];
                if (id == -1) continue;
                const Rec rc = receptor(p, a, id);
                const double ddx = p->tube ? rc.x - xi : min_image_signed(rc.x - xi, p->lx, xi, rc.x);
                const double ddy = rc.y - yi;
                const double ddz = min_image_signed(rc.z - zi, p->lz, zi, rc.z);
                const double r = std::sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                if (a->wlc) {
                    const double LL = a->lcont * dx;
                    const double dr = 0.25 / ((1.0 - r / LL) * (1.0 - r / LL)) + r / LL - 0.25;
                    if (std::fabs(ddy) < a->l0 * dx) {
                        const double ex = std::exp(-a->alpha2 * std::fabs(ddy));
                        repulf = a->alpha1 / (a->gs * a->ra) * a->alpha2 * ex / (1.0 - ex) * ddy / std::fabs(ddy);
                    }
                    f[0] += a->f_wlc / (a->gs * a->ra) * dr * ddx / r;
                    f[1] += a->f_wlc / (a->gs * a->ra) * dr * ddy / r + repulf;
                    f[2] += a->f_wlc / (a->gs * a->ra) * dr * ddz / r;
                    continue;
                }
                const double dr = r - a->l0 * dx;
                f[0] += a->k_ad / a->gs * dr * ddx / r;
                f[1] += a->k_ad / a->gs * dr * ddy / r;
                f[2] += a->k_ad / a->gs * dr * ddz / r;
            }
        }
        for (int c = 0; c < 3; c++) a->f_adh[3 * v + c] = f[c];
    }
    // CytoAdhesionBELL (adhesion.cu:730-770): search, then the Bell model
    for (int v = 0; v < nv; v++) search(p, s, a, v);
    std::vector<int> active;
    for (int v = 0; v < nv; v++) {
        if (!a->is_knob[v]) continue;
        for (int k = 0; k < NS; k++) {
            const int sid = v * NS + k, id = a->attach[sid];
            if (id == -1) { active.push_back(sid); continue; }
            // detachment (adhesion.cu:584-612): z image only
            const Rec rc = receptor(p, a, id);
            const double ddx = rc.x - s->x[v], ddy = rc.y - s->y[v];
            const double ddz = min_image_signed(rc.z - s->z[v], p->lz, s->z[v], rc.z);
            const double r = std::sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
            const double dr = bell_extension(p, a, r);
            const double f_ad = bell_force(a, dr);
            const double k_off = a->k_off0 / a->shear_rate * std::exp(f_ad * a->chi_beta / a->kbt);
            const double p_off = 1.0 - std::exp(-k_off * p->dt * (double)p->m);
            if (p_off <= orc_adh_uniform(a->seed, iter, sid)) continue;
            a->is_attached[id] = 0;
            a->attach[sid] = -1;
            a->r_ij[sid] = dr;
        }
    }
    attach_rounds(p, s, a, active, false, iter);
}

}  // extern "C"
