// Host-side preprocessing of the membrane meshes: icosphere and Evans-Fung
// templates (replacing the reference's ./cell/*.dat inputs, init1.cu:490-519),
// seeded cell placement (replacing the time-seeded rotation of
// init1.cu:565-579) and the FEM topology of init1.cu:626-853 and
// init2.cu:12-400.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <map>
#include <unordered_map>
#include <utility>
#include <vector>

#include "lbmibm.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
};

void icosahedron(std::vector<std::array<double, 3>> &v, std::vector<std::array<int, 3>> &f) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    f = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
         {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
         {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (auto &p : v) {
        const double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (auto &c : p) c /= r;
    }
}

void subdivide(std::vector<std::array<double, 3>> &v, std::vector<std::array<int, 3>> &f) {
    std::map<std::pair<int, int>, int> mid;
    auto midpoint = [&](int a, int b) {
        auto key = std::make_pair(std::min(a, b), std::max(a, b));
        auto it = mid.find(key);
        if (it != mid.end()) return it->second;
        std::array<double, 3> m{(v[a][0] + v[b][0]) * 0.5, (v[a][1] + v[b][1]) * 0.5, (v[a][2] + v[b][2]) * 0.5};
        const double r = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        for (auto &c : m) c /= r;
        v.push_back(m);
        const int id = (int)v.size() - 1;
        mid[key] = id;
        return id;
    };
    std::vector<std::array<int, 3>> nf;
    nf.reserve(f.size() * 4);
    for (const auto &t : f) {
        const int a = midpoint(t[0], t[1]), b = midpoint(t[1], t[2]), c = midpoint(t[2], t[0]);
        nf.push_back({t[0], a, c});
        nf.push_back({t[1], b, a});
        nf.push_back({t[2], c, b});
        nf.push_back({a, b, c});
    }
    f.swap(nf);
}

// Morton code of a point of [-1,1]^3 quantised to 10 bits per axis.
uint32_t morton3(const std::array<double, 3> &p) {
    uint32_t key = 0;
    uint32_t q[3];
    for (int d = 0; d < 3; d++) {
        const double u = std::min(std::max(0.5 * (p[d] + 1.0), 0.0), 1.0);
        q[d] = std::min(1023u, (uint32_t)(u * 1024.0));
    }
    for (int b = 9; b >= 0; b--)
        for (int d = 0; d < 3; d++) key = (key << 1) | ((q[d] >> b) & 1u);
    return key;
}

// Renumber vertices along a Morton curve of their positions and sort the
// triangles by their lowest vertex id, keeping each triangle's cyclic order:
// vertices and elements that are close on the surface get close ids, so the
// per-vertex and per-element kernels read neighbouring data from the same
// cache lines.
void locality_order(std::vector<std::array<double, 3>> &v, std::vector<std::array<int, 3>> &f) {
    const int nv = (int)v.size();
    std::vector<int> order(nv), rank(nv);
    std::vector<uint32_t> key(nv);
    for (int i = 0; i < nv; i++) { order[i] = i; key[i] = morton3(v[i]); }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
    std::vector<std::array<double, 3>> nvx(nv);
    for (int i = 0; i < nv; i++) { rank[order[i]] = i; nvx[i] = v[order[i]]; }
    v.swap(nvx);
    for (auto &t : f) {
        for (auto &c : t) c = rank[c];
        // rotate so that the lowest id comes first (same orientation)
        while (t[0] != std::min(t[0], std::min(t[1], t[2]))) t = {t[1], t[2], t[0]};
    }
    std::stable_sort(f.begin(), f.end(), [](const std::array<int, 3> &a, const std::array<int, 3> &b) {
        return a[0] != b[0] ? a[0] < b[0] : (a[1] != b[1] ? a[1] < b[1] : a[2] < b[2]);
    });
}

}  // namespace

extern "C" {

lbmibm_status lbmibm_mesh_icosphere(int level, int *nv, int *nt, double *xyz, int *tri) {
    if (level < 0 || level > 7 || !nv || !nt) return LBMIBM_ERR_ARG;
    const int nvv = 10 * (1 << (2 * level)) + 2, ntt = 20 * (1 << (2 * level));
    *nv = nvv;
    *nt = ntt;
    if (!xyz && !tri) return LBMIBM_OK;
    std::vector<std::array<double, 3>> v;
    std::vector<std::array<int, 3>> f;
    icosahedron(v, f);
    for (int l = 0; l < level; l++) subdivide(v, f);
    if ((int)v.size() != nvv || (int)f.size() != ntt) return LBMIBM_ERR_TOPOLOGY;
    locality_order(v, f);
    if (xyz)
        for (int i = 0; i < nvv; i++)
            for (int d = 0; d < 3; d++) xyz[3 * i + d] = v[i][d];
    if (tri)
        for (int i = 0; i < ntt; i++)
            for (int d = 0; d < 3; d++) tri[3 * i + d] = f[i][d];
    return LBMIBM_OK;
}

lbmibm_status lbmibm_mesh_biconcave(int nv, double *xyz) {
    // Evans & Fung (1972) profile at unit radius, symmetry axis z:
    //   z = +-0.5 sqrt(1 - rho^2) (c0 + c1 rho^2 + c2 rho^4),  rho^2 = x^2 + y^2.
    if (nv < 0 || (nv > 0 && !xyz)) return LBMIBM_ERR_ARG;
    const double c0 = 0.207161, c1 = 2.002558, c2 = -1.122762;
    for (int i = 0; i < nv; i++) {
        const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        const double r2 = std::min(1.0, x * x + y * y);
        const double h = 0.5 * std::sqrt(std::max(0.0, 1.0 - r2)) * (c0 + c1 * r2 + c2 * r2 * r2);
        xyz[3 * i + 2] = (z >= 0.0) ? h : -h;
    }
    return LBMIBM_OK;
}

lbmibm_status lbmibm_place_cells(int nc, double r, double lxd, double lyd, double lzd, int periodic_y,
                                 unsigned long long seed, int mode, int gx, int gy, int gz,
                                 double *centres, double *rot) {
    if (nc < 0 || r <= 0.0 || !centres || !rot) return LBMIBM_ERR_ARG;
    SplitMix64 rng(seed);
    auto min_img = [](double d, double L) {
        if (d > 0.5 * L) d -= L;
        else if (d < -0.5 * L) d += L;
        return d;
    };
    if (mode == 0) {
        const double ylo = periodic_y ? -0.5 * lyd : -0.5 * lyd + r + 1.0;
        const double yhi = periodic_y ? 0.5 * lyd : 0.5 * lyd - r - 1.0;
        if (yhi <= ylo) return LBMIBM_ERR_ARG;
        for (int c = 0; c < nc; c++) {
            int tries = 0;
            for (;;) {
                if (++tries > 200000) return LBMIBM_ERR_ARG;
                const double x = (rng.uniform() - 0.5) * lxd;
                const double y = ylo + rng.uniform() * (yhi - ylo);
                const double z = (rng.uniform() - 0.5) * lzd;
                bool ok = true;
                for (int o = 0; o < c && ok; o++) {
                    const double dx = min_img(x - centres[3 * o], lxd);
                    const double dy = periodic_y ? min_img(y - centres[3 * o + 1], lyd) : y - centres[3 * o + 1];
                    const double dz = min_img(z - centres[3 * o + 2], lzd);
                    ok = dx * dx + dy * dy + dz * dz >= (2.0 * r + 1.0) * (2.0 * r + 1.0);
                }
                if (ok) { centres[3 * c] = x; centres[3 * c + 1] = y; centres[3 * c + 2] = z; break; }
            }
        }
    } else if (mode == 1) {
        if (gx <= 0 || gy <= 0 || gz <= 0 || (long long)gx * gy * gz < nc) return LBMIBM_ERR_ARG;
        const double sx = lxd / gx, sy = lyd / gy, sz = lzd / gz;
        if (std::min(sx, std::min(sy, sz)) < 2.0 * r + 1.0) return LBMIBM_ERR_ARG;
        const double jx = 0.5 * (sx - 2.0 * r - 1.0), jz = 0.5 * (sz - 2.0 * r - 1.0);
        const double jy = 0.5 * (sy - 2.0 * r - 1.0);
        for (int c = 0; c < nc; c++) {
            const int ix = c % gx, iy = (c / gx) % gy, iz = c / (gx * gy);
            double y = -0.5 * lyd + (iy + 0.5) * sy + (2.0 * rng.uniform() - 1.0) * jy;
            if (!periodic_y) {
                const double lim = 0.5 * lyd - r - 1.0;
                if (std::fabs(y) > lim) y = std::copysign(lim, y);
            }
            centres[3 * c] = -0.5 * lxd + (ix + 0.5) * sx + (2.0 * rng.uniform() - 1.0) * jx;
            centres[3 * c + 1] = y;
            centres[3 * c + 2] = -0.5 * lzd + (iz + 0.5) * sz + (2.0 * rng.uniform() - 1.0) * jz;
        }
    } else {
        return LBMIBM_ERR_ARG;
    }
    for (int c = 0; c < nc; c++) {
        // Shoemake (1992) uniform random unit quaternion.
        const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
        const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
        const double qx = a * std::sin(2.0 * M_PI * u2), qy = a * std::cos(2.0 * M_PI * u2);
        const double qz = b * std::sin(2.0 * M_PI * u3), qw = b * std::cos(2.0 * M_PI * u3);
        double *R = rot + 9 * c;
        R[0] = 1 - 2 * (qy * qy + qz * qz); R[1] = 2 * (qx * qy - qz * qw); R[2] = 2 * (qx * qz + qy * qw);
        R[3] = 2 * (qx * qy + qz * qw); R[4] = 1 - 2 * (qx * qx + qz * qz); R[5] = 2 * (qy * qz - qx * qw);
        R[6] = 2 * (qx * qz - qy * qw); R[7] = 2 * (qy * qz + qx * qw); R[8] = 1 - 2 * (qx * qx + qy * qy);
    }
    return LBMIBM_OK;
}

lbmibm_status lbmibm_build_topology(int nc, const int *voff, const int *eoff, const double *xr,
                                    const int *ele, int *nln, int *ln, int *lo, int *nle, int *le,
                                    double *nref, int *ec, double *c0, double *s0) {
    if (nc < 0 || !voff || !eoff || !xr || !ele || !nln || !ln || !lo || !nle || !le || !nref || !ec ||
        !c0 || !s0)
        return LBMIBM_ERR_ARG;
    const int nv = voff[nc], ne = eoff[nc];
    std::fill(nln, nln + nv, 0);
    std::fill(nle, nle + nv, 0);
    std::fill(ln, ln + 7 * (size_t)nv, 0);
    std::fill(lo, lo + 7 * (size_t)nv, 0);
    std::fill(le, le + 6 * (size_t)nv, 0);
    for (int c = 0; c < nc; c++) {
        const int v0 = voff[c], v1 = voff[c + 1], e0 = eoff[c], e1 = eoff[c + 1];
        if (v1 <= v0 || e1 <= e0) return LBMIBM_ERR_TOPOLOGY;
        // the per-cell system must have at least 256 rows after rounding 3V up
        // to a power of two (init2.cu:286-287): meshes of 86+ vertices
        int mn = 1;
        while (mn < 3 * (v1 - v0)) mn *= 2;
        if (mn < 256) return LBMIBM_ERR_TOPOLOGY;
        // Neighbour lists in first-seen order: for each element in order the
        // directed pairs (a,b),(a,c),(b,c),(b,a),(c,a),(c,b) (init2.cu:332-337);
        // incident elements in element order (init2.cu:339-341).
        for (int e = e0; e < e1; e++) {
            const int a = ele[3 * e], b = ele[3 * e + 1], cc = ele[3 * e + 2];
            if (a < v0 || a >= v1 || b < v0 || b >= v1 || cc < v0 || cc >= v1) return LBMIBM_ERR_TOPOLOGY;
            const int pairs[6][2] = {{a, b}, {a, cc}, {b, cc}, {b, a}, {cc, a}, {cc, b}};
            for (const auto &pr : pairs) {
                int *l = ln + 7 * (size_t)pr[0];
                bool seen = false;
                for (int k = 0; k < nln[pr[0]]; k++) seen |= (l[k] == pr[1]);
                if (!seen) {
                    if (nln[pr[0]] >= 7) return LBMIBM_ERR_TOPOLOGY;
                    l[nln[pr[0]]++] = pr[1];
                }
            }
            for (int vtx : {a, b, cc}) {
                if (nle[vtx] >= 6) return LBMIBM_ERR_TOPOLOGY;
                le[6 * (size_t)vtx + nle[vtx]++] = e;
            }
        }
        long long links = 0;
        for (int j = v0; j < v1; j++) links += nln[j];
        if (links != 2LL * ((v1 - v0) + (e1 - e0) - 2)) return LBMIBM_ERR_TOPOLOGY;  // init2.cu:353-358
        // lo: neighbours and the vertex itself in ascending order (init2.cu:169-203).
        for (int j = v0; j < v1; j++) {
            int *o = lo + 7 * (size_t)j;
            for (int k = 0; k < nln[j]; k++) o[k] = ln[7 * (size_t)j + k];
            o[nln[j]] = j;
            std::sort(o, o + nln[j] + 1);
        }
        // Ring order: start at the first-seen neighbour, then each next
        // neighbour is the one linked to the previous whose cross product
        // points away from the cell centroid (init2.cu:92-166).
        double g[3] = {0.0, 0.0, 0.0};
        for (int j = v0; j < v1; j++)
            for (int d = 0; d < 3; d++) g[d] += xr[3 * (size_t)j + d];
        for (int d = 0; d < 3; d++) g[d] /= (double)(v1 - v0);
        for (int j = v0; j < v1; j++) {
            const int cnt = nln[j];
            const int *nb = ln + 7 * (size_t)j;
            int ring[7];
            ring[0] = nb[0];
            const double *xj = xr + 3 * (size_t)j;
            for (int k = 1; k < cnt; k++) {
                const int a = ring[k - 1];
                const double va[3] = {xr[3 * (size_t)a] - xj[0], xr[3 * (size_t)a + 1] - xj[1], xr[3 * (size_t)a + 2] - xj[2]};
                int found = -1, nfound = 0;
                for (int l = 0; l < cnt; l++) {
                    const int b = nb[l];
                    if (b == a) continue;
                    bool linked = false;
                    for (int m = 0; m < nln[a]; m++) linked |= (ln[7 * (size_t)a + m] == b);
                    if (!linked) continue;
                    const double vb[3] = {xr[3 * (size_t)b] - xj[0], xr[3 * (size_t)b + 1] - xj[1], xr[3 * (size_t)b + 2] - xj[2]};
                    const double cr[3] = {va[1] * vb[2] - va[2] * vb[1], va[2] * vb[0] - va[0] * vb[2],
                                          va[0] * vb[1] - va[1] * vb[0]};
                    const double s = cr[0] * (xj[0] - g[0]) + cr[1] * (xj[1] - g[1]) + cr[2] * (xj[2] - g[2]);
                    if (s > 0) { found = b; nfound++; }
                }
                if (nfound != 1) return LBMIBM_ERR_TOPOLOGY;
                ring[k] = found;
            }
            // Overwriting this vertex's list in place is safe: later vertices
            // only test membership in it, which the reordering keeps.
            std::copy(ring, ring + cnt, ln + 7 * (size_t)j);
        }
        // Edge-adjacent elements, ascending (init1.cu:741-795).
        std::unordered_map<unsigned long long, std::array<int, 2>> edges;
        edges.reserve(3 * (size_t)(e1 - e0));
        auto key = [](int a, int b) {
            if (a > b) std::swap(a, b);
            return ((unsigned long long)(unsigned)a << 32) | (unsigned)b;
        };
        for (int e = e0; e < e1; e++)
            for (int k = 0; k < 3; k++) {
                auto &slot = edges.try_emplace(key(ele[3 * e + k], ele[3 * e + (k + 1) % 3]),
                                               std::array<int, 2>{-1, -1}).first->second;
                if (slot[0] < 0) slot[0] = e;
                else if (slot[1] < 0) slot[1] = e;
                else return LBMIBM_ERR_TOPOLOGY;
            }
        for (int e = e0; e < e1; e++) {
            int nbr[3];
            for (int k = 0; k < 3; k++) {
                const auto &slot = edges[key(ele[3 * e + k], ele[3 * e + (k + 1) % 3])];
                if (slot[1] < 0) return LBMIBM_ERR_TOPOLOGY;
                nbr[k] = (slot[0] == e) ? slot[1] : slot[0];
            }
            std::sort(nbr, nbr + 3);
            for (int k = 0; k < 3; k++) {
                ec[3 * e + k] = nbr[k];
                c0[3 * e + k] = 1.0;  // flat reference angle (init1.cu:847-850)
                s0[3 * e + k] = 0.0;
            }
        }
    }
    // Reference contravariant metric (init2.cu:206-255).
    for (int e = 0; e < ne; e++) {
        const double *a = xr + 3 * (size_t)ele[3 * e], *b = xr + 3 * (size_t)ele[3 * e + 1],
                     *c = xr + 3 * (size_t)ele[3 * e + 2];
        const double v1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double v2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const double l1 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
        const double l2 = v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2];
        const double i12 = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
        const double det = l1 * l2 - i12 * i12;
        if (!(det > 0.0)) return LBMIBM_ERR_TOPOLOGY;
        nref[4 * e + 0] = l2 / det;
        nref[4 * e + 1] = -i12 / det;
        nref[4 * e + 2] = -i12 / det;
        nref[4 * e + 3] = l1 / det;
    }
    return LBMIBM_OK;
}

}  // extern "C"
