"""Cell sets for the membrane path: mesh templates, seeded placement and the
CSR of cells (voff/eoff) consumed by ``lbmibm_create``.

Replaces the reference's mesh files ``./cell/icosa_div/icosa_div%d_cal.dat``
and ``./cell/rbc_div/rbc_div%d_cal.dat`` (init1.cu:490-519) and its placement
(init1.cu:565-675: rotate about x, y, z, scale by the radius, shift to the
centre) with a deterministic, seeded generator.
"""
from __future__ import annotations

import numpy as np

from . import native as N

_templates: dict = {}


def icosphere(level: int):
    """Unit icosphere (10*4^level+2 vertices), outward triangles."""
    L = N.lib()
    nv, nt = C_int(), C_int()
    N.check(L.lbmibm_mesh_icosphere(level, nv.ref, nt.ref, None, None))
    xyz = np.zeros((nv.value, 3))
    tri = np.zeros((nt.value, 3), np.int32)
    N.check(L.lbmibm_mesh_icosphere(level, nv.ref, nt.ref, N.dp(xyz), N.ip(tri)))
    return xyz, tri


def biconcave(xyz):
    """Evans-Fung biconcave map of unit-sphere points (axis z)."""
    out = np.ascontiguousarray(xyz, np.float64).copy()
    N.check(N.lib().lbmibm_mesh_biconcave(len(out), N.dp(out)))
    return out


def template(level: int, shape: str = "biconcave"):
    key = (level, shape)
    if key not in _templates:
        xyz, tri = icosphere(level)
        if shape == "biconcave":
            xyz = biconcave(xyz)
        elif shape != "sphere":
            raise ValueError(shape)
        _templates[key] = (xyz, tri)
    return _templates[key]


def place(nc, radius, lx, ly, lz, seed, mode="random", grid=(0, 0, 0), periodic_y=False):
    centres = np.zeros((nc, 3))
    rot = np.zeros((nc, 3, 3))
    m = {"random": 0, "grid": 1}[mode]
    N.check(N.lib().lbmibm_place_cells(nc, float(radius), float(lx), float(ly), float(lz), int(periodic_y),
                                       int(seed), m, int(grid[0]), int(grid[1]), int(grid[2]),
                                       N.dp(centres), N.dp(rot.reshape(-1))))
    return centres, rot


def assemble(levels, centres, rot, radius, shape="biconcave"):
    """CSR cell set from per-cell mesh levels, centres [nc,3], rotations [nc,3,3]."""
    xs, eles, voff, eoff = [], [], [0], [0]
    for c, lev in enumerate(levels):
        xyz, tri = template(int(lev), shape)
        pts = (xyz * radius) @ rot[c].T + centres[c]
        eles.append(tri + voff[-1])
        xs.append(pts)
        voff.append(voff[-1] + len(xyz))
        eoff.append(eoff[-1] + len(tri))
    X = np.concatenate(xs) if xs else np.zeros((0, 3))
    return dict(
        voff=np.array(voff, np.int32), eoff=np.array(eoff, np.int32),
        x=np.ascontiguousarray(X[:, 0]), y=np.ascontiguousarray(X[:, 1]), z=np.ascontiguousarray(X[:, 2]),
        ele=np.ascontiguousarray(np.concatenate(eles).reshape(-1) if eles else np.zeros(0, np.int32), np.int32),
        levels=np.array(levels, np.int32), centres=centres, radius=radius,
    )


def mesh_mix(nc, fractions, levels, seed):
    """Seeded assignment of mesh levels to cells with the given fractions."""
    counts = [int(round(f * nc)) for f in fractions]
    counts[-1] = nc - sum(counts[:-1])
    lv = np.concatenate([np.full(k, l, np.int32) for k, l in zip(counts, levels)])
    rng = np.random.default_rng(seed)
    rng.shuffle(lv)
    return lv


def cell_volume(cells):
    """Polyhedral volume per cell (divergence theorem), for reporting."""
    X = np.stack([cells["x"], cells["y"], cells["z"]], 1)
    tri = cells["ele"].reshape(-1, 3)
    a, b, c = X[tri[:, 0]], X[tri[:, 1]], X[tri[:, 2]]
    dv = np.einsum("ij,ij->i", a, np.cross(b, c)) / 6.0
    return np.add.reduceat(dv, cells["eoff"][:-1])


class C_int:
    def __init__(self):
        import ctypes
        self._v = ctypes.c_int(0)
        self.ref = ctypes.byref(self._v)

    @property
    def value(self):
        return self._v.value
