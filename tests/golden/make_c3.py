"""Writes tests/golden/c3_channel_64rbc_f64.npz: FP64 oracle outputs of the
128x64x64 capsule channel with 64 red blood cells, at full size.

Configuration: 128x64x64 fluid nodes, half-way bounce-back walls at y = +-32,
periodic x and z, tau = 1, Guo body force 2 P = 1e-6 along x, 64 Evans-Fung
biconcave cells of radius 4 lattice units on 2562-vertex / 5120-triangle
icosphere meshes, placed by the library's seeded non-overlapping placement
(seed 2026, clear of the periodic faces) with uniform random orientations,
Skalak Gs = 5e-4 (B = Gs / 2), C = 100, BR = 0, RK4, M = 1, volume
correction on, RESIDUAL 1e-12.  The fluid
starts in the steady Poiseuille profile u(y) = 2P (H^2/4 - y^2) / (2 nu) so
that the cells deform from the first step.  STEPS whole steps of the oracle
are recorded.

The fixture keeps what pins the run at this size in about 1 MB:
  - per-cell volume, area and centroid (all 64 cells);
  - positions and nodal forces of every 32nd vertex;
  - the M-averaged velocity on the plane z = 0 (all 128 x 64 nodes);
  - the post-stream populations of 2048 seeded fluid nodes;
  - per-cell BiCGSTAB iteration counts and the volume-correction passes;
  - a checksum of the initial positions, so that a changed mesh generator or
    placement is caught before the comparison.

The reference ships no golden vectors; the fixture freezes the oracle's
outputs.  Regenerate only when the oracle's algorithm is deliberately
changed (about a minute on 8 cores):  python tests/golden/make_c3.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

SHAPE = (128, 64, 64)
NC, RADIUS, LEVEL, SEED = 64, 4.0, 4, 2026
GS = 5e-4
PARAMS = dict(tau=1.0, p=0.5e-6, e=(1.0, 0.0, 0.0), b=GS / 2.0, m=1, c=100.0, residual=1e-12, itrmax=1000,
              runge_kutta=4, br=0.0, volume_correction=True)
STEPS = 40
VSTRIDE = 32
NSAMPLE = 2048
PATH = os.path.join(HERE, "c3_channel_64rbc_f64.npz")


def make_cells():
    """The seeded placement in the box shrunk by RADIUS + 1 on the periodic
    axes (x, z), so that no cell straddles a periodic face at the start.  A
    straddling cell's initial vertices lie outside the box, where the
    reference's padded frame has no nodes for part of their stencils (it
    drops them until the first volume correction wraps the positions back),
    while the library wraps every stencil node periodically."""
    from preprocess_src_amd import mesh
    lx, ly, lz = SHAPE
    m = 2.0 * (RADIUS + 1.0)
    centres, rot = mesh.place(NC, RADIUS, lx - m, ly, lz - m, SEED)
    return mesh.assemble(np.full(NC, LEVEL), centres, rot, RADIUS)


def position_checksum(cells):
    return np.array([cells["x"].sum(), cells["y"].sum(), cells["z"].sum(),
                     float(np.dot(cells["x"], np.arange(len(cells["x"])) % 97))])


def poiseuille(ycoord):
    """Steady channel profile of the Guo-forced BGK lattice, walls at +-LY/2."""
    ly = SHAPE[1]
    nu = (PARAMS["tau"] - 0.5) / 3.0
    u = 2.0 * PARAMS["p"] * (ly * ly / 4.0 - ycoord * ycoord) / (2.0 * nu)
    return np.where(np.abs(ycoord) < ly / 2.0, u, 0.0)


def sample_nodes(comp_idx):
    rng = np.random.default_rng(SEED)
    return np.sort(rng.choice(comp_idx, NSAMPLE, replace=False))


def plane_z0(a):
    """The 128 x 64 nodes of the fluid plane nearest z = 0 from a padded [n] field."""
    lx, ly, lz = SHAPE
    return a.reshape(lz + 4, ly + 4, lx + 4)[2 + lz // 2, 2:-2, 2:-2].copy()


def record(fn, macros, cells_state, volumes, iters):
    """The fixture's fields from a run's outputs in reference layouts."""
    comp = np.flatnonzero(comp_mask())
    idx = sample_nodes(comp)
    out = dict(f_sample=fn.reshape(-1, 19)[idx].copy(), node_sample=idx,
               um_z0=plane_z0(macros["um"]), vm_z0=plane_z0(macros["vm"]), wm_z0=plane_z0(macros["wm"]))
    sel = slice(0, None, VSTRIDE)
    for k in ("x", "y", "z", "fx", "fy", "fz"):
        out[k] = cells_state[k][sel].copy()
    out.update(volume=volumes["volume"].copy(), area=volumes["area"].copy(), centre=volumes["gx"].copy(),
               bicg_iters=np.asarray(iters, np.int32).copy())
    return out


def comp_mask():
    lx, ly, lz = SHAPE
    m = np.zeros((lz + 4, ly + 4, lx + 4), bool)
    m[2:-2, 2:-2, 2:-2] = True
    return m.reshape(-1)


def run_oracle(cells, steps=STEPS):
    from tests.pairs import oracle_of
    lx, ly, lz = SHAPE
    o = oracle_of(lx, ly, lz, cells=cells, **PARAMS)
    o.volume_init()
    X, Y, Z = o.node_coords()
    u = poiseuille(Y)
    u[o.bc == 2] = 0.0
    o.init_fluid(None, u, None, None)
    passes = []
    for it in range(1, steps + 1):
        rc, k = o.step(it)
        assert rc == 0
        passes.append(k)
    st = dict(x=o.cells["x"], y=o.cells["y"], z=o.cells["z"], fx=o.cells["cfx"], fy=o.cells["cfy"],
              fz=o.cells["cfz"])
    vols = dict(volume=o.cells["volume"], area=o.cells["area"], gx=o.cells["gx"])
    out = record(o.fn, dict(um=o.um, vm=o.vm, wm=o.wm), st, vols, o.cells["bicg_iters"])
    out["vol_passes"] = np.array(passes, np.int32)
    return out


def main():
    cells = make_cells()
    t0 = time.time()
    out = run_oracle(cells)
    np.savez_compressed(PATH, checksum=position_checksum(cells), **out)
    print(f"wrote {PATH} ({os.path.getsize(PATH)} bytes, {time.time() - t0:.1f} s)")


if __name__ == "__main__":
    main()
