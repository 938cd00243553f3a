"""Writes tests/golden/channel_cell_f64.npz: inputs and FP64 oracle outputs of a
small coupled channel (16x14x12 lattice, slab walls, body force, one 162-vertex
Evans-Fung cell of radius 3 next to the periodic x boundary, sheared start,
6 steps).  Initial positions must lie inside the box, as the reference's
cell placement guarantees: spreading of positions beyond the halo layers is
not defined by the reference.

The reference ships no golden vectors; this fixture freezes the oracle's
outputs so that later changes to the oracle or to the HIP path are checked
against the same numbers.  Regenerate only when the oracle's algorithm is
deliberately changed:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

PARAMS = dict(tau=0.9, p=5e-6, e=(1.0, 0.0, 0.0), b=5e-3, m=1, c=100.0, residual=1e-12, itrmax=1000,
              runge_kutta=4, br=0.0)
SHAPE = (16, 14, 12)
STEPS = 6


def make_cells():
    from preprocess_src_amd import mesh
    lx, ly, lz = SHAPE
    centres = np.array([[lx / 2 - 3.6, 0.4, 0.3]])
    _, rot = mesh.place(1, 3.0, lx, ly, lz, seed=77)
    return mesh.assemble([2], centres, rot, 3.0)


def run_oracle(cells):
    from tests.pairs import oracle_of
    lx, ly, lz = SHAPE
    o = oracle_of(lx, ly, lz, cells=cells, **PARAMS)
    o.volume_init()
    X, Y, Z = o.node_coords()
    vel = 2e-3 * Y / (ly / 2)
    vel[o.bc == 2] = 0.0
    o.init_fluid(None, vel, None, None)
    vols = []
    for it in range(1, STEPS + 1):
        rc, k = o.step(it)
        assert rc == 0
        vols.append(k)
    comp = o.bc == 0
    return dict(fn_comp=o.fn.reshape(-1, 19)[comp].copy(), um=o.um[comp].copy(), vm=o.vm[comp].copy(),
                wm=o.wm[comp].copy(), x=o.cells["x"].copy(), y=o.cells["y"].copy(), z=o.cells["z"].copy(),
                fx=o.cells["cfx"].copy(), fy=o.cells["cfy"].copy(), fz=o.cells["cfz"].copy(),
                volume=o.cells["volume"].copy(), vol_iters=np.array(vols, np.int32))


def main():
    cells = make_cells()
    out = run_oracle(cells)
    path = os.path.join(HERE, "channel_cell_f64.npz")
    np.savez_compressed(path, voff=cells["voff"], eoff=cells["eoff"], cx=cells["x"], cy=cells["y"], cz=cells["z"],
                        ele=cells["ele"], **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
