"""The committed oracle fixture (tests/golden/channel_cell_f64.npz, written by
tests/golden/make_golden.py): the oracle must still reproduce it on the CPU,
and the HIP path must match it on the GPU (FP64 1e-11, FP32 1e-5)."""
import os

import numpy as np
import pytest

from tests.golden.make_golden import PARAMS, SHAPE, STEPS, run_oracle
from tests.pairs import gpu_of, rel

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "channel_cell_f64.npz"))


def fixture_cells():
    return dict(voff=G["voff"], eoff=G["eoff"], x=G["cx"], y=G["cy"], z=G["cz"], ele=G["ele"])


def test_oracle_reproduces_the_fixture():
    out = run_oracle(fixture_cells())
    for k in ("fn_comp", "um", "x", "y", "z", "fx", "volume"):
        assert rel(out[k], G[k]) < 1e-13, k
    assert np.array_equal(out["vol_iters"], G["vol_iters"])


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("f64", 1e-11), ("f32", 1e-5)])
def test_hip_path_matches_the_fixture(prec, tol):
    lx, ly, lz = SHAPE
    print(f"[golden] prec={prec}", flush=True)
    g = gpu_of(lx, ly, lz, prec, cells=fixture_cells(), **PARAMS)
    g.volume_init()
    from tests.pairs import oracle_of
    o = oracle_of(lx, ly, lz, cells=None, **PARAMS)  # geometry only: node coordinates and masks
    X, Y, Z = o.node_coords()
    vel = 2e-3 * Y / (ly / 2)
    vel[o.bc == 2] = 0.0
    g.set_equilibrium(None, vel, None, None)
    for it in range(1, STEPS + 1):
        g.step(it)
    comp = o.bc == 0
    assert rel(g.populations().reshape(-1, 19)[comp], G["fn_comp"]) < tol
    s = g.cells_state()
    for k in "xyz":
        assert np.max(np.abs(s[k] - G[k])) < tol * lx, k
    for a in ("fx", "fy", "fz"):
        assert rel(s[a], G[a]) < tol * 10, a
    mac = g.macros()
    assert rel(mac["um"][comp], G["um"]) < tol * 10
    assert rel(g.volumes()["volume"], G["volume"]) < tol


def test_oracle_reproduces_the_c3_fixture():
    """The full-size 128x64x64 channel with 64 RBCs (tests/golden/make_c3.py):
    the oracle still writes the committed numbers (about half a minute)."""
    from tests.golden import make_c3 as M
    C3 = np.load(os.path.join(HERE, "golden", "c3_channel_64rbc_f64.npz"))
    cells = M.make_cells()
    assert np.allclose(M.position_checksum(cells), C3["checksum"], rtol=1e-14, atol=0)
    out = M.run_oracle(cells)
    for k in ("f_sample", "um_z0", "x", "fx", "volume", "area", "centre"):
        assert rel(out[k], C3[k]) < 1e-12, k
    assert np.array_equal(out["vol_passes"], C3["vol_passes"])
    assert np.array_equal(out["bicg_iters"], C3["bicg_iters"])
