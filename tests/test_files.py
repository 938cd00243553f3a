"""The reference's binary files: restart.bin, prof%04d.bin, cell%04d.bin.

CPU: the oracle's restatement of the layouts (oracle/reffiles.py) is pinned
by the ELL padding sizes the reference's set-up produces and by closed-form
sizes, and parses back to the state it was written from.  GPU: the
library's files against the oracle's for the same state (same sizes, ints
exact, values within the parity tolerances), and a run resumed from a
restart file against the uninterrupted run.
"""
import math
import os

import numpy as np
import pytest

from oracle import reffiles as RF
from tests.pairs import Q, gpu_of, oracle_of, rel, small_cells

TOL = {"f64": 1e-11, "f32": 1e-5}
# tp (principal tensions, stress.cu:101-118) takes sqrt((i1+2)^2 - 4(i2+1)) =
# |l1^2 - l2^2|, a cancelling difference near isotropic stretch: a 1-ulp
# difference in i1, i2 (FMA contraction on the GPU, none in the oracle)
# becomes ~1e-8 absolute, so tp is compared at 1e-5 of its maximum
TOL_TP = 1e-5


def test_expand2pow2_matches_the_ell_padding_and_the_closed_form():
    # mn of the 642-, 2562- and 10242-vertex meshes (init2.cu:286)
    assert [RF.expand2pow2(3 * v) for v in (642, 2562, 10242)] == [2048, 8192, 32768]
    for n in range(1, 5000):
        assert RF.expand2pow2(n) == (1 if n == 1 else 2 ** math.ceil(math.log2(n))), n


LX, LY, LZ = 24, 16, 20
KW = dict(tau=0.9, p=5e-6, b=5e-3, fileoutput=2)


def _cells(levels=(2, 2), seed=31):
    # inside the box: the oracle's two-node halo holds stencils of positions
    # within it only
    return small_cells(LX, LY, LZ, nc=len(levels), radius=3.0, levels=list(levels), seed=seed)


def _shear(o):
    # a sheared start, so that the cells deform and q, qi, tp are not zero
    X, Y, Z = o.node_coords()
    vel = 2e-3 * Y / (LY / 2)
    vel[o.bc == 2] = 0.0
    return vel


def _oracle_case(nsteps=3):
    o = oracle_of(LX, LY, LZ, cells=_cells(), **KW)
    o.volume_init()
    o.init_fluid(None, _shear(o), None, None)
    for it in range(1, nsteps + 1):
        rc, _ = o.step(it)
        assert rc == 0
    return o


def test_oracle_restart_bytes_parse_back_to_the_state():
    o = _oracle_case()
    c = o.cells
    nc, V, E = len(c["voff"]) - 1, int(c["voff"][1]), int(c["eoff"][1])
    buf = RF.restart_bytes(o, 3, 7)
    assert len(buf) == RF.restart_size(o.n, nc, V, E)
    r = RF.parse_restart(buf, o.n, nc, V, E)
    assert (r["iter"], r["fnum"]) == (3, 7)
    assert np.array_equal(r["fn"], o.fn)
    for k in ("dm", "um", "vm", "wm"):
        assert np.array_equal(r[k], getattr(o, k)), k
    assert not r["vf"].any() and not r["vf2"].any()
    for k in ("x", "y", "z", "u", "v", "w", "tp"):
        assert np.array_equal(r[k], c[k]), k
    assert np.array_equal(r["qi"][:3 * nc * V], c["qi"]) and not r["qi"][3 * nc * V:].any()
    assert np.allclose(r["rx"], c["x"].reshape(nc, V).mean(1))
    assert c["qi"].any() and c["tp"].any()  # the accumulators ran (FILEOUTPUT = 2)


def test_oracle_prof_and_cell_bytes_parse_back():
    o = _oracle_case(nsteps=2)
    p = RF.parse_prof(RF.prof_bytes(o, 2), o.n)
    assert p["iter"] == 2 and np.array_equal(p["bc"], o.bc) and np.array_equal(p["um"], o.um)
    cl = RF.parse_cell(RF.cell_bytes(o))
    nc = len(o.cells["voff"]) - 1
    assert (cl["n"], cl["vertex"], cl["element"]) == (nc, int(o.cells["voff"][1]), int(o.cells["eoff"][1]))
    assert np.array_equal(cl["ele"], o.cells["ele"]) and np.array_equal(cl["qi"], o.cells["qi"])


def _pair(prec, **extra):
    cells = _cells()
    kw = dict(KW, **extra)
    o = oracle_of(LX, LY, LZ, cells=cells, **kw)
    g = gpu_of(LX, LY, LZ, prec, cells=cells, **kw)
    o.volume_init()
    g.volume_init()
    vel = _shear(o)
    o.init_fluid(None, vel, None, None)
    g.set_equilibrium(None, vel, None, None)
    return o, g, cells, kw


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_restart_file_matches_the_oracle(prec, tmp_path):
    print(f"[restart file] prec={prec}", flush=True)
    o, g, _, _ = _pair(prec)
    for it in range(1, 4):
        assert o.step(it)[0] == 0
        g.step(it)
    path = str(tmp_path / "restart.bin")
    g.write_restart(path, 3, 7)
    buf = open(path, "rb").read()
    c = o.cells
    nc, V, E = len(c["voff"]) - 1, int(c["voff"][1]), int(c["eoff"][1])
    assert len(buf) == RF.restart_size(o.n, nc, V, E)
    a = RF.parse_restart(buf, o.n, nc, V, E)
    b = RF.parse_restart(RF.restart_bytes(o, 3, 7), o.n, nc, V, E)
    assert (a["iter"], a["fnum"]) == (3, 7)
    comp, buff = o.bc == 0, o.bc == 1
    fa, fb = a["fn"].reshape(-1, Q), b["fn"].reshape(-1, Q)
    assert rel(fa[comp], fb[comp]) < TOL[prec]
    assert not fa[~comp].any()  # only COMP populations are state (lbm.cu:1031)
    for k in ("dm", "um", "vm", "wm"):
        m = comp | buff
        assert rel(a[k][m], b[k][m]) < TOL[prec] * 10, k
    for k in ("x", "y", "z"):
        assert np.max(np.abs(a[k] - b[k])) < TOL[prec] * 20, k
    for k in ("rx", "ry", "rz"):
        assert np.max(np.abs(a[k] - b[k])) < TOL[prec] * 20, k
    for k in ("u", "v", "w", "qi"):
        assert rel(a[k], b[k]) < TOL[prec] * 10, k
    assert rel(a["tp"], b["tp"]) < max(TOL_TP, TOL[prec] * 10)
    assert not a["qi"][3 * nc * V:].any()


@pytest.mark.gpu
def test_prof_and_cell_files_match_the_oracle(tmp_path):
    o, g, _, _ = _pair("f64")
    for it in range(1, 3):
        assert o.step(it)[0] == 0
        g.step(it)
    g.write_prof(str(tmp_path / "prof0000.bin"), 2)
    g.write_cell(str(tmp_path / "cell0000.bin"))
    p = RF.parse_prof(open(tmp_path / "prof0000.bin", "rb").read(), o.n)
    q = RF.parse_prof(RF.prof_bytes(o, 2), o.n)
    assert p["iter"] == 2 and np.array_equal(p["bc"], q["bc"]) and np.array_equal(g.bc_mask(), o.bc)
    comp = o.bc == 0
    for k in ("dm", "um", "vm", "wm"):
        assert rel(p[k][comp], q[k][comp]) < TOL["f64"] * 10, k
    a = RF.parse_cell(open(tmp_path / "cell0000.bin", "rb").read())
    b = RF.parse_cell(RF.cell_bytes(o))
    assert (a["n"], a["vertex"], a["element"]) == (b["n"], b["vertex"], b["element"])
    assert np.array_equal(a["ele"], b["ele"])
    for k in ("x", "y", "z", "u", "v", "w", "qi"):
        assert rel(a[k], b[k]) < TOL["f64"] * 10, k
    assert rel(a["tp"], b["tp"]) < TOL_TP


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-6)])
def test_resumed_run_continues_the_uninterrupted_one(prec, tol, tmp_path):
    """restart_file semantics: a fresh context that reads restart.bin written
    at step 4 and runs steps 5-8 ends where the run that never stopped
    does (populations rounded through FP64 f = g + w, so not bitwise).  The
    volume correction is off: it moves cel->x only, and restart_file sets
    fem->x from cel->x (etc.cu:254-262), so with it a resumed run's first
    membrane force sees the corrected positions, in the reference too."""
    print(f"[resume] prec={prec}", flush=True)
    _, g, cells, kw = _pair(prec, volume_correction=False)
    path = str(tmp_path / "restart.bin")
    for it in range(1, 9):
        g.step(it)
        if it == 4:
            g.write_restart(path, it, 2)
    h = gpu_of(LX, LY, LZ, prec, cells=cells, **kw)
    h.volume_init()
    assert h.read_restart(path) == (4, 2)
    for it in range(5, 9):
        h.step(it)
    comp = g.comp_mask()
    assert rel(h.populations().reshape(-1, Q)[comp], g.populations().reshape(-1, Q)[comp]) < tol
    sa, sb = h.cells_state(), g.cells_state()
    for k in ("x", "y", "z"):
        assert np.max(np.abs(sa[k] - sb[k])) < tol * 20, k
    assert rel(h.membrane()["tp"], g.membrane()["tp"]) < tol * 10


@pytest.mark.gpu
def test_restart_of_another_size_and_mixed_meshes_are_refused(tmp_path):
    o, g, _, _ = _pair("f64")
    path = str(tmp_path / "restart.bin")
    g.write_restart(path, 0, 0)
    with open(path, "r+b") as f:
        f.truncate(os.path.getsize(path) - 8)
    with pytest.raises(RuntimeError):
        g.read_restart(path)
    mixed = gpu_of(LX, LY, LZ, "f64", cells=_cells(levels=(2, 3)))
    with pytest.raises(RuntimeError):
        mixed.write_restart(str(tmp_path / "mixed.bin"), 0, 0)
    mixed.write_prof(str(tmp_path / "prof.bin"), 0)  # no cell section: any meshes


@pytest.mark.gpu
def test_output_data_names_files_as_the_reference(tmp_path):
    _, g, _, _ = _pair("f64")
    os.makedirs(tmp_path / "cell")
    fnum = 0
    for it in (0, 2, 4):
        if it:
            g.step(it - 1)
            g.step(it)
        fnum = g.output_data(str(tmp_path), it, fnum)
    names = sorted(os.listdir(tmp_path)) + sorted(os.listdir(tmp_path / "cell"))
    assert names == ["cell", "prof0000.bin", "restart.bin", "restart0000.bin",
                     "cell0000.bin", "cell0001.bin", "cell0002.bin"], names
    assert fnum == 3


def test_adhesion_file_layout_round_trip():
    """The restated adhe%05d.bin layout (adhesion.cu:785-810): size formula and
    a parse of the written bytes give the arrays back."""
    from oracle import reffiles
    rng = np.random.default_rng(5)
    nv, ns, nrec = 37, 2, 53
    a = dict(f_adh=rng.standard_normal(3 * nv), attach=rng.integers(-1, nrec, nv * ns).astype(np.int32),
             is_knob=(rng.random(nv) < 0.3).astype(np.uint8), is_attached=rng.integers(0, 2, nrec).astype(np.int32),
             r_ij=rng.random(nv * ns))
    buf = reffiles.adhesion_bytes(3, 41, **a)
    assert len(buf) == 4 + 4 + 8 * 3 * nv + 4 * nv * ns + nv + 4 * nrec + 8 * nv * ns
    p = reffiles.parse_adhesion(buf, nv, ns, nrec)
    assert p["fnum_ad"] == 3 and p["iter"] == 41
    for k, v in a.items():
        assert np.array_equal(p[k], v), k
