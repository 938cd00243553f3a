"""HIP path vs the FP64 oracle, stage by stage and for whole steps.

Every call goes through the C ABI (liblbmibm.so via ctypes).  Tolerances:
FP64 path 1e-11 relative (same algorithm, different summation order);
FP32 path 1e-5 relative; FP16 storage path reported
against a 2e-3 bound on the populations.
"""
import numpy as np
import pytest

from tests.pairs import (Q, gpu_of, oracle_of, perturb_cells, pop_err, random_lattice_force, random_populations,
                         rel, small_cells)

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-11, "f32": 1e-5, "f16": 2e-3}


def comp_pops(o, fn):
    m = o.bc == 0
    return fn.reshape(-1, Q)[m]


@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
@pytest.mark.parametrize("walls,hasf,m", [(True, True, 1), (False, True, 1), (True, False, 1), (True, True, 3)])
def test_lbm_parity(prec, walls, hasf, m):
    lx, ly, lz = 12, 10, 8
    kw = dict(tau=0.7, p=2e-5, e=(0.6, 0.0, 0.8), b=0.3, m=m, periodic_y=not walls, cr=0.8, dx=0.5)
    print(f"[lbm parity] prec={prec} walls={walls} hasf={hasf} m={m}", flush=True)
    o = oracle_of(lx, ly, lz, **kw)
    g = gpu_of(lx, ly, lz, prec, **kw)
    fn = random_populations(o, seed=3)
    g.set_populations(fn)
    assert rel(comp_pops(o, g.populations()), comp_pops(o, fn)) < (1e-15 if prec == "f64" else TOL[prec])
    if hasf:
        F = random_lattice_force(o, seed=4)
        o.fx[:], o.fy[:], o.fz[:] = F
        g.set_lattice_force(*F)
    for it in range(1, 13):
        o.lbm_d3q19(it)
        g.lbm_d3q19(it)
    comp = o.bc == 0
    assert rel(comp_pops(o, g.populations()), comp_pops(o, o.fn)) < TOL[prec]
    mac = g.macros(instantaneous=True)
    buff = o.bc == 1
    for k in ("um", "vm", "wm", "un", "vn", "wn"):
        ref = getattr(o, k)
        tol = TOL[prec] if prec != "f16" else 5e-2
        assert rel(mac[k][comp], ref[comp]) < tol, k
        assert rel(mac[k][buff], ref[buff]) < tol, k
    assert rel(mac["dn"][comp], o.dn[comp]) < TOL[prec]
    assert rel(mac["dm"][comp], o.dm[comp]) < TOL[prec]


def _cells_across_boundaries(lx, ly, lz, radius=3.0, periodic_y=False):
    from preprocess_src_amd import mesh
    y1 = -ly / 2 + 0.4 if periodic_y else -ly / 2 + radius + 1.2
    centres = np.array([[lx / 2 - 1.0, 0.3, -lz / 2 + 1.3], [0.7, y1, 0.4]])
    _, rot = mesh.place(2, radius, lx, ly, lz, seed=5)
    return mesh.assemble([2, 2], centres, rot, radius)


def _wrapped(cells, lx, ly, lz, periodic_y=False):
    """Vertex positions wrapped into the box as ibm_st2a leaves them (ibm.cu:629-632)."""
    out = {k: cells[k].copy() for k in ("x", "y", "z")}
    for k, L in (("x", lx), ("z", lz)) + ((("y", ly),) if periodic_y else ()):
        a = out[k]
        a[a > L / 2] -= L
        a[a < -L / 2] += L
    out["xd"] = np.stack([out["x"], out["y"], out["z"]], 1).reshape(-1).copy()
    return out


def _set_positions(o, g, pos):
    for k in ("x", "y", "z", "xd"):
        o.cells[k][:] = pos[k]
    g.set_cells_state(x=pos["x"], y=pos["y"], z=pos["z"], xd=pos["xd"])


# lattices with full tiles, with 1-node partial tiles and fully periodic
SHAPES = [(16, 14, 12, False), (33, 17, 9, False), (18, 9, 10, True)]


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("lx,ly,lz,py", SHAPES)
def test_ibm_spread_parity(prec, lx, ly, lz, py, form, monkeypatch):
    """Both spreading forms (source bins per tile; home boxes + gather, the
    autotuner picks one) against the oracle, on lattices with partial tiles,
    boxes that wrap onto themselves, walls and periodic y."""
    monkeypatch.setenv("LBMIBM_SPREAD_FORM", str(form))
    print(f"[spread parity] prec={prec} shape={lx}x{ly}x{lz} periodic_y={py} form={form}", flush=True)
    cells = _cells_across_boundaries(lx, ly, lz, periodic_y=py)
    o = oracle_of(lx, ly, lz, cells=cells, periodic_y=py)
    g = gpu_of(lx, ly, lz, prec, cells=cells, periodic_y=py)
    _set_positions(o, g, _wrapped(cells, lx, ly, lz, py))
    rng = np.random.default_rng(7)
    fx, fy, fz = (1e-3 * rng.standard_normal(o.cells["x"].size) for _ in range(3))
    o.cells["cfx"][:], o.cells["cfy"][:], o.cells["cfz"][:] = fx, fy, fz
    g.set_cells_state(fx=fx, fy=fy, fz=fz)
    o.ibm_st1()
    g.ibm_st1()
    F = g.lattice_force()
    comp = o.bc == 0
    for a, b in zip(F, (o.fx, o.fy, o.fz)):
        assert rel(a[comp], b[comp]) < TOL[prec]
    # total force is conserved by the spreading (sum_nodes delta dx^3 = 1)
    assert abs(F[0][comp].sum() - fx.sum()) < 1e-9 + 1e-5 * np.abs(fx).sum()


def _random_velocity_field(o, seed, amp=1e-2):
    """Random (um,vm,wm) on COMP nodes, BUFF nodes = periodic images, WALL = 0."""
    from oracle.pyoracle import lib
    rng = np.random.default_rng(seed)
    n = o.n
    out = []
    comp = o.bc == 0
    for _ in range(3):
        a = np.zeros(n)
        a[comp] = amp * rng.standard_normal(comp.sum())
        out.append(a)
    o.um[:], o.vm[:], o.wm[:] = out
    # fill BUFF nodes from their images with one boundary copy of the oracle:
    # copy through a zero-step by reusing the halo logic of the LBM boundary
    nx, ny, nz = o.nx, o.ny, o.nz
    idx = np.arange(n)
    iz, r = np.divmod(idx, nx * ny)
    iy, ix = np.divmod(r, nx)
    w = lambda a, m: np.where(a < 2, a + m - 4, np.where(a > m - 3, a - m + 4, a))
    img = w(ix, nx) + nx * (w(iy, ny) if o.P.periodic_y else iy) + nx * ny * w(iz, nz)
    buff = o.bc == 1
    for a in (o.um, o.vm, o.wm):
        a[buff] = a[img[buff]]
    return o.um.copy(), o.vm.copy(), o.wm.copy()


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("rk,m,dx,shape", [(4, 1, 1.0, SHAPES[0]), (4, 2, 0.5, SHAPES[0]), (2, 1, 1.0, SHAPES[0]),
                                           (1, 3, 1.0, SHAPES[0]), (4, 1, 1.0, SHAPES[1]), (4, 1, 1.0, SHAPES[2])])
def test_ibm_interp_parity(prec, rk, m, dx, shape, form, monkeypatch):
    """Both interpolation forms (LDS-staged tiles, rows read from memory; the
    autotuner picks one) against the oracle, with cells across the periodic
    and wall boundaries."""
    monkeypatch.setenv("LBMIBM_INTERP_FORM", str(form))
    lx, ly, lz, py = shape
    print(f"[interp parity] prec={prec} rk={rk} m={m} dx={dx} shape={shape} form={form}", flush=True)
    cells = _cells_across_boundaries(lx * dx, ly * dx, lz * dx, radius=3.0 * dx, periodic_y=py)
    kw = dict(runge_kutta=rk, m=m, dx=dx, cr=0.7, periodic_y=py)
    o = oracle_of(lx, ly, lz, cells=cells, **kw)
    g = gpu_of(lx, ly, lz, prec, cells=cells, **kw)
    _set_positions(o, g, _wrapped(cells, lx * dx, ly * dx, lz * dx, py))
    um, vm, wm = _random_velocity_field(o, seed=9)
    g.set_macros(um, vm, wm)
    o.ibm_st2()
    g.ibm_st2()
    s = g.cells_state()
    for k in ("u", "v", "w"):
        assert rel(s[k], o.cells[k]) < TOL[prec], k
    for k in ("x", "y", "z"):
        assert np.max(np.abs(s[k] - o.cells[k])) < TOL[prec] * max(lx, ly, lz) * dx, k
    assert np.max(np.abs(s["xd"] - o.cells["xd"])) < TOL[prec] * max(lx, ly, lz) * dx


# mesh levels 2-5 (162 / 642 / 2562 / 10242 vertices) cover the three solver
# classes: on chip 4 x 256, on chip 6 x 512 (FP32), vectors in global memory
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("levels", [[2, 2], [2, 3], [4], [5]])
def test_stress_parity(prec, levels):
    lx, ly, lz = 24, 20, 24
    print(f"[stress parity] prec={prec} levels={levels}", flush=True)
    cells = small_cells(lx, ly, lz, nc=len(levels), radius=4.0, levels=levels, seed=21)
    kw = dict(c=100.0, br=0.0, fileoutput=2)
    o = oracle_of(lx, ly, lz, cells=cells, **kw)
    g = gpu_of(lx, ly, lz, prec, cells=cells, **kw)
    for it in (1, 2, 3):
        d = perturb_cells(cells, seed=22 + it, amp=0.05)
        xd = np.stack([d["x"], d["y"], d["z"]], 1).reshape(-1).copy()
        o.cells["xd"][:] = xd
        g.set_cells_state(xd=xd)
        assert o.stress(it) == 0
        g.stress(it)
        mb = g.membrane()
        assert rel(mb["rg"], o.cells["rg"]) < TOL[prec]
        assert rel(mb["t"], o.cells["t"]) < TOL[prec] * 10
        assert rel(mb["q"], o.cells["q"]) < TOL[prec] * 10
        # FILEOUTPUT accumulators (stress.cu:101-118, 299-300), reset at (iter-1) % 2 == 0
        assert rel(mb["qi"], o.cells["qi"]) < TOL[prec] * 10
        assert rel(mb["tp"], o.cells["tp"]) < TOL[prec] * 10
        s = g.cells_state()
        for a, b in (("fx", "cfx"), ("fy", "cfy"), ("fz", "cfz")):
            assert rel(s[a], o.cells[b]) < TOL[prec] * 10, a
        assert np.all(mb["bicg_iters"] == o.cells["bicg_iters"]) or prec != "f64"


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_bending_parity(prec):
    lx, ly, lz = 24, 20, 24
    cells = small_cells(lx, ly, lz, nc=2, radius=4.0, levels=[2, 2], seed=23)
    kw = dict(br=0.05)
    o = oracle_of(lx, ly, lz, cells=cells, **kw)
    g = gpu_of(lx, ly, lz, prec, cells=cells, **kw)
    d = perturb_cells(cells, seed=24, amp=0.08)
    xd = np.stack([d["x"], d["y"], d["z"]], 1).reshape(-1).copy()
    for k in ("x", "y", "z"):
        o.cells[k][:] = d[k]
    o.cells["xd"][:] = xd
    g.set_cells_state(x=d["x"], y=d["y"], z=d["z"], xd=xd)
    assert o.stress(1) == 0
    g.stress(1)
    s = g.cells_state()
    for a, b in (("fx", "cfx"), ("fy", "cfy"), ("fz", "cfz")):
        assert rel(s[a], o.cells[b]) < TOL[prec] * 10, a


@pytest.mark.parametrize("inflate", ["both", "one"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_volume_parity(prec, inflate):
    """Volume correction against the oracle for cells across the periodic
    faces: both cells inflated (each runs correction passes), or one
    inflated and the other only translated across the faces (its first
    sums are within tolerance: no pass, but its positions are unwrapped,
    summed for the centroid and wrapped back)."""
    lx, ly, lz = 24, 20, 24
    cells = _cells_across_boundaries(lx, ly, lz, radius=3.0)
    o = oracle_of(lx, ly, lz, cells=cells)
    g = gpu_of(lx, ly, lz, prec, cells=cells)
    o.volume_init()
    g.volume_init()
    v = g.volumes()
    assert rel(v["volume0"], o.cells["volume0"]) < TOL[prec]
    assert np.max(np.abs(v["gx"] - o.cells["gx"])) < TOL[prec] * lx
    # inflate every cell by ~1% of volume about its centroid, then wrap the
    # vertices as the advection does
    x, y, z = (o.cells[k].copy() for k in ("x", "y", "z"))
    for c in range(2):
        sl = slice(cells["voff"][c], cells["voff"][c + 1])
        for a, gc in ((x, 0), (y, 1), (z, 2)):
            ref = o.cells["gx"][3 * c + gc]
            L = (lx, ly, lz)[gc]
            d = a[sl] - ref
            if gc != 1:
                d = np.where(d > L / 2, d - L, np.where(d < -L / 2, d + L, d))
            if c == 0 and inflate == "one":
                a[sl] = ref + d + (0.45, 0.0, -0.37)[gc]
            else:
                a[sl] = ref + 1.0033 * d
    pos = _wrapped(dict(x=x, y=y, z=z), lx, ly, lz)
    for k in "xyz":
        o.cells[k][:] = pos[k]
    g.set_cells_state(x=pos["x"], y=pos["y"], z=pos["z"])
    k_o = o.volume_correction()
    k_g = g.volume_correction()
    assert k_o > 0 and abs(k_o - k_g) <= (0 if prec == "f64" else 1)
    s = g.cells_state()
    for k in "xyz":
        assert np.max(np.abs(s[k] - o.cells[k])) < TOL[prec] * lx, k
    v = g.volumes()
    assert rel(v["volume"], o.cells["volume"]) < TOL[prec]
    assert np.all(np.abs(v["volume"] - v["volume0"]) / v["volume0"] < 1e-5)
    assert np.max(np.abs(v["gx"] - o.cells["gx"])) < TOL[prec] * lx


def _step_pair(prec, m=1, nsteps=10, **extra):
    lx, ly, lz = 24, 16, 20
    cells = small_cells(lx, ly, lz, nc=2, radius=3.0, levels=[2, 2], seed=31)
    kw = dict(tau=0.9, p=5e-6, e=(1.0, 0.0, 0.0), b=5e-3, m=m, c=100.0, residual=1e-12)
    kw.update(extra)
    o = oracle_of(lx, ly, lz, cells=cells, **kw)
    g = gpu_of(lx, ly, lz, prec, cells=cells, **kw)
    o.volume_init()
    g.volume_init()
    # start from a sheared state so the cells deform in the first steps
    X, Y, Z = o.node_coords()
    vel = 2e-3 * Y / (ly / 2)
    vel[o.bc == 2] = 0.0  # wall nodes hold zero velocity (channel start from rest)
    o.init_fluid(None, vel, None, None)
    g.set_equilibrium(None, vel, None, None)
    return o, g, nsteps


@pytest.mark.parametrize("prec,m", [("f64", 1), ("f64", 2), ("f32", 1)])
def test_step_parity(prec, m):
    o, g, nsteps = _step_pair(prec, m=m)
    print(f"[step parity] prec={prec} m={m}", flush=True)
    for it in range(1, nsteps + 1):
        rc, _ = o.step(it)
        assert rc == 0
        g.step(it)
    comp = o.bc == 0
    assert rel(comp_pops(o, g.populations()), comp_pops(o, o.fn)) < TOL[prec]
    s = g.cells_state()
    for k in "xyz":
        assert np.max(np.abs(s[k] - o.cells[k])) < TOL[prec] * 24, k
    for a, b in (("fx", "cfx"), ("fy", "cfy"), ("fz", "cfz")):
        assert rel(s[a], o.cells[b]) < TOL[prec] * 10, a
    mac = g.macros()
    assert rel(mac["um"][comp], o.um[comp]) < TOL[prec] * 10


@pytest.mark.parametrize("warm", [1, 2])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_warm_started_solves_match_the_oracle(prec, warm):
    """BiCGSTAB started from the previous step's loads (bicg_warm_start 1) or
    from their linear extrapolation (2) stops at the same relative residual
    as the reference's start from 0, so whole steps keep the tolerances of
    the cold start.  The cells start unstressed in a shear flow, so their
    loads grow from zero: the last load is a poor start there (its relative
    error is about 1/n at step n) and the extrapolation a good one."""
    o, g, nsteps = _step_pair(prec, m=1, bicg_warm_start=warm)
    _, h, _ = _step_pair(prec, m=1)
    print(f"[warm start] prec={prec} warm={warm}", flush=True)
    its_w, its_c = [], []
    for it in range(1, nsteps + 1):
        assert o.step(it)[0] == 0
        g.step(it)
        h.step(it)
        its_w.append(int(g.membrane()["bicg_iters"].sum()))
        its_c.append(int(h.membrane()["bicg_iters"].sum()))
    print(f"iterations per step, warm {warm}: {its_w}; cold: {its_c}", flush=True)
    comp = o.bc == 0
    assert pop_err(comp_pops(o, g.populations()), comp_pops(o, o.fn)) < TOL[prec]
    s = g.cells_state()
    for k in "xyz":
        assert np.max(np.abs(s[k] - o.cells[k])) < TOL[prec] * 24, k
    for a, b in (("fx", "cfx"), ("fy", "cfy"), ("fz", "cfz")):
        assert rel(s[a], o.cells[b]) < TOL[prec] * 10, a
    assert its_w[0] == its_c[0]  # q = 0 before the first solve
    assert sum(its_w) <= 1.1 * sum(its_c)
    if warm == 2:  # from the third step on the extrapolation saves iterations
        assert sum(its_w[2:]) < 0.85 * sum(its_c[2:]), (its_w, its_c)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_run_graph_matches_steps(prec):
    o, g1, nsteps = _step_pair(prec, m=1)
    _, g2, _ = _step_pair(prec, m=1)
    # 20 steps: two replays of the 8-step graph and 4 eager steps, with each
    # volume correction overlapping the next membrane force, against single steps
    for it in range(1, 21):
        g1.step(it)
    g2.run(0, 20, graph=True)
    p1, p2 = g1.populations(), g2.populations()
    comp = o.bc == 0
    assert rel(comp_pops(o, p2), comp_pops(o, p1)) < (1e-13 if prec == "f64" else 1e-6)
    s1, s2 = g1.cells_state(), g2.cells_state()
    for k in "xyz":
        assert np.max(np.abs(s1[k] - s2[k])) < (1e-12 if prec == "f64" else 1e-5)


def test_bicgstab_not_converged_is_reported():
    from preprocess_src_amd.native import LbmibmError
    lx, ly, lz = 24, 20, 24
    cells = small_cells(lx, ly, lz, nc=1, radius=4.0, levels=[2], seed=41)
    g = gpu_of(lx, ly, lz, "f64", cells=cells, itrmax=1, residual=1e-30)
    d = perturb_cells(cells, seed=42, amp=0.05)
    g.set_cells_state(xd=np.stack([d["x"], d["y"], d["z"]], 1).reshape(-1).copy())
    with pytest.raises(LbmibmError) as e:
        g.stress(1)
    assert e.value.code == 4


@pytest.mark.parametrize("form", [0, 1])
def test_spread_is_bitwise_reproducible(form, monkeypatch):
    """Fixed-point LDS accumulation (and, in form 1, a fixed summation order
    over the boxes): repeated spreads give identical bits."""
    monkeypatch.setenv("LBMIBM_SPREAD_FORM", str(form))
    lx, ly, lz = 24, 20, 24
    cells = small_cells(lx, ly, lz, nc=3, radius=4.0, levels=[3, 2, 3], seed=51)
    g = gpu_of(lx, ly, lz, "f32", cells=cells)
    rng = np.random.default_rng(52)
    n = cells["x"].size
    g.set_cells_state(fx=rng.standard_normal(n), fy=rng.standard_normal(n), fz=rng.standard_normal(n))
    g.ibm_st1()
    F1 = [a.copy() for a in g.lattice_force()]
    g.ibm_st1()
    F2 = g.lattice_force()
    for a, b in zip(F1, F2):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("shapes", ["0,0,0", "1,1,1", "2,2,2"])
def test_every_solver_shape_matches_the_oracle(shapes, monkeypatch):
    """Each candidate work-group shape of the on-chip solver classes and each
    form of the large-cell solver (the autotuner picks one of each) gives the
    oracle's loads and forces."""
    monkeypatch.setenv("LBMIBM_BICG_SHAPES", shapes)
    lx, ly, lz = 28, 24, 28
    cells = small_cells(lx, ly, lz, nc=4, radius=4.0, levels=[3, 4, 2, 5], seed=61)
    kw = dict(c=100.0, br=0.0)
    o = oracle_of(lx, ly, lz, cells=cells, **kw)
    g = gpu_of(lx, ly, lz, "f32", cells=cells, **kw)
    assert "LBMIBM_BICG_SHAPES" in g.autotune_report()
    d = perturb_cells(cells, seed=62, amp=0.05)
    xd = np.stack([d["x"], d["y"], d["z"]], 1).reshape(-1).copy()
    o.cells["xd"][:] = xd
    g.set_cells_state(xd=xd)
    assert o.stress(1) == 0
    g.stress(1)
    assert rel(g.membrane()["q"], o.cells["q"]) < 1e-4
    s = g.cells_state()
    for a, b in (("fx", "cfx"), ("fy", "cfy"), ("fz", "cfz")):
        assert rel(s[a], o.cells[b]) < 1e-4, a


def test_autotune_reports_a_choice_per_class():
    lx, ly, lz = 24, 20, 24
    cells = small_cells(lx, ly, lz, nc=2, radius=4.0, levels=[3, 4], seed=63)
    g = gpu_of(lx, ly, lz, "f32", cells=cells)
    rep = g.autotune_report()
    assert "class 0: using" in rep and "class 1: using" in rep, rep
    assert "interp: using" in rep and "spread: using" in rep, rep


def test_autotune_covers_the_large_cell_forms():
    lx, ly, lz = 24, 20, 24
    cells = small_cells(lx, ly, lz, nc=1, radius=4.0, levels=[5], seed=64)
    g = gpu_of(lx, ly, lz, "f32", cells=cells)
    assert "class 2: using" in g.autotune_report()


@pytest.mark.parametrize("prec,tol", [("f64", 2e-3), ("f32", 2e-3)])
def test_poiseuille_channel_64x32x32_against_the_transient_series(prec, tol):
    """The plain channel configuration: 64x32x32, half-way walls at y = 0.5 and
    30.5 of the padded frame (H = 32 fluid nodes here, walls at +-LY/2), tau =
    0.8, Guo body force 1e-6 along x, 1000 steps from rest; u_x(y) against the
    start-up Fourier series (the same closed form that pins the oracle), and
    against the FP64 oracle at 1e-10 (FP64) / 1e-5 (FP32) relative."""
    from tests.test_oracle import _poiseuille_series
    lx, ly, lz, tau, a = 64, 32, 32, 0.8, 1e-6
    g = gpu_of(lx, ly, lz, prec, tau=tau, p=a / 2, volume_correction=False)
    g.set_equilibrium()
    g.run(0, 1000, graph=True)
    u = g.macros()["um"].reshape(lz + 4, ly + 4, lx + 4)[2:-2, 2:-2, 2:-2]
    assert np.ptp(u, axis=(0, 2)).max() <= 1e-12 * np.abs(u).max()
    nu = (tau - 0.5) / 3
    ua = _poiseuille_series(np.arange(ly) + 0.5, 1000, a, nu, ly)
    assert np.max(np.abs(u[0, :, 0] - ua)) / ua.max() < tol
    o = oracle_of(4, ly, 4, tau=tau, p=a / 2, volume_correction=False)
    for it in range(1, 1001):
        o.lbm_d3q19(it)
    uo = o.un.reshape(o.nz, o.ny, o.nx)[2, 2:-2, 2]
    assert np.max(np.abs(u[0, :, 0] - uo)) / np.abs(uo).max() < (1e-10 if prec == "f64" else 1e-5)
