"""Pins of the FP64 host oracle (CPU only).

The reference ships no golden vectors or known-answer tests, so every stage of
the oracle is pinned here by what the mathematics fixes: closed-form flows,
conservation laws, lattice symmetry, the delta function's moment identities,
energy gradients by central finite differences, a dense direct solve and the
sphere volume.  Each check is chosen so that a plausible slip in the
restatement (a dropped forcing term, a wrong sign, a transposed index) fails.
"""
import numpy as np
import pytest

from oracle.pyoracle import OracleSim, bicgstab, delta_weight, lib
from tests.pairs import perturb_cells, small_cells

Q = 19
# D3Q19 directions in the reference order (lbm.cu:1-20)
CQ = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0],
               [-1, -1, 0], [-1, 1, 0], [1, -1, 0], [0, 1, 1], [0, -1, -1], [0, -1, 1], [0, 1, -1], [1, 0, 1],
               [-1, 0, -1], [1, 0, -1], [-1, 0, 1]])


def comp_view(o, a):
    return a.reshape(o.nz, o.ny, o.nx, *a.shape[1:])[2:-2, 2:-2, 2:-2]


# ---- delta function ----------------------------------------------------------

def test_delta_partition_of_unity_and_parity_sums():
    rng = np.random.default_rng(0)
    for X in rng.uniform(-3, 3, 50):
        j = np.arange(np.floor(X) - 1, np.floor(X) + 3)
        w = np.array([delta_weight(X - jj) for jj in j])
        assert abs(w.sum() - 1.0) < 1e-15
        even = w[(j % 2) == 0].sum()
        assert abs(even - 0.5) < 1e-15
    assert delta_weight(0.0) == 0.5 and delta_weight(2.0) == 0.0 and abs(delta_weight(1.0) - 0.25) < 1e-16


def test_delta_first_moment_closed_form():
    """sum_j (j - X) phi(X - j) = -(t - 1/2 + (cos(pi t/2) - sin(pi t/2))/2),
    t = frac(X): not zero (the cosine kernel is not first-moment exact),
    -1.5627e-2 at X = 0.1."""
    for X in (0.1, 0.81, -2.37, 5.5):
        j = np.arange(np.floor(X) - 1, np.floor(X) + 3)
        m1 = sum((jj - X) * delta_weight(X - jj) for jj in j)
        t = X - np.floor(X)
        closed = -(t - 0.5 + (np.cos(np.pi * t / 2) - np.sin(np.pi * t / 2)) / 2)
        assert abs(m1 - closed) < 1e-15
    assert abs(-(0.1 - 0.5 + (np.cos(0.05 * np.pi) - np.sin(0.05 * np.pi)) / 2) + 1.5627e-2) < 1e-6


# ---- lattice -------------------------------------------------------------------

def _poiseuille_series(y, t, a, nu, H, nmax=801):
    u = a / (2 * nu) * y * (H - y)
    for n in range(1, nmax, 2):
        u -= 4 * a * H ** 2 / (nu * np.pi ** 3 * n ** 3) * np.sin(n * np.pi * y / H) * np.exp(-n * n * np.pi ** 2 * nu * t / H ** 2)
    return u


def test_poiseuille_start_up_follows_the_transient_series():
    """Body force 1e-6 (Guo force 2 rho P e, P=5e-7) between half-way walls,
    tau=0.8, from rest: u_x(y) after 1000 steps against the Fourier series
    (x and z are periodic and the flow is invariant along them)."""
    tau, a, H = 0.8, 1e-6, 30
    o = OracleSim(4, H, 4, tau=tau, p=a / 2, e=(1, 0, 0))
    o.init_fluid()
    for it in range(1, 1001):
        o.lbm_d3q19(it)
    ux = comp_view(o, o.un)[:, :, :]
    assert np.ptp(ux, axis=(0, 2)).max() < 1e-18  # invariant along x and z
    nu = (tau - 0.5) / 3
    ua = _poiseuille_series(np.arange(H) + 0.5, 1000, a, nu, H)
    err = np.max(np.abs(ux[0, :, 0] - ua)) / ua.max()
    assert err < 2e-3, err
    # a missing half-force velocity shift or forcing prefactor is far outside
    assert 0.3 * a * H * H / (8 * nu) < ua.max()


def test_taylor_green_decay_rate():
    tau, u0, n = 0.6, 0.01, 32
    o = OracleSim(n, n, 2, tau=tau, periodic_y=True)
    X, Y, _ = o.node_coords()
    k = 2 * np.pi / n
    o.init_fluid(1.0 - 0.75 * u0 * u0 * (np.cos(2 * k * X) + np.cos(2 * k * Y)),
                 u0 * np.cos(k * X) * np.sin(k * Y), -u0 * np.sin(k * X) * np.cos(k * Y), None)
    steps = 200
    for it in range(1, steps + 1):
        o.lbm_d3q19(it)
    m = o.comp_mask()
    basis = (np.cos(k * X) * np.sin(k * Y))[m]
    amp = (o.un[m] * basis).sum() / (basis * basis).sum()
    nu = (tau - 0.5) / 3
    # O(k^2) lattice dispersion: 0.4% at 32 nodes per wavelength
    assert abs(amp / u0 - np.exp(-2 * nu * k * k * steps)) < 1e-2 * np.exp(-2 * nu * k * k * steps)


def _random_state(o, seed, amp=2e-3):
    rng = np.random.default_rng(seed)
    n = o.n
    o.init_fluid(1 + amp * rng.standard_normal(n), *(amp * rng.standard_normal(n) for _ in range(3)))
    o.fn[:] += 0.1 * amp * rng.standard_normal(o.fn.shape) / Q


def test_mass_and_momentum_balance_periodic():
    """Fully periodic: sum rho is conserved and the total momentum changes by
    exactly the summed Guo force 2 rho (P e + B F) per step."""
    P, B = 3e-5, 0.4
    o = OracleSim(6, 5, 4, tau=0.7, p=P, e=(0.6, 0.0, 0.8), b=B, periodic_y=True)
    _random_state(o, 1)
    rng = np.random.default_rng(2)
    o.fx[:], o.fy[:], o.fz[:] = (1e-3 * rng.standard_normal(o.n) for _ in range(3))
    m = o.comp_mask()
    for it in range(1, 4):
        f = o.fn.reshape(-1, Q)[m]
        rho = f.sum(1)
        J = f @ CQ
        G = 2 * rho[:, None] * (P * np.array([0.6, 0.0, 0.8]) + B * np.stack([o.fx[m], o.fy[m], o.fz[m]], 1))
        o.lbm_d3q19(it)
        f2 = o.fn.reshape(-1, Q)[m]
        assert abs(f2.sum() - f.sum()) < 1e-12 * f.sum()
        dJ = (f2 @ CQ).sum(0) - J.sum(0)
        assert np.max(np.abs(dJ - G.sum(0))) < 1e-13 * rho.sum()


def test_uniform_equilibrium_is_a_fixed_point():
    o = OracleSim(5, 4, 3, tau=0.9, periodic_y=True)
    o.init_fluid(np.full(o.n, 1.02), np.full(o.n, 0.03), np.full(o.n, -0.01), np.full(o.n, 0.02))
    f0 = o.fn.copy()
    for it in range(1, 3):
        o.lbm_d3q19(it)
    m = o.comp_mask()
    assert np.max(np.abs(o.fn.reshape(-1, Q)[m] - f0.reshape(-1, Q)[m])) < 1e-15


def test_channel_is_symmetric_under_x_z_exchange():
    """Exchanging the x and z axes (both periodic) of state and force commutes
    with the step: catches wrong neighbour offsets or direction tables."""
    perm = np.array([int(np.where((CQ == c[[2, 1, 0]]).all(1))[0][0]) for c in CQ])
    kw = dict(tau=0.75, p=2e-5, b=0.3)
    a = OracleSim(6, 5, 6, e=(0.6, 0.0, 0.8), **kw)
    b = OracleSim(6, 5, 6, e=(0.8, 0.0, 0.6), **kw)
    _random_state(a, 3)
    rng = np.random.default_rng(4)
    a.fx[:], a.fy[:], a.fz[:] = (1e-3 * rng.standard_normal(a.n) for _ in range(3))
    sw = lambda v: v.reshape(a.nz, a.ny, a.nx, *v.shape[1:]).swapaxes(0, 2).reshape(v.shape)
    b.fn[:] = sw(a.fn.reshape(-1, Q))[:, perm].reshape(-1)
    b.fx[:], b.fy[:], b.fz[:] = sw(a.fz), sw(a.fy), sw(a.fx)
    for it in range(1, 5):
        a.lbm_d3q19(it)
        b.lbm_d3q19(it)
    m = a.comp_mask()
    fa = sw(a.fn.reshape(-1, Q))[:, perm]
    assert np.max(np.abs(fa[sw(m)] - b.fn.reshape(-1, Q)[sw(m)])) < 1e-15
    assert np.max(np.abs(sw(a.un) - b.wn)[m]) < 1e-15


def test_wall_bounce_back_keeps_mass_and_no_flux():
    """Half-way bounce-back: mass is conserved and the wall-normal momentum
    through the wall planes vanishes in the mean."""
    o = OracleSim(4, 6, 4, tau=0.8, p=1e-5, e=(0.0, 1.0, 0.0))
    _random_state(o, 5)
    m = o.comp_mask()
    M0 = o.fn.reshape(-1, Q)[m].sum()
    for it in range(1, 30):
        o.lbm_d3q19(it)
    assert abs(o.fn.reshape(-1, Q)[m].sum() - M0) < 1e-12 * M0


# ---- immersed boundary -----------------------------------------------------------

def _velocity_field(o, seed, amp=1e-2, const=None):
    rng = np.random.default_rng(seed)
    comp = o.bc == 0
    out = []
    for d in range(3):
        a = np.zeros(o.n)
        a[comp] = amp * rng.standard_normal(comp.sum()) if const is None else const[d]
        out.append(a)
    idx = np.arange(o.n)
    iz, r = np.divmod(idx, o.nx * o.ny)
    iy, ix = np.divmod(r, o.nx)
    w = lambda a, m: np.where(a < 2, a + m - 4, np.where(a > m - 3, a - m + 4, a))
    img = w(ix, o.nx) + o.nx * (w(iy, o.ny) if o.P.periodic_y else iy) + o.nx * o.ny * w(iz, o.nz)
    buff = o.bc == 1
    for a in out:
        a[buff] = a[img[buff]]
    o.um[:], o.vm[:], o.wm[:] = out
    return out


def _cells_near_boundaries(lx, ly, lz, radius=3.0):
    """Two cells, one across the x/z periodic boundaries and one near a wall,
    unwrapped (as set up), plus the positions wrapped into the box."""
    from preprocess_src_amd import mesh
    centres = np.array([[lx / 2 - 1.0, 0.3, -lz / 2 + 1.3], [0.7, -ly / 2 + radius + 1.2, 0.4]])
    _, rot = mesh.place(2, radius, lx, ly, lz, seed=5)
    cells = mesh.assemble([2, 2], centres, rot, radius)
    wrapped = {k: cells[k].copy() for k in "xyz"}
    for k, L in (("x", lx), ("z", lz)):
        a = wrapped[k]
        a[a > L / 2] -= L
        a[a < -L / 2] += L
    return cells, wrapped


def _set_wrapped(o, wrapped):
    for k in "xyz":
        o.cells[k][:] = wrapped[k]
    o.cells["xd"][:] = np.stack([wrapped[k] for k in "xyz"], 1).reshape(-1)


def test_spread_and_interpolation_are_adjoint():
    """sum_nodes u . S(f) dx^3 = sum_v f . I(u) for the cosine delta, with the
    halo fold and the periodic images in place (RK1 so that I is one
    interpolation; dx=1 as the reference's RK1 branch assumes)."""
    lx, ly, lz = 16, 14, 12
    cells, wrapped = _cells_near_boundaries(lx, ly, lz)
    o = OracleSim(lx, ly, lz, tau=0.8, runge_kutta=1, cells=cells)
    _set_wrapped(o, wrapped)
    rng = np.random.default_rng(8)
    nv = cells["x"].size
    f = [rng.standard_normal(nv) for _ in range(3)]
    o.cells["cfx"][:], o.cells["cfy"][:], o.cells["cfz"][:] = f
    o.ibm_st1()
    u = _velocity_field(o, 9)
    o.ibm_st2()
    comp = o.bc == 0
    lhs = sum((u[d][comp] * F[comp]).sum() for d, F in enumerate((o.fx, o.fy, o.fz)))
    rhs = sum((f[d] * o.cells[k]).sum() for d, k in enumerate("uvw"))
    assert abs(lhs - rhs) < 1e-12 * abs(rhs)
    # total force is conserved by the spreading (interior cells, no wall loss)
    for d, F in enumerate((o.fx, o.fy, o.fz)):
        assert abs(F[comp].sum() - f[d].sum()) < 1e-12 * np.abs(f[d]).sum()


def test_uniform_field_is_interpolated_and_advected_exactly():
    lx, ly, lz = 16, 14, 12
    cells, wrapped = _cells_near_boundaries(lx, ly, lz)
    x0 = wrapped
    U = (0.013, -0.004, 0.021)
    o = OracleSim(lx, ly, lz, tau=0.8, runge_kutta=4, m=2, cr=0.7, dx=1.0, cells=cells)
    _set_wrapped(o, wrapped)
    _velocity_field(o, 0, const=U)
    o.ibm_st2()
    dt = o.P.dt
    for d, k in enumerate("uvw"):
        assert np.max(np.abs(o.cells[k] - U[d])) < 1e-15
    for d, (k, L) in enumerate((("x", lx), ("y", None), ("z", lz))):
        moved = x0[k] + U[d] * 2 * dt
        if L is not None:
            moved = np.where(moved > L / 2, moved - L, np.where(moved < -L / 2, moved + L, moved))
        assert np.max(np.abs(o.cells[k] - moved)) < 1e-13
    assert np.array_equal(o.cells["xd"].reshape(-1, 3)[:, 0], o.cells["x"])


# ---- membrane ------------------------------------------------------------------

def _cell(level=2, radius=4.0, seed=3, **kw):
    cells = small_cells(24, 20, 24, nc=1, radius=radius, levels=[level], seed=seed)
    o = OracleSim(24, 20, 24, tau=0.8, cells=cells, **kw)
    return cells, o


def _xd(d):
    return np.stack([d["x"], d["y"], d["z"]], 1).reshape(-1).copy()


def test_tension_vanishes_in_the_reference_state():
    cells, o = _cell()
    o.calc_tension(1)
    assert np.max(np.abs(o.cells["t"])) < 1e-12


def test_equibiaxial_tension_matches_the_closed_form():
    """x -> lam x about the centre: g = lam^2 G, I1 = 2(lam^2-1),
    I2 = lam^4-1, J = lam^2, so t = [(2 lam^2-1)/lam^2 + C(lam^4-1)-1] G^-1
    with G^-1 the reference contravariant metric (Skalak law, rgs=1)."""
    cells, o = _cell(c=100.0)
    lam = 1.01
    cx = [cells[k].mean() for k in "xyz"]
    d = {k: cx[i] + lam * (cells[k] - cx[i]) for i, k in enumerate("xyz")}
    o.cells["xd"][:] = _xd(d)
    o.calc_tension(1)
    coef = (2 * lam ** 2 - 1) / lam ** 2 + 100.0 * (lam ** 4 - 1) - 1
    assert np.max(np.abs(o.cells["t"] - coef * o.cells["nref"])) < 1e-10 * np.abs(coef * o.cells["nref"]).max()
    # rg = 2 x deformed area = lam^2 x reference
    X = _xd(cells).reshape(-1, 3)
    tri = cells["ele"].reshape(-1, 3)
    A0 = 0.5 * np.linalg.norm(np.cross(X[tri[:, 1]] - X[tri[:, 0]], X[tri[:, 2]] - X[tri[:, 0]]), axis=1)
    assert np.max(np.abs(o.cells["rg"] - 2 * lam ** 2 * A0)) < 1e-12


def _jacobi(o, cells):
    V = int(cells["voff"][1])
    rg, le, nle = o.cells["rg"], o.cells["le"].reshape(-1, 6), o.cells["nle"]
    return np.array([rg[le[j, :nle[j]]].sum() / 12.0 for j in range(V)])


def test_rhs_is_the_gradient_of_the_discrete_skalak_energy():
    cells, o = _cell(c=100.0)
    o.cells["xd"][:] = _xd(perturb_cells(cells, 5, amp=0.05))
    o.calc_tension(1)
    ptr, index, value, b = o.assemble(0)
    V = int(cells["voff"][1])
    bun = b[:3 * V].reshape(V, 3) * _jacobi(o, cells)[:, None]
    h = 1e-6
    rng = np.random.default_rng(1)
    for j in rng.choice(V, 12, replace=False):
        for d in range(3):
            o.cells["xd"][3 * j + d] += h
            Ep = o.skalak_energy()
            o.cells["xd"][3 * j + d] -= 2 * h
            Em = o.skalak_energy()
            o.cells["xd"][3 * j + d] += h
            g = (Ep - Em) / (2 * h)
            assert abs(bun[j, d] - g) < 1e-6 * np.abs(bun).max()


def test_mass_matrix_row_sums_are_nodal_areas_and_symmetric():
    cells, o = _cell()
    d = perturb_cells(cells, 6, amp=0.05)
    o.cells["xd"][:] = _xd(d)
    o.calc_tension(1)
    ptr, index, value, b = o.assemble(0)
    V = int(cells["voff"][1])
    mn = len(b)
    jac = _jacobi(o, cells)
    idx = index.reshape(7, mn)
    val = value.reshape(7, mn)
    A = np.zeros((V, V))
    for r in range(0, 3 * V, 3):
        for k in range(ptr[r]):
            A[r // 3, idx[k, r] // 3] = val[k, r] * jac[r // 3]
    X = d["x"], d["y"], d["z"]
    X = np.stack(X, 1)
    tri = cells["ele"].reshape(-1, 3)
    At = 0.5 * np.linalg.norm(np.cross(X[tri[:, 1]] - X[tri[:, 0]], X[tri[:, 2]] - X[tri[:, 0]]), axis=1)
    nodal = np.zeros(V)
    np.add.at(nodal, tri.reshape(-1), np.repeat(At / 3, 3))
    assert np.max(np.abs(A.sum(1) - nodal)) < 1e-13 * nodal.max()
    assert np.max(np.abs(A - A.T)) < 1e-15 * nodal.max()
    # pad rows hold one self entry and a zero right-hand side (stress.cu:173-176)
    assert np.all(ptr[3 * V:] == 1) and np.all(idx[0, 3 * V:] == np.arange(3 * V, mn)) and np.all(b[3 * V:] == 0)
    # unit diagonal after the Jacobi scaling
    diag = np.array([val[k, r] for r in range(3 * V) for k in range(ptr[r]) if idx[k, r] == r])
    assert diag.size == 3 * V and np.all(diag == 1.0)


def test_bicgstab_matches_a_dense_solve():
    cells, o = _cell(level=2)
    o.cells["xd"][:] = _xd(perturb_cells(cells, 7, amp=0.05))
    o.calc_tension(1)
    ptr, index, value, b = o.assemble(0)
    mn = len(b)
    assert mn == 512  # 162 vertices -> expand2pow2(486)
    A = np.zeros((mn, mn))
    idx, val = index.reshape(7, mn), value.reshape(7, mn)
    for r in range(mn):
        for k in range(ptr[r]):
            A[r, idx[k, r]] += val[k, r]
    n3 = 3 * int(cells["voff"][1])  # pad rows are zero rows with zero right-hand side
    x_dense = np.linalg.solve(A[:n3, :n3], b[:n3])
    rc, x, its = bicgstab(ptr, index, value, b, residual=1e-13)
    assert rc == 0 and 0 < its < 100
    assert np.max(np.abs(x[:n3] - x_dense)) < 1e-11 * np.abs(x_dense).max() and np.all(x[n3:] == 0)
    rc2, _, its2 = bicgstab(ptr, index, value, b, itrmax=2, residual=1e-13)
    assert rc2 == 2 and its2 == 2
    rc3, _, _ = bicgstab(ptr, index, value, np.zeros(mn))
    assert rc3 == 1  # (r, r0) = 0 breakdown for a zero right-hand side


def test_expand2pow2():
    f = lib().orc_expand2pow2
    assert [f(n) for n in (1, 2, 3, 256, 257, 1926, 7686, 30726)] == [1, 2, 4, 256, 512, 2048, 8192, 32768]


def test_nodal_forces_have_no_net_force_and_the_rhs_no_torque():
    """Row sums of the mass matrix are the q2f nodal areas, so
    sum_j ds_j (M^-1 b)_j = sum_j b_j = 0: the nodal forces carry no net force.
    b is the gradient of a rotation-invariant energy: no net torque (f itself
    is not torque-free, the consistent-mass solve mixes the nodes)."""
    cells, o = _cell(c=100.0, br=0.05)
    d = perturb_cells(cells, 8, amp=0.05)
    for k in "xyz":
        o.cells[k][:] = d[k]
    o.cells["xd"][:] = _xd(d)
    assert o.stress(1) == 0
    f = np.stack([o.cells["cfx"], o.cells["cfy"], o.cells["cfz"]], 1)
    assert np.max(np.abs(f.sum(0))) < 1e-12 * np.abs(f).sum()
    ptr, index, value, b = o.assemble(0)
    V = int(cells["voff"][1])
    bun = b[:3 * V].reshape(V, 3) * _jacobi(o, cells)[:, None]
    X = np.stack([d["x"], d["y"], d["z"]], 1)
    assert np.max(np.abs(bun.sum(0))) < 1e-12 * np.abs(bun).sum()
    assert np.max(np.abs(np.cross(X - X.mean(0), bun).sum(0))) < 1e-11 * np.abs(bun).sum() * 5.0


def test_bending_forces_are_the_energy_gradient():
    cells, o = _cell(br=0.1)
    d = perturb_cells(cells, 5, amp=0.1)
    for k in "xyz":
        o.cells[k][:] = d[k]
    o.cells["xd"][:] = _xd(cells)  # undeformed in-plane: q = 0, only bending remains
    assert o.stress(1) == 0
    f = np.stack([o.cells["cfx"], o.cells["cfy"], o.cells["cfz"]], 1).copy()
    h = 1e-6
    for j in range(0, len(f), 11):
        for dd, k in enumerate("xyz"):
            o.cells[k][j] += h
            Ep = o.bending_energy()
            o.cells[k][j] -= 2 * h
            Em = o.bending_energy()
            o.cells[k][j] += h
            assert abs(f[j, dd] + (Ep - Em) / (2 * h)) < 1e-6 * np.abs(f).max()
    X = np.stack([d["x"], d["y"], d["z"]], 1)
    assert np.max(np.abs(f.sum(0))) < 1e-12 * np.abs(f).sum()
    assert np.max(np.abs(np.cross(X - X.mean(0), f).sum(0))) < 1e-11 * np.abs(f).sum() * 5.0


def test_q2f_ring_areas_sum_to_the_surface_area():
    """f_j = -q_j ds_j with ds_j a third of the incident triangle areas, so
    sum_j ds_j (recovered as -f/q) is the membrane area."""
    cells, o = _cell(c=100.0)
    d = perturb_cells(cells, 12, amp=0.05)
    for k in "xyz":
        o.cells[k][:] = d[k]
    o.cells["xd"][:] = _xd(d)
    assert o.stress(1) == 0
    q = o.cells["q"].reshape(-1, 3)
    f = np.stack([o.cells["cfx"], o.cells["cfy"], o.cells["cfz"]], 1)
    k = np.argmax(np.abs(q), axis=1)
    r = np.arange(len(q))
    ds = -f[r, k] / q[r, k]
    _, area = o.calc_volume()
    assert abs(ds.sum() - area[0]) < 1e-12 * area[0]
    assert np.max(np.abs(f + q * ds[:, None])) < 1e-14 * np.abs(f).max()


# ---- volume --------------------------------------------------------------------

def test_sphere_volume_and_area_converge():
    from preprocess_src_amd import mesh
    r = 5.0
    errs = []
    for lev in (2, 3, 4):
        cells = mesh.assemble([lev], np.zeros((1, 3)), np.eye(3)[None], r, shape="sphere")
        o = OracleSim(24, 20, 24, tau=0.8, cells=cells)
        V, A = o.calc_volume()
        X = np.stack([cells["x"], cells["y"], cells["z"]], 1)
        tri = cells["ele"].reshape(-1, 3)
        tet = np.einsum("ij,ij->i", X[tri[:, 0]], np.cross(X[tri[:, 1]], X[tri[:, 2]])).sum() / 6
        assert abs(V[0] - tet) < 1e-12 * tet
        errs.append(abs(V[0] - 4 / 3 * np.pi * r ** 3) / (4 / 3 * np.pi * r ** 3))
        assert abs(A[0] - 4 * np.pi * r * r) / (4 * np.pi * r * r) < 0.05
    # second order in the edge length: the error drops 4x per subdivision
    assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5 and errs[2] < 3e-3


def test_volume_correction_restores_the_reference_volume():
    cells = small_cells(24, 20, 24, nc=2, radius=4.0, levels=[2, 3], seed=9)
    o = OracleSim(24, 20, 24, tau=0.8, cells=cells)
    o.volume_init()
    V0 = o.cells["volume0"].copy()
    for c in range(2):
        sl = slice(cells["voff"][c], cells["voff"][c + 1])
        for i, k in enumerate("xyz"):
            g = o.cells["gx"][3 * c + i]
            o.cells[k][sl] = g + 1.004 * (o.cells[k][sl] - g)
    k = o.volume_correction()
    assert k > 0
    # volumes of the final pass (computed on positions unwrapped about the centroid)
    assert np.all(np.abs(o.cells["volume"] - V0) / V0 < 1e-5)
    # the same from scratch, unwrapping by hand
    for c in range(2):
        sl = slice(cells["voff"][c], cells["voff"][c + 1])
        X = np.stack([o.cells[k][sl] for k in "xyz"], 1)
        g = o.cells["gx"][3 * c:3 * c + 3]
        for i, L in ((0, 24.0), (2, 24.0)):
            dlt = X[:, i] - g[i]
            X[:, i] = g[i] + np.where(dlt > L / 2, dlt - L, np.where(dlt < -L / 2, dlt + L, dlt))
        tri = cells["ele"][3 * cells["eoff"][c]:3 * cells["eoff"][c + 1]].reshape(-1, 3) - cells["voff"][c]
        vol = np.einsum("ij,ij->i", X[tri[:, 0]] - g, np.cross(X[tri[:, 1]] - g, X[tri[:, 2]] - g)).sum() / 6
        assert abs(vol - V0[c]) / V0[c] < 1e-5


# ---- body-force field and wall-shear objective (gradient checks) ---------------

def test_uniform_body_force_field_equals_the_body_force():
    """A field g = P e on every node reproduces the run with body force P e."""
    lx, ly, lz, P = 8, 6, 6, 2e-4
    a = OracleSim(lx, ly, lz, tau=0.8, p=P, volume_correction=False)
    b = OracleSim(lx, ly, lz, tau=0.8, p=0.0, volume_correction=False)
    a.init_fluid()
    b.init_fluid()
    b.set_body_force_field(np.full(b.n, P), np.zeros(b.n), np.zeros(b.n))
    Ja = a.wall_shear(0, 30, 10)
    Jb = b.wall_shear(0, 30, 10)
    assert Ja > 0 and abs(Ja - Jb) <= 1e-14 * abs(Ja)
    fa, fb = comp_view(a, a.fn.reshape(-1, Q)), comp_view(b, b.fn.reshape(-1, Q))
    assert np.max(np.abs(fa - fb)) <= 1e-15


def test_wall_shear_is_the_mean_first_node_velocity():
    """The objective is 2 Df (dx/dt) e.u at the two wall-adjacent node layers,
    averaged over them and over the window (checked against the macros)."""
    lx, ly, lz = 6, 8, 4
    o = OracleSim(lx, ly, lz, tau=0.9, p=1e-4, e=(0.6, 0.0, 0.8), volume_correction=False)
    o.init_fluid()
    J = o.wall_shear(0, 5, 1)
    df = (2 * 0.9 - 1) / 6
    u = comp_view(o, o.um)
    w = comp_view(o, o.wm)
    walls = [0, ly - 1]
    expect = 2 * df * np.mean([0.6 * u[:, iy, :].mean() + 0.8 * w[:, iy, :].mean() for iy in walls])
    assert abs(J - expect) <= 1e-13 * abs(expect)


# ---- pins of the RK stage offsets, the modulated force and the neo-Hookean law ----

def _m1(t):
    """Closed-form first moment of the cosine delta, sum_j (j - X) phi(X - j)
    with t = frac(X) (in units of the spacing)."""
    return -(t - 0.5 + (np.cos(np.pi * t / 2) - np.sin(np.pi * t / 2)) / 2)


def _linear_interp(X, a, B, lo, h):
    """Interpolation of the linear field u(x) = a + B x by the cosine delta:
    a + B (X + h m1(t)), t the fraction of X in the node frame (nodes at
    lo + j h); exact because sum phi = 1 and the first moment is m1."""
    t = np.mod((X - lo[:, None]) / h, 1.0)
    return a[:, None] + B @ (X + h * _m1(t))


def _linear_field_oracle(rk, dx, m, cr):
    lx, ly, lz = 16 * dx, 14 * dx, 12 * dx
    cells = small_cells(lx, ly, lz, nc=1, radius=3.0 * dx, levels=[2], seed=71)
    o = OracleSim(lx, ly, lz, dx=dx, tau=0.8, runge_kutta=rk, m=m, cr=cr, cells=cells)
    rng = np.random.default_rng(72)
    nv = o.cells["x"].size
    # vertices anywhere in the interior, well clear of the padded edges
    for k, L in (("x", lx), ("y", ly), ("z", lz)):
        o.cells[k][:] = rng.uniform(-L / 2 + 3 * dx, L / 2 - 3 * dx, nv)
    a = np.array([0.011, -0.007, 0.005])
    B = np.array([[0.004, 0.021, -0.006], [-0.013, 0.003, 0.009], [0.008, -0.017, 0.002]]) / dx
    X, Y, Z = o.node_coords()
    o.um[:], o.vm[:], o.wm[:] = a[:, None] + B @ np.stack([X, Y, Z])  # every padded node
    lo = np.array([-lx / 2 - 1.5 * dx, -ly / 2 - 1.5 * dx, -lz / 2 - 1.5 * dx])
    return o, a, B, lo


@pytest.mark.parametrize("dx,m,cr", [(1.0, 2, 0.7), (0.5, 3, 0.9)])
def test_rk4_stages_on_a_linear_field(dx, m, cr):
    """RK4 with the field frozen (ibm.cu:430-517): stages at X, X + k1 M dt/2,
    X + k2 M dt/2, X + k3 M dt, then X + M dt (k1 + 2k2 + 2k3 + k4)/6.  The
    field is linear, so each stage's velocity is a + B (X + h m1(t)) in
    closed form; a slipped stage offset (k M dt for the second stage, k1 for
    the third) moves the result by ~1e-4, far above the 1e-13 bound."""
    o, a, B, lo = _linear_field_oracle(4, dx, m, cr)
    X0 = np.stack([o.cells[k].copy() for k in "xyz"])
    o.ibm_st2()
    h, Mdt = dx, m * o.P.dt
    k1 = _linear_interp(X0, a, B, lo, h)
    k2 = _linear_interp(X0 + k1 * Mdt / 2, a, B, lo, h)
    k3 = _linear_interp(X0 + k2 * Mdt / 2, a, B, lo, h)
    k4 = _linear_interp(X0 + k3 * Mdt, a, B, lo, h)
    um = (k1 + 2 * k2 + 2 * k3 + k4) / 6
    X1 = X0 + um * Mdt
    for d, k in enumerate("xyz"):
        assert np.max(np.abs(o.cells[k] - X1[d])) < 1e-13, k
        assert np.max(np.abs(o.cells["uvw"[d]] - um[d])) < 1e-14
    # the slips named above are visible at this size
    k3s = _linear_interp(X0 + k2 * Mdt, a, B, lo, h)
    assert np.max(np.abs((k1 + 2 * k2 + 2 * k3s + k4) / 6 - um)) * Mdt > 1e-6


def test_rk2_predictor_on_a_linear_field():
    """RK2 (ibm.cu:520-564): the predictor is X + k1 M with no dt and the
    stencil assumes unit spacing; the update is X + M dt (k1 + k2)/2."""
    o, a, B, lo = _linear_field_oracle(2, 1.0, 2, 0.7)
    X0 = np.stack([o.cells[k].copy() for k in "xyz"])
    o.ibm_st2()
    M, dt = 2, o.P.dt
    k1 = _linear_interp(X0, a, B, lo, 1.0)
    k2 = _linear_interp(X0 + k1 * M, a, B, lo, 1.0)
    X1 = X0 + M * dt * (k1 + k2) / 2
    for d, k in enumerate("xyz"):
        assert np.max(np.abs(o.cells[k] - X1[d])) < 1e-13, k
    k2dt = _linear_interp(X0 + k1 * M * dt, a, B, lo, 1.0)
    assert np.max(np.abs(k2dt - k2)) * M * dt > 1e-6


@pytest.mark.parametrize("flowstart", [0, 3])
def test_modulated_body_force_follows_e_of_t(flowstart):
    """force_modulated = 1 (lbm.cu:1032-1035): the force is P e (1 + 0.5
    sin(pi t / FREQ)) with t = dt iter once iter > FLOWSTART, and 0 before.
    In a fully periodic box the total momentum grows each step by exactly
    sum 2 rho P e amp(iter), which fixes the gate, the time and the phase."""
    P, e, freq, cr = 3e-5, np.array([0.6, 0.0, 0.8]), 7.0, 0.8
    o = OracleSim(6, 5, 4, tau=0.7, p=P, e=tuple(e), cr=cr, periodic_y=True, force_modulated=1,
                  flowstart=flowstart, freq=freq)
    _random_state(o, 11)
    m = o.comp_mask()
    dt = o.P.dt
    for it in range(1, 9):
        f = o.fn.reshape(-1, Q)[m]
        rho = f.sum(1)
        J = (f @ CQ).sum(0)
        o.lbm_d3q19(it)
        dJ = (o.fn.reshape(-1, Q)[m] @ CQ).sum(0) - J
        amp = 1.0 + 0.5 * np.sin(np.pi * dt * it / freq) if it > flowstart else 0.0
        assert np.max(np.abs(dJ - 2 * rho.sum() * P * amp * e)) < 1e-13 * rho.sum(), it


@pytest.mark.parametrize("rgs", [1.0, 1.7])
def test_neohookean_equibiaxial_tension_matches_the_closed_form(rgs):
    """NEOHOOKEAN (stress.cu:97-99): t = rgs (G^-1/J - g^-1 J/(I2+1)^2).  Under
    x -> lam x, g = lam^2 G, J = lam^2, I2 + 1 = lam^4, so
    t = rgs (1/lam^2 - 1/lam^8) G^-1."""
    cells, o = _cell(neohookean=True, rgs=rgs)
    lam = 1.013
    cx = [cells[k].mean() for k in "xyz"]
    o.cells["xd"][:] = _xd({k: cx[i] + lam * (cells[k] - cx[i]) for i, k in enumerate("xyz")})
    o.calc_tension(1)
    ref = rgs * (1 / lam ** 2 - 1 / lam ** 8) * o.cells["nref"]
    assert np.max(np.abs(o.cells["t"] - ref)) < 1e-10 * np.abs(ref).max()


def test_skalak_shear_term_alone_is_scaled_by_rgs():
    """SKALAK (stress.cu:96): rgs multiplies only the (I1 + 1)/J term, not the
    area term (C I2 - 1) J g^-1, so under x -> lam x,
    t = [rgs (2 lam^2 - 1)/lam^2 + C (lam^4 - 1) - 1] G^-1."""
    rgs, C_, lam = 1.7, 100.0, 1.01
    cells, o = _cell(c=C_, rgs=rgs)
    cx = [cells[k].mean() for k in "xyz"]
    o.cells["xd"][:] = _xd({k: cx[i] + lam * (cells[k] - cx[i]) for i, k in enumerate("xyz")})
    o.calc_tension(1)
    coef = rgs * (2 * lam ** 2 - 1) / lam ** 2 + C_ * (lam ** 4 - 1) - 1
    ref = coef * o.cells["nref"]
    assert np.max(np.abs(o.cells["t"] - ref)) < 1e-10 * np.abs(ref).max()


# ---- sliding walls (CapsuleSHEARFLOW / SHEARFLOW) -----------------------------------

def _couette_series(eta, t, U, nu, H, top_only, nmax=2001):
    """Start-up Couette flow between walls at eta = 0 and H, from rest, by the
    sine series of u_t = nu u_yy: allwall (u = -U at 0, +U at H) or topwall
    (0 at 0, U at H)."""
    n = np.arange(1, nmax)[:, None]
    dec = np.exp(-nu * (n * np.pi / H) ** 2 * t) * np.sin(n * np.pi * eta[None, :] / H)
    if top_only:
        steady = U * eta / H
        b = 2 * U * (-1.0) ** n / (n * np.pi)
    else:
        steady = -U + 2 * U * eta / H
        b = np.where(n % 2 == 0, 4 * U / (n * np.pi), 0.0)
    return steady + (b * dec).sum(0)


@pytest.mark.parametrize("mode", [1, 2])
def test_couette_start_up_follows_the_series(mode):
    """Walls at y = +-LY/2 sliding along z (Ladd's term, lbm.cu:951-970), from
    rest: u_z(y) after 600 steps against the series, to 1% of U.  A wrong sign
    or a missing Cr in the wall term reverses or rescales the profile."""
    H, tau, cr, U = 24, 0.8, 0.5, 0.04
    o = OracleSim(4, H, 4, tau=tau, cr=cr, wall_motion=mode, wall_velocity=U)
    o.init_fluid()
    steps = 600
    for it in range(1, steps + 1):
        o.lbm_d3q19(it)
    uz = comp_view(o, o.wn)
    assert np.ptp(uz, axis=(0, 2)).max() < 1e-16
    assert np.max(np.abs(comp_view(o, o.un))) < 1e-16
    nu = (tau - 0.5) / 3
    # lattice time: nu in lattice units, steps in lattice steps
    ua = _couette_series(np.arange(H) + 0.5, steps, U, nu, H, top_only=(mode == 2))
    assert np.max(np.abs(uz[0, :, 0] - ua)) < 1e-2 * U, np.max(np.abs(uz[0, :, 0] - ua)) / U


@pytest.mark.parametrize("mode", [1, 2])
def test_couette_steady_profile_is_linear(mode):
    """Half-way bounce-back with Ladd's term gives the exact linear profile of
    plane Couette flow: from the sheared start of flu_init (init1.cu:367-384,
    u_z = sr y for allwall, sr (y + LY/2) for topwall), after long relaxation
    u_z is linear to round-off and meets +-U at the half-way wall planes."""
    H, tau, cr, sr = 12, 0.9, 0.5, 0.004
    U = sr * H / 2 if mode == 1 else sr * H
    o = OracleSim(4, H, 4, tau=tau, cr=cr, wall_motion=mode, wall_velocity=U)
    _, Y, _ = o.node_coords()
    o.init_fluid(None, None, None, sr * Y if mode == 1 else sr * (Y + H / 2))
    for it in range(1, 3001):
        o.lbm_d3q19(it)
    y = np.arange(H) + 0.5 - H / 2
    ref = sr * y if mode == 1 else sr * (y + H / 2)
    uz = comp_view(o, o.wn)[0, :, 0]
    assert np.max(np.abs(uz - ref)) < 1e-12 * U


def test_walls_rest_before_flowstart():
    """sra = 0 while iter < FLOWSTART (lbm.cu:755): the fluid stays at rest
    through step FLOWSTART - 1, so after step n the state equals that of the
    run with FLOWSTART = 0 after step n - FLOWSTART + 1."""
    kw = dict(tau=0.8, cr=0.5, wall_motion=1, wall_velocity=0.03)
    a = OracleSim(4, 10, 4, flowstart=5, **kw)
    b = OracleSim(4, 10, 4, flowstart=0, **kw)
    a.init_fluid()
    b.init_fluid()
    for it in range(1, 5):
        a.lbm_d3q19(it)
    m = a.comp_mask()
    assert np.max(np.abs(comp_view(a, a.wn))) == 0.0
    for it in range(5, 15):
        a.lbm_d3q19(it)
    for it in range(1, 11):
        b.lbm_d3q19(it)
    assert np.max(np.abs(a.fn.reshape(-1, Q)[m] - b.fn.reshape(-1, Q)[m])) < 1e-15


def test_sliding_walls_conserve_mass_and_shear_stress_balances():
    """Ladd's term moves momentum, not mass: sum f is conserved.  In the
    steady state the z-momentum the top wall puts in per step equals what the
    bottom wall takes out, so total z-momentum stops changing."""
    H = 10
    o = OracleSim(4, H, 4, tau=0.8, cr=0.5, wall_motion=1, wall_velocity=0.02)
    _, Y, _ = o.node_coords()
    o.init_fluid(None, None, None, 0.02 * Y / (H / 2))
    m = o.comp_mask()
    M0 = o.fn.reshape(-1, Q)[m].sum()
    for it in range(1, 2001):
        o.lbm_d3q19(it)
    f = o.fn.reshape(-1, Q)[m]
    assert abs(f.sum() - M0) < 1e-12 * M0
    Jz = (f @ CQ)[:, 2].sum()
    o.lbm_d3q19(2001)
    Jz2 = (o.fn.reshape(-1, Q)[m] @ CQ)[:, 2].sum()
    assert abs(Jz2 - Jz) < 1e-13 * M0


def test_tube_staircase_is_half_way_bounce_back_on_the_mask():
    """lbm_CapsuleCHANNELFLOW's tube (lbm.cu:1050-1094): every link into a
    WALL node of the disc mask is bounced half-way, with no interpolation:
    mass is conserved exactly, the flow keeps the square lattice's symmetry
    x <-> y, and Poiseuille flow at R = 7.5 comes within 7% of the circular
    profile's centreline, and further from it than the interpolated form
    (the staircase wall sits inside r = R: 0.947 against 0.987 here)."""
    R, L, tau, a = 7.5, 18, 0.8, 1e-6
    o = OracleSim(L, L, 4, tau=tau, p=a / 2, e=(0, 0, 1), tube_radius=R, tube_staircase=True)
    o.init_fluid()
    m = o.comp_mask()
    M0 = o.fn.reshape(-1, Q)[m].sum()
    for it in range(1, 3001):
        o.lbm_d3q19(it)
    assert abs(o.fn.reshape(-1, Q)[m].sum() - M0) < 1e-12 * M0
    w = o.wn.reshape(o.nz, o.ny, o.nx)[2]
    assert np.max(np.abs(w - w.T)) < 1e-15
    nu = (tau - 0.5) / 3
    wmax = a * R * R / (4 * nu)
    assert 0.93 < w.max() / wmax < 0.97
    o2 = OracleSim(L, L, 4, tau=tau, p=a / 2, e=(0, 0, 1), tube_radius=R)
    o2.init_fluid()
    for it in range(1, 3001):
        o2.lbm_d3q19(it)
    w2 = o2.wn.reshape(o2.nz, o2.ny, o2.nx)[2]
    assert abs(w2.max() / wmax - 1) < abs(w.max() / wmax - 1) / 3
