"""VOF indicator: front tracking, its red-black SOR solve and the THINC/WLIC
advection (fronttrack.cu, matrix_solver.cu:112-216, thincwlic.cu), and the
viscosity it sets (lbm.cu:1131).

CPU: the oracle's restatement is pinned by what the mathematics fixes -- the
THINC flux is the integral of the tanh profile whose cell mean is the
volume fraction (quadrature, root-finding), the normals of a linear field
are its gradient, the conservative update conserves the COMP sum, a planar
front reduces to the 1-D scheme, the spread of one triangle is its area
vector with the cosine weights, the SOR iterate is the direct solution of
the reference's Jacobi-scaled matrix (scipy), and the indicator of a sphere
is 1 inside, 0 outside with the sphere's volume.  GPU: the library's
front tracking and THINC/WLIC against the oracle, and whole steps with a
viscosity contrast.
"""
import math

import numpy as np
import pytest
import scipy.integrate
import scipy.optimize
import scipy.sparse
import scipy.sparse.linalg

from oracle import pyoracle as PO
from tests.pairs import Q, gpu_of, oracle_of, rel

BETA = 2.0


def _tanh_profile(c, alpha):
    """H(x) = (1 + alpha tanh(beta (x - xc)))/2 on [0, 1] with mean c."""
    mean = lambda xc: 0.5 + 0.5 * alpha / BETA * (math.log(math.cosh(BETA * (1 - xc)))
                                                   - math.log(math.cosh(BETA * xc))) - c
    xc = scipy.optimize.brentq(mean, -20.0, 20.0, xtol=1e-14)
    return lambda x: 0.5 * (1.0 + alpha * math.tanh(BETA * (x - xc)))


@pytest.mark.parametrize("c", [0.05, 0.3, 0.5, 0.77, 0.96])
@pytest.mark.parametrize("alpha", [1.0, -1.0])
@pytest.mark.parametrize("courant", [0.31, 0.04, -0.22])
def test_thinc_flux_is_the_integral_of_the_tanh_profile(c, alpha, courant):
    # flux through the face during M dt: the upwind cell's profile over the
    # segment that crosses it (below the face for uc > 0, above for uc < 0)
    dx, dt, m = 0.5, 0.25, 2
    uc = courant * dx / (m * dt)
    H = _tanh_profile(c, alpha)
    if courant > 0:
        exact = dx * scipy.integrate.quad(H, 1.0 - courant, 1.0, epsabs=1e-14)[0]
        isgn = 1
    else:
        exact = -dx * scipy.integrate.quad(H, 0.0, -courant, epsabs=1e-14)[0]
        isgn = 0
    got = PO.thinc_flux(c, alpha, uc, isgn, m, dt, dx)
    # expf / coshf of the reference carry float rounding
    assert abs(got - exact) < 2e-6 * dx, (got, exact)


LX, LY, LZ = 16, 12, 14


def _oracle(**kw):
    return oracle_of(LX, LY, LZ, **kw)


def test_normals_of_a_linear_field_are_its_gradient():
    o = _oracle(dx=0.5)
    X, Y, Z = o.node_coords()
    g = (0.3, -1.1, 0.45)
    n1 = o.vof_normals(g[0] * X + g[1] * Y + g[2] * Z)
    inner = np.zeros((o.nz, o.ny, o.nx), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    inner = inner.reshape(-1)
    for d in range(3):
        assert np.max(np.abs(n1[d][inner] - g[d])) < 1e-12


def _periodic_fill(o, a):
    """BUFF nodes take their image's value, WALL nodes 0."""
    nx, ny, nz = o.nx, o.ny, o.nz
    a = a.reshape(nz, ny, nx).copy()
    ix = np.r_[np.arange(nx - 4, nx - 2), np.arange(2, nx - 2), np.arange(2, 4)]
    iz = np.r_[np.arange(nz - 4, nz - 2), np.arange(2, nz - 2), np.arange(2, 4)]
    a = a[iz][:, :, ix]
    a = a.reshape(-1)
    a[o.bc == 2] = 0.0
    return a


def _smooth_field(o, seed, lo=0.0, hi=1.0):
    rng = np.random.default_rng(seed)
    X, Y, Z = o.node_coords()
    f = np.zeros(o.n)
    for _ in range(3):
        k = rng.integers(1, 3, size=3) * 2 * np.pi / np.array([LX, LY, LZ]) * o.P.dx ** 0
        f += rng.uniform(0.5, 1.0) * np.sin(k[0] * X + rng.uniform(0, 6)) * np.cos(k[2] * Z + rng.uniform(0, 6)) \
            * np.cos(np.pi * Y / LY)
    f = (f - f.min()) / (f.max() - f.min())
    return _periodic_fill(o, lo + (hi - lo) * f)


def _velocity(o, seed, amp=0.08):
    rng = np.random.default_rng(seed)
    X, Y, Z = o.node_coords()
    out = []
    for d in range(3):
        k = 2 * np.pi * rng.integers(1, 3, size=2) / np.array([LX, LZ])
        out.append(_periodic_fill(o, amp * (np.sin(k[0] * X + d) * np.cos(k[1] * Z) + 0.5 * rng.uniform(-1, 1))
                                  * np.cos(np.pi * Y / (LY + 1))))
    return out


def test_thinc_conserves_the_comp_sum():
    o = _oracle(m=2)
    o.vf[:] = _smooth_field(o, 3, 0.0, 1.0)
    o.um[:], o.vm[:], o.wm[:] = _velocity(o, 4)
    comp = o.bc == 0
    before = o.vf[comp].sum()
    o.thinc_wlic()
    assert abs(o.vf[comp].sum() - before) < 1e-12 * before
    assert not np.array_equal(o.vf, _smooth_field(o, 3, 0.0, 1.0))


def test_thinc_of_a_planar_front_is_the_one_dimensional_scheme():
    o = _oracle(m=1)
    X, Y, Z = o.node_coords()
    x = X.reshape(o.nz, o.ny, o.nx)[0, 0]
    # a periodic front profile in x only, with plateaus at 0 and 1
    prof = 0.5 * (np.tanh(1.3 * np.sin(2 * np.pi * (x + 0.37) / LX) * 3) + 1.0)
    vf = np.broadcast_to(prof, (o.nz, o.ny, o.nx)).reshape(-1).copy()
    vf[o.bc == 2] = 0.0
    o.vf[:] = vf
    u0 = 0.21
    o.um[:] = u0
    o.vm[:] = 0.0
    o.wm[:] = 0.0
    o.thinc_wlic()
    # 1-D: faces i (between i-1 and i), central-difference normal, alpha from
    # the upwind cell's neighbours, THINC weight 1 where the normal is not 0
    dx, dt = o.P.dx, o.P.dt
    flux = np.zeros(o.nx)
    for i in range(2, o.nx - 1):
        ii = i - 1  # u0 > 0
        c = prof[ii]
        up = c * u0 * dt
        nn = abs(prof[ii + 1] - prof[ii - 1]) / (2 * dx)
        if c >= 1 - 1e-6 or c <= 1e-6 or nn <= 1e-6:
            flux[i] = up
        else:
            alpha = 1.0 if prof[ii - 1] <= prof[ii + 1] else -1.0
            flux[i] = PO.thinc_flux(c, alpha, u0, 1, 1, dt, dx)
    new = prof.copy()
    new[2:o.nx - 2] = prof[2:o.nx - 2] - (flux[3:o.nx - 1] - flux[2:o.nx - 2]) / dx
    got = o.vf.reshape(o.nz, o.ny, o.nx)
    # rows whose normals see no wall (the wall layers hold vf = 0)
    for iz in (2, o.nz // 2):
        for iy in (3, o.ny // 2):
            assert np.max(np.abs(got[iz, iy, 2:-2] - new[2:-2])) < 1e-13


def _sphere(center, radius, level=3):
    from preprocess_src_amd import mesh
    return mesh.assemble([level], np.array([center], float), np.eye(3)[None], radius, shape="sphere")


def _one_element(o, pts):
    """Restrict the oracle's cell set to element 0 with vertices pts (3x3)."""
    c = o.cells
    for k, v in enumerate(c["ele"][:3]):
        c["x"][v], c["y"][v], c["z"][v] = pts[k]
    o._eoff1 = np.array([0, 1], np.int32)
    o.S.eoff = PO.ip(o._eoff1)
    o.S.nc = 1


def _weights(a):
    k = np.arange(-1, 3)
    return (1.0 + np.cos(np.pi * (k - a) / 2)) / 4.0


@pytest.mark.parametrize("frac", [(0.0, 0.0, 0.0), (0.3, 0.55, 0.9)])
def test_front_tracking_spread_of_one_triangle(frac):
    dx = 0.5
    o = _oracle(dx=dx, cells=_sphere((0.0, 0.0, 0.0), 2.0, level=2))
    # centroid at padded node (7, 6, 8) plus frac
    cen = np.array([(7 + frac[0]) * dx - (o.nx - 1) * dx / 2, (6 + frac[1]) * dx - (o.ny - 1) * dx / 2,
                    (8 + frac[2]) * dx - (o.nz - 1) * dx / 2])
    tri = cen + np.array([[0.31, -0.12, 0.05], [-0.2, 0.27, -0.11], [0.0, 0.0, 0.0]])
    tri[2] = 3 * cen - tri[0] - tri[1]
    _one_element(o, tri)
    gx, gy, gz = o.ft_gradient()
    e1, e2 = tri[1] - tri[0], tri[2] - tri[0]
    area_vec = -np.cross(e1, e2) / 2.0  # inward for outward-wound triangles
    G = np.stack([gx, gy, gz], 1).reshape(o.nz, o.ny, o.nx, 3)
    assert np.allclose(G.reshape(-1, 3).sum(0) * dx ** 3, area_vec, rtol=0, atol=1e-14)
    wx, wy, wz = (_weights(a) for a in frac)
    box = G[8 - 1:8 + 3, 6 - 1:6 + 3, 7 - 1:7 + 3]
    expect = (wz[:, None, None] * wy[None, :, None] * wx[None, None, :])[..., None] * area_vec / dx ** 3
    assert np.max(np.abs(box - expect)) < 1e-14 / dx ** 3
    assert np.count_nonzero(np.abs(G).sum(-1)) == np.count_nonzero(np.abs(expect).sum(-1))


def test_front_tracking_spread_folds_the_halo_onto_the_images():
    dx = 1.0
    o = _oracle(dx=dx, cells=_sphere((0.0, 0.0, 0.0), 2.0, level=2))
    # an element whose box straddles the +x and -z ends
    cen = np.array([LX / 2 - 0.4, 0.7, -LZ / 2 + 0.6])
    tri = cen + np.array([[0.3, -0.2, 0.1], [-0.25, 0.3, -0.15], [0.0, 0.0, 0.0]])
    tri[2] = 3 * cen - tri[0] - tri[1]
    _one_element(o, tri)
    gx, gy, gz = o.ft_gradient()
    area_vec = -np.cross(tri[1] - tri[0], tri[2] - tri[0]) / 2.0
    comp = o.bc == 0
    G = np.stack([gx, gy, gz], 1)
    assert np.allclose(G[comp].sum(0) * dx ** 3, area_vec, rtol=0, atol=1e-14)
    buff = np.flatnonzero(o.bc == 1)
    img = _periodic_fill(o, np.arange(o.n, dtype=float)).astype(np.int64)
    assert np.array_equal(G[buff], G[img[buff]])


def _reference_matrix(o):
    """front_tracking_matrix (fronttrack.cu:27-238) as an explicit CSR matrix
    with the Jacobi scaling, for the slab (CapsuleSHEARFLOW) connectivity."""
    n = o.n
    mn = PO.lib().orc_expand2pow2(n)
    dx = o.P.dx
    nx, nxy = o.nx, o.nx * o.ny
    img = _periodic_fill(o, np.arange(n, dtype=float)).astype(np.int64)
    rows, cols, vals = [], [], []
    for i in range(mn):
        if i >= n or o.bc[i] == 2:
            rows.append(i); cols.append(i); vals.append(1.0)
        elif o.bc[i] == 0:
            for j, v in ((i - nxy, 1), (i - nx, 1), (i - 1, 1), (i, -6), (i + 1, 1), (i + nx, 1), (i + nxy, 1)):
                rows.append(i); cols.append(j); vals.append((v / dx ** 2) / (-6.0 / dx ** 2))
        else:
            rows += [i, i]; cols += [i, int(img[i])]; vals += [1.0, -1.0]
    return scipy.sparse.csr_matrix((vals, (rows, cols)), shape=(mn, mn))


def test_red_black_sor_reaches_the_direct_solution():
    o = _oracle(dx=0.5, sor_itrmax=20000, sor_residual=1e-12)
    rng = np.random.default_rng(5)
    b = np.where(o.bc == 0, rng.standard_normal(o.n), 0.0)
    rc, x, its = o.ft_sor(b, np.zeros(o.n))
    assert rc == 0 and its > 10
    A = _reference_matrix(o)
    mn = A.shape[0]
    xd = scipy.sparse.linalg.spsolve(A.tocsc(), np.r_[b, np.zeros(mn - o.n)])
    assert rel(x, xd[:o.n]) < 1e-9
    # a start from x = 3 (the retry guess) converges to the same solution
    rc, x3, _ = o.ft_sor(b, np.full(o.n, 3.0), 3.0)
    assert rc == 0 and rel(x3, xd[:o.n]) < 1e-9
    # ITRMAX reached: flag 2
    o.P.sor_itrmax = 5
    assert o.ft_sor(b, np.zeros(o.n))[0] == 2


def test_front_tracking_indicator_of_a_sphere():
    R = 3.6
    o = oracle_of(24, 20, 22, cells=_sphere((0.37, -0.61, 0.23), R), sor_itrmax=6000, sor_residual=1e-9)
    rc, its, att = o.front_tracking()
    assert rc == 0 and att == 1 and its > 100
    X, Y, Z = o.node_coords()
    r = np.sqrt((X - 0.37) ** 2 + (Y + 0.61) ** 2 + (Z - 0.23) ** 2)
    comp = o.bc == 0
    assert np.all(np.abs(o.vf[comp & (r < R - 2.2)] - 1.0) < 0.03)
    assert np.all(np.abs(o.vf[comp & (r > R + 2.2)]) < 0.03)
    c = o.cells
    p = np.stack([c["x"], c["y"], c["z"]], 1)
    t = c["ele"].reshape(-1, 3)
    V = np.einsum("ij,ij->i", p[t[:, 0]], np.cross(p[t[:, 1]], p[t[:, 2]])).sum() / 6.0
    assert abs(o.vf[comp].sum() - V) < 0.01 * V
    # ITRMAX too small for every start: error 5 after the eleventh attempt
    o.P.sor_itrmax = 3
    rc, _, att = o.front_tracking()
    assert rc == 5 and att == 11


# ---- GPU: the library against the oracle ------------------------------------

GL = (24, 20, 22)
VOF_KW = dict(sor_itrmax=6000, sor_residual=1e-9)


def _two_cells():
    a = _sphere((0.37, -0.61, 0.23), 3.6)
    # one mesh for both (the file layouts assume it): the second copy just
    # inside the +x and -z ends, so that its delta stencils reach the halo
    shift = np.array([GL[0] / 2 - 3.9, 0.4, -GL[2] / 2 + 3.85])
    out = dict(voff=np.array([0, len(a["x"]), 2 * len(a["x"])], np.int32),
               eoff=np.array([0, len(a["ele"]) // 3, 2 * (len(a["ele"]) // 3)], np.int32),
               ele=np.r_[a["ele"], a["ele"] + len(a["x"])].astype(np.int32))
    for k, sh in zip("xyz", shift):
        out[k] = np.r_[a[k], a[k] - a[k].mean() + sh]
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_front_tracking_matches_the_oracle(prec):
    print(f"[front tracking] prec={prec}", flush=True)
    cells = _two_cells()
    o = oracle_of(*GL, cells=cells, **VOF_KW)
    g = gpu_of(*GL, prec, cells=cells, vof_period=4, **VOF_KW)
    rc, its, att = o.front_tracking()
    assert rc == 0
    gits, gatt = g.front_tracking()
    assert gatt == att and abs(gits - its) <= 1, (gits, its)
    vo, vg = o.vf, g.vf()
    keep = o.bc != 2
    assert rel(vg[keep], vo[keep]) < 1e-8
    assert not vg[o.bc == 2].any()
    # the same again: deterministic
    g.front_tracking()
    assert np.array_equal(g.vf(), vg)


@pytest.mark.gpu
def test_front_tracking_failure_after_ten_starts():
    from preprocess_src_amd import native as N
    g = gpu_of(*GL, "f64", cells=_two_cells(), vof_period=4, sor_itrmax=2, sor_residual=1e-9)
    with pytest.raises(N.LbmibmError) as e:
        g.front_tracking()
    assert e.value.code == 8


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_thinc_wlic_matches_the_oracle(prec):
    print(f"[thinc] prec={prec}", flush=True)
    cells = _two_cells()
    o = oracle_of(*GL, cells=cells, m=2, **VOF_KW)
    g = gpu_of(*GL, prec, cells=cells, m=2, vof_period=4, **VOF_KW)
    global LX, LY, LZ
    saved = (LX, LY, LZ)
    LX, LY, LZ = GL
    try:
        vf = _smooth_field(o, 7, -0.02, 1.02)
        vel = _velocity(o, 8, amp=0.15)
    finally:
        LX, LY, LZ = saved
    if prec == "f32":  # the library keeps um..wm in the lattice precision
        vel = [v.astype(np.float32).astype(np.float64) for v in vel]
    o.vf[:] = vf
    o.um[:], o.vm[:], o.wm[:] = vel
    g.set_vf(vf)
    g.set_macros(*vel)
    for _ in range(2):
        o.thinc_wlic()
        g.thinc_wlic()
    keep = o.bc != 2
    assert rel(g.vf()[keep], o.vf[keep]) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("f64", 1e-7), ("f32", 2e-5)])
def test_steps_with_a_viscosity_contrast_match_the_oracle(prec, tol):
    print(f"[vof steps] prec={prec}", flush=True)
    cells = _two_cells()
    kw = dict(tau=0.9, p=5e-6, b=5e-3, lambda1=5.0, vof_period=2, **VOF_KW)
    o = oracle_of(*GL, cells=cells, **kw)
    g = gpu_of(*GL, prec, cells=cells, **kw)
    o.volume_init()
    g.volume_init()
    X, Y, Z = o.node_coords()
    vel = 2e-3 * Y / (GL[1] / 2)
    vel[o.bc == 2] = 0.0
    o.init_fluid(None, vel, None, None)
    g.set_equilibrium(None, vel, None, None)
    assert o.front_tracking()[0] == 0
    g.front_tracking()
    for it in range(1, 5):
        assert o.step(it)[0] == 0
        g.step(it)
    comp = o.bc == 0
    assert rel(g.vf()[comp], o.vf[comp]) < max(tol, 1e-6)
    fo, fg = o.fn.reshape(-1, Q)[comp], g.populations().reshape(-1, Q)[comp]
    assert rel(fg - 1.0 / np.array([3] + [18] * 6 + [36] * 12), fo - 1.0 / np.array([3] + [18] * 6 + [36] * 12)) < tol * 50
    so, sg = o.cells, g.cells_state()
    for k in ("x", "y", "z"):
        assert np.max(np.abs(sg[k] - so[k])) < tol * 10, k


@pytest.mark.gpu
def test_vof_arguments_and_states_are_checked(tmp_path):
    from preprocess_src_amd import native as N
    cells = _two_cells()
    with pytest.raises(N.LbmibmError):
        gpu_of(*GL, "f64", cells=cells, lambda1=3.0)  # a contrast needs the indicator
    with pytest.raises(N.LbmibmError):
        gpu_of(*GL, "f64", cells=cells, vof_period=2, periodic_y=True)
    with pytest.raises(N.LbmibmError):
        gpu_of(*GL, "f64", vof_period=2)  # no cells
    plain = gpu_of(*GL, "f64", cells=cells)
    with pytest.raises(N.LbmibmError):
        plain.front_tracking()
    assert not plain.vf().any()
    g = gpu_of(*GL, "f64", cells=cells, vof_period=2, **VOF_KW)
    with pytest.raises(N.LbmibmError):
        g.adjoint_reset()
        g.adj_lbm_d3q19(1)
    g.front_tracking()
    path = str(tmp_path / "restart.bin")
    g.write_restart(path, 0, 0)
    from oracle import reffiles as RF
    nc, V, E = 2, len(cells["x"]) // 2, len(cells["ele"]) // 6
    r = RF.parse_restart(open(path, "rb").read(), g.np_, nc, V, E)
    assert np.array_equal(r["vf"], g.vf()) and r["vf"].any() and not r["vf2"].any()
    h = gpu_of(*GL, "f64", cells=cells, vof_period=2, **VOF_KW)
    h.read_restart(path)
    assert np.array_equal(h.vf(), g.vf())


TUBE_VOF_CENTRES = {
    "axis": (0.4, -0.3, 0.2),
    # stencils and indicator reach the +z halo: the z-only fold, the z BUFF
    # copies of THINC and the periodic SOR rows act on vf > 0
    "z end": (0.4, -0.3, 7.0 - 3.3),
    # 0.4 lu from the wall: the THINC face closure and the WALL rows of the
    # SOR meet the interface
    "wall": (4.1, 0.3, 0.2),
}


@pytest.mark.gpu
@pytest.mark.parametrize("where", list(TUBE_VOF_CENTRES))
@pytest.mark.parametrize("prec,tol", [("f64", 1e-7), ("f32", 2e-5)])
def test_vof_in_the_tube_matches_the_oracle(prec, tol, where):
    """CapsuleCHANNELFLOW: the indicator of a cell in the tube (walls outside
    the disc, z periodic; the fold, the BUFF rows and the THINC copies in z
    only) and steps with a viscosity contrast, for a cell on the axis, one
    whose indicator crosses the z period and one next to the wall."""
    print(f"[vof tube] prec={prec} cell={where}", flush=True)
    cells = _sphere(TUBE_VOF_CENTRES[where], 3.0, level=3)
    kw = dict(tube_radius=7.5, tau=0.9, p=2e-5, e=(0.0, 0.0, 1.0), force_modulated=2, flowstart=0, b=5e-3,
              lambda1=5.0, vof_period=2, **VOF_KW)
    o = oracle_of(18, 18, 14, cells=cells, **kw)
    g = gpu_of(18, 18, 14, prec, cells=cells, **kw)
    o.volume_init()
    g.volume_init()
    assert o.front_tracking()[0] == 0
    g.front_tracking()
    keep = o.bc != 2
    assert rel(g.vf()[keep], o.vf[keep]) < 1e-8
    assert not g.vf()[o.bc == 2].any()
    comp = o.bc == 0
    # next to the wall the WALL rows hold vf = 0 and take about 13% of the volume
    assert abs(o.vf[comp].sum() / (4 / 3 * np.pi * 27) - 1) < (0.15 if where == "wall" else 0.03)
    if where == "z end":
        nz_f = o.vf.reshape(o.nz, o.ny, o.nx)
        assert nz_f[2].max() > 0.05 and nz_f[-3].max() > 0.3  # the indicator wraps through z
    if where == "wall":
        r = np.hypot(*o.node_coords()[:2])
        assert o.vf[comp & (r > 6.4)].max() > 0.05  # the interface reaches the wall's last fluid ring
    for it in range(1, 5):
        assert o.step(it)[0] == 0
        g.step(it)
    assert rel(g.vf()[comp], o.vf[comp]) < max(tol, 1e-6)
    w = 1.0 / np.array([3] + [18] * 6 + [36] * 12)
    fo, fg = o.fn.reshape(-1, Q)[comp] - w, g.populations().reshape(-1, Q)[comp] - w
    assert rel(fg, fo) < tol * 50
