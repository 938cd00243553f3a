"""Reverse mode: stage VJPs against central finite differences of the GPU
forward in FP64, the checkpointed wall-shear gradient against finite
differences of the FP64 oracle and of the GPU forward, the FP32 gradient
against the FP64 one, and the torch.autograd.Function wrapper.

Every check is a directional one: for random seeds wbar on the outputs and a
random direction d on the inputs, <VJP(wbar), d> must equal
(L(x + eps d) - L(x - eps d)) / (2 eps) with L = <wbar, outputs>.
Tolerances are relative to the magnitude of the directional derivative.
"""
from __future__ import annotations

import numpy as np
import pytest

from preprocess_src_amd import native as N
from tests.pairs import gpu_of, oracle_of, random_populations, small_cells

pytestmark = pytest.mark.gpu

Q = 19


def comp(g, a, k=1):
    """Zero every non-COMP entry of a padded field of k components per node."""
    m = g.comp_mask()
    a = a.reshape(-1, g.np_)
    return (a * m[None, :]).reshape(-1) if k == 1 else a


def comp_pop(g, a):
    m = np.repeat(g.comp_mask(), Q)
    return a * m


def fd_check(analytic, f_plus, f_minus, eps, tol):
    fd = (f_plus - f_minus) / (2 * eps)
    scale = max(abs(fd), abs(analytic), 1e-300)
    assert abs(analytic - fd) / scale < tol, (analytic, fd)


@pytest.mark.parametrize("walls", [True, False])
def test_lbm_vjp(walls):
    lx, ly, lz = 8, 6, 8
    kw = dict(tau=0.8, p=2e-5, b=0.7, periodic_y=not walls, store_macros=True, volume_correction=False)
    o = oracle_of(lx, ly, lz, **kw)
    g = gpu_of(lx, ly, lz, "f64", **kw)
    rng = np.random.default_rng(7)
    fn0 = random_populations(o, seed=8)
    F0 = [comp(g, 1e-3 * rng.standard_normal(g.np_)) for _ in range(3)]
    H0 = [comp(g, 1e-4 * rng.standard_normal(g.np_)) for _ in range(3)]

    def outputs(fn, F, H):
        g.set_populations(fn)
        g.set_lattice_force(*F)
        g.set_body_force_field(*H)
        g.lbm_d3q19(1)
        m = g.macros()
        return g.populations(), np.concatenate([m["um"], m["vm"], m["wm"]])

    pb = comp_pop(g, rng.standard_normal(g.np_ * Q))
    ub = np.concatenate([comp(g, rng.standard_normal(g.np_)) for _ in range(3)])
    g.set_populations(fn0)
    g.set_lattice_force(*F0)
    g.set_body_force_field(*H0)
    g.adjoint_reset()
    g.set_adjoint(N.ADJ_POPULATIONS, pb)
    g.set_adjoint(N.ADJ_VELOCITY, ub)
    g.adj_lbm_d3q19(1)
    gin = g.adjoint(N.ADJ_POPULATIONS)
    Fb = g.adjoint(N.ADJ_LATTICE_FORCE)
    Hb = g.adjoint(N.ADJ_GRAD_BODY_FORCE)

    def L(fn, F, H):
        p, u = outputs(fn, F, H)
        return float(pb @ p + ub @ u)

    df = comp_pop(g, 1e-3 * rng.standard_normal(g.np_ * Q))
    dF = [comp(g, rng.standard_normal(g.np_)) for _ in range(3)]
    dH = [comp(g, rng.standard_normal(g.np_)) for _ in range(3)]
    eps = 1e-5
    fd_check(float(gin @ df), L(fn0 + eps * df, F0, H0), L(fn0 - eps * df, F0, H0), eps, 1e-6)
    fd_check(float(Fb @ np.concatenate(dF)),
             L(fn0, [a + eps * d for a, d in zip(F0, dF)], H0), L(fn0, [a - eps * d for a, d in zip(F0, dF)], H0),
             eps, 1e-6)
    fd_check(float(Hb @ np.concatenate(dH)),
             L(fn0, F0, [a + eps * d for a, d in zip(H0, dH)]), L(fn0, F0, [a - eps * d for a, d in zip(H0, dH)]),
             eps, 1e-6)


def _cell_case(seed, levels=(2, 2), lx=16, ly=14, lz=16, **kw):
    cells = small_cells(lx, ly, lz, nc=len(levels), radius=3.0, levels=list(levels), seed=seed)
    base = dict(tau=0.8, b=0.5, c=10.0, residual=1e-13, volume_correction=False)
    base.update(kw)
    g = gpu_of(lx, ly, lz, "f64", cells=cells, **base)
    return g, cells


def _set_positions(g, X):
    x, y, z = X.reshape(3, -1)
    g.set_cells_state(x=x, y=y, z=z, xd=np.stack([x, y, z], 1).reshape(-1).copy())


def test_spread_vjp():
    g, cells = _cell_case(21)
    rng = np.random.default_rng(22)
    X0 = np.concatenate([cells["x"], cells["y"], cells["z"]]) + 0.05 * rng.standard_normal(3 * g.nv)
    f0 = 1e-2 * rng.standard_normal(3 * g.nv)

    def F_of(X, f):
        _set_positions(g, X)
        g.set_cells_state(fx=f[:g.nv], fy=f[g.nv:2 * g.nv], fz=f[2 * g.nv:])
        g.ibm_st1()
        return np.concatenate([comp(g, a) for a in g.lattice_force()])

    Fb = np.concatenate([comp(g, rng.standard_normal(g.np_)) for _ in range(3)])
    F_of(X0, f0)
    g.adjoint_reset()
    g.set_adjoint(N.ADJ_LATTICE_FORCE, Fb)
    g.adj_ibm_st1()
    fb = g.adjoint(N.ADJ_NODAL_FORCE)
    xb = g.adjoint(N.ADJ_POSITIONS)
    L = lambda X, f: float(Fb @ F_of(X, f))
    dX, df = rng.standard_normal(3 * g.nv), rng.standard_normal(3 * g.nv)
    eps = 1e-6
    fd_check(float(fb @ df), L(X0, f0 + eps * df), L(X0, f0 - eps * df), eps, 1e-8)
    fd_check(float(xb @ dX), L(X0 + eps * dX, f0), L(X0 - eps * dX, f0), eps, 1e-5)


def test_advection_vjp():
    g, cells = _cell_case(31)
    rng = np.random.default_rng(32)
    X0 = np.concatenate([cells["x"], cells["y"], cells["z"]]) + 0.05 * rng.standard_normal(3 * g.nv)
    U0 = [comp(g, 0.05 * rng.standard_normal(g.np_)) for _ in range(3)]

    def X_of(X, U):
        _set_positions(g, X)
        g.set_macros(*U)
        g.ibm_st2()
        s = g.cells_state()
        return np.concatenate([s["x"], s["y"], s["z"]])

    xbo = rng.standard_normal(3 * g.nv)
    X_of(X0, U0)  # leaves the state after the advection; restore the inputs
    _set_positions(g, X0)
    g.set_macros(*U0)
    g.adjoint_reset()
    g.set_adjoint(N.ADJ_POSITIONS, xbo)
    g.adj_ibm_st2()
    xb = g.adjoint(N.ADJ_POSITIONS)
    ub = g.adjoint(N.ADJ_VELOCITY)
    L = lambda X, U: float(xbo @ X_of(X, U))
    dX = rng.standard_normal(3 * g.nv)
    dU = [comp(g, rng.standard_normal(g.np_)) for _ in range(3)]
    eps = 1e-6
    fd_check(float(xb @ dX), L(X0 + eps * dX, U0), L(X0 - eps * dX, U0), eps, 1e-6)
    fd_check(float(ub @ np.concatenate(dU)), L(X0, [a + eps * d for a, d in zip(U0, dU)]),
             L(X0, [a - eps * d for a, d in zip(U0, dU)]), eps, 1e-6)


@pytest.mark.parametrize("neo", [False, True])
def test_stress_vjp(neo):
    g, cells = _cell_case(41, levels=(2, 3), neohookean=neo)
    rng = np.random.default_rng(42)
    X0 = np.concatenate([cells["x"], cells["y"], cells["z"]]) + 0.03 * rng.standard_normal(3 * g.nv)
    s0 = np.array([1.0, 0.7])

    def f_of(X, sc):
        _set_positions(g, X)
        g.set_cell_force_scale(sc)
        g.stress(1)
        s = g.cells_state()
        return np.concatenate([s["fx"], s["fy"], s["fz"]])

    fbar = rng.standard_normal(3 * g.nv)
    f_of(X0, s0)
    g.adjoint_reset()
    g.set_adjoint(N.ADJ_NODAL_FORCE, fbar)
    g.adj_stress(1)
    xb = g.adjoint(N.ADJ_POSITIONS)
    sb = g.adjoint(N.ADJ_GRAD_FORCE_SCALE)
    L = lambda X, sc: float(fbar @ f_of(X, sc))
    dX = rng.standard_normal(3 * g.nv)
    eps = 1e-6
    fd_check(float(xb @ dX), L(X0 + eps * dX, s0), L(X0 - eps * dX, s0), eps, 1e-5)
    ds = np.array([0.3, -0.8])
    fd_check(float(sb @ ds), L(X0, s0 + eps * ds), L(X0, s0 - eps * ds), eps, 1e-7)


# ---- whole runs -----------------------------------------------------------------

CASE = dict(lx=16, ly=12, lz=14, nsteps=12, window=6)


def _run_case(prec, seed=51):
    lx, ly, lz = CASE["lx"], CASE["ly"], CASE["lz"]
    cells = small_cells(lx, ly, lz, nc=2, radius=3.0, levels=[2, 2], seed=seed)
    kw = dict(tau=0.8, p=1e-4, b=5e-2, c=10.0, residual=1e-13 if prec == "f64" else 1e-9,
              volume_correction=False)
    return cells, kw, gpu_of(lx, ly, lz, prec, cells=cells, **kw)


def _reset(g, cells):
    g.set_equilibrium()
    _set_positions(g, np.concatenate([cells["x"], cells["y"], cells["z"]]))


def _oracle_J(cells, kw, H, scale):
    lx, ly, lz = CASE["lx"], CASE["ly"], CASE["lz"]
    c = dict(cells)
    c["gs_scale"] = scale
    o = oracle_of(lx, ly, lz, cells=c, **kw)
    o.set_body_force_field(*H)
    return o.wall_shear(0, CASE["nsteps"], CASE["window"])


def test_wall_shear_gradient_matches_oracle_fd():
    """FP64 adjoint vs central differences of the FP64 oracle (reference gradient)."""
    cells, kw, g = _run_case("f64")
    n, w = CASE["nsteps"], CASE["window"]
    rng = np.random.default_rng(52)
    zero = [np.zeros(g.np_) for _ in range(3)]
    _reset(g, cells)
    J, gH, gs = g.adjoint_wall_shear(0, n, w, 4)
    Jo = _oracle_J(cells, kw, zero, np.ones(2))
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    dH = [comp(g, 1e-5 * rng.standard_normal(g.np_)) for _ in range(3)]
    eps = 1.0
    Jp = _oracle_J(cells, kw, [eps * d for d in dH], np.ones(2))
    Jm = _oracle_J(cells, kw, [-eps * d for d in dH], np.ones(2))
    fd_check(float(sum(a @ d for a, d in zip(gH, dH))), Jp, Jm, eps, 1e-6)
    ds = np.array([0.5, -0.25])
    eps = 1e-4
    fd_check(float(gs @ ds), _oracle_J(cells, kw, zero, 1 + eps * ds), _oracle_J(cells, kw, zero, 1 - eps * ds),
             eps, 1e-5)


def test_wall_shear_gradient_matches_gpu_fd_and_checkpoint_interval():
    cells, kw, g = _run_case("f64")
    n, w = CASE["nsteps"], CASE["window"]
    rng = np.random.default_rng(53)
    _reset(g, cells)
    J1, gH1, gs1 = g.adjoint_wall_shear(0, n, w, 4)
    _reset(g, cells)
    J2, gH2, gs2 = g.adjoint_wall_shear(0, n, w, 5)  # uneven last segment
    assert abs(J1 - J2) <= 1e-13 * abs(J1)  # the objective sums with atomics: order-dependent last bits
    for a, b in zip(gH1, gH2):
        assert np.max(np.abs(a - b)) <= 1e-10 * np.max(np.abs(a))
    assert np.max(np.abs(gs1 - gs2)) <= 1e-10 * np.max(np.abs(gs1))
    dH = [comp(g, 1e-5 * rng.standard_normal(g.np_)) for _ in range(3)]

    def J_of(H):
        _reset(g, cells)
        g.set_body_force_field(*H)
        return g.wall_shear(0, n, w)

    fd_check(float(sum(a @ d for a, d in zip(gH1, dH))), J_of(dH), J_of([-d for d in dH]), 1.0, 1e-6)


def test_fp32_gradient_matches_fp64():
    cells, _, g64 = _run_case("f64")
    _, _, g32 = _run_case("f32")
    n, w = CASE["nsteps"], CASE["window"]
    _reset(g64, cells)
    _reset(g32, cells)
    J64, gH64, gs64 = g64.adjoint_wall_shear(0, n, w, 4)
    J32, gH32, gs32 = g32.adjoint_wall_shear(0, n, w, 4)
    assert abs(J32 - J64) <= 1e-5 * abs(J64)
    a, b = np.concatenate(gH32), np.concatenate(gH64)
    assert np.max(np.abs(a - b)) <= 1e-3 * np.max(np.abs(b))
    assert np.max(np.abs(gs32 - gs64)) <= 1e-3 * np.max(np.abs(gs64))


def test_autograd_function():
    torch = pytest.importorskip("torch")
    from tests.adjoint_fn import WallShear

    cells, kw, g = _run_case("f64")
    n, w = CASE["nsteps"], CASE["window"]
    H = torch.zeros(3, g.np_, dtype=torch.float64, requires_grad=True)
    s = torch.ones(2, dtype=torch.float64, requires_grad=True)
    reset = lambda: _reset(g, cells)
    J = WallShear.apply(H, s, g, reset, n, w, 4)
    (2.0 * J).backward()
    rng = np.random.default_rng(54)
    dH = np.stack([comp(g, 1e-5 * rng.standard_normal(g.np_)) for _ in range(3)])
    with torch.no_grad():
        Jp = WallShear.apply(torch.from_numpy(dH), s, g, reset, n, w, 4)
        Jm = WallShear.apply(torch.from_numpy(-dH), s, g, reset, n, w, 4)
    fd_check(float((H.grad.numpy() * dH).sum()), 2 * float(Jp), 2 * float(Jm), 1.0, 1e-6)
    assert s.grad is not None and s.grad.shape == (2,)
