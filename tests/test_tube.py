"""The tube geometry of CapsuleCHANNELFLOW: the mask (init1.cu:190-236), the
wall-crossing fractions (lbm_cylinder.cu:118-190) and the interpolated
bounce-back (lbm_cylinder.cu:61-85, 193-235) with the z-periodic push.

CPU: the oracle's mask is the closed-form disc, its fractions match the
chord of the circle found independently (closed form along the axes,
bisection along the diagonals), and a body-force-driven flow reaches the
Poiseuille profile u = g (R^2 - r^2) / (4 nu).  GPU: the library's tube
against the oracle, steps with cells inside the tube included.
"""
import numpy as np
import pytest

from oracle import pyoracle as PO
from tests.pairs import Q, gpu_of, oracle_of, rel

CXYZ = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                 [1, 1, 0], [-1, -1, 0], [-1, 1, 0], [1, -1, 0], [0, 1, 1], [0, -1, -1], [0, -1, 1],
                 [0, 1, -1], [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1]])


def _tube(lx=14, ly=14, lz=4, R=5.5, **kw):
    return oracle_of(lx, ly, lz, tube_radius=R, **kw)


def test_tube_mask_is_the_disc():
    o = _tube(dx=0.5, R=2.6)
    X, Y, Z = o.node_coords()
    inside = np.sqrt(X ** 2 + Y ** 2) <= 2.6
    zin = np.abs(Z) < o.P.lz / 2
    assert np.array_equal(o.bc == 0, inside & zin)
    assert np.array_equal(o.bc == 1, inside & ~zin)
    assert np.array_equal(o.bc == 2, ~inside)


def test_tube_fractions_are_the_chord_of_the_circle():
    R = 5.5
    o = _tube(R=R)
    X, Y, Z = o.node_coords()
    checked = 0
    for i in np.flatnonzero(o.bc == 0):
        for q in range(1, 7 + 4):  # axis and xy-diagonal links
            c = CXYZ[q]
            j = i + c[0] + o.nx * c[1] + o.nx * o.ny * c[2]
            if o.bc[j] != 2:
                continue
            r = PO.lib().orc_tube_fraction(o.P, int(i), q)
            p0, d = np.array([X[i], Y[i]]), c[:2].astype(float)
            if c[0] == 0 or c[1] == 0:  # along an axis: sqrt(R^2 - other^2)
                k = 0 if c[0] else 1
                w = np.sqrt(R * R - p0[1 - k] ** 2) * np.sign(d[k])
                expect = (w - p0[k]) / d[k]
            else:  # bisection on |p0 + t d| = R
                lo, hi = 0.0, 1.0
                for _ in range(80):
                    mid = 0.5 * (lo + hi)
                    lo, hi = (mid, hi) if np.hypot(*(p0 + mid * d)) <= R else (lo, mid)
                expect = 0.5 * (lo + hi)
            assert abs(r - expect) < 1e-12, (i, q, r, expect)
            assert 0.0 < r < 1.0
            checked += 1
    assert checked > 50


def test_tube_flow_reaches_poiseuille():
    R, P, tau = 5.5, 1e-5, 1.0
    o = _tube(R=R, tau=tau, p=P, e=(0.0, 0.0, 1.0), force_modulated=2, flowstart=0)
    for it in range(1, 1801):
        o.lbm_d3q19(it)
    X, Y, Z = o.node_coords()
    comp = o.bc == 0
    r2 = X[comp] ** 2 + Y[comp] ** 2
    nu = (tau - 0.5) / 3.0
    exact = 2.0 * P / (4.0 * nu) * (R * R - r2)  # Guo force 2 rho P e
    w = o.wm[comp]
    assert np.max(np.abs(o.um[comp])) < 1e-3 * w.max()
    assert np.sqrt(np.mean((w - exact) ** 2)) / exact.max() < 0.01
    assert abs(w.max() / exact.max() - 1.0) < 0.01


# ---- GPU: the library's tube against the oracle ------------------------------

def _tube_cell(R=2.2):
    from preprocess_src_amd import mesh
    return mesh.assemble([2], np.array([[0.31, -0.22, 0.4]]), np.eye(3)[None], R, shape="biconcave")


TUBE_KW = dict(tube_radius=5.5, tau=0.9, p=2e-5, e=(0.0, 0.0, 1.0), force_modulated=2, flowstart=2, b=5e-3)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("f64", 1e-11), ("f32", 2e-5)])
def test_tube_steps_match_the_oracle(prec, tol):
    print(f"[tube steps] prec={prec}", flush=True)
    cells = _tube_cell()
    o = oracle_of(14, 14, 8, cells=cells, **TUBE_KW)
    g = gpu_of(14, 14, 8, prec, cells=cells, **TUBE_KW)
    o.volume_init()
    g.volume_init()
    assert np.array_equal(g.bc_mask(), o.bc)
    for it in range(1, 25):
        assert o.step(it)[0] == 0
        g.step(it)
    comp = o.bc == 0
    m = g.macros()
    for k in ("um", "vm", "wm"):
        assert rel(m[k][comp], getattr(o, k)[comp]) < tol * 10, k
    w = 1.0 / np.array([3] + [18] * 6 + [36] * 12)
    fo, fg = o.fn.reshape(-1, Q)[comp] - w, g.populations().reshape(-1, Q)[comp] - w
    assert rel(fg, fo) < tol * 10
    so, sg = o.cells, g.cells_state()
    for k in ("x", "y", "z"):
        assert np.max(np.abs(sg[k] - so[k])) < tol * 100, k
    assert np.abs(o.wm[comp]).max() > 1e-6  # the flow started at FLOWSTART


@pytest.mark.gpu
def test_tube_flow_reaches_poiseuille_on_the_gpu():
    R, P, tau = 10.5, 4e-6, 1.0
    g = gpu_of(24, 24, 4, "f32", tube_radius=R, tau=tau, p=P, e=(0.0, 0.0, 1.0), force_modulated=2, flowstart=0)
    g.run(0, 5000, graph=True)
    m = g.macros()
    comp = g.bc_mask() == 0
    nx, ny = 28, 28
    ix = np.arange(g.np_) % nx
    iy = (np.arange(g.np_) // nx) % ny
    x, y = -12 - 1.5 + ix, -12 - 1.5 + iy
    r2 = (x * x + y * y)[comp]
    exact = 2.0 * P / (4.0 * (tau - 0.5) / 3.0) * (R * R - r2)
    w = m["wm"][comp]
    assert np.sqrt(np.mean((w - exact) ** 2)) / exact.max() < 0.01
    assert abs(w.max() / exact.max() - 1.0) < 0.01


@pytest.mark.gpu
def test_tube_arguments_are_checked():
    from preprocess_src_amd import native as N
    with pytest.raises(N.LbmibmError):
        gpu_of(14, 14, 8, "f64", tube_radius=6.6)  # the halo columns would be fluid
    with pytest.raises(N.LbmibmError):
        gpu_of(14, 14, 8, "f64", tube_radius=5.5, periodic_y=True)
    g = gpu_of(14, 14, 8, "f64", cells=_tube_cell(), tube_radius=5.5)
    with pytest.raises(N.LbmibmError):
        g.adjoint_reset()
        g.adj_lbm_d3q19(1)


@pytest.mark.gpu
def test_tube_fp16_storage_matches_the_oracle():
    """FP16 population storage (deviations scaled by 1/w) through the tube's
    interpolated bounce-back.  Errors per field against the FP64 oracle,
    each relative to that field's own scale: the population deviations
    f - w_q and the axial velocity (a run left at rest would score 1)."""
    print("[tube f16]", flush=True)
    kw = dict(tube_radius=5.5, tau=0.8, p=2e-5, e=(0.0, 0.0, 1.0), force_modulated=2, flowstart=1)
    o = oracle_of(14, 14, 8, **kw)
    g = gpu_of(14, 14, 8, "f16", **kw)
    for it in range(1, 21):
        o.lbm_d3q19(it)
        g.lbm_d3q19(it)
    comp = o.bc == 0
    w = 1.0 / np.array([3] + [18] * 6 + [36] * 12)
    fo, fg = o.fn.reshape(-1, Q)[comp] - w, g.populations().reshape(-1, Q)[comp] - w
    m = g.macros()
    err = {"f - w": rel(fg, fo), "wm": rel(m["wm"][comp], o.wm[comp])}
    print(f"[tube f16] errors {err}, max|f - w| {np.abs(fo).max():.3e}, max|wm| {np.abs(o.wm[comp]).max():.3e}",
          flush=True)
    assert np.abs(o.wm[comp]).max() > 1e-5 and np.abs(fo).max() > 1e-5
    assert err["f - w"] < 2e-2
    assert err["wm"] < 2e-2
