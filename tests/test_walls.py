"""Sliding walls and the staircase tube on the GPU against the oracle.

- lbm_CapsuleSHEARFLOW's walls moving along z with Ladd's term
  (lbm.cu:951-970), both walls (allwall) or the top one (topwall), at rest
  before FLOWSTART (lbm.cu:755);
- lbm_SHEARFLOW (lbm.cu:282-477): both walls at ub = 1 from its sheared start
  (init1.cu:319-326);
- lbm_CapsuleCHANNELFLOW's tube with the half-way bounce-back on the mask
  (lbm.cu:1050-1094) and the body force e(t) = 1 + 0.5 sin(pi t/FREQ) after
  FLOWSTART (lbm.cu:1032-1035).

Errors are relative to each field's own scale: population deviations
f - w_q against max |f - w_q|, each velocity (instantaneous, M-averaged) as
one vector field against its largest component.  FP64 is
held to 1e-11 and FP32 to 1e-5; the FP16-storage error is printed per field
and held below 1% of the field.  The closed-form Couette pins of the same
terms are in tests/test_oracle.py.
"""
import numpy as np
import pytest

from tests.pairs import Q, gpu_of, oracle_of, pop_err, random_lattice_force, random_populations, rel
from tests.test_oracle import _couette_series

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-11, "f32": 1e-5, "f16": 1e-2}


def _fields(o, g):
    comp = o.bc == 0
    out = {"f": pop_err(g.populations().reshape(-1, Q)[comp], o.fn.reshape(-1, Q)[comp])}
    m = g.macros(instantaneous=True)
    # each velocity as one vector field: in the tube the transverse components
    # are set by the cell alone and are orders of magnitude below the axial one
    for name, ks in (("u", ("un", "vn", "wn")), ("u averaged", ("um", "vm", "wm"))):
        out[name] = rel(np.concatenate([m[k][comp] for k in ks]), np.concatenate([getattr(o, k)[comp] for k in ks]))
    # rho is kept as rho (not rho - 1) in the arithmetic type, so its own
    # scale is 1; rho - 1 is carried by the population deviations above
    out["rho"] = rel(m["dn"][comp], o.dn[comp])
    return out


@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
@pytest.mark.parametrize("mode,flowstart,modulated", [(1, 0, 0), (2, 0, 0), (1, 5, 0), (0, 3, 1)])
def test_sliding_wall_steps_match_the_oracle(prec, mode, flowstart, modulated):
    lx, ly, lz = 12, 10, 8
    kw = dict(tau=0.7, cr=0.6, p=2e-5, e=(0.6, 0.0, 0.8), b=0.3, wall_motion=mode, wall_velocity=0.03,
              flowstart=flowstart, force_modulated=modulated, freq=5.0)
    print(f"[sliding walls] prec={prec} mode={mode} flowstart={flowstart} modulated={modulated}", flush=True)
    o = oracle_of(lx, ly, lz, **kw)
    g = gpu_of(lx, ly, lz, prec, **kw)
    fn = random_populations(o, seed=3)
    g.set_populations(fn)
    F = random_lattice_force(o, seed=4)
    o.fx[:], o.fy[:], o.fz[:] = F
    g.set_lattice_force(*F)
    for it in range(1, 13):
        o.lbm_d3q19(it)
        g.lbm_d3q19(it)
    err = _fields(o, g)
    print(f"  errors relative to each field: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    for k, v in err.items():
        assert v < TOL[prec], (k, v)


@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
@pytest.mark.parametrize("mode", [1, 2])
def test_couette_start_up_on_the_gpu(prec, mode):
    """600 graph-replayed steps from rest between sliding walls: u_z(y) against
    the start-up series (1% of U) and against the FP64 oracle."""
    H, tau, cr, U = 24, 0.8, 0.5, 0.04
    kw = dict(tau=tau, cr=cr, wall_motion=mode, wall_velocity=U, volume_correction=False)
    print(f"[couette] prec={prec} mode={mode}", flush=True)
    g = gpu_of(4, H, 4, prec, **kw)
    g.set_equilibrium()
    g.run(0, 600, graph=True)
    uz = g.macros()["wm"].reshape(8, H + 4, 8)[2:-2, 2:-2, 2:-2]
    assert np.ptp(uz, axis=(0, 2)).max() <= 1e-6 * U
    nu = (tau - 0.5) / 3
    ua = _couette_series(np.arange(H) + 0.5, 600, U, nu, H, top_only=(mode == 2))
    assert np.max(np.abs(uz[0, :, 0] - ua)) < 1e-2 * U
    o = oracle_of(4, H, 4, **kw)
    for it in range(1, 601):
        o.lbm_d3q19(it)
    uo = o.wm.reshape(o.nz, o.ny, o.nx)[2, 2:-2, 2]
    e = np.max(np.abs(uz[0, :, 0] - uo)) / np.abs(uo).max()
    print(f"  u_z error against the oracle after 600 steps: {e:.2e} of max |u_z|", flush=True)
    assert e < {"f64": 1e-11, "f32": 1e-5, "f16": 1e-2}[prec]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_shearflow_keeps_its_linear_start(prec):
    """lbm_SHEARFLOW: both walls at ub = 1 (lattice speed Cr), no force, from
    the sheared start u_z = y/(LY/2) of flu_init (init1.cu:319-326): plane
    Couette flow is exact for half-way bounce-back with Ladd's term, so the
    profile stays linear; checked against the oracle step by step."""
    H, cr = 10, 0.02
    kw = dict(tau=0.8, cr=cr, wall_motion=1, wall_velocity=1.0, volume_correction=False)
    o = oracle_of(6, H, 4, **kw)
    g = gpu_of(6, H, 4, prec, **kw)
    _, Y, _ = o.node_coords()
    w = Y / (H / 2)
    o.init_fluid(None, None, None, w)
    g.set_equilibrium(None, None, None, w)
    for it in range(1, 301):
        o.lbm_d3q19(it)
    g.run(0, 300, graph=True)
    comp = o.bc == 0
    wm = g.macros()["wm"]
    assert rel(wm[comp], o.wm[comp]) < TOL[prec]
    assert np.max(np.abs(o.wm[comp] - w[comp])) < 1e-12


TUBE_KW = dict(tube_radius=5.5, tube_staircase=True, tau=0.9, p=2e-5, e=(0.0, 0.0, 1.0), force_modulated=1,
               flowstart=2, freq=9.0, b=5e-3)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_staircase_tube_steps_with_a_cell_match_the_oracle(prec):
    from tests.test_tube import _tube_cell
    print(f"[staircase tube steps] prec={prec}", flush=True)
    cells = _tube_cell()
    o = oracle_of(14, 14, 8, cells=cells, **TUBE_KW)
    g = gpu_of(14, 14, 8, prec, cells=cells, **TUBE_KW)
    o.volume_init()
    g.volume_init()
    assert np.array_equal(g.bc_mask(), o.bc)
    for it in range(1, 25):
        assert o.step(it)[0] == 0
        g.step(it)
    err = _fields(o, g)
    so, sg = o.cells, g.cells_state()
    err["x"] = max(np.max(np.abs(sg[k] - so[k])) for k in "xyz") / 5.5
    print(f"  errors relative to each field: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    for k, v in err.items():
        assert v < TOL[prec], (k, v)
    assert np.abs(o.wm[o.bc == 0]).max() > 1e-6  # the flow started at FLOWSTART


@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
def test_staircase_tube_lattice_matches_the_oracle(prec):
    kw = {k: v for k, v in TUBE_KW.items() if k != "b"}
    print(f"[staircase tube lattice] prec={prec}", flush=True)
    o = oracle_of(14, 14, 8, **kw)
    g = gpu_of(14, 14, 8, prec, **kw)
    fn = random_populations(o, seed=5)
    fn.reshape(-1, Q)[o.bc == 2] = 0.0
    g.set_populations(fn)
    for it in range(1, 21):
        o.lbm_d3q19(it)
        g.lbm_d3q19(it)
    err = _fields(o, g)
    print(f"  errors relative to each field: " + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    for k, v in err.items():
        assert v < TOL[prec], (k, v)


def test_sliding_wall_arguments_are_checked():
    from preprocess_src_amd import native as N
    with pytest.raises(N.LbmibmError):
        gpu_of(8, 8, 8, "f64", wall_motion=1, wall_velocity=0.01, periodic_y=True)
    with pytest.raises(N.LbmibmError):
        gpu_of(14, 14, 8, "f64", wall_motion=1, wall_velocity=0.01, tube_radius=5.5)
    with pytest.raises(N.LbmibmError):
        gpu_of(8, 8, 8, "f64", wall_motion=3)
    g = gpu_of(8, 8, 8, "f64", wall_motion=1, wall_velocity=0.01)
    with pytest.raises(N.LbmibmError):
        g.adjoint_reset()
        g.adj_lbm_d3q19(1)
