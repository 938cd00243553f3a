"""Matched oracle / HIP set-ups for the parity tests (test infrastructure).

Both sides receive identical parameters and initial data in the reference's
layouts; results are compared on the reference's COMP nodes.
"""
from __future__ import annotations

import numpy as np

from oracle.pyoracle import OracleSim

Q = 19


def common_params(**kw):
    p = dict(dx=1.0, cr=1.0, uc=1.0, tau=0.8, p=0.0, e=(1.0, 0.0, 0.0), b=0.0, m=1, c=100.0, br=0.0,
             itrmax=1000, residual=1e-12, runge_kutta=4, volume_correction=True, periodic_y=False)
    p.update(kw)
    return p


def oracle_of(lx, ly, lz, cells=None, **kw):
    p = common_params(**kw)
    o = OracleSim(lx * p["dx"], ly * p["dx"], lz * p["dx"], dx=p["dx"], cr=p["cr"], uc=p["uc"], tau=p["tau"], p=p["p"], e=p["e"],
                  b=p["b"], m=p["m"], c=p["c"], br=p["br"], itrmax=p["itrmax"], residual=p["residual"],
                  runge_kutta=p["runge_kutta"], volume_correction=p["volume_correction"],
                  periodic_y=p["periodic_y"], force_modulated=p.get("force_modulated", False),
                  flowstart=p.get("flowstart", 0), freq=p.get("freq", 1.0),
                  fileoutput=p.get("fileoutput", 0) or 1000, cells=cells, lambda1=p.get("lambda1", 1.0),
                  vof_period=p.get("vof_period", 0), sor_itrmax=p.get("sor_itrmax"),
                  sor_residual=p.get("sor_residual"), tube_radius=p.get("tube_radius", 0.0),
                  tube_staircase=p.get("tube_staircase", False), wall_motion=p.get("wall_motion", 0),
                  wall_velocity=p.get("wall_velocity", 0.0))
    o.init_fluid()
    return o


def gpu_of(lx, ly, lz, precision, cells=None, store_macros=True, **kw):
    from preprocess_src_amd.simulation import Simulation
    p = common_params(**kw)
    tau = p.pop("tau")
    return Simulation(lx, ly, lz, cells=cells, precision=precision, tau=tau, store_macros=store_macros, **p)


def random_populations(o, seed, amp=2e-3):
    """Equilibria of a random small (rho, u) field plus non-equilibrium noise."""
    rng = np.random.default_rng(seed)
    n = o.n
    rho = 1.0 + amp * rng.standard_normal(n)
    u, v, w = (amp * rng.standard_normal(n) for _ in range(3))
    o.init_fluid(rho, u, v, w)
    fn = o.fn
    fn += 0.1 * amp * rng.standard_normal(fn.shape) / 19.0
    return fn.copy()


def random_lattice_force(o, seed, amp=1e-3):
    rng = np.random.default_rng(seed)
    return [amp * rng.standard_normal(o.n) for _ in range(3)]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


# D3Q19 weights (lbm.cu:219-237): populations are compared as deviations
# f - w_q rho0 (rho0 = 1), the part that carries the flow, so that the rest
# weight of max|f| ~ 1/3 does not dilute the error measure
WQ = 1.0 / np.array([3.0] + [18.0] * 6 + [36.0] * 12)


def pop_err(fa, fb):
    """Max error of the population deviations of fa against fb ([n, 19]
    arrays of the same nodes), relative to max |fb - w_q|."""
    return rel(np.asarray(fa).reshape(-1, Q) - WQ, np.asarray(fb).reshape(-1, Q) - WQ)


def small_cells(lx, ly, lz, nc=2, radius=3.0, levels=None, seed=11):
    from preprocess_src_amd import mesh
    levels = levels if levels is not None else [2] * nc
    c, r = mesh.place(nc, radius, lx, ly, lz, seed)
    return mesh.assemble(levels, c, r, radius)


def perturb_cells(cells, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    out = dict(cells)
    for k in ("x", "y", "z"):
        out[k] = cells[k] + amp * rng.standard_normal(cells[k].shape)
    return out


def dev_err(a, b, base=0.0):
    """Error of a against b measured on the deviation b - base: max-abs error,
    that over max |b - base| (rel) and the L2 ratio (l2_rel)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    d = a - b
    s = b - base
    ms = float(np.max(np.abs(s))) if s.size else 0.0
    ns = float(np.linalg.norm(s))
    md = float(np.max(np.abs(d))) if d.size else 0.0
    return dict(max_abs=md, rel=md / ms if ms > 0 else md, l2_rel=float(np.linalg.norm(d)) / ns if ns > 0 else md)


def record_errors(case, errors):
    """Appends a case's per-field errors as one JSON line to the file named by
    LBMIBM_FIELD_ERRORS (nothing when unset), and prints them."""
    import json
    import os
    line = json.dumps({"case": case, "errors": errors}, default=float)
    print(line, flush=True)
    path = os.environ.get("LBMIBM_FIELD_ERRORS")
    if path:
        with open(path, "a") as f:
            f.write(line + "\n")
