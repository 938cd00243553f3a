#!/usr/bin/env python3
"""Throughput of the coupled D3Q19 LBM + IBM + Skalak membrane timestep on one
MI355X (MLUPS), with the lattice kernel's achieved HBM bandwidth against the
8 TB/s peak and the FP64 host oracle timed beside it.

Default workload: 256 x 128 x 128 slab channel (half-way bounce-back walls in
y, periodic x and z, body force 1e-6 along x), 1024 Evans-Fung cells with a
seeded mesh mix of 50% 642-, 40% 2562- and 10% 10242-vertex meshes, radius
7 lattice units on a jittered 16x8x8 grid, Skalak C=100, Gs=5e-4, FP32 with
populations stored as deviations and collided in FP64, BiCGSTAB started from
the extrapolated loads, whole steps replayed from a hipGraph; the timed
steps follow 1000 untimed steps from rest (--spinup).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--spinup S] [--precision f32|f64|f16]
                    [--workload suspension|tgv|adjoint|vof|tube|adhesion] [--cpu-seconds S]
                    [--no-cpu]

--workload adjoint times the checkpointed wall-shear gradient on a 128x64x64
channel with 32 RBCs (500 steps, last 200 averaged, checkpoints every 50)
against the forward run and checks it against a central difference.
--workload vof times front tracking (red-black SOR), THINC/WLIC and steps
with a viscosity contrast on the default suspension; --workload tube the
lattice step of the interpolated-bounce-back tube; --workload adhesion the
suspension with cytoadhesion to the lower wall.
"""
from __future__ import annotations

import argparse
import json
import re
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X
Q = 19
STORE_BYTES = {"f64": 8, "f32": 4, "f16": 2}
ARITH_BYTES = {"f64": 8, "f32": 4, "f16": 4}


def suspension_workload(nc=1024, lx=256, ly=128, lz=128, radius=7.0, seed=2024):
    from preprocess_src_amd import mesh
    levels = mesh.mesh_mix(nc, [0.5, 0.4, 0.1], [3, 4, 5], seed)
    centres, rot = mesh.place(nc, radius, lx, ly, lz, seed, mode="grid", grid=(16, 8, 8))
    cells = mesh.assemble(levels, centres, rot, radius)
    gs = 5e-4
    params = dict(tau=1.0, p=0.5e-6, e=(1.0, 0.0, 0.0), b=gs / 2.0, m=1, c=100.0, runge_kutta=4,
                  residual=1e-10, itrmax=1000, volume_correction=True)
    ht = float(mesh.cell_volume(cells).sum() / (lx * ly * lz))
    desc = (f"batched suspension {lx}x{ly}x{lz} channel, {nc} cells "
            f"({int((levels == 3).sum())}x642 + {int((levels == 4).sum())}x2562 + {int((levels == 5).sum())}x10242 "
            f"vertices), r={radius:g} lu, Ht={100 * ht:.1f}%")
    return (lx, ly, lz), cells, params, desc, ht


def adjoint_workload(nc=32, lx=128, ly=64, lz=64, radius=4.0, seed=2025):
    """128x64x64 channel with 32 biconcave RBCs of 2562 vertices, as the
    128x64x64 suspension, for the wall-shear gradient."""
    from preprocess_src_amd import mesh
    levels = np.full(nc, 4)
    centres, rot = mesh.place(nc, radius, lx, ly, lz, seed)
    cells = mesh.assemble(levels, centres, rot, radius)
    gs = 5e-4
    params = dict(tau=1.0, p=0.5e-6, e=(1.0, 0.0, 0.0), b=gs / 2.0, m=1, c=100.0, runge_kutta=4,
                  residual=1e-10, itrmax=1000, volume_correction=False)
    ht = float(mesh.cell_volume(cells).sum() / (lx * ly * lz))
    desc = (f"adjoint: {lx}x{ly}x{lz} channel, {nc} RBCs x 2562 vertices, r={radius:g} lu; J = mean wall shear "
            f"over the last 200 of 500 steps; gradients w.r.t. the per-node body-force field and per-cell Gs")
    return (lx, ly, lz), cells, params, desc, ht


def adjoint_measure(precision, nsteps):
    """Cost of one checkpointed gradient of J relative to the forward run, and
    its agreement with a central finite difference along a random direction
    of the body-force field (10% of P)."""
    from preprocess_src_amd.simulation import Simulation
    shape, cells, params, desc, ht = adjoint_workload()
    lx, ly, lz = shape
    window, K = (2 * nsteps) // 5, min(50, nsteps)
    sim = Simulation(lx, ly, lz, cells=cells, precision=precision, **params)
    x0 = np.concatenate([cells["x"], cells["y"], cells["z"]])

    def reset():
        sim.set_equilibrium()
        x, y, z = x0.reshape(3, -1)
        sim.set_cells_state(x=x, y=y, z=z, xd=np.stack([x, y, z], 1).reshape(-1).copy())

    print(f"[bench] {desc}", file=sys.stderr, flush=True)
    reset()
    sim.wall_shear(0, 20, 10)  # warm-up
    reset()
    t0 = time.perf_counter()
    J = sim.wall_shear(0, nsteps, window)
    t_fwd = time.perf_counter() - t0
    print(f"[bench] forward {nsteps} steps: {t_fwd:.3f} s", file=sys.stderr, flush=True)
    reset()
    t0 = time.perf_counter()
    J2, gH, gs = sim.adjoint_wall_shear(0, nsteps, window, K)
    t_adj = time.perf_counter() - t0
    print(f"[bench] forward + adjoint: {t_adj:.3f} s", file=sys.stderr, flush=True)
    rng = np.random.default_rng(7)
    m = sim.comp_mask()
    d = [0.1 * params["p"] * rng.standard_normal(sim.np_) * m for _ in range(3)]
    fd = []
    for sgn in (1.0, -1.0):
        reset()
        sim.set_body_force_field(*[sgn * v for v in d])
        fd.append(sim.wall_shear(0, nsteps, window))
    sim.set_body_force_field()
    dd = (fd[0] - fd[1]) / 2.0
    ad = float(sum(g @ v for g, v in zip(gH, d)))
    n = lx * ly * lz
    out = dict(workload=desc, precision=precision, steps=nsteps, window=window, checkpoint_every=K,
               cells=sim.nc, vertices=sim.nv, hematocrit=round(ht, 4),
               adjoint_over_forward=round((t_adj - t_fwd) / t_fwd, 3),
               forward_s=round(t_fwd, 4), forward_plus_adjoint_s=round(t_adj, 4),
               forward_ms_per_step=round(1e3 * t_fwd / nsteps, 4),
               forward_mlups=round(n * nsteps / t_fwd / 1e6, 1), J=J, J_in_adjoint_call=J2,
               fd_check={"directional_adjoint": ad, "central_difference": dd,
                         "rel_err": abs(ad - dd) / max(abs(dd), 1e-300)},
               grad_force_scale_range=[float(np.min(gs)), float(np.max(gs))])
    sim.close()
    return out


def adjoint_bench(a):
    r = adjoint_measure(a.precision, a.adj_steps)
    out = {
        "metric": "adjoint/forward cost (wall-shear gradient, checkpointed)",
        "value": r["adjoint_over_forward"], "unit": "x forward", "n_gpus": 1, "steps": r["steps"], "warmup": 20,
        "ms_per_step": r["forward_ms_per_step"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": a.precision, "data": "synthetic (seeded cell placement, fluid at rest start)",
        "config": {"workload": r["workload"], "lattice": [128, 64, 64], "cells": r["cells"],
                   "vertices": r["vertices"], "window": r["window"], "checkpoint_every": r["checkpoint_every"],
                   "hematocrit": r["hematocrit"]},
        **{k: r[k] for k in ("forward_s", "forward_plus_adjoint_s", "forward_mlups", "J", "J_in_adjoint_call",
                             "fd_check", "grad_force_scale_range")},
    }
    print(json.dumps(out), flush=True)


def tgv_workload(n=256):
    return (n, n, n), None, dict(tau=0.6, periodic_y=True), f"Taylor-Green vortex {n}^3 fully periodic", 0.0


def tgv_init(shape, u0=0.01):
    lx, ly, lz = shape
    k = 2 * np.pi / lx
    xp = np.arange(lx + 4) - 1.5 - lx / 2
    yp = np.arange(ly + 4) - 1.5 - ly / 2
    X, Y = np.meshgrid(xp, yp, indexing="xy")
    u2 = u0 * np.cos(k * X) * np.sin(k * Y)
    v2 = -u0 * np.sin(k * X) * np.cos(k * Y)
    r2 = 1.0 - 0.75 * u0 * u0 * (np.cos(2 * k * X) + np.cos(2 * k * Y))
    rep = lambda a: np.broadcast_to(a, (lz + 4,) + a.shape).reshape(-1)
    return rep(r2), rep(u2), rep(v2)


def tgv_field_errors(sim, shape, params, steps, precision):
    """Per-field error of this run against an FP64 run of the same steps
    (max-abs and L2-relative for rho, u_x, u_y, u_z and the populations), and
    the decay of the velocity mode against exp(-2 nu k^2 t)."""
    from preprocess_src_amd.simulation import Simulation
    lx, ly, lz = shape
    cq = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0],
                   [-1, -1, 0], [-1, 1, 0], [1, -1, 0], [0, 1, 1], [0, -1, -1], [0, -1, 1], [0, 1, -1], [1, 0, 1],
                   [-1, 0, -1], [1, 0, -1], [-1, 0, 1]], float)

    def fields(s):
        # moments of the post-stream populations (no force: u = j / rho)
        f = s.populations().reshape(-1, Q)[s.comp_mask()]
        rho = f.sum(axis=1)
        j = f @ cq
        return {"rho": rho, "u_x": j[:, 0] / rho, "u_y": j[:, 1] / rho, "u_z": j[:, 2] / rho, "f_q": f}
    mine = fields(sim)
    if precision == "f64":
        ref = mine
    else:
        r = Simulation(lx, ly, lz, precision="f64", **params)
        r.set_equilibrium(*tgv_init(shape))
        r.run(0, steps, graph=True)
        ref = fields(r)
        r.close()
    err = {}
    for k in mine:
        d = mine[k] - ref[k]
        base = ref[k] - (1.0 if k == "rho" else 0.0)
        err[k] = {"max_abs": float(np.max(np.abs(d))),
                  "l2_rel": float(np.linalg.norm(d) / max(np.linalg.norm(base), 1e-300))}
    k = 2 * np.pi / lx
    xp = np.arange(lx) - lx / 2 + 0.5
    yp = np.arange(ly) - ly / 2 + 0.5
    X, Y = np.meshgrid(xp, yp, indexing="xy")
    basis = np.broadcast_to(np.cos(k * X) * np.sin(k * Y), (lz, ly, lx)).reshape(-1)
    amp = float((mine["u_x"] * basis).sum() / (basis * basis).sum())
    nu = (params["tau"] - 0.5) / 3
    err["decay"] = {"amplitude_over_u0": amp / 0.01, "analytic": float(np.exp(-2 * nu * k * k * steps))}
    return err


def lattice_bytes_per_node(precision, coupled):
    s = STORE_BYTES[precision]
    t = ARITH_BYTES[precision]
    return 2 * Q * s + (3 * t * 2 if coupled else 3 * t)  # populations + (F read, u write) or u write


def step_bytes(n, nv, nt, precision, coupled):
    """Algorithmic HBM bytes of one whole step: populations read and written,
    the lattice force written by the spreading and read by the collision, the
    velocity written (12 B/node each in FP32), plus the membrane's 64 B per
    vertex and 28 B per triangle."""
    return n * (lattice_bytes_per_node(precision, coupled) + (3 * ARITH_BYTES[precision] if coupled else 0)) + \
        (64 * nv + 28 * nt if coupled else 0)


def traffic_from_profile(precision):
    """HBM bytes per launch of each kernel measured by rocprofv3 FETCH_SIZE
    (x2.0, the copy-kernel calibration) and WRITE_SIZE on this workload, from
    the summary committed in profiles/ (FP32 only; {} when absent)."""
    path = os.path.join(ROOT, "profiles", "suspension_256x128x128_f32_hbm_traffic.txt")
    out = {}
    if precision != "f32" or not os.path.exists(path):
        return out
    table = False
    for line in open(path):
        if line.startswith("kernel "):
            table = True
            continue
        if table and line.strip() and not line.startswith(" "):
            parts = line.split()
            if len(parts) >= 6 and parts[-1].replace(".", "").isdigit():
                name = " ".join(parts[:-5])
                out[name] = round((float(parts[-3]) + float(parts[-1])) * 1e6)
        elif table and not line.strip():
            table = False
    return out


# kernel-table names -> kernel names in the rocprof summaries (FP32 workload)
PROF_NAMES = {"k_lbm": "k_lbm<float, float, double, true, false, 1>", "k_spread_box": "k_spread_box<float>",
              "k_spread_gather": "k_spread_gather<float>", "k_interp_tiles": "k_interp_tiles<float>",
              "k_bicg_res": "k_bicg_res<11>", "k_assemble": "k_assemble<float>",
              "k_bicg_onchip class 0": "k_bicg_onchip<float, 2, 512>",
              "k_bicg_onchip class 1": "k_bicg_onchip<float, 6, 512>", "k_vol_sums": "k_vol_sums",
              "k_vol_center": "k_vol_center", "k_volume<256>": "k_volume<256>", "k_volume<1024>": "k_volume<1024>",
              # the tile counts come with the volume correction: the step's binning is the scatter
              "binning": "k_bin_scatter<float>"}


def step_traffic(kernels):
    """HBM bytes of one step from the rocprof bytes per launch of the timed
    kernels (launches per step x bytes per launch); None unless every kernel
    of the table above 1% of the step's kernel time has a measurement."""
    total = sum(k["launches_per_step"] * k["avg_ms"] for k in kernels) or 1.0
    out = 0.0
    for k in kernels:
        b = k.get("rocprof_bytes")
        if b is None:
            if k["launches_per_step"] * k["avg_ms"] > 0.01 * total:
                return None
            continue
        out += k["launches_per_step"] * b
    return round(out)


def kernel_table(kt, n, nv, nt, precision, prof, autotune=None):
    """Per-kernel entries: launches per step and average duration (HIP events
    on the launch's stream, eager steps), algorithmic bytes per launch where
    the kernel has its own share of the step's dataflow, achieved bandwidth
    and fraction of the HBM peak, and the rocprof-measured bytes per launch."""
    t = ARITH_BYTES[precision]
    algo = {
        "k_lbm": n * lattice_bytes_per_node(precision, True),
        # spreading: positions (24 B) and forces (12 B) of each vertex in, F out
        "k_spread_box": 36 * nv, "k_spread_gather": 3 * t * n, "k_spread_tiles": 36 * nv + 3 * t * n,
        # interpolation + advection: positions in (24 B), SoA and AoS positions
        # out (48 B), the velocity field in
        "k_interp_tiles": 72 * nv + 3 * t * n, "k_interp_vert": 72 * nv + 3 * t * n,
    }
    prof_name = dict(PROF_NAMES)
    # the on-chip classes run the shape the autotuner chose for this run
    for line in autotune or []:
        m = re.match(r"bicg class ([01]): using (\d+)x(\d+)", line)
        if m:
            prof_name[f"k_bicg_onchip class {m.group(1)}"] = f"k_bicg_onchip<float, {m.group(2)}, {m.group(3)}>"
    rows = []
    for name, (per_step, ms) in sorted(kt.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        b = algo.get(name)
        row = {"name": name, "launches_per_step": round(per_step, 3), "avg_ms": round(ms, 5),
               "algorithmic_bytes": b}
        if b:
            row["achieved_GBps"] = round(b / (ms * 1e-3) / 1e9, 1)
            row["frac"] = round(b / (ms * 1e-3) / HBM_PEAK, 4)
        pb = prof.get(prof_name.get(name, ""))
        if pb:
            row["rocprof_bytes"] = pb
        rows.append(row)
    return rows


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return dict(cpu_model=model, nproc=avail, omp_num_threads=os.environ.get("OMP_NUM_THREADS"))


def cpu_baseline(shape, cells, params, seconds, nv=0, nt=0):
    """FP64 host oracle on the same workload for a bounded number of steps."""
    from oracle import pyoracle
    lx, ly, lz = shape
    kw = {k: v for k, v in params.items() if k not in ("store_macros", "bicg_warm_start", "residual_floor")}
    t0 = time.time()
    o = pyoracle.OracleSim(lx, ly, lz, cells=cells, **kw)
    o.init_fluid()
    if cells is not None:
        o.volume_init()
    setup = time.time() - t0
    steps, t_run = 0, 0.0
    while t_run < seconds or steps == 0:
        t1 = time.time()
        rc, _ = o.step(steps + 1)
        t_run += time.time() - t1
        steps += 1
        if rc != 0:
            raise RuntimeError(f"oracle step failed ({rc})")
    n = lx * ly * lz
    b = step_bytes(n, nv, nt, "f64", cells is not None) * steps / t_run
    return dict(value=n * steps / t_run / 1e6, unit="MLUPS", cores=int(pyoracle.lib().orc_num_threads()),
                kind="port", **cpu_info(), achieved_GBps=round(b / 1e9, 3),
                roofline_frac=b / HBM_PEAK,
                sample=f"{steps} full step(s) of the same workload in the FP64 OpenMP oracle "
                       f"({t_run:.1f} s timed, {setup:.1f} s set-up excluded); GB/s from the FP64 step's "
                       f"algorithmic bytes, fraction of the GPU's 8 TB/s HBM peak")


def copy_calibration(nbytes=256 << 20, reps=3):
    """A device-to-device copy of known size: under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE per dispatch against nbytes give the factors that
    turn those counters into HBM bytes."""
    import torch
    a = torch.ones(nbytes // 4, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "copy", "bytes_read": nbytes, "bytes_written": nbytes, "dispatches": reps}),
          flush=True)


def vof_bench(a):
    """The VOF indicator on the default suspension: one front tracking
    (spread, divergence, red-black SOR to sor_residual), THINC/WLIC sweeps
    and whole steps with a viscosity contrast (LAMBDA1 = 5), which run eagerly
    (front tracking waits for the host)."""
    from preprocess_src_amd.simulation import Simulation
    shape, cells, params, desc, ht = suspension_workload()
    lx, ly, lz = shape
    n = lx * ly * lz
    npad = (lx + 4) * (ly + 4) * (lz + 4)
    kw = dict(params, lambda1=5.0, vof_period=a.vof_period, sor_itrmax=a.sor_itrmax, sor_residual=a.sor_residual)
    sim = Simulation(lx, ly, lz, cells=cells, precision=a.precision, **kw)
    sim.volume_init()
    sim.synchronize()
    print("[bench] front tracking", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    its, att = sim.front_tracking()
    t_ft = time.perf_counter() - t0
    print(f"[bench] front tracking {t_ft:.3f} s, {its} SOR iterations, {att} start(s)", file=sys.stderr, flush=True)
    reps = 20
    sim.thinc_wlic()
    sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.thinc_wlic()
    sim.synchronize()
    t_th = (time.perf_counter() - t0) / reps
    steps = min(a.steps, a.vof_period - 1)
    sim.run(0, 2, graph=False)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.run(2, steps, graph=False)
    sim.synchronize()
    t_step = (time.perf_counter() - t0) / steps
    vf = sim.vf()
    sor_bytes = 32 * npad  # per iteration: x read (8 B) + b and x written on each colour's half, both colours
    th_bytes = 3 * n * (16 + ARITH_BYTES[a.precision])  # per call: 3 sweeps of vf read + write + one velocity
    out = {
        "workload": "vof",
        "config": {"workload": desc + ", LAMBDA1=5", "lattice": [lx, ly, lz], "precision": a.precision,
                   "vof_period": a.vof_period, "sor_itrmax": a.sor_itrmax, "sor_residual": a.sor_residual},
        "front_tracking_ms": round(1e3 * t_ft, 3),
        "sor_iterations": its, "sor_starts": att,
        "sor_us_per_iteration": round(1e6 * t_ft / max(its, 1), 3),
        "sor_roofline": {"bound": "hbm", "achieved": round(sor_bytes * its / t_ft / 1e9, 1),
                         "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": round(sor_bytes * its / t_ft / HBM_PEAK, 4), "bytes_per_iteration": sor_bytes},
        "thinc_ms": round(1e3 * t_th, 4),
        "thinc_roofline": {"bound": "hbm", "achieved": round(th_bytes / t_th / 1e9, 1), "peak": HBM_PEAK / 1e9,
                           "unit": "GB/s", "frac": round(th_bytes / t_th / HBM_PEAK, 4), "bytes_per_call": th_bytes},
        "step_ms_with_vof": round(1e3 * t_step, 4), "steps": steps,
        "indicator_volume": round(float(vf.reshape(lz + 4, ly + 4, lx + 4)[2:-2, 2:-2, 2:-2].sum()), 1),
        "cell_volume": round(float(ht * n), 1),
    }
    print(json.dumps(out), flush=True)
    sim.close()


def tube_bench(a):
    """The CapsuleCHANNELFLOW tube: a 128 x 128 x 512 box holding a tube of
    radius 63 lattice units along z, body force along z from the first step,
    no cells (the lattice kernel with the interpolated bounce-back alone)."""
    from preprocess_src_amd.simulation import Simulation
    lx, ly, lz, R = 128, 128, 512, 63.0
    sim = Simulation(lx, ly, lz, precision=a.precision, tau=1.0, p=1e-6, e=(0.0, 0.0, 1.0), force_modulated=2,
                     flowstart=0, tube_radius=R)
    fluid = int((sim.bc_mask() == 0).sum())
    sim.run(0, a.warmup, graph=True)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.run(a.warmup, a.steps, graph=True)
    sim.synchronize()
    dt = time.perf_counter() - t0
    stages = sim.time_stages(a.warmup + a.steps, a.stage_steps)
    t_lbm = stages["lbm_d3q19"][0] * 1e-3
    bpn = lattice_bytes_per_node(a.precision, False)
    out = {
        "workload": "tube",
        "config": {"workload": f"tube r={R:g} lu in a {lx}x{ly}x{lz} box (z periodic, interpolated bounce-back), "
                               f"body force along z, no cells", "precision": a.precision, "fluid_nodes": fluid},
        "mlups": round(fluid * a.steps / dt / 1e6, 1), "ms_per_step": round(1e3 * dt / a.steps, 4),
        "roofline": {"bound": "hbm", "kernel": "k_lbm (tube)", "achieved": round(fluid * bpn / t_lbm / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(fluid * bpn / t_lbm / HBM_PEAK, 4),
                     "bytes_per_fluid_node": bpn, "avg_ms": round(stages["lbm_d3q19"][0], 5)},
    }
    print(json.dumps(out), flush=True)
    sim.close()


def adhesion_bench(a):
    """The default suspension with cytoadhesion to the lower wall: receptors
    every 0.5 lu, 40 knobs per cell drawn from the seed, 2 springs per knob,
    16 receptors searched per knob within 2 lu; whole steps from the graph."""
    from preprocess_src_amd.simulation import Simulation
    shape, cells, params, desc, ht = suspension_workload()
    lx, ly, lz = shape
    n = lx * ly * lz
    sim = Simulation(lx, ly, lz, cells=cells, precision=a.precision, **params)
    sim.volume_init()
    adh = dict(receptor_density=4.0, rc=1.0, k_ad=2e-3, l0=0.5, k_on0=0.6, k_off0=0.25, chi_beta=0.3, kbt=1.0,
               shear_rate=1.0, gs=1.0, limit=2, max_neighwall=16, springs_per_knob=2, knob=40.0, seed=2024)
    sim.adhesion_setup(adh)
    nrec = int(lx / 0.5 + 1e-3) * int(lz / 0.5 + 1e-3)
    sim.run(0, a.warmup, graph=True)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.run(a.warmup, a.steps, graph=True)
    sim.synchronize()
    dt = time.perf_counter() - t0
    st = sim.adhesion_state(nrec)
    out = {
        "workload": "adhesion",
        "config": {"workload": desc + ", cytoadhesion to the lower wall", "precision": a.precision, "adhesion": adh},
        "mlups": round(n * a.steps / dt / 1e6, 2), "ms_per_step": round(1e3 * dt / a.steps, 4),
        "knobs": int(st["is_knob"].sum()), "attached_springs": int((st["attach"] >= 0).sum()),
        "receptors": nrec,
    }
    print(json.dumps(out), flush=True)
    sim.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="f32", choices=["f32", "f64", "f16"])
    ap.add_argument("--workload", default="suspension", choices=["suspension", "tgv", "adjoint", "copy", "vof", "tube", "adhesion"],
                    help="copy: a plain 256 MiB device copy, the calibration run of the FETCH_SIZE / "
                         "WRITE_SIZE counters")
    ap.add_argument("--stage-steps", type=int, default=6)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cold-start", action="store_true",
                    help="BiCGSTAB from x0 = 0 every step, as the reference (default: --warm-start)")
    ap.add_argument("--warm-start", type=int, default=2, choices=[1, 2],
                    help="BiCGSTAB start: 1 the previous loads, 2 their linear extrapolation from the last two")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the FP64 run and the adjoint measurement of the default workload")
    ap.add_argument("--adj-steps", type=int, default=500, help="steps of the adjoint workload (window 40%%)")
    ap.add_argument("--spinup", type=int, default=1000,
                    help="untimed steps from rest before the warm-up (suspension workload): the timed steps then "
                         "sit in the middle of the 2000-step run")
    ap.add_argument("--vof-period", type=int, default=1000, help="K of the vof workload")
    ap.add_argument("--sor-itrmax", type=int, default=200000)
    ap.add_argument("--sor-residual", type=float, default=1e-10)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the path is single-GPU, "
              "multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    if a.workload == "copy":
        copy_calibration()
        return

    from preprocess_src_amd.simulation import Simulation

    if a.workload == "adjoint":
        adjoint_bench(a)
        return
    if a.workload == "vof":
        vof_bench(a)
        return
    if a.workload == "tube":
        tube_bench(a)
        return
    if a.workload == "adhesion":
        adhesion_bench(a)
        return
    if a.workload == "suspension":
        shape, cells, params, desc, ht = suspension_workload()
        params["bicg_warm_start"] = 0 if a.cold_start else a.warm_start
    else:
        shape, cells, params, desc, ht = tgv_workload()
    lx, ly, lz = shape
    n = lx * ly * lz
    graph = not a.no_graph
    # the timed steps sit after `spinup` untimed steps from rest: with the
    # solver started from the extrapolated loads, a step's cost falls as the
    # flow develops, and step 1000 is the middle of the 2000-step
    # suspension run (its mean step time is within 1 % of step 1000's)
    spinup = a.spinup if a.workload == "suspension" else 0

    def timed(precision):
        t_setup = time.time()
        sim = Simulation(lx, ly, lz, cells=cells, precision=precision, **params)
        if a.workload == "tgv":
            sim.set_equilibrium(*tgv_init(shape))
        if cells is not None:
            sim.volume_init()
        print(f"[bench] set-up {time.time() - t_setup:.1f} s: {desc}, precision {precision}", file=sys.stderr,
              flush=True)
        if spinup:
            # in blocks of 200 steps (whole graph blocks), each reported, so a
            # slow run (under a profiler's counters) shows its progress
            done = 0
            while done < spinup:
                k = min(200, spinup - done)
                sim.run(done, k, graph=graph)
                sim.synchronize()
                done += k
                print(f"[bench] {precision}: spin-up {done}/{spinup} steps", file=sys.stderr, flush=True)
        sim.run(spinup, a.warmup, graph=graph)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.run(spinup + a.warmup, a.steps, graph=graph)
        sim.synchronize()
        dt = time.perf_counter() - t0
        print(f"[bench] {precision}: timed {a.steps} steps in {dt:.3f} s", file=sys.stderr, flush=True)
        return sim, dt

    sim, dt = timed(a.precision)
    ms_per_step = 1e3 * dt / a.steps
    mlups = n * a.steps / dt / 1e6
    coupled = cells is not None
    nv, nt = sim.nv, sim.nt
    sbytes = step_bytes(n, nv, nt, a.precision, coupled)
    # per-stage and per-kernel HIP-event timing on the library's streams,
    # eager steps right after the timed region
    it0 = spinup + a.warmup + a.steps
    stages = sim.time_stages(it0, a.stage_steps)
    kt = sim.time_kernels(it0 + a.stage_steps, a.stage_steps)
    diag = {}
    if coupled:
        its = sim.membrane()["bicg_iters"]
        nvc = np.diff(np.asarray(cells["voff"]))
        diag["bicg_iters_mean_by_vertices"] = {int(k): round(float(its[nvc == k].mean()), 2) for k in np.unique(nvc)}
        diag["autotune"] = [l for l in sim.autotune_report().splitlines() if l]
    prof = traffic_from_profile(a.precision) if a.workload == "suspension" else {}
    kernels = kernel_table(kt, n, nv, nt, a.precision, prof, diag.get("autotune"))
    dominant = kernels[0]["name"] if kernels else None
    lbm = kt.get("k_lbm", (1, stages["lbm_d3q19"][0]))[1]
    bpn = lattice_bytes_per_node(a.precision, coupled)
    out = {
        "metric": "MLUPS + % HBM roofline (D3Q19+IBM step, FP32 vs FP64); adjoint/forward cost",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": a.precision,
        "data": "synthetic (seeded cell placement, fluid at rest start)",
        "config": {"workload": desc, "lattice": [lx, ly, lz], "cells": sim.nc, "vertices": nv, "triangles": nt,
                   "hematocrit": round(ht, 4), "precision": a.precision, "graph": graph,
                   "timed_steps": [spinup + a.warmup + 1, spinup + a.warmup + a.steps],
                   "collide": "f64" if a.precision == "f64" or (a.precision == "f32" and not params.get("fp32_collide"))
                   else "f32",
                   "params": {k: (list(v) if isinstance(v, tuple) else v) for k, v in params.items()}},
        "roofline": {"bound": "hbm", "scope": "whole step (D3Q19 + IBM + membrane)",
                     "achieved": round(sbytes / (ms_per_step * 1e-3) / 1e9, 1), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(sbytes / (ms_per_step * 1e-3) / HBM_PEAK, 4),
                     "traffic": step_traffic(kernels), "algorithmic_bytes_per_step": sbytes,
                     "bytes_per_node": bpn + (3 * ARITH_BYTES[a.precision] if coupled else 0),
                     "membrane_bytes": (64 * nv + 28 * nt) if coupled else 0},
        "lattice_kernel_roofline": {"kernel": "k_lbm", "achieved": round(n * bpn / (lbm * 1e-3) / 1e9, 1),
                                    "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                                    "frac": round(n * bpn / (lbm * 1e-3) / HBM_PEAK, 4), "avg_ms": round(lbm, 5),
                                    "bytes_per_node": bpn,
                                    "traffic": prof.get(PROF_NAMES["k_lbm"])},
        "dominant_kernel": dominant,
        "kernels": kernels,
        "stages_ms": {k: round(v[0], 5) for k, v in stages.items()},
        "diagnostics": diag,
    }
    sim.close()
    if a.workload == "tgv":
        print("[bench] per-field errors against FP64", file=sys.stderr, flush=True)
        s2, _ = timed(a.precision)
        out["field_errors"] = tgv_field_errors(s2, shape, params, a.warmup + a.steps, a.precision)
        s2.close()
    if a.workload == "suspension" and not a.no_extra:
        if a.precision != "f64":
            s64, dt64 = timed("f64")
            ms64 = 1e3 * dt64 / a.steps
            b64 = step_bytes(n, nv, nt, "f64", coupled)
            out["fp64"] = {"mlups": round(n * a.steps / dt64 / 1e6, 2), "ms_per_step": round(ms64, 4),
                           "roofline_frac": round(b64 / (ms64 * 1e-3) / HBM_PEAK, 4),
                           "algorithmic_bytes_per_step": b64, "fp32_over_fp64": round(ms64 / ms_per_step, 3)}
            s64.close()
        print("[bench] adjoint (128x64x64, 32 RBCs)", file=sys.stderr, flush=True)
        out["adjoint"] = adjoint_measure(a.precision, a.adj_steps)
    if not a.no_cpu:
        print("[bench] cpu baseline (FP64 oracle)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(shape, cells, params, a.cpu_seconds, nv if coupled else 0,
                                           nt if coupled else 0)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
