"""The reference-style adaptor (include/lbmibm_refcompat.hpp) compiles against
stand-in structs, links liblbmibm.so and, on a GPU, drives main.cu's loop."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "preprocess_src_amd")


def _build(tmp_path):
    exe = str(tmp_path / "refcompat_driver")
    cmd = ["g++", "-std=c++17", "-O1", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(ROOT, "tests", "refcompat_driver.cpp"), f"-L{PKG}", "-llbmibm", f"-Wl,-rpath,{PKG}",
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_adaptor_compiles_links_and_runs_host_side(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "adaptor ok" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
def test_adaptor_drives_the_reference_loop(tmp_path):
    exe = _build(tmp_path)
    work = tmp_path / "work"
    os.makedirs(work / "result" / "cell")
    r = subprocess.run([exe, "run"], capture_output=True, text=True, timeout=120, cwd=work)
    assert r.returncode == 0, r.stdout + r.stderr
    assert float(r.stdout.split()[-1]) > 0.0
    # output_data's files, named as the reference names them
    assert sorted(os.listdir(work / "result")) == ["cell", "prof0000.bin", "restart.bin", "restart0000.bin"]


@pytest.mark.gpu
def test_adaptor_drives_the_vof_stages(tmp_path):
    """vof_init, then main.cu's loop with thinc_wlic and front_tracking, for
    one spherical cell with LAMBDA1 = 4: the indicator is 1 at the centre
    and integrates to the sphere's volume."""
    exe = _build(tmp_path)
    r = subprocess.run([exe, "vof"], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "vf max" in r.stdout


def _write_case(path, shape, prec, steps, par, cells):
    import numpy as np
    voff, eoff = np.asarray(cells["voff"], np.int32), np.asarray(cells["eoff"], np.int32)
    with open(path, "wb") as f:
        np.array([*shape, len(voff) - 1, prec, steps], np.int32).tofile(f)
        np.array([par["tau"], par["p"], par["b"], par["c"], par["residual"]], np.float64).tofile(f)
        voff.tofile(f)
        eoff.tofile(f)
        for k in "xyz":
            np.asarray(cells[k], np.float64).tofile(f)
        np.asarray(cells["ele"], np.int32).tofile(f)


def _read_state(path, n, nv, nc):
    import numpy as np
    a = np.fromfile(path, np.float64, count=19 * n + 3 * n + 6 * nv + 2 * nc)
    passes = int(np.fromfile(path, np.int32, offset=a.nbytes)[0])
    o = 0
    out = {}
    for k, m in (("fn", 19 * n), ("um", n), ("vm", n), ("wm", n), ("x", nv), ("y", nv), ("z", nv), ("fx", nv),
                 ("fy", nv), ("fz", nv), ("volume", nc), ("area", nc)):
        out[k] = a[o:o + m]
        o += m
    out["passes"] = passes
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [("f64", 1e-9), ("f32", 1e-5)])
def test_adaptor_drives_cells_with_mixed_meshes_against_the_oracle(tmp_path, prec, tol):
    """Two cells of different meshes (162 and 642 vertices), volume
    correction on, 12 steps of main.cu's loop through the adaptor and the C
    ABI, against the same 12 steps of the FP64 oracle: populations (as
    deviations), averaged velocities, vertex displacements, nodal forces,
    volumes and areas, and the last correction's pass count."""
    import numpy as np

    from tests.pairs import dev_err, oracle_of, small_cells

    print(f"[adaptor cells] prec={prec}", flush=True)
    exe = _build(tmp_path)
    shape, steps = (24, 16, 20), 12
    par = dict(tau=0.9, p=5e-6, b=5e-3, c=100.0, residual=1e-12)
    cells = small_cells(*shape, nc=2, radius=3.0, levels=[2, 3], seed=31)
    case, state = str(tmp_path / "case.bin"), str(tmp_path / "state.bin")
    _write_case(case, shape, {"f64": 0, "f32": 1}[prec], steps, par, cells)
    r = subprocess.run([exe, "cells", case, state], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    o = oracle_of(*shape, cells=cells, tau=par["tau"], p=par["p"], b=par["b"], c=par["c"],
                  residual=par["residual"], volume_correction=True)
    o.volume_init()
    passes = 0
    for it in range(1, steps + 1):
        rc, passes = o.step(it)
        assert rc == 0
    nv, nc = len(cells["x"]), 2
    g = _read_state(state, o.n, nv, nc)
    comp = o.bc == 0
    W = 1.0 / np.array([3.0] + [18.0] * 6 + [36.0] * 12)
    e = {"f_q - w_q": dev_err(g["fn"].reshape(-1, 19)[comp], o.fn.reshape(-1, 19)[comp], W)}
    # the velocity as one field (u_y, u_z are a thousandth of u_x here)
    e["u"] = dev_err(np.concatenate([g[k][comp] for k in ("um", "vm", "wm")]),
                     np.concatenate([getattr(o, k)[comp] for k in ("um", "vm", "wm")]))
    x0 = np.concatenate([cells[k] for k in "xyz"])
    e["displacement"] = dev_err(np.concatenate([g[k] for k in "xyz"]) - x0,
                                np.concatenate([o.cells[k] for k in "xyz"]) - x0)
    e["nodal force"] = dev_err(np.concatenate([g[k] for k in ("fx", "fy", "fz")]),
                               np.concatenate([o.cells[k] for k in ("cfx", "cfy", "cfz")]))
    print(e, flush=True)
    for k in ("f_q - w_q", "u", "displacement"):
        assert e[k]["rel"] < tol, k
    # the cells barely deform in 12 steps from rest: their nodal forces are
    # small differences of the corner terms (about 7e-9 of max |f| in FP64,
    # 3e-4 in FP32)
    assert e["nodal force"]["rel"] < (1e-7 if prec == "f64" else 1e-3)
    assert np.max(np.abs(g["volume"] / o.cells["volume"] - 1)) < tol
    assert np.max(np.abs(g["area"] / o.cells["area"] - 1)) < tol
    assert g["passes"] == passes
