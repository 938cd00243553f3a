"""Host-side checks of liblbmibm.so that need no GPU: the C ABI's exported
symbols against include/lbmibm.h, the mesh generators, the topology builder
against the oracle's restatement of fem_init, seeded placement, parameter
defaults and the loud failure of the GPU path when no device is present."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from preprocess_src_amd import mesh
from preprocess_src_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "lbmibm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:lbmibm_status|void|const char\s*\*|int)\s*\**\s*(lbmibm_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    L = C.CDLL(N.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python bindings cover exactly the declared set
    assert sorted(N.SIGNATURES) == names


def test_default_params_follow_the_reference_relations():
    p = N.Params()
    N.lib().lbmibm_default_params(C.byref(p), 64, 32, 32)
    assert (p.lx, p.ly, p.lz) == (64, 32, 32)
    assert p.dx == 1.0 and p.lambda1 == 1.0 and p.m >= 1 and p.runge_kutta == 4
    assert abs(p.dt - p.cr * p.dx) < 1e-15  # dt = Cr dx / Uc with Uc = 1 (init1.cu:139)
    assert p.precision == N.F32


@pytest.mark.parametrize("level", [0, 1, 2, 3, 4])
def test_icosphere_counts_orientation_and_radius(level):
    xyz, tri = mesh.icosphere(level)
    assert len(xyz) == 10 * 4 ** level + 2 and len(tri) == 20 * 4 ** level
    assert np.allclose(np.linalg.norm(xyz, axis=1), 1.0, atol=1e-14)
    # closed, consistently oriented outward: signed volume positive, each
    # directed edge appears exactly once
    X = xyz[tri]
    vol = np.einsum("ij,ij->i", X[:, 0], np.cross(X[:, 1], X[:, 2])).sum() / 6
    assert vol > 0
    edges = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    assert len(np.unique(edges, axis=0)) == len(edges)
    assert len(edges) == 3 * len(tri)
    valence = np.bincount(tri.reshape(-1), minlength=len(xyz))
    assert valence.max() <= 6 and valence.min() >= 5


def test_biconcave_profile():
    """Evans-Fung: z = +-0.5 sqrt(1-r^2)(c0 + c1 r^2 + c2 r^4) on the unit disc."""
    xyz, _ = mesh.icosphere(3)
    b = mesh.biconcave(xyz)
    r2 = b[:, 0] ** 2 + b[:, 1] ** 2
    assert np.allclose(r2, xyz[:, 0] ** 2 + xyz[:, 1] ** 2)
    h = 0.5 * np.sqrt(np.clip(1 - r2, 0, None)) * (0.207161 + 2.002558 * r2 - 1.122762 * r2 * r2)
    assert np.allclose(np.abs(b[:, 2]), h, atol=1e-12)
    side = np.abs(xyz[:, 2]) > 1e-12
    assert np.all(np.sign(b[side, 2]) == np.sign(xyz[side, 2]))


@pytest.mark.parametrize("levels", [[2], [2, 3], [3, 2, 4]])
def test_topology_matches_the_oracle_bit_for_bit(levels):
    from oracle.pyoracle import topology
    c, r = mesh.place(len(levels), 4.0, 40, 30, 40, seed=3)
    cells = mesh.assemble(levels, c, r, 4.0)
    nc, nv, nt = len(levels), int(cells["voff"][-1]), int(cells["eoff"][-1])
    xr = np.stack([cells["x"], cells["y"], cells["z"]], 1).reshape(-1).copy()
    out = dict(nln=np.zeros(nv, np.int32), ln=np.zeros(7 * nv, np.int32), lo=np.zeros(7 * nv, np.int32),
               nle=np.zeros(nv, np.int32), le=np.zeros(6 * nv, np.int32), nref=np.zeros(4 * nt),
               ec=np.zeros(3 * nt, np.int32), c0=np.zeros(3 * nt), s0=np.zeros(3 * nt))
    N.check(N.lib().lbmibm_build_topology(nc, N.ip(cells["voff"]), N.ip(cells["eoff"]), N.dp(xr), N.ip(cells["ele"]),
                                          *[(N.ip if v.dtype == np.int32 else N.dp)(v) for v in out.values()]))
    ref = topology(cells["voff"], cells["eoff"], xr, cells["ele"])
    for k, v in out.items():
        assert np.array_equal(v, ref[k]), k


def test_topology_rejects_meshes_below_the_minimum_system_size():
    c, r = mesh.place(1, 3.0, 20, 20, 20, seed=1)
    cells = mesh.assemble([1], c, r, 3.0)  # 42 vertices: 3V=126 -> 128 rows < 256
    nv, nt = 42, 80
    xr = np.stack([cells["x"], cells["y"], cells["z"]], 1).reshape(-1).copy()
    bufs = [np.zeros(nv, np.int32), np.zeros(7 * nv, np.int32), np.zeros(7 * nv, np.int32), np.zeros(nv, np.int32),
            np.zeros(6 * nv, np.int32), np.zeros(4 * nt), np.zeros(3 * nt, np.int32), np.zeros(3 * nt), np.zeros(3 * nt)]
    rc = N.lib().lbmibm_build_topology(1, N.ip(cells["voff"]), N.ip(cells["eoff"]), N.dp(xr), N.ip(cells["ele"]),
                                       *[(N.ip if v.dtype == np.int32 else N.dp)(v) for v in bufs])
    assert rc == 6


def test_placement_is_seeded_non_overlapping_and_inside_the_walls():
    lx, ly, lz, r = 64, 48, 64, 4.0
    c1, r1 = mesh.place(24, r, lx, ly, lz, seed=11)
    c2, r2 = mesh.place(24, r, lx, ly, lz, seed=11)
    assert np.array_equal(c1, c2) and np.array_equal(r1, r2)
    c3, _ = mesh.place(24, r, lx, ly, lz, seed=12)
    assert not np.array_equal(c1, c3)
    # minimum-image centre distances >= 2r+1, wall distance >= r+1
    d = c1[:, None, :] - c1[None, :, :]
    for i, L in ((0, lx), (2, lz)):
        d[..., i] -= L * np.round(d[..., i] / L)
    dist = np.linalg.norm(d, axis=2) + np.eye(len(c1)) * 1e9
    assert dist.min() >= 2 * r + 1 - 1e-12
    assert np.all(np.abs(c1[:, 1]) <= ly / 2 - (r + 1) + 1e-12)
    # rotations are proper orthogonal
    for R in r1:
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-14) and abs(np.linalg.det(R) - 1) < 1e-14


def test_grid_placement_of_the_suspension_workload():
    c, r = mesh.place(1024, 7.0, 256, 128, 128, seed=2024, mode="grid", grid=(16, 8, 8))
    d = c[:, None, :] - c[None, :, :]
    for i, L in ((0, 256), (2, 128)):
        d[..., i] -= L * np.round(d[..., i] / L)
    dist = np.linalg.norm(d, axis=2) + np.eye(1024) * 1e9
    # jitter bounded by half the slack of 16-lu slots: bounding spheres never overlap
    assert dist.min() >= 2 * 7.0 + 1 - 1e-12
    assert np.all(np.abs(c[:, 1]) <= 64 - 8.0 + 1e-12)


def test_mesh_mix_counts():
    lev = mesh.mesh_mix(1024, [0.5, 0.4, 0.1], [3, 4, 5], seed=2024)
    assert [(lev == k).sum() for k in (3, 4, 5)] == [512, 410, 102]


def test_gpu_path_fails_loudly_without_a_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from preprocess_src_amd.simulation import Simulation
    with pytest.raises(N.LbmibmError) as e:
        Simulation(8, 8, 8, precision="f32", tau=0.8)
    assert e.value.code == 2  # LBMIBM_ERR_HIP
