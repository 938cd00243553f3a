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
