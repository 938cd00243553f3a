"""lbmibm_wall_shear_noalloc, the wall-shear objective with its partial sums
in the library's static device array, against lbmibm_wall_shear on twin
contexts: two calls in a row, the second on a smaller lattice (so stale
partial sums of the first would show), in FP64 and FP32."""
import numpy as np
import pytest

from tests.pairs import gpu_of, small_cells

pytestmark = pytest.mark.gpu


def _noalloc(g, iter0, nsteps, window):
    from preprocess_src_amd import native as N
    J = np.zeros(1)
    g._ck(g.L.lbmibm_wall_shear_noalloc(g.ctx, int(iter0), int(nsteps), int(window), N.dp(J)))
    return float(J[0])


def _twins(shape, prec, seed):
    cells = small_cells(*shape, nc=2, radius=3.0, levels=[2, 2], seed=seed)
    kw = dict(tau=0.8, p=1e-4, b=5e-2, c=10.0, residual=1e-13 if prec == "f64" else 1e-9, volume_correction=False)
    return [gpu_of(*shape, prec, cells=cells, **kw) for _ in range(2)]


@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("f32", 1e-6)])
def test_noalloc_wall_shear_equals_the_allocating_entry_point(prec, tol, monkeypatch):
    # the same interpolation and spreading forms in both contexts (the
    # autotuner picks by timing), so that their steps are bitwise equal
    monkeypatch.setenv("LBMIBM_INTERP_FORM", "0")
    monkeypatch.setenv("LBMIBM_SPREAD_FORM", "1")
    for shape, nsteps, window, seed in (((20, 14, 18), 12, 6, 61), ((16, 12, 14), 8, 3, 62)):
        print(f"[noalloc] prec={prec} shape={shape}", flush=True)
        a, b = _twins(shape, prec, seed)
        Ja = a.wall_shear(0, nsteps, window)
        Jb = _noalloc(b, 0, nsteps, window)
        assert Ja != 0.0
        assert abs(Jb - Ja) <= tol * abs(Ja), (Ja, Jb)
        # the populations the two runs leave behind are the same
        fa, fb = a.populations(), b.populations()
        assert np.max(np.abs(fa - fb)) == 0.0
        a.close()
        b.close()
