"""Host-side mirror of the reference's driver for the coupled timestep.

``Simulation`` owns one ``lbmibm_ctx`` and exposes the reference's host
functions under their own names (``stress``, ``ibm_st1``, ``lbm_d3q19``,
``ibm_st2``, ``volume_init``, ``volume_correction``; stress.cu:580,
ibm.cu:299, lbm.cu:1463, ibm.cu:642, volume_correction.cu:427/453) so a
Python driver reads like main.cu:73-110.  Arrays crossing this boundary use
the reference layouts (padded lattice, AoS populations, SoA/AoS vertices).
"""
from __future__ import annotations

import ctypes as C

import os

import numpy as np

from . import native as N

Q = 19
PRECISIONS = {"f64": N.F64, "f32": N.F32, "f16": N.F16}


def make_params(lx, ly, lz, *, precision="f32", tau=None, df=None, dx=1.0, cr=1.0, uc=1.0, p=0.0,
                e=(1.0, 0.0, 0.0), force_modulated=False, flowstart=0, freq=1.0, b=0.0, m=1, c=100.0,
                rgs=1.0, br=0.0, itrmax=1000, residual=1e-10, runge_kutta=4, neohookean=False,
                volume_correction=True, store_macros=False, fileoutput=0, periodic_y=False, lambda1=1.0,
                vof_period=0, sor_itrmax=None, sor_residual=None, tube_radius=0.0, tube_staircase=False,
                wall_motion=0, wall_velocity=0.0, bicg_warm_start=False, residual_floor=1e-7,
                fp32_collide=False):
    P = N.Params()
    N.lib().lbmibm_default_params(C.byref(P), int(lx), int(ly), int(lz))
    if df is None:
        df = (2.0 * (0.8 if tau is None else tau) - 1.0) / 6.0  # tau = (6Df+1)/2 (init1.cu:140)
    P.periodic_y = int(periodic_y)
    P.dx, P.cr = float(dx), float(cr)
    P.dt = cr * dx / uc  # init1.cu:139
    P.df, P.lambda1, P.p = float(df), float(lambda1), float(p)
    P.ex, P.ey, P.ez = map(float, e)
    P.force_modulated, P.flowstart, P.freq = int(force_modulated), int(flowstart), float(freq)
    P.b, P.m, P.c, P.rgs, P.br = float(b), int(m), float(c), float(rgs), float(br)
    P.itrmax, P.residual, P.runge_kutta = int(itrmax), float(residual), int(runge_kutta)
    P.neohookean, P.volume_correction = int(neohookean), int(volume_correction)
    P.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    P.store_macros, P.fileoutput = int(store_macros), int(fileoutput)
    P.vof_period = int(vof_period)
    P.sor_itrmax = int(itrmax if sor_itrmax is None else sor_itrmax)
    P.sor_residual = float(residual if sor_residual is None else sor_residual)
    P.tube_radius = float(tube_radius)
    P.tube_staircase = int(tube_staircase)
    P.wall_motion, P.wall_velocity = int(wall_motion), float(wall_velocity)
    P.bicg_warm_start, P.residual_floor = int(bicg_warm_start), float(residual_floor)
    P.fp32_collide = int(fp32_collide)
    return P


class Simulation:
    def __init__(self, lx, ly, lz, cells=None, **kw):
        self.L = N.lib()
        self.P = make_params(lx, ly, lz, **kw)
        self.lx, self.ly, self.lz = int(lx), int(ly), int(lz)
        self.nxp, self.nyp, self.nzp = self.lx + 4, self.ly + 4, self.lz + 4
        self.np_ = self.nxp * self.nyp * self.nzp
        self._keep = []
        cp = None
        if cells is not None and len(cells["voff"]) > 1:
            arrs = dict(voff=np.ascontiguousarray(cells["voff"], np.int32),
                        eoff=np.ascontiguousarray(cells["eoff"], np.int32),
                        x=np.ascontiguousarray(cells["x"], np.float64), y=np.ascontiguousarray(cells["y"], np.float64),
                        z=np.ascontiguousarray(cells["z"], np.float64), ele=np.ascontiguousarray(cells["ele"], np.int32))
            self._keep.append(arrs)
            cp = N.Cells(len(arrs["voff"]) - 1, N.ip(arrs["voff"]), N.ip(arrs["eoff"]), N.dp(arrs["x"]),
                         N.dp(arrs["y"]), N.dp(arrs["z"]), N.ip(arrs["ele"]))
        h = C.c_void_p()
        code = self.L.lbmibm_create(C.byref(self.P), C.byref(cp) if cp is not None else None, C.byref(h))
        if code != 0:
            raise N.LbmibmError(code, "lbmibm_create failed")
        self.ctx = h
        n = C.c_longlong(0)
        nv, nt, nc = C.c_int(0), C.c_int(0), C.c_int(0)
        self.L.lbmibm_sizes(self.ctx, C.byref(n), C.byref(nv), C.byref(nt), C.byref(nc))
        self.nv, self.nt, self.nc = nv.value, nt.value, nc.value

    def close(self):
        if getattr(self, "ctx", None):
            self.L.lbmibm_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, code):
        N.check(code, self.ctx)

    # ---- the reference's host functions --------------------------------
    def lbm_d3q19(self, it):
        self._ck(self.L.lbmibm_lbm_d3q19(self.ctx, int(it)))

    def ibm_st1(self):
        self._ck(self.L.lbmibm_ibm_st1(self.ctx))

    def ibm_st2(self):
        self._ck(self.L.lbmibm_ibm_st2(self.ctx))

    def stress(self, it):
        self._ck(self.L.lbmibm_stress(self.ctx, int(it)))

    def volume_init(self):
        self._ck(self.L.lbmibm_volume_init(self.ctx))

    def volume_correction(self):
        k = C.c_int(0)
        self._ck(self.L.lbmibm_volume_correction(self.ctx, C.byref(k)))
        return k.value

    def step(self, it):
        self._ck(self.L.lbmibm_step(self.ctx, int(it)))

    def run(self, iter0, nsteps, graph=True):
        self._ck(self.L.lbmibm_run(self.ctx, int(iter0), int(nsteps), int(bool(graph))))

    STAGES = ("stress", "ibm_st1", "lbm_d3q19", "ibm_st2", "volume_correction", "step")

    def autotune_report(self):
        """Timings and choices of the solver-shape autotuning done at creation."""
        return self.L.lbmibm_autotune_report(self.ctx).decode()

    def time_stages(self, iter0, nsteps):
        """Average ms per call of each stage, timed with HIP events on the library stream."""
        ms = np.zeros(6)
        calls = np.zeros(6, np.int32)
        self._ck(self.L.lbmibm_time_stages(self.ctx, int(iter0), int(nsteps), N.dp(ms), N.ip(calls)))
        return {k: (float(ms[i]), int(calls[i])) for i, k in enumerate(self.STAGES)}

    def time_kernels(self, iter0, nsteps):
        """{kernel: (launches per step, average ms per launch)} of eager steps,
        timed with HIP events on each launch's stream."""
        buf = C.create_string_buffer(8192)
        self._ck(self.L.lbmibm_time_kernels(self.ctx, int(iter0), int(nsteps), buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, n, ms = line.split("\t")
            out[name] = (int(n) / nsteps, float(ms))
        return out

    def synchronize(self):
        self._ck(self.L.lbmibm_synchronize(self.ctx))

    # ---- reference-layout exchange --------------------------------------
    def set_equilibrium(self, rho=None, u=None, v=None, w=None):
        conv = lambda a: None if a is None else np.ascontiguousarray(a, np.float64).reshape(-1)
        arrs = [conv(a) for a in (rho, u, v, w)]
        for a in arrs:
            assert a is None or a.size == self.np_
        self._ck(self.L.lbmibm_set_equilibrium(self.ctx, *[N.dp(a) for a in arrs]))

    def populations(self):
        fn = np.zeros(self.np_ * Q)
        self._ck(self.L.lbmibm_download_populations(self.ctx, N.dp(fn)))
        return fn

    def set_populations(self, fn):
        fn = np.ascontiguousarray(fn, np.float64).reshape(-1)
        assert fn.size == self.np_ * Q
        self._ck(self.L.lbmibm_upload_populations(self.ctx, N.dp(fn)))

    def macros(self, instantaneous=False):
        names = ["dn", "un", "vn", "wn", "dm", "um", "vm", "wm"]
        out = {k: np.zeros(self.np_) for k in names}
        want = names if instantaneous else ["um", "vm", "wm"]
        self._ck(self.L.lbmibm_download_macros(self.ctx, *[N.dp(out[k]) if k in want else None for k in names]))
        return {k: out[k] for k in want}

    def set_macros(self, um, vm, wm):
        a = [np.ascontiguousarray(v, np.float64).reshape(-1) for v in (um, vm, wm)]
        self._ck(self.L.lbmibm_upload_macros(self.ctx, *[N.dp(v) for v in a]))

    def lattice_force(self):
        out = [np.zeros(self.np_) for _ in range(3)]
        self._ck(self.L.lbmibm_download_lattice_force(self.ctx, *[N.dp(v) for v in out]))
        return out

    def set_lattice_force(self, fx, fy, fz):
        a = [np.ascontiguousarray(v, np.float64).reshape(-1) for v in (fx, fy, fz)]
        self._ck(self.L.lbmibm_upload_lattice_force(self.ctx, *[N.dp(v) for v in a]))

    def cells_state(self):
        nv = self.nv
        names = ["x", "y", "z", "u", "v", "w", "fx", "fy", "fz"]
        out = {k: np.zeros(nv) for k in names}
        out["xd"] = np.zeros(3 * nv)
        self._ck(self.L.lbmibm_download_cells(self.ctx, *[N.dp(out[k]) for k in names], N.dp(out["xd"])))
        return out

    def set_cells_state(self, x=None, y=None, z=None, xd=None, fx=None, fy=None, fz=None):
        conv = lambda a: None if a is None else np.ascontiguousarray(a, np.float64).reshape(-1)
        arrs = [conv(a) for a in (x, y, z, xd, fx, fy, fz)]
        self._ck(self.L.lbmibm_upload_cells(self.ctx, *[N.dp(a) for a in arrs]))

    def membrane(self):
        out = dict(q=np.zeros(3 * self.nv), t=np.zeros(4 * self.nt), rg=np.zeros(self.nt),
                   tp=np.zeros(2 * self.nt), qi=np.zeros(3 * self.nv), bicg_iters=np.zeros(self.nc, np.int32))
        self._ck(self.L.lbmibm_download_membrane(self.ctx, N.dp(out["q"]), N.dp(out["t"]), N.dp(out["rg"]),
                                                 N.dp(out["tp"]), N.dp(out["qi"]), N.ip(out["bicg_iters"])))
        return out

    def volumes(self):
        out = dict(volume0=np.zeros(self.nc), volume=np.zeros(self.nc), area=np.zeros(self.nc),
                   gx=np.zeros(3 * self.nc))
        self._ck(self.L.lbmibm_download_volume(self.ctx, *[N.dp(out[k]) for k in ("volume0", "volume", "area", "gx")]))
        return out

    def set_cell_force_scale(self, scale=None):
        s = None if scale is None else np.ascontiguousarray(scale, np.float64)
        self._ck(self.L.lbmibm_set_cell_force_scale(self.ctx, N.dp(s)))

    # -- the reference's binary files (etc.cu) -------------------------------
    def bc_mask(self):
        """bc codes of the padded lattice (0 COMP, 1 BUFF, 2 WALL)."""
        bc = np.zeros(self.nxp * self.nyp * self.nzp, np.int32)
        self._ck(self.L.lbmibm_bc_mask(self.ctx, N.ip(bc)))
        return bc

    def write_restart(self, path, iter, fnum):
        """restart.bin of output_data (etc.cu:738-789)."""
        self._ck(self.L.lbmibm_write_restart(self.ctx, os.fsencode(path), int(iter), int(fnum)))

    def read_restart(self, path):
        """restart_file (etc.cu:164-323); returns (iter, fnum)."""
        it, fn = np.zeros(1, np.int32), np.zeros(1, np.int32)
        self._ck(self.L.lbmibm_read_restart(self.ctx, os.fsencode(path), N.ip(it), N.ip(fn)))
        return int(it[0]), int(fn[0])

    def write_prof(self, path, iter):
        """prof%04d.bin (etc.cu:547-560)."""
        self._ck(self.L.lbmibm_write_prof(self.ctx, os.fsencode(path), int(iter)))

    def write_cell(self, path):
        """cell%04d.bin (etc.cu:565-611)."""
        self._ck(self.L.lbmibm_write_cell(self.ctx, os.fsencode(path)))

    def output_data(self, directory, iter, fnum):
        """output_data's files with the reference's names; returns fnum + 1."""
        f = np.array([fnum], np.int32)
        self._ck(self.L.lbmibm_output_data(self.ctx, os.fsencode(directory), int(iter), N.ip(f)))
        return int(f[0])

    def front_tracking(self):
        """vof_init / front_tracking; returns (SOR iterations, starts used)."""
        its, att = C.c_int(0), C.c_int(0)
        self._ck(self.L.lbmibm_front_tracking(self.ctx, C.byref(its), C.byref(att)))
        return its.value, att.value

    def thinc_wlic(self):
        self._ck(self.L.lbmibm_thinc_wlic(self.ctx))

    def vf(self):
        a = np.zeros(self.np_)
        self._ck(self.L.lbmibm_download_vf(self.ctx, N.dp(a)))
        return a

    def set_vf(self, vf):
        self._ck(self.L.lbmibm_upload_vf(self.ctx, N.dp(np.ascontiguousarray(vf, np.float64))))

    def adhesion_setup(self, params, is_knob=None):
        """adhesion_init + wall_init; params: dict of AdhesionParams fields;
        is_knob [Nv] or None (drawn with probability knob / cell vertices)."""
        A = N.AdhesionParams()
        for k, v in params.items():
            setattr(A, k, v)
        self._adh = A
        kn = None
        if is_knob is not None:
            self._knob = np.ascontiguousarray(is_knob, np.uint8)
            kn = self._knob.ctypes.data_as(C.POINTER(C.c_ubyte))
        self._ck(self.L.lbmibm_adhesion_setup(self.ctx, C.byref(A), kn))

    def adhesion(self):
        self._ck(self.L.lbmibm_adhesion(self.ctx))

    def write_adhesion(self, directory, it, fnum_ad):
        """adhe%05d.bin and restart_adhe.dat under directory/adhesion; returns
        the next file number."""
        f = C.c_int(int(fnum_ad))
        self._ck(self.L.lbmibm_write_adhesion(self.ctx, str(directory).encode(), int(it), C.byref(f)))
        return f.value

    def adhesion_state(self, nrec):
        A = self._adh
        nv, mn, ns = self.nv, A.max_neighwall, A.springs_per_knob
        out = dict(is_knob=np.zeros(nv, np.uint8), id_wall=np.zeros(nv * mn, np.int32), r_wall=np.zeros(nv * mn),
                   attach=np.zeros(nv * ns, np.int32), r_ij=np.zeros(nv * ns), is_attached=np.zeros(nrec, np.int32),
                   f_adh=np.zeros(3 * nv))
        self._ck(self.L.lbmibm_adhesion_state(self.ctx, out["is_knob"].ctypes.data_as(C.POINTER(C.c_ubyte)),
                                              N.ip(out["id_wall"]), N.dp(out["r_wall"]), N.ip(out["attach"]),
                                              N.dp(out["r_ij"]), N.ip(out["is_attached"]), N.dp(out["f_adh"])))
        return out

    def comp_mask(self):
        """COMP nodes of the padded reference lattice (init1.cu:164-178)."""
        m = np.zeros((self.nzp, self.nyp, self.nxp), bool)
        m[2:-2, 2:-2, 2:-2] = True
        return m.reshape(-1)

    # ---- reverse mode ------------------------------------------------------
    def _adj_size(self, which):
        return {N.ADJ_POPULATIONS: self.np_ * Q, N.ADJ_VELOCITY: 3 * self.np_, N.ADJ_LATTICE_FORCE: 3 * self.np_,
                N.ADJ_POSITIONS: 3 * self.nv, N.ADJ_NODAL_FORCE: 3 * self.nv,
                N.ADJ_GRAD_BODY_FORCE: 3 * self.np_, N.ADJ_GRAD_FORCE_SCALE: self.nc}[which]

    def set_body_force_field(self, gx=None, gy=None, gz=None):
        """Per-node body-force field added to P e (padded [n] arrays); None removes it."""
        if gx is None:
            self._ck(self.L.lbmibm_set_body_force_field(self.ctx, None, None, None))
            return
        a = [np.ascontiguousarray(v, np.float64).reshape(-1) for v in (gx, gy, gz)]
        self._ck(self.L.lbmibm_set_body_force_field(self.ctx, *[N.dp(v) for v in a]))

    def adjoint_reset(self):
        self._ck(self.L.lbmibm_adjoint_reset(self.ctx))

    def adjoint(self, which):
        out = np.zeros(self._adj_size(which))
        self._ck(self.L.lbmibm_adjoint_get(self.ctx, int(which), N.dp(out)))
        return out

    def set_adjoint(self, which, a):
        a = np.ascontiguousarray(a, np.float64).reshape(-1)
        assert a.size == self._adj_size(which)
        self._ck(self.L.lbmibm_adjoint_set(self.ctx, int(which), N.dp(a)))

    def adj_lbm_d3q19(self, it):
        self._ck(self.L.lbmibm_adj_lbm_d3q19(self.ctx, int(it)))

    def adj_ibm_st1(self):
        self._ck(self.L.lbmibm_adj_ibm_st1(self.ctx))

    def adj_ibm_st2(self):
        self._ck(self.L.lbmibm_adj_ibm_st2(self.ctx))

    def adj_stress(self, it):
        self._ck(self.L.lbmibm_adj_stress(self.ctx, int(it)))

    def wall_shear(self, iter0, nsteps, window):
        J = np.zeros(1)
        self._ck(self.L.lbmibm_wall_shear(self.ctx, int(iter0), int(nsteps), int(window), N.dp(J)))
        return float(J[0])

    def adjoint_wall_shear(self, iter0, nsteps, window, checkpoint_every):
        """J and dJ/d(body-force field) (3 padded [n] arrays) and dJ/d(per-cell force scale)."""
        J = np.zeros(1)
        g = [np.zeros(self.np_) for _ in range(3)]
        ds = np.zeros(max(self.nc, 1))
        self._ck(self.L.lbmibm_adjoint_wall_shear(self.ctx, int(iter0), int(nsteps), int(window),
                                                  int(checkpoint_every), N.dp(J), *[N.dp(v) for v in g],
                                                  N.dp(ds)))
        return float(J[0]), g, ds[:self.nc]
