"""ctypes front end of the FP64 host oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this module.  It owns numpy arrays in the reference's padded layout and
drives ``liblbmibm_oracle.so`` (built from oracle.cpp by oracle/Makefile).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblbmibm_oracle.so")
Q = 19

c_double_p = C.POINTER(C.c_double)
c_int_p = C.POINTER(C.c_int)


class OrcParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("periodic_y", C.c_int),
        ("dx", C.c_double), ("dt", C.c_double), ("cr", C.c_double),
        ("df", C.c_double), ("lambda1", C.c_double), ("p", C.c_double),
        ("ex", C.c_double), ("ey", C.c_double), ("ez", C.c_double),
        ("force_modulated", C.c_int), ("flowstart", C.c_int), ("freq", C.c_double),
        ("b", C.c_double), ("m", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("c", C.c_double), ("rgs", C.c_double), ("br", C.c_double),
        ("itrmax", C.c_int), ("residual", C.c_double), ("runge_kutta", C.c_int),
        ("neohookean", C.c_int), ("fileoutput", C.c_int), ("volume_correction", C.c_int),
        ("vof_period", C.c_int), ("sor_itrmax", C.c_int), ("sor_residual", C.c_double),
        ("tube", C.c_int), ("tube_radius", C.c_double), ("tube_staircase", C.c_int),
        ("wall_motion", C.c_int), ("wall_velocity", C.c_double),
    ]


class OrcState(C.Structure):
    _fields_ = [
        ("fn", c_double_p), ("fm", c_double_p),
        ("dn", c_double_p), ("un", c_double_p), ("vn", c_double_p), ("wn", c_double_p),
        ("dm", c_double_p), ("um", c_double_p), ("vm", c_double_p), ("wm", c_double_p),
        ("fx", c_double_p), ("fy", c_double_p), ("fz", c_double_p),
        ("vf", c_double_p), ("bc", c_int_p),
        ("nc", C.c_int), ("voff", c_int_p), ("eoff", c_int_p),
        ("x", c_double_p), ("y", c_double_p), ("z", c_double_p),
        ("u", c_double_p), ("v", c_double_p), ("w", c_double_p),
        ("cfx", c_double_p), ("cfy", c_double_p), ("cfz", c_double_p),
        ("xd", c_double_p), ("ele", c_int_p), ("ec", c_int_p),
        ("c0", c_double_p), ("s0", c_double_p), ("nref", c_double_p),
        ("t", c_double_p), ("rg", c_double_p), ("tp", c_double_p),
        ("nln", c_int_p), ("ln", c_int_p), ("lo", c_int_p), ("nle", c_int_p), ("le", c_int_p),
        ("q", c_double_p), ("qi", c_double_p), ("gs_scale", c_double_p),
        ("volume0", c_double_p), ("volume", c_double_p), ("area", c_double_p), ("gx", c_double_p),
        ("bicg_iters", c_int_p),
        ("gbx", c_double_p), ("gby", c_double_p), ("gbz", c_double_p),
        ("adh", C.c_void_p),
    ]


class OrcAdhesion(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("receptor_density", "rc", "k_ad", "l0", "k_on0", "k_off0", "chi_beta",
                                          "kbt", "shear_rate", "gs")] + [
        ("limit", C.c_int), ("max_neighwall", C.c_int), ("springs_per_knob", C.c_int), ("seed", C.c_ulonglong),
        ("wlc", C.c_int), ("lcont", C.c_double), ("f_wlc", C.c_double), ("ra", C.c_double), ("alpha1", C.c_double),
        ("alpha2", C.c_double),
        ("nxw", C.c_int), ("nzw", C.c_int), ("dxw", C.c_double), ("dzw", C.c_double),
        ("is_knob", C.POINTER(C.c_ubyte)), ("id_wall", c_int_p), ("r_wall", c_double_p), ("attach", c_int_p),
        ("r_ij", c_double_p), ("is_attached", c_int_p), ("f_adh", c_double_p),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ only)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        P, S = C.POINTER(OrcParams), C.POINTER(OrcState)
        sig = {
            "orc_bc_init": (None, [P, c_int_p]),
            "orc_flu_init": (None, [P, S, c_double_p, c_double_p, c_double_p, c_double_p]),
            "orc_lbm_step": (None, [P, S, C.c_int]),
            "orc_ibm_st1": (None, [P, S]),
            "orc_ibm_st2": (None, [P, S]),
            "orc_delta_weight": (C.c_double, [C.c_double]),
            "orc_stress": (C.c_int, [P, S, C.c_int]),
            "orc_calc_tension": (None, [P, S, C.c_int]),
            "orc_assemble_cell": (C.c_int, [P, S, C.c_int, c_int_p, c_int_p, c_double_p, c_double_p]),
            "orc_bicgstab": (C.c_int, [c_int_p, c_int_p, c_double_p, c_double_p, c_double_p,
                                       C.c_int, C.c_int, C.c_double, c_int_p]),
            "orc_expand2pow2": (C.c_int, [C.c_int]),
            "orc_calc_volume": (None, [P, S, c_double_p, c_double_p]),
            "orc_volume_init": (None, [P, S]),
            "orc_volume_correction": (C.c_int, [P, S]),
            "orc_fem_topology": (C.c_int, [C.c_int, c_int_p, c_int_p, c_double_p, c_int_p,
                                           c_int_p, c_int_p, c_int_p, c_int_p, c_int_p,
                                           c_double_p, c_int_p, c_double_p, c_double_p]),
            "orc_step": (C.c_int, [P, S, C.c_int, c_int_p]),
            "orc_skalak_energy": (C.c_double, [P, S]),
            "orc_bending_energy": (C.c_double, [P, S]),
            "orc_num_threads": (C.c_int, []),
            "orc_ft_gradient": (None, [P, S, c_double_p, c_double_p, c_double_p]),
            "orc_ft_rhs": (None, [P, c_int_p, c_double_p, c_double_p, c_double_p, c_double_p]),
            "orc_ft_sor": (C.c_int, [P, c_int_p, c_double_p, c_double_p, C.c_double, c_int_p]),
            "orc_front_tracking": (C.c_int, [P, S, c_int_p, c_int_p]),
            "orc_vof_normals": (None, [P, c_double_p, c_double_p, c_double_p, c_double_p]),
            "orc_thinc_flux": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                            C.c_double, C.c_double]),
            "orc_thinc_wlic": (None, [P, S]),
            "orc_tube_fraction": (C.c_double, [P, C.c_int, C.c_int]),
            "orc_adhesion_grid": (C.c_int, [P, C.POINTER(OrcAdhesion)]),
            "orc_adhesion_init": (None, [P, S, C.POINTER(OrcAdhesion)]),
            "orc_adhesion_step": (None, [P, S, C.POINTER(OrcAdhesion), C.c_int]),
            "orc_adh_uniform": (C.c_double, [C.c_ulonglong, C.c_longlong, C.c_longlong]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(c_double_p)


def ip(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(c_int_p)


def topology(voff, eoff, xr_aos, ele):
    """Reference topology rules (init1.cu:626-853, init2.cu:12-400)."""
    voff = np.ascontiguousarray(voff, np.int32)
    eoff = np.ascontiguousarray(eoff, np.int32)
    nc = len(voff) - 1
    nv, ne = int(voff[-1]), int(eoff[-1])
    out = dict(
        nln=np.zeros(nv, np.int32), ln=np.zeros(7 * nv, np.int32), lo=np.zeros(7 * nv, np.int32),
        nle=np.zeros(nv, np.int32), le=np.zeros(6 * nv, np.int32), nref=np.zeros(4 * ne),
        ec=np.zeros(3 * ne, np.int32), c0=np.zeros(3 * ne), s0=np.zeros(3 * ne),
    )
    rc = lib().orc_fem_topology(nc, ip(voff), ip(eoff), dp(np.ascontiguousarray(xr_aos, np.float64)),
                                ip(np.ascontiguousarray(ele, np.int32)), ip(out["nln"]), ip(out["ln"]),
                                ip(out["lo"]), ip(out["nle"]), ip(out["le"]), dp(out["nref"]),
                                ip(out["ec"]), dp(out["c0"]), dp(out["s0"]))
    if rc != 0:
        raise RuntimeError(f"oracle topology check failed (error {rc})")
    return out


class OracleSim:
    """Reference-layout FP64 state plus the five stage calls of main.cu."""

    def __init__(self, lx, ly, lz, *, dx=1.0, cr=1.0, uc=1.0, tau=None, df=None, lambda1=1.0,
                 p=0.0, e=(1.0, 0.0, 0.0), force_modulated=False, flowstart=0, freq=1.0,
                 b=0.0, m=1, c=100.0, rgs=1.0, br=0.0, itrmax=1000, residual=1e-12,
                 runge_kutta=4, neohookean=False, fileoutput=1000, volume_correction=True,
                 periodic_y=False, cells=None, vof_period=0, sor_itrmax=None, sor_residual=None,
                 tube_radius=0.0, tube_staircase=False, wall_motion=0, wall_velocity=0.0):
        if df is None:
            df = (2.0 * tau - 1.0) / 6.0  # tau = (6 Df + 1)/2 (init1.cu:140)
        P = OrcParams()
        P.nx, P.ny, P.nz = int(round(lx / dx)) + 4, int(round(ly / dx)) + 4, int(round(lz / dx)) + 4
        P.periodic_y = int(periodic_y)
        P.dx, P.cr = dx, cr
        P.dt = cr * dx / uc  # init1.cu:139
        P.df, P.lambda1, P.p = df, lambda1, p
        P.ex, P.ey, P.ez = e
        P.force_modulated, P.flowstart, P.freq = int(force_modulated), flowstart, freq
        P.b, P.m = b, m
        P.lx, P.ly, P.lz = lx, ly, lz
        P.c, P.rgs, P.br = c, rgs, br
        P.itrmax, P.residual, P.runge_kutta = itrmax, residual, runge_kutta
        P.neohookean, P.fileoutput, P.volume_correction = int(neohookean), fileoutput, int(volume_correction)
        P.vof_period = int(vof_period)
        P.tube, P.tube_radius = int(tube_radius > 0), float(tube_radius)
        P.tube_staircase = int(tube_staircase)
        P.wall_motion, P.wall_velocity = int(wall_motion), float(wall_velocity)
        P.sor_itrmax = int(itrmax if sor_itrmax is None else sor_itrmax)
        P.sor_residual = float(residual if sor_residual is None else sor_residual)
        self.P = P
        self.nx, self.ny, self.nz = P.nx, P.ny, P.nz
        n = P.nx * P.ny * P.nz
        self.n = n
        L = lib()
        self.bc = np.zeros(n, np.int32)
        L.orc_bc_init(C.byref(P), ip(self.bc))
        self.bufs = {k: np.zeros(n * Q) for k in ("fa", "fb")}
        for k in ("dn", "un", "vn", "wn", "dm", "um", "vm", "wm", "fx", "fy", "fz", "vf"):
            setattr(self, k, np.zeros(n))
        S = OrcState()
        S.fn, S.fm = dp(self.bufs["fa"]), dp(self.bufs["fb"])
        for k in ("dn", "un", "vn", "wn", "dm", "um", "vm", "wm", "fx", "fy", "fz"):
            setattr(S, k, dp(getattr(self, k)))
        S.vf = dp(self.vf)
        S.bc = ip(self.bc)
        S.nc = 0
        self.S = S
        self.cells = None
        if cells is not None:
            self.set_cells(cells)

    # -- lattice -----------------------------------------------------------
    @property
    def fn(self):
        """Current post-stream populations (reference fn, AoS [n*19])."""
        a = C.addressof(self.S.fn.contents)
        for buf in self.bufs.values():
            if buf.ctypes.data == a:
                return buf
        raise RuntimeError("fn points to no owned buffer")

    def comp_mask(self):
        return self.bc == 0

    def init_fluid(self, rho=None, u=None, v=None, w=None):
        conv = lambda a: None if a is None else np.ascontiguousarray(a, np.float64).reshape(-1)
        self._init = [conv(rho), conv(u), conv(v), conv(w)]
        lib().orc_flu_init(C.byref(self.P), C.byref(self.S), *[dp(a) for a in self._init])

    def node_coords(self):
        """Physical coordinates of the padded nodes (init1.cu:169-171)."""
        dx = self.P.dx
        ix = np.arange(self.nx)
        iy = np.arange(self.ny)
        iz = np.arange(self.nz)
        x = -self.P.lx / 2 - 1.5 * dx + ix * dx
        y = -self.P.ly / 2 - 1.5 * dx + iy * dx
        z = -self.P.lz / 2 - 1.5 * dx + iz * dx
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        return X.reshape(-1), Y.reshape(-1), Z.reshape(-1)

    def lbm_d3q19(self, it):
        lib().orc_lbm_step(C.byref(self.P), C.byref(self.S), it)

    # -- cells ---------------------------------------------------------------
    def set_cells(self, cells):
        """cells: dict with voff, eoff, x, y, z (SoA), ele [3Nt]."""
        voff = np.ascontiguousarray(cells["voff"], np.int32)
        eoff = np.ascontiguousarray(cells["eoff"], np.int32)
        nc, nv, ne = len(voff) - 1, int(voff[-1]), int(eoff[-1])
        x = np.ascontiguousarray(cells["x"], np.float64).copy()
        y = np.ascontiguousarray(cells["y"], np.float64).copy()
        z = np.ascontiguousarray(cells["z"], np.float64).copy()
        ele = np.ascontiguousarray(cells["ele"], np.int32)
        xd = np.stack([x, y, z], axis=1).reshape(-1).copy()
        topo = topology(voff, eoff, xd, ele)
        c = dict(voff=voff, eoff=eoff, x=x, y=y, z=z, ele=ele, xd=xd,
                 u=np.zeros(nv), v=np.zeros(nv), w=np.zeros(nv),
                 cfx=np.zeros(nv), cfy=np.zeros(nv), cfz=np.zeros(nv),
                 t=np.zeros(4 * ne), rg=np.zeros(ne), tp=np.zeros(2 * ne),
                 q=np.zeros(3 * nv), qi=np.zeros(3 * nv),
                 volume0=np.zeros(nc), volume=np.zeros(nc), area=np.zeros(nc), gx=np.zeros(3 * nc),
                 bicg_iters=np.zeros(nc, np.int32), **topo)
        if "gs_scale" in cells and cells["gs_scale"] is not None:
            c["gs_scale"] = np.ascontiguousarray(cells["gs_scale"], np.float64)
        else:
            c["gs_scale"] = None
        self.cells = c
        S = self.S
        S.nc = nc
        S.voff, S.eoff, S.ele, S.ec = ip(voff), ip(eoff), ip(ele), ip(c["ec"])
        for k in ("x", "y", "z", "u", "v", "w", "cfx", "cfy", "cfz", "xd", "c0", "s0", "nref",
                  "t", "rg", "tp", "q", "qi", "volume0", "volume", "area", "gx"):
            setattr(S, k, dp(c[k]))
        for k in ("nln", "ln", "lo", "nle", "le", "bicg_iters"):
            setattr(S, k, ip(c[k]))
        S.gs_scale = dp(c["gs_scale"]) if c["gs_scale"] is not None else None

    def stress(self, it):
        return lib().orc_stress(C.byref(self.P), C.byref(self.S), it)

    def calc_tension(self, it=1):
        lib().orc_calc_tension(C.byref(self.P), C.byref(self.S), it)

    def ibm_st1(self):
        lib().orc_ibm_st1(C.byref(self.P), C.byref(self.S))

    def ibm_st2(self):
        lib().orc_ibm_st2(C.byref(self.P), C.byref(self.S))

    def volume_init(self):
        lib().orc_volume_init(C.byref(self.P), C.byref(self.S))

    def volume_correction(self):
        return lib().orc_volume_correction(C.byref(self.P), C.byref(self.S))

    def calc_volume(self):
        nc = self.S.nc
        V, A = np.zeros(nc), np.zeros(nc)
        lib().orc_calc_volume(C.byref(self.P), C.byref(self.S), dp(V), dp(A))
        return V, A

    def assemble(self, cell):
        V = int(self.cells["voff"][cell + 1] - self.cells["voff"][cell])
        mn = lib().orc_expand2pow2(3 * V)
        ptr, index = np.zeros(mn, np.int32), np.zeros(7 * mn, np.int32)
        value, b = np.zeros(7 * mn), np.zeros(mn)
        lib().orc_assemble_cell(C.byref(self.P), C.byref(self.S), cell, ip(ptr), ip(index), dp(value), dp(b))
        return ptr, index, value, b

    def skalak_energy(self):
        return lib().orc_skalak_energy(C.byref(self.P), C.byref(self.S))

    def bending_energy(self):
        return lib().orc_bending_energy(C.byref(self.P), C.byref(self.S))

    def step(self, it):
        k = C.c_int(0)
        rc = lib().orc_step(C.byref(self.P), C.byref(self.S), it, C.byref(k))
        return rc, k.value

    # -- VOF indicator ---------------------------------------------------------
    def front_tracking(self):
        """vof_init / front_tracking: returns (status, SOR iterations, attempts)."""
        its, att = C.c_int(0), C.c_int(0)
        rc = lib().orc_front_tracking(C.byref(self.P), C.byref(self.S), C.byref(its), C.byref(att))
        return rc, its.value, att.value

    def thinc_wlic(self):
        lib().orc_thinc_wlic(C.byref(self.P), C.byref(self.S))

    def ft_gradient(self):
        g = [np.zeros(self.n) for _ in range(3)]
        lib().orc_ft_gradient(C.byref(self.P), C.byref(self.S), *[dp(a) for a in g])
        return g

    def ft_rhs(self, gx, gy, gz):
        b = np.zeros(self.n)
        lib().orc_ft_rhs(C.byref(self.P), ip(self.bc), *[dp(np.ascontiguousarray(a, np.float64)) for a in (gx, gy, gz)],
                         dp(b))
        return b

    def ft_sor(self, b, x0, pad0=0.0):
        x = np.ascontiguousarray(x0, np.float64).copy()
        its = C.c_int(0)
        rc = lib().orc_ft_sor(C.byref(self.P), ip(self.bc), dp(np.ascontiguousarray(b, np.float64)), dp(x),
                              float(pad0), C.byref(its))
        return rc, x, its.value

    def vof_normals(self, vf):
        out = [np.zeros(self.n) for _ in range(3)]
        lib().orc_vof_normals(C.byref(self.P), dp(np.ascontiguousarray(vf, np.float64)), *[dp(a) for a in out])
        return out

    # -- cytoadhesion ----------------------------------------------------------
    def adhesion_setup(self, params, is_knob):
        """adhesion_init with the given constants (dict of OrcAdhesion
        fields) and knob flags [Nv]; orc_step then runs Adhesion()."""
        A = OrcAdhesion()
        for k, v in params.items():
            setattr(A, k, v)
        nrec = lib().orc_adhesion_grid(C.byref(self.P), C.byref(A))
        nv = int(self.cells["voff"][-1])
        ns, mn = A.springs_per_knob, A.max_neighwall
        self._adh = dict(is_knob=np.ascontiguousarray(is_knob, np.uint8).copy(), id_wall=np.zeros(nv * mn, np.int32),
                         r_wall=np.zeros(nv * mn), attach=np.zeros(nv * ns, np.int32), r_ij=np.zeros(nv * ns),
                         is_attached=np.zeros(nrec, np.int32), f_adh=np.zeros(3 * nv))
        A.is_knob = self._adh["is_knob"].ctypes.data_as(C.POINTER(C.c_ubyte))
        for k in ("id_wall", "attach", "is_attached"):
            setattr(A, k, ip(self._adh[k]))
        for k in ("r_wall", "r_ij", "f_adh"):
            setattr(A, k, dp(self._adh[k]))
        self._adhs = A
        lib().orc_adhesion_init(C.byref(self.P), C.byref(self.S), C.byref(A))
        self.S.adh = C.addressof(A)
        return self._adh

    def adhesion_step(self, it):
        lib().orc_adhesion_step(C.byref(self.P), C.byref(self.S), C.byref(self._adhs), it)

    def set_body_force_field(self, gx=None, gy=None, gz=None):
        """Per-node body-force field (padded [n]) added to P e; None removes it."""
        if gx is None:
            self._gb = None
            self.S.gbx = self.S.gby = self.S.gbz = None
            return
        self._gb = [np.ascontiguousarray(a, np.float64).reshape(-1).copy() for a in (gx, gy, gz)]
        self.S.gbx, self.S.gby, self.S.gbz = (dp(a) for a in self._gb)

    def wall_shear(self, iter0, nsteps, window):
        """Mean over the last `window` steps of tau_w = 2 Df (dx/dt) (e . u) at
        the first fluid nodes of both y walls (main.cu loop with M = 1)."""
        P = self.P
        e = np.array([P.ex, P.ey, P.ez], float)
        e = e / np.linalg.norm(e) if np.linalg.norm(e) > 0 else np.array([1.0, 0.0, 0.0])
        shape = (self.nz, self.ny, self.nx)
        coef = 2.0 * P.df * P.dx / P.dt / (2.0 * (self.nx - 4) * (self.nz - 4)) / window
        J = 0.0
        for k in range(nsteps):
            rc, _ = self.step(iter0 + k + 1)
            if rc != 0:
                raise RuntimeError(f"oracle step failed ({rc})")
            if k >= nsteps - window:
                for iy in (2, self.ny - 3):
                    u = sum(e[d] * getattr(self, f)[...].reshape(shape)[2:-2, iy, 2:-2].sum()
                            for d, f in enumerate(("um", "vm", "wm")))
                    J += coef * u
        return J


def bicgstab(ptr, index, value, b, x0=None, itrmax=1000, residual=1e-12):
    n = len(b)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    its = C.c_int(0)
    rc = lib().orc_bicgstab(ip(ptr), ip(index), dp(value), dp(x), dp(np.ascontiguousarray(b, np.float64)),
                            n, itrmax, residual, C.byref(its))
    return rc, x, its.value


def thinc_flux(c, alpha, uc, isgn, m, dt, dx):
    return lib().orc_thinc_flux(float(c), float(alpha), float(uc), int(isgn), int(m), float(dt), float(dx))


def delta_weight(r):
    return lib().orc_delta_weight(float(r))
