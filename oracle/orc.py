"""ctypes binding of the CPU FP64 oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in __graft_entry__.py and by
the cpu_baseline leg of bench.py.  The product (the HIP library behind include/lbm_rb.h and
its Python host wrapper) never imports this module.

The oracle restates 2DRB/3DRB.F90 and 2DRB/2DRB.F90 (see rb_oracle.c for the per-routine
citations).  Arrays come back in the reference's Fortran order: populations as
(nz, ny, nx, q) C-order views of f(0:q-1, nx, ny, nz), macros as (nz, ny, nx).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

VEL_NOSLIP, VEL_PERIODIC = 0, 1
T_CONST, T_ADIABATIC, T_PERIODIC = 0, 1, 2


class _Cfg(C.Structure):
    _fields_ = [
        ("dims", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("rayleigh", C.c_double), ("prandtl", C.c_double), ("mach", C.c_double),
        ("thot", C.c_double), ("tcold", C.c_double),
        ("side_heated", C.c_int), ("vel_x", C.c_int), ("vel_z", C.c_int),
        ("temp_x", C.c_int), ("temp_y", C.c_int), ("temp_z", C.c_int), ("use_g", C.c_int),
        ("legacy", C.c_int),
    ]


class _Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "L", "Tref", "tauf", "viscosity", "diffusivity", "gbeta", "time_unit", "velocity_unit",
        "Se", "Seps", "Snu", "Spi", "Sq", "Sm", "taug", "Qk", "Qnu", "SG", "heat_flux_scale",
        "paraA")]


class _Diag(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "nu_vol", "re_vol", "nu_global", "nu_hot", "nu_cold", "nu_middle", "nu_hot_max",
        "nu_hot_max_pos", "nu_hot_min", "nu_hot_min_pos", "mass", "heat", "kinetic")]


class _Mid(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "u_max", "u_pos1", "u_pos2", "u_plane", "v_max", "v_pos1", "v_pos2", "v_plane",
        "w_max", "w_pos1", "w_pos2", "w_plane")]


def build() -> str:
    """Compile liborc.so with the committed Makefile (gcc, -ffp-contract=off, OpenMP)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_params_compute.argtypes = [C.POINTER(_Cfg), C.POINTER(_Params)]
        L.orc_create.argtypes = [C.POINTER(_Cfg)]
        L.orc_create.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_init_exact.argtypes = [C.c_void_p, dp]
        L.orc_init_seeded.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_double]
        L.orc_noise.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_noise.restype = C.c_double
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_itc.argtypes = [C.c_void_p]
        L.orc_itc.restype = C.c_long
        L.orc_get.argtypes = [C.c_void_p] + [dp] * 10
        L.orc_set_fg.argtypes = [C.c_void_p, dp, dp]
        L.orc_set_fg.restype = C.c_int
        L.orc_check.argtypes = [C.c_void_p, dp, dp]
        L.orc_diagnostics.argtypes = [C.c_void_p, C.POINTER(_Diag)]
        L.orc_midplane.argtypes = [C.c_void_p, C.POINTER(_Mid)]
        L.orc_midplane_slice.argtypes = [C.c_void_p, C.c_int] + [dp] * 6
        L.orc_psi_vort_2d.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_spline_eval.argtypes = [C.c_int, dp, dp, C.c_int, dp, dp]
        L.orc_lagrange4.argtypes = [C.c_double, dp, dp]
        L.orc_lagrange4.restype = C.c_double
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_threads.restype = C.c_int
        L.orc_fit_adiabatic_wall_T4.argtypes = [C.c_double, dp, dp, dp]
        L.orc_fit_parabola_ls5.argtypes = [dp, dp, C.c_int, dp, dp]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class Config:
    """Mirror of the reference's compile-time switches and parameters."""
    dims: int = 3
    nx: int = 16
    ny: int = 16
    nz: int = 16
    rayleigh: float = 1.0e6
    prandtl: float = 0.71
    mach: float = 0.1
    thot: float = 0.5
    tcold: float = -0.5
    side_heated: int = 0
    vel_x: int = VEL_NOSLIP
    vel_z: int = VEL_NOSLIP
    temp_x: int = T_ADIABATIC
    temp_y: int = T_CONST
    temp_z: int = T_ADIABATIC
    use_g: int = 1
    legacy: int = 0

    def c(self) -> _Cfg:
        nz = self.nz if self.dims == 3 else 1
        return _Cfg(self.dims, self.nx, self.ny, nz, self.rayleigh, self.prandtl, self.mach,
                    self.thot, self.tcold, self.side_heated, self.vel_x, self.vel_z,
                    self.temp_x, self.temp_y, self.temp_z, self.use_g, self.legacy)

    @property
    def qf(self):
        return 19 if self.dims == 3 else 9

    @property
    def qt(self):
        return 7 if self.dims == 3 else 5

    @property
    def shape(self):
        return (self.nz if self.dims == 3 else 1, self.ny, self.nx)


def params(cfg: Config) -> dict:
    p = _Params()
    c = cfg.c()
    lib().orc_params_compute(C.byref(c), C.byref(p))
    return {n: getattr(p, n) for n, _ in _Params._fields_}


def set_threads(n: int) -> int:
    """OpenMP thread count of the oracle's sweeps; returns the count in effect."""
    return lib().orc_set_threads(int(n))


def spline_eval(x, y, xq):
    x, y = np.ascontiguousarray(x, float), np.ascontiguousarray(y, float)
    xq = np.ascontiguousarray(xq, float)
    out = np.empty_like(xq)
    lib().orc_spline_eval(len(x), _ptr(x), _ptr(y), len(xq), _ptr(xq), _ptr(out))
    return out


def lagrange4(xq, xk, fk):
    xk, fk = np.ascontiguousarray(xk, float), np.ascontiguousarray(fk, float)
    return lib().orc_lagrange4(float(xq), _ptr(xk), _ptr(fk))


def noise(seed: int, instance: int, idx: int) -> float:
    return lib().orc_noise(seed, instance, idx)


def fit_adiabatic_wall_T4(y0, y, tt):
    y = np.ascontiguousarray(y, dtype=np.float64)
    tt = np.ascontiguousarray(tt, dtype=np.float64)
    out = np.zeros(1)
    lib().orc_fit_adiabatic_wall_T4(float(y0), _ptr(y), _ptr(tt), _ptr(out))
    return float(out[0])


def fit_parabola_ls5(y, f, mode):
    y = np.ascontiguousarray(y, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    fs, ys = np.zeros(1), np.zeros(1)
    lib().orc_fit_parabola_ls5(_ptr(y), _ptr(f), int(mode), _ptr(fs), _ptr(ys))
    return float(fs[0]), float(ys[0])


class Oracle:
    """One reference-equivalent simulation held by the C oracle."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self._c = cfg.c()
        self.h = lib().orc_create(C.byref(self._c))
        if not self.h:
            raise ValueError(f"oracle rejected config {cfg}")

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init_exact(self, T0=None):
        if T0 is not None:
            T0 = np.ascontiguousarray(T0, dtype=np.float64).reshape(self.cfg.shape)
        lib().orc_init_exact(self.h, _ptr(T0))

    def init_seeded(self, seed: int, amp: float = 1e-3, instance: int = 0):
        lib().orc_init_seeded(self.h, seed, instance, amp)

    def step(self, n: int = 1):
        lib().orc_step(self.h, int(n))

    @property
    def itc(self) -> int:
        return lib().orc_itc(self.h)

    def state(self) -> dict:
        sh, qf, qt = self.cfg.shape, self.cfg.qf, self.cfg.qt
        out = {"f": np.empty(sh + (qf,)), "g": np.empty(sh + (qt,))}
        for k in ("rho", "u", "v", "w", "T", "bx", "by", "bz"):
            out[k] = np.empty(sh)
        lib().orc_get(self.h, *[_ptr(out[k]) for k in
                                ("f", "g", "rho", "u", "v", "w", "T", "bx", "by", "bz")])
        return out

    def set_fg(self, f, g) -> int:
        f = np.ascontiguousarray(f, dtype=np.float64)
        g = np.ascontiguousarray(g, dtype=np.float64)
        return lib().orc_set_fg(self.h, _ptr(f), _ptr(g))

    def check(self):
        eu, et = np.zeros(1), np.zeros(1)
        lib().orc_check(self.h, _ptr(eu), _ptr(et))
        return float(eu[0]), float(et[0])

    def diagnostics(self) -> dict:
        d = _Diag()
        lib().orc_diagnostics(self.h, C.byref(d))
        return {n: getattr(d, n) for n, _ in _Diag._fields_}

    def midplane_slice(self, axis: int) -> dict:
        """3D mid-plane slice across axis 0/1/2: u, v, w (scaled), T, psi, vort as (n2, n1)."""
        n = (self.cfg.nx, self.cfg.ny, self.cfg.nz)
        n1 = n[1] if axis == 0 else n[0]
        n2 = n[1] if axis == 2 else n[2]
        out = {k: np.empty((n2, n1)) for k in ("u", "v", "w", "T", "psi", "vort")}
        lib().orc_midplane_slice(self.h, int(axis), *[_ptr(out[k]) for k in
                                                      ("u", "v", "w", "T", "psi", "vort")])
        return out

    def psi_vorticity_2d(self, stats: bool = True) -> dict:
        """2D stream function, vorticity (ny, nx) and the |psi| statistics."""
        sh = (self.cfg.ny, self.cfg.nx)
        psi, vort, st = np.empty(sh), np.empty(sh), np.zeros(5)
        lib().orc_psi_vort_2d(self.h, _ptr(psi), _ptr(vort), _ptr(st) if stats else None)
        out = {"psi": psi, "vort": vort}
        if stats:
            out.update(zip(("psi_center_coarse", "psi_center_fine", "psi_abs_max", "x_at_max",
                            "y_at_max"), st.tolist()))
        return out

    def midplane(self) -> dict:
        """Mid-plane velocity maxima (scaled by L / kappa) and their positions."""
        m = _Mid()
        lib().orc_midplane(self.h, C.byref(m))
        return {n: getattr(m, n) for n, _ in _Mid._fields_}
