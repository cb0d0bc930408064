"""Python host binding of the gfx950 solver (ctypes over include/lbm_rb.h).

This is the Python-side mirror of the reference's driver interface: a `Solver` owns one
domain and exposes the reference's phases by name (`initial`, `step`, `check`, `calNuRe` /
`diagnostics`, `backupData` / `gather`, `reload` / `scatter`).  Every call goes to the HIP
library librblbm.so; there is no CPU fallback.  If the library is missing the import of a
Solver fails with an error naming the build command.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, asdict

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "librblbm.so")

VEL_NOSLIP, VEL_PERIODIC = 0, 1
T_CONST, T_ADIABATIC, T_PERIODIC = 0, 1, 2
FP64, FP32_SHIFT, FP16_SHIFT = 0, 1, 2
KERNEL_AUTO, KERNEL_TWO_PASS, KERNEL_ONE_PASS = 0, 1, 2
LAYOUT_TWO_ARRAY, LAYOUT_AA, LAYOUT_AUTO = 0, 1, 2
TILE_32X16, TILE_16X16, TILE_32X8, TILE_32X32, TILE_34X15, TILE_66X15 = 0, 1, 2, 3, 4, 5
PREFETCH_NONE, PREFETCH_REGS, PREFETCH_STAGED = 0, 1, 2
KCHUNK_AUTO = -1
ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_STATE, ERR_UNSUPPORTED, ERR_RHO = -1, -2, -3, -4, -5, -6

EXPORTS = ("lbm_create", "lbm_destroy", "lbm_init_exact", "lbm_init_seeded", "lbm_step",
           "lbm_step_timed", "lbm_sync", "lbm_itc", "lbm_get_params", "lbm_macros", "lbm_check",
           "lbm_diagnostics", "lbm_midplane_extrema", "lbm_gather", "lbm_scatter", "lbm_bytes_per_update",
           "lbm_set_kernel_variant", "lbm_set_fused_tile", "lbm_get_kernel_form", "lbm_get_layout",
           "lbm_last_error", "lbm_error_string")


class LbmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lbm error {code}: {msg}")
        self.code = code


class _Config(C.Structure):
    _fields_ = [("dims", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("rayleigh", C.c_double), ("prandtl", C.c_double), ("mach", C.c_double),
                ("thot", C.c_double), ("tcold", C.c_double), ("side_heated", C.c_int32),
                ("vel_x", C.c_int32), ("vel_z", C.c_int32), ("temp_x", C.c_int32),
                ("temp_y", C.c_int32), ("temp_z", C.c_int32), ("use_g", C.c_int32),
                ("precision", C.c_int32), ("device", C.c_int32), ("legacy_thermal", C.c_int32),
                ("layout", C.c_int32)]


class _Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "L", "tref", "tauf", "viscosity", "diffusivity", "gbeta", "time_unit", "velocity_unit",
        "s_e", "s_eps", "s_nu", "s_pi", "s_q", "s_m", "taug", "q_k", "q_nu", "sg", "para_a")]


class _Diag(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "nu_vol", "re_vol", "nu_global", "nu_hot", "nu_cold", "nu_middle", "nu_hot_max",
        "nu_hot_max_pos", "nu_hot_min", "nu_hot_min_pos", "mass", "heat", "kinetic")]


class _Mid(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "u_max", "u_pos1", "u_pos2", "u_plane", "v_max", "v_pos1", "v_pos2", "v_plane",
        "w_max", "w_pos1", "w_pos2", "w_plane")]


_lib = None


def library(path: str = LIB_PATH):
    """Load librblbm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `python ./-2d-3d-_amd/build.py` "
            "(or __graft_entry__.build()); there is no fallback path")
    L = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    h = C.c_void_p
    L.lbm_create.argtypes = [C.POINTER(_Config), C.POINTER(h)]
    L.lbm_destroy.argtypes = [h]
    L.lbm_init_exact.argtypes = [h, dp]
    L.lbm_init_seeded.argtypes = [h, C.c_uint64, C.c_double]
    L.lbm_step.argtypes = [h, C.c_int32]
    L.lbm_step_timed.argtypes = [h, C.c_int32, dp]
    L.lbm_sync.argtypes = [h]
    L.lbm_itc.argtypes = [h, C.POINTER(C.c_int64)]
    L.lbm_get_params.argtypes = [h, C.POINTER(_Params)]
    L.lbm_macros.argtypes = [h, dp, dp, dp, dp, dp]
    L.lbm_check.argtypes = [h, dp, dp]
    L.lbm_diagnostics.argtypes = [h, C.POINTER(_Diag)]
    L.lbm_midplane_extrema.argtypes = [h, C.POINTER(_Mid)]
    L.lbm_gather.argtypes = [h, dp, dp, dp]
    L.lbm_scatter.argtypes = [h, dp, dp]
    L.lbm_bytes_per_update.argtypes = [h, dp]
    L.lbm_set_kernel_variant.argtypes = [h, C.c_int32, C.c_int32]
    L.lbm_set_fused_tile.argtypes = [h, C.c_int32, C.c_int32]
    ip = C.POINTER(C.c_int32)
    L.lbm_get_kernel_form.argtypes = [h, ip, ip, ip, ip]
    L.lbm_get_layout.argtypes = [h, ip]
    L.lbm_last_error.argtypes = [h]
    L.lbm_last_error.restype = C.c_char_p
    L.lbm_error_string.argtypes = [C.c_int]
    L.lbm_error_string.restype = C.c_char_p
    for name in EXPORTS:
        if name not in ("lbm_last_error", "lbm_error_string"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class Config:
    """Run-time form of the reference's compile-time switches (2DRB/3DRB.F90:9-136)."""
    dims: int = 3
    nx: int = 128
    ny: int = 128
    nz: int = 128
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
    precision: int = FP64
    device: int = 0
    legacy: int = 0           # EnableLegacyThermalScheme (FP64 form only)
    layout: int = 0           # LAYOUT_TWO_ARRAY / LAYOUT_AA / LAYOUT_AUTO

    @property
    def shape(self):
        return (self.nz if self.dims == 3 else 1, self.ny, self.nx)

    @property
    def qf(self):
        return 19 if self.dims == 3 else 9

    @property
    def qt(self):
        return 7 if self.dims == 3 else 5

    def c(self) -> _Config:
        d = asdict(self)
        d["legacy_thermal"] = d.pop("legacy")
        return _Config(**d)


class Solver:
    """One GPU domain.  Methods follow the reference driver's phases."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self.lib = library()
        self._h = C.c_void_p()
        self._c = cfg.c()
        rc = self.lib.lbm_create(C.byref(self._c), C.byref(self._h))
        if rc != 0:
            raise LbmError(rc, self.lib.lbm_error_string(rc).decode())

    def _chk(self, rc):
        if rc != 0:
            msg = self.lib.lbm_last_error(self._h).decode() or self.lib.lbm_error_string(rc).decode()
            raise LbmError(rc, msg)

    def close(self):
        if self._h:
            self.lib.lbm_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- initial (2DRB/3DRB.F90:342-643) -----------------------------------------------------
    def initial(self, T0=None):
        if T0 is not None:
            T0 = np.ascontiguousarray(T0, dtype=np.float64).reshape(self.cfg.shape)
        self._chk(self.lib.lbm_init_exact(self._h, _p(T0)))

    def initial_seeded(self, seed: int, amp: float = 1.0e-3):
        self._chk(self.lib.lbm_init_seeded(self._h, seed, amp))

    # -- time loop body ---------------------------------------------------------------------
    def step(self, n: int = 1):
        self._chk(self.lib.lbm_step(self._h, int(n)))

    def step_timed(self, n: int) -> float:
        ms = np.zeros(1)
        self._chk(self.lib.lbm_step_timed(self._h, int(n), _p(ms)))
        return float(ms[0])

    def sync(self):
        self._chk(self.lib.lbm_sync(self._h))

    @property
    def itc(self) -> int:
        v = C.c_int64()
        self._chk(self.lib.lbm_itc(self._h, C.byref(v)))
        return v.value

    def params(self) -> dict:
        p = _Params()
        self._chk(self.lib.lbm_get_params(self._h, C.byref(p)))
        return {n: getattr(p, n) for n, _ in _Params._fields_}

    def set_kernel_variant(self, variant: int, kchunk: int = 0):
        """KERNEL_AUTO / KERNEL_TWO_PASS / KERNEL_ONE_PASS; kchunk: z-planes per block."""
        self._chk(self.lib.lbm_set_kernel_variant(self._h, int(variant), int(kchunk)))

    def set_fused_tile(self, tile: int, prefetch: int = 0):
        """Thread tile of the one-pass form: TILE_32X16 / TILE_16X16 / TILE_32X8 / TILE_34X15, and
        TILE_32X32 / TILE_66X15 (FP32, FP16 only); prefetch PREFETCH_NONE / PREFETCH_REGS /
        PREFETCH_STAGED (FP32 with TILE_34X15 or TILE_66X15 only)."""
        self._chk(self.lib.lbm_set_fused_tile(self._h, int(tile), int(prefetch)))

    def kernel_form(self) -> dict:
        """The kernel form in effect: variant, kchunk, tile, prefetch."""
        v = [C.c_int32() for _ in range(4)]
        self._chk(self.lib.lbm_get_kernel_form(self._h, *[C.byref(x) for x in v]))
        return dict(zip(("variant", "kchunk", "tile", "prefetch"), (x.value for x in v)))

    @property
    def layout(self) -> int:
        """The population layout in effect: LAYOUT_TWO_ARRAY or LAYOUT_AA."""
        v = C.c_int32()
        self._chk(self.lib.lbm_get_layout(self._h, C.byref(v)))
        return v.value

    def autotune(self, cache: str | None = None, steps: int = 10) -> float:
        """Pick the kernel form and z-chunk for this shape (lbm_autotune); returns ms/step."""
        L = self.lib
        L.lbm_autotune.argtypes = [C.c_void_p, C.c_char_p, C.c_int32, C.POINTER(C.c_double)]
        ms = C.c_double()
        self._chk(L.lbm_autotune(self._h, cache.encode() if cache else None, int(steps),
                                 C.byref(ms)))
        return ms.value

    def bytes_per_update(self) -> float:
        b = np.zeros(1)
        self._chk(self.lib.lbm_bytes_per_update(self._h, _p(b)))
        return float(b[0])

    # -- observation ------------------------------------------------------------------------
    def macros(self) -> dict:
        out = {k: np.empty(self.cfg.shape) for k in ("rho", "u", "v", "w", "T")}
        self._chk(self.lib.lbm_macros(self._h, *[_p(out[k]) for k in ("rho", "u", "v", "w", "T")]))
        return out

    def check(self):
        eu, et = np.zeros(1), np.zeros(1)
        self._chk(self.lib.lbm_check(self._h, _p(eu), _p(et)))
        return float(eu[0]), float(et[0])

    def diagnostics(self) -> dict:
        d = _Diag()
        self._chk(self.lib.lbm_diagnostics(self._h, C.byref(d)))
        return {n: getattr(d, n) for n, _ in _Diag._fields_}

    def run_diagnostics(self, nsteps: int, every: int) -> list:
        """nsteps steps with the diagnostics every `every` steps, the periods replayed from a
        HIP graph (lbm_run_diagnostics); one dict per sample."""
        L = self.lib
        L.lbm_run_diagnostics.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(_Diag)]
        ns = nsteps // every if every > 0 else 0
        out = (_Diag * max(ns, 1))()
        self._chk(L.lbm_run_diagnostics(self._h, int(nsteps), int(every), out))
        return [{n: getattr(d, n) for n, _ in _Diag._fields_} for d in out[:ns]]

    def midplane_slice(self, axis: int) -> dict:
        """3D mid-plane slice across axis 0/1/2 (lbm_midplane_slice): u, v, w scaled by
        L / kappa, T, stream function and vorticity, each as (n2, n1) arrays."""
        n = (self.cfg.nx, self.cfg.ny, self.cfg.nz)
        n1 = n[1] if axis == 0 else n[0]
        n2 = n[1] if axis == 2 else n[2]
        out = {k: np.empty((n2, n1)) for k in ("u", "v", "w", "T", "psi", "vort")}
        L = self.lib
        if not getattr(L, "_slice_bound", False):
            L.lbm_midplane_slice.argtypes = [C.c_void_p, C.c_int32] + [C.POINTER(C.c_double)] * 6
            L.lbm_midplane_slice.restype = C.c_int
            L._slice_bound = True
        self._chk(L.lbm_midplane_slice(self._h, int(axis), *[_p(out[k]) for k in
                                                            ("u", "v", "w", "T", "psi", "vort")]))
        return out

    def psi_vorticity_2d(self, stats: bool = True) -> dict:
        """2D stream function and vorticity (ny, nx) and the |psi| statistics
        (lbm_psi_vorticity_2d)."""
        L = self.lib
        if not getattr(L, "_psi_bound", False):
            L.lbm_psi_vorticity_2d.argtypes = [C.c_void_p, C.POINTER(C.c_double),
                                               C.POINTER(C.c_double), C.POINTER(C.c_double)]
            L.lbm_psi_vorticity_2d.restype = C.c_int
            L._psi_bound = True
        sh = (self.cfg.ny, self.cfg.nx)
        psi, vort, st = np.empty(sh), np.empty(sh), np.zeros(5)
        self._chk(L.lbm_psi_vorticity_2d(self._h, _p(psi), _p(vort), _p(st) if stats else None))
        out = {"psi": psi, "vort": vort}
        if stats:
            out.update(zip(("psi_center_coarse", "psi_center_fine", "psi_abs_max", "x_at_max",
                            "y_at_max"), st.tolist()))
        return out

    def midplane(self) -> dict:
        """Mid-plane velocity maxima (scaled by L / kappa) and their positions."""
        m = _Mid()
        self._chk(self.lib.lbm_midplane_extrema(self._h, C.byref(m)))
        return {n: getattr(m, n) for n, _ in _Mid._fields_}

    def gather(self, with_bprev: bool = True) -> dict:
        sh = self.cfg.shape
        f = np.empty(sh + (self.cfg.qf,))
        g = np.empty(sh + (self.cfg.qt,))
        b = np.empty((3,) + sh) if with_bprev else None
        self._chk(self.lib.lbm_gather(self._h, _p(f), _p(g), _p(b)))
        out = {"f": f, "g": g}
        if b is not None:
            out["bx"], out["by"], out["bz"] = b[0], b[1], b[2]
        return out

    def scatter(self, f, g):
        f = np.ascontiguousarray(f, dtype=np.float64)
        g = np.ascontiguousarray(g, dtype=np.float64)
        self._chk(self.lib.lbm_scatter(self._h, _p(f), _p(g)))


BATCH_PER_DOMAIN, BATCH_ONE_LAUNCH, BATCH_GRAPH = 0, 1, 2


class _Domain(Solver):
    """A domain owned by a Batch: the ordinary per-domain calls, no own lifetime."""

    def __init__(self, cfg: Config, handle: C.c_void_p, lib):
        self.cfg, self.lib, self._h = cfg, lib, handle

    def close(self):
        self._h = C.c_void_p()


class Batch:
    """Independent cavities advancing together on one device (lbm_batch_* entry points)."""

    def __init__(self, cfgs):
        self.lib = library()
        L = self.lib
        if not hasattr(L, "_batch_bound"):
            h = C.c_void_p
            L.lbm_batch_create.argtypes = [C.POINTER(_Config), C.c_int32, C.POINTER(h)]
            L.lbm_batch_destroy.argtypes = [h]
            L.lbm_batch_domain.argtypes = [h, C.c_int32, C.POINTER(h)]
            L.lbm_batch_init_seeded.argtypes = [h, C.c_uint64, C.c_double]
            L.lbm_batch_set_mode.argtypes = [h, C.c_int32]
            L.lbm_batch_step.argtypes = [h, C.c_int32]
            L.lbm_batch_step_timed.argtypes = [h, C.c_int32, C.POINTER(C.c_double)]
            L.lbm_batch_nodes.argtypes = [h, C.POINTER(C.c_int64)]
            L._batch_bound = True
        self.cfgs = list(cfgs)
        arr = (_Config * len(self.cfgs))(*[c.c() for c in self.cfgs])
        self._b = C.c_void_p()
        rc = L.lbm_batch_create(arr, len(self.cfgs), C.byref(self._b))
        if rc != 0:
            raise LbmError(rc, L.lbm_error_string(rc).decode())
        self.domains = []
        for i, c in enumerate(self.cfgs):
            hd = C.c_void_p()
            L.lbm_batch_domain(self._b, i, C.byref(hd))
            self.domains.append(_Domain(c, hd, L))

    def _chk(self, rc):
        if rc != 0:
            raise LbmError(rc, self.lib.lbm_error_string(rc).decode())

    def close(self):
        if self._b:
            for d in self.domains:
                d.close()
            self.lib.lbm_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initial_seeded(self, seed: int, amp: float = 1.0e-3):
        self._chk(self.lib.lbm_batch_init_seeded(self._b, seed, amp))

    def set_mode(self, mode: int):
        self._chk(self.lib.lbm_batch_set_mode(self._b, int(mode)))

    def set_kernel_variant(self, variant: int):
        L = self.lib
        L.lbm_batch_set_kernel_variant.argtypes = [C.c_void_p, C.c_int32]
        self._chk(L.lbm_batch_set_kernel_variant(self._b, int(variant)))

    def run_diagnostics(self, nsteps: int, every: int) -> list:
        """lbm_batch_run_diagnostics: result[k][i] is sample k of domain i."""
        L = self.lib
        L.lbm_batch_run_diagnostics.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(_Diag)]
        ns, n = (nsteps // every if every > 0 else 0), len(self.domains)
        out = (_Diag * max(ns * n, 1))()
        self._chk(L.lbm_batch_run_diagnostics(self._b, int(nsteps), int(every), out))
        fields = [f for f, _ in _Diag._fields_]
        return [[{f: getattr(out[k * n + i], f) for f in fields} for i in range(n)] for k in range(ns)]

    def autotune(self, cache: str | None = None, steps: int = 10):
        L = self.lib
        L.lbm_batch_autotune.argtypes = [C.c_void_p, C.c_char_p, C.c_int32]
        self._chk(L.lbm_batch_autotune(self._b, cache.encode() if cache else None, int(steps)))

    def step(self, n: int = 1):
        self._chk(self.lib.lbm_batch_step(self._b, int(n)))

    def step_timed(self, n: int) -> float:
        ms = C.c_double()
        self._chk(self.lib.lbm_batch_step_timed(self._b, int(n), C.byref(ms)))
        return ms.value

    @property
    def nodes(self) -> int:
        v = C.c_int64()
        self._chk(self.lib.lbm_batch_nodes(self._b, C.byref(v)))
        return v.value
