"""The full-size benchmark configurations, checked through properties that do not need a CPU run
of the same size:

  512^3, Ra = 1e8, FP64 and FP32-shift, 300 steps:
    - the one-pass, two-pass and in-place two-pass forms give bitwise equal macros;
    - mass of the closed cavity conserved to 1e-12 (FP64);
    - an x-symmetric start stays x-symmetric (u antisymmetric; rho, v, w, T symmetric) to 1e-12
      of the free-fall velocity, of dT and of rho (measured: 3e-14 and below) (the sums of the collision associate
      differently under the mirror, so the symmetry holds to rounding, not bitwise);
    - the FP32-shift form tracks the FP64 form: Nu_vol, Nu_hot, Nu_cold and Re within 1e-5
      relative after 300 steps (measured: 1.0e-6, 1.2e-6, 1.2e-6, 1.8e-6).
  the seeded 64-cavity ensemble (edges 32-128, Ra log-uniform 1e4..1e7, Pr 0.71 / 7) in FP32-shift,
  HIP-graph table launches, 100 steps:
    - every cavity conserves mass to 1e-9;
    - each of the 22 cavities of edge 32 and 48 tracks the FP64 oracle within the FP32 bars of
      test_gpu_lowprec (u 1e-4 of max|u|, T 2e-5 of dT).
  1024^3, Ra = 1e10, FP16-shift, in place, 100 steps from the seeded start:
    - finite diagnostics, mass within 1e-6 (FP16 rounding of the deviations is not conservative);
    - Nu_vol, Nu_hot, Nu_cold and Re within 1e-3 of the FP32-shift form at the same resolution
      (the storage precision alone);
  1024^3, FP16-shift, in place, from the conduction profile at Ra = 0.01, Ma = 0.001:
    - Nu_hot = Nu_cold = Nu_global = 1 within 1e-3 after 100 steps.
Each domain is freed before the next is created.
"""
import gc

import numpy as np
import pytest

import ensemble
import orc
import rblbm

pytestmark = pytest.mark.gpu

SEED = 20261203
N3 = 512


def _cavity(n, ra, mach=0.1, **kw):
    return rblbm.Config(dims=3, nx=n, ny=n, nz=n, rayleigh=ra, prandtl=0.71, mach=mach, **kw)


def _mirror_T0(cfg):
    k, j, i = np.ogrid[0:cfg.nz, 0:cfg.ny, 0:cfg.nx]
    x, y, z = (i + 0.5) / cfg.nx, (j + 0.5) / cfg.ny, (k + 0.5) / cfg.nz
    dT = cfg.thot - cfg.tcold
    T0 = cfg.thot - y * dT + 2e-3 * dT * np.cos(2 * np.pi * x) * np.sin(np.pi * y) * (1 + 0.5 * np.cos(2 * np.pi * z))
    return np.ascontiguousarray(np.broadcast_to(T0, cfg.shape))


def test_c3_512_fp64_forms_mass_and_mirror_symmetry():
    cfg = _cavity(N3, 1e8)
    T0 = _mirror_T0(cfg)
    forms = [("one-pass", rblbm.LAYOUT_TWO_ARRAY, rblbm.KERNEL_AUTO),
             ("two-pass", rblbm.LAYOUT_TWO_ARRAY, rblbm.KERNEL_TWO_PASS),
             ("in-place", rblbm.LAYOUT_AA, rblbm.KERNEL_TWO_PASS)]
    ref = None
    for name, layout, variant in forms:
        c = _cavity(N3, 1e8, layout=layout)
        sol = rblbm.Solver(c)
        sol.set_kernel_variant(variant)
        if name == "one-pass":
            assert sol.kernel_form()["variant"] == rblbm.KERNEL_ONE_PASS
        sol.initial(T0)
        sol_params = sol.params()
        m0 = sol.diagnostics()["mass"]
        sol.step(300)
        d = sol.diagnostics()
        m = sol.macros()
        sol.close()
        del sol
        gc.collect()
        print(name, "mass rel change", (d["mass"] - m0) / m0, "nu_vol", d["nu_vol"], flush=True)
        assert abs(d["mass"] - m0) <= 1e-12 * m0, name
        if ref is None:
            ref = m
            # mirror x -> nx - 1 - x: u odd, the other fields even
            uff = sol_params["velocity_unit"]
            for key, sign, scale in (("u", -1.0, uff), ("v", 1.0, uff), ("w", 1.0, uff),
                                     ("T", 1.0, cfg.thot - cfg.tcold), ("rho", 1.0, 1.0)):
                a = m[key]
                err = np.abs(a - sign * a[:, :, ::-1]).max() / scale
                print("mirror", key, err, "max", np.abs(a).max(), flush=True)
                assert err <= 1e-12, (key, err)
            assert np.abs(m["u"]).max() > 0
        else:
            for key in ("rho", "u", "v", "w", "T"):
                np.testing.assert_array_equal(m[key], ref[key], err_msg=f"{name} {key}")
        del m
        gc.collect()


def test_c3_512_fp32_one_pass_forms_match_two_pass():
    """At 512^3 the FP32-shift form runs the one-pass kernel with the 66x15 tile and no
    prefetching by default; 100 steps of it, and of the staged-load form, equal 100 steps of the
    two-pass form bit for bit."""
    ref = None
    for variant, prefetch in ((rblbm.KERNEL_AUTO, None), (rblbm.KERNEL_ONE_PASS, rblbm.PREFETCH_STAGED),
                              (rblbm.KERNEL_TWO_PASS, None)):
        sol = rblbm.Solver(_cavity(N3, 1e8, precision=rblbm.FP32_SHIFT))
        sol.set_kernel_variant(variant)
        if variant == rblbm.KERNEL_AUTO:
            assert sol.kernel_form() == dict(variant=rblbm.KERNEL_ONE_PASS, kchunk=32,
                                             tile=rblbm.TILE_66X15,
                                             prefetch=rblbm.PREFETCH_NONE)
        if prefetch is not None:
            sol.set_fused_tile(rblbm.TILE_66X15, prefetch)
        sol.initial_seeded(SEED, 1e-3)
        sol.step(100)
        m = sol.macros()
        sol.close()
        del sol
        gc.collect()
        if ref is None:
            ref = m
        else:
            for key in ("rho", "u", "v", "w", "T"):
                np.testing.assert_array_equal(m[key], ref[key], err_msg=key)
        del m
        gc.collect()


def test_c3_512_fp32_shift_tracks_fp64():
    out = {}
    for prec in (rblbm.FP64, rblbm.FP32_SHIFT):
        sol = rblbm.Solver(_cavity(N3, 1e8, precision=prec))
        sol.initial_seeded(SEED, 1e-3)
        sol.step(300)
        out[prec] = sol.diagnostics()
        sol.close()
        del sol
        gc.collect()
    a, b = out[rblbm.FP32_SHIFT], out[rblbm.FP64]
    print({k: (a[k], b[k]) for k in ("nu_vol", "nu_hot", "nu_cold", "re_vol")}, flush=True)
    for k in ("nu_vol", "nu_hot", "nu_cold"):
        assert abs(a[k] - b[k]) <= 1e-5 * abs(b[k]), (k, a[k], b[k])
    assert abs(a["re_vol"] - b["re_vol"]) <= 1e-5 * b["re_vol"]
    assert abs(a["mass"] - b["mass"]) <= 1e-9 * b["mass"]


def test_c4_64_cavity_ensemble_fp32_graph():
    cfgs = ensemble.ensemble(64, SEED, rblbm.FP32_SHIFT)
    b = rblbm.Batch(cfgs)
    b.set_mode(rblbm.BATCH_GRAPH)
    b.initial_seeded(SEED, 1e-3)
    m0 = [d.diagnostics()["mass"] for d in b.domains]
    b.step(100)
    small = 0
    for i, (c, d) in enumerate(zip(cfgs, b.domains)):
        dg = d.diagnostics()
        assert abs(dg["mass"] - m0[i]) <= 1e-9 * m0[i], (i, c.nx)
        assert np.isfinite(dg["nu_vol"]) and np.isfinite(dg["re_vol"])
        if c.nx > 48:
            continue
        small += 1
        o = orc.Oracle(orc.Config(dims=3, nx=c.nx, ny=c.ny, nz=c.nz, rayleigh=c.rayleigh,
                                  prandtl=c.prandtl, mach=c.mach))
        o.init_seeded(SEED, 1e-3, instance=i)
        o.step(100)
        s, m = o.state(), d.macros()
        umax = max(np.abs(s[k]).max() for k in ("u", "v", "w"))
        du = max(np.abs(m[k] - s[k]).max() for k in ("u", "v", "w")) / umax
        dT = np.abs(m["T"] - s["T"]).max() / (c.thot - c.tcold)
        assert du <= 1e-4 and dT <= 2e-5, (i, c.nx, du, dT)
        o.close()
    assert small == 22
    b.close()


def _c5_run(prec, seeded=True, **kw):
    cfg = _cavity(1024, kw.pop("ra", 1e10), precision=prec, layout=rblbm.LAYOUT_AA, **kw)
    sol = rblbm.Solver(cfg)
    assert sol.layout == rblbm.LAYOUT_AA
    if seeded:
        sol.initial_seeded(SEED, 1e-3)
    else:   # conduction profile between the isothermal walls
        y = (np.arange(cfg.ny) + 0.5) / cfg.ny
        sol.initial(np.ascontiguousarray(np.broadcast_to(
            (cfg.thot - y * (cfg.thot - cfg.tcold))[None, :, None], cfg.shape)))
    m0 = sol.diagnostics()["mass"]
    sol.step(100)
    d = sol.diagnostics()
    sol.close()
    del sol
    gc.collect()
    return m0, d


def test_c5_1024_fp16_in_place_tracks_fp32():
    # the storage precision alone: FP16-shift against FP32-shift at the same resolution (FP64 at
    # 1024^3 would need 550 GB); measured after 100 steps: Nu 1.8e-5 - 3.6e-5, Re 1.6e-5 relative
    # (the 10,000-step accuracy study at this configuration: time-averaged Nu 1.2e-3 / 1.7e-3)
    m0, h = _c5_run(rblbm.FP16_SHIFT)
    _, f = _c5_run(rblbm.FP32_SHIFT)
    keys = ("nu_vol", "nu_hot", "nu_cold", "re_vol", "kinetic")
    print("fp16", {k: h[k] for k in keys}, "\nfp32", {k: f[k] for k in keys}, flush=True)
    assert all(np.isfinite(h[k]) for k in keys)
    # FP16 deviations round without conserving mass exactly (4e-8 after 100 steps measured)
    assert abs(h["mass"] - m0) <= 1e-6 * m0
    for k in ("nu_vol", "nu_hot", "nu_cold"):
        assert abs(h[k] - f[k]) <= 1e-3 * abs(f[k]), (k, h[k], f[k])
    assert abs(h["re_vol"] - f["re_vol"]) <= 1e-3 * f["re_vol"]


def test_c5_1024_fp16_conduction_shift_keeps_unit_nusselt():
    # The FP16 form stores deviations from the conduction profile, so that profile is held to
    # rounding level at full size.  Ra and Ma far below onset keep the discrete hydrostatic flow
    # negligible (at Ra = 1e10 the y-varying buoyancy of the profile drives a transient that
    # moves Nu_hot to 1.2 within 100 steps in any precision).
    m0, d = _c5_run(rblbm.FP16_SHIFT, seeded=False, ra=0.01, mach=0.001)
    print("conduction", {k: d[k] for k in ("nu_hot", "nu_cold", "nu_global", "mass")}, flush=True)
    for k in ("nu_hot", "nu_cold", "nu_global"):
        assert d[k] == pytest.approx(1.0, abs=1e-3), (k, d[k])
    assert abs(d["mass"] - m0) <= 1e-6 * m0
