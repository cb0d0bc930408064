"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Bar (FP64 form): relative max-norm <= 1e-12 on f, g, u, T (max |a - b| / max |b|) and
relative error <= 1e-10 on the Nusselt / Reynolds diagnostics, on the same seeded inputs.
"""
import numpy as np
import pytest

import orc
import rblbm

pytestmark = pytest.mark.gpu

REL_FIELDS = 1e-12
REL_DIAG = 1e-10


def _cfgs(dims, nx, ny, nz, **kw):
    o = orc.Config(dims=dims, nx=nx, ny=ny, nz=nz, **kw)
    g = rblbm.Config(dims=dims, nx=nx, ny=ny, nz=nz, **kw)
    return o, g


def _relmax(a, b):
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den > 0 else 1.0)


def _compare_state(sol, ora, label, tol=REL_FIELDS):
    s = ora.state()
    m = sol.macros()
    st = sol.gather()
    errs = {
        "f": _relmax(st["f"], s["f"]), "g": _relmax(st["g"], s["g"]),
        "u": _relmax(m["u"], s["u"]), "v": _relmax(m["v"], s["v"]),
        "T": _relmax(m["T"], s["T"]), "rho": _relmax(m["rho"], s["rho"]),
        "bx": _relmax(st["bx"], s["bx"]), "by": _relmax(st["by"], s["by"]),
    }
    if sol.cfg.dims == 3:
        errs["w"] = _relmax(m["w"], s["w"])
    print(label, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= tol, (label, k, v)


def _run_pair(dims, nx, ny, nz, steps, init="seeded", T0=None, **kw):
    oc, gc = _cfgs(dims, nx, ny, nz, **kw)
    ora = orc.Oracle(oc)
    sol = rblbm.Solver(gc)
    if init == "seeded":
        ora.init_seeded(20261203, 1e-3)
        sol.initial_seeded(20261203, 1e-3)
    else:
        ora.init_exact(T0)
        sol.initial(T0)
    done = 0
    for n in steps:
        ora.step(n - done)
        sol.step(n - done)
        done = n
        _compare_state(sol, ora, f"{dims}D {nx}x{ny}x{nz} {kw} step {n}")
    return sol, ora


CASES_3D = {
    "rb_noslip": dict(),
    "periodic_xz": dict(vel_x=1, vel_z=1, temp_x=2, temp_z=2),
    "side_heated": dict(side_heated=1, temp_x=0, temp_y=1, temp_z=1),
    "periodic_x": dict(vel_x=1, temp_x=2),
    "no_m1g": dict(use_g=0),
}


def test_seeded_initial_state_bitwise():
    oc, gc = _cfgs(3, 12, 10, 8, rayleigh=1e6)
    ora, sol = orc.Oracle(oc), rblbm.Solver(gc)
    ora.init_seeded(7, 1e-3)
    sol.initial_seeded(7, 1e-3)
    s, m, st = ora.state(), sol.macros(), sol.gather()
    np.testing.assert_array_equal(m["T"], s["T"])
    np.testing.assert_array_equal(st["g"], s["g"])
    np.testing.assert_array_equal(st["f"], s["f"])


@pytest.mark.parametrize("name", list(CASES_3D))
def test_3d_small_matches_oracle(name):
    rng = np.random.default_rng(1)
    T0 = rng.uniform(-0.3, 0.3, size=(8, 10, 12))
    sol, ora = _run_pair(3, 12, 10, 8, (1, 2, 10, 60), init="exact", T0=T0, rayleigh=1e5,
                         mach=0.3, **CASES_3D[name])
    # The FP64 form keeps the reference's expression order without contraction: bitwise equal.
    np.testing.assert_array_equal(sol.gather()["f"], ora.state()["f"])


CASES_2D = {
    "rb_periodic": dict(vel_x=1, temp_x=2),
    "rb_walls": dict(vel_x=0, temp_x=1),
    "side_heated": dict(side_heated=1, vel_x=0, temp_x=0, temp_y=1),
    "no_m1g": dict(vel_x=1, temp_x=2, use_g=0),
}


@pytest.mark.parametrize("name", list(CASES_2D))
def test_2d_small_matches_oracle(name):
    rng = np.random.default_rng(2)
    T0 = rng.uniform(0.2, 0.8, size=(1, 10, 17))
    sol, ora = _run_pair(2, 17, 10, 1, (1, 2, 10, 60), init="exact", T0=T0, rayleigh=1e5,
                         mach=0.3, thot=1.0, tcold=0.0, **CASES_2D[name])
    np.testing.assert_array_equal(sol.gather()["f"], ora.state()["f"])


def test_reference_init_profile_and_perturbation():
    # Ra <= 1e4 triggers the reference's cos*sin perturbation of the linear profile.
    _run_pair(3, 16, 12, 10, (1, 20), init="exact", rayleigh=1e4)
    _run_pair(2, 24, 12, 1, (1, 20), init="exact", rayleigh=5e3, vel_x=1, temp_x=2)


def _compare_diag(sol, ora):
    d, e = sol.diagnostics(), ora.diagnostics()
    for k in ("nu_vol", "re_vol", "nu_global", "nu_hot", "nu_cold", "nu_middle", "nu_hot_max",
              "nu_hot_min", "mass", "heat", "kinetic"):
        den = abs(e[k]) if abs(e[k]) > 1e-300 else 1.0
        assert abs(d[k] - e[k]) / den <= REL_DIAG, (k, d[k], e[k])
    for k in ("nu_hot_max_pos", "nu_hot_min_pos"):
        assert d[k] == pytest.approx(e[k], rel=1e-9, abs=1e-12), k
    eu, et = sol.check()
    fu, ft = ora.check()
    assert eu == pytest.approx(fu, rel=REL_DIAG) and et == pytest.approx(ft, rel=REL_DIAG)


def test_diagnostics_match_oracle_3d():
    sol, ora = _run_pair(3, 32, 24, 16, (300,), rayleigh=1e6)
    _compare_diag(sol, ora)
    sol.step(50)
    ora.step(50)
    _compare_diag(sol, ora)


def test_diagnostics_match_oracle_2d():
    sol, ora = _run_pair(2, 64, 33, 1, (400,), rayleigh=1e5, vel_x=1, temp_x=2,
                         thot=1.0, tcold=0.0)
    _compare_diag(sol, ora)


@pytest.mark.parametrize("nx", [31, 32])
def test_side_heated_diagnostics_match_oracle(nx):
    """Side-heated cell (hot wall x = 0): Nu_vol from u T, the x-direction Nu_global, and the
    hot-wall profile along y with its own minimum window; odd and even nx (mid-plane rule)."""
    kw = dict(side_heated=1, temp_x=0, temp_y=1, temp_z=1, rayleigh=1e6)
    sol, ora = _run_pair(3, nx, 24, 12, (300,), **kw)
    _compare_diag(sol, ora)
    sol2, ora2 = _run_pair(2, nx, 40, 1, (400,), rayleigh=1e5, side_heated=1, vel_x=0, temp_x=0,
                           temp_y=1, thot=1.0, tcold=0.0)
    _compare_diag(sol2, ora2)


def _compare_midplane(sol, ora):
    a, b = sol.midplane(), ora.midplane()
    for k in b:
        if k.endswith("_max"):
            assert a[k] == pytest.approx(b[k], rel=REL_DIAG, abs=1e-300), (k, a[k], b[k])
        else:       # positions and plane coordinates: grid values or fitted vertices
            assert a[k] == pytest.approx(b[k], rel=1e-9, abs=1e-12), (k, a[k], b[k])


@pytest.mark.parametrize("case", ["rb_3d", "side_3d", "rb_2d_periodic", "rb_2d_walls", "side_2d_odd",
                                  "side_2d_even"])
def test_midplane_maxima_match_oracle(case):
    runs = {
        "rb_3d": (3, 20, 17, 12, (200,), dict(rayleigh=1e6)),
        "side_3d": (3, 19, 16, 10, (200,), dict(side_heated=1, temp_x=0, temp_y=1, temp_z=1,
                                               rayleigh=1e6)),
        "rb_2d_periodic": (2, 64, 33, 1, (400,), dict(rayleigh=1e5, vel_x=1, temp_x=2, thot=1.0,
                                                      tcold=0.0)),
        "rb_2d_walls": (2, 48, 32, 1, (400,), dict(rayleigh=1e5, vel_x=0, temp_x=1, thot=1.0,
                                                   tcold=0.0)),
        "side_2d_odd": (2, 33, 40, 1, (400,), dict(rayleigh=1e5, side_heated=1, vel_x=0,
                                                  temp_x=0, temp_y=1, thot=1.0, tcold=0.0)),
        "side_2d_even": (2, 32, 41, 1, (400,), dict(rayleigh=1e5, side_heated=1, vel_x=0,
                                                   temp_x=0, temp_y=1, thot=1.0, tcold=0.0)),
    }
    dims, nx, ny, nz, steps, kw = runs[case]
    sol, ora = _run_pair(dims, nx, ny, nz, steps, **kw)
    _compare_midplane(sol, ora)


def test_c2_128cubed_1000_steps():
    """128^3, Ra=1e6, Pr=0.71, FP64: relative max-norm <= 1e-12 after 1000 steps."""
    sol, ora = _run_pair(3, 128, 128, 128, (1000,), rayleigh=1e6)
    _compare_diag(sol, ora)


def test_c2_128cubed_5000_steps():
    """The full 128^3 run: checked at 1000 and 5000 steps, bitwise on f and g."""
    sol, ora = _run_pair(3, 128, 128, 128, (1000, 5000), rayleigh=1e6)
    np.testing.assert_array_equal(sol.gather()["g"], ora.state()["g"])
    _compare_diag(sol, ora)


def test_c1_2d_257x129_20000_steps():
    """2D, 257 x 129, Ra=1e5, Pr=0.71, Thot=1 / Tcold=0, periodic sides, 20,000 steps."""
    sol, ora = _run_pair(2, 257, 129, 1, (20000,), rayleigh=1e5, vel_x=1, temp_x=2,
                         thot=1.0, tcold=0.0)
    np.testing.assert_array_equal(sol.gather()["f"], ora.state()["f"])
    _compare_diag(sol, ora)


def test_2d_rb_ra1e4_nusselt_matches_published_value():
    # Wang et al. (2013) Table 12 (2DRB/_paper_extract.txt:1123): 2D periodic RB, Ra=1e4,
    # Pr=0.71, Ma=0.01: <Nu> = 2.6551312.  128 x 64 lattice at Ma=0.1, run to steady state.
    cfg = rblbm.Config(dims=2, nx=128, ny=64, nz=1, rayleigh=1e4, vel_x=1, temp_x=2,
                       thot=1.0, tcold=0.0)
    sol = rblbm.Solver(cfg)
    sol.initial_seeded(20261203, 1e-3)
    steps = int(150 * sol.params()["time_unit"])
    sol.step(steps)
    d = sol.diagnostics()
    print("Nu", d)
    assert d["nu_hot"] == pytest.approx(2.6551312, rel=0.02)
    assert d["nu_global"] == pytest.approx(2.6551312, rel=0.02)
    assert d["nu_vol"] == pytest.approx(2.6551312, rel=0.02)


def test_restart_scatter_matches_oracle_reload():
    oc, gc = _cfgs(3, 12, 10, 8, rayleigh=1e5, mach=0.3)
    ora = orc.Oracle(oc)
    ora.init_seeded(3, 0.1)
    ora.step(25)
    s = ora.state()
    sol = rblbm.Solver(gc)
    sol.scatter(s["f"], s["g"])
    rl = orc.Oracle(oc)
    rl.init_exact()
    assert rl.set_fg(s["f"], s["g"]) == 0
    _compare_state(sol, rl, "restart t=0")
    for n in (1, 30):
        sol.step(n)
        rl.step(n)
        _compare_state(sol, rl, f"restart +{n}")


def test_gather_scatter_roundtrip_and_rho_error():
    gc = rblbm.Config(dims=3, nx=12, ny=10, nz=8, rayleigh=1e5)
    sol = rblbm.Solver(gc)
    sol.initial_seeded(11, 0.01)
    sol.step(7)
    st = sol.gather()
    sol2 = rblbm.Solver(gc)
    sol2.scatter(st["f"], st["g"])
    st2 = sol2.gather()
    np.testing.assert_array_equal(st2["f"], st["f"])
    np.testing.assert_array_equal(st2["g"], st["g"])
    bad = st["f"].copy()
    bad[3, 4, 5, :] = -0.01
    with pytest.raises(rblbm.LbmError) as ei:
        sol2.scatter(bad, st["g"])
    assert ei.value.code == -6


def test_rejects_bad_configs_and_order():
    with pytest.raises(rblbm.LbmError):
        rblbm.Solver(rblbm.Config(dims=3, nx=2, ny=10, nz=8))
    with pytest.raises(rblbm.LbmError):
        rblbm.Solver(rblbm.Config(dims=4))
    sol = rblbm.Solver(rblbm.Config(dims=3, nx=8, ny=8, nz=8))
    with pytest.raises(rblbm.LbmError) as ei:
        sol.step(1)
    assert ei.value.code == -4


def test_mass_conservation_and_symmetry_512_slab():
    # Full 512 x 512 plane, 64 deep: mass is conserved by the closed cavity, and the x-mirror of
    # an x-symmetric initial state stays x-symmetric (populations permuted by the mirror).
    gc = rblbm.Config(dims=3, nx=512, ny=512, nz=64, rayleigh=1e8)
    sol = rblbm.Solver(gc)
    y = (np.arange(512) + 0.5) / 512
    x = (np.arange(512) + 0.5) / 512
    T = (0.5 - y)[None, :, None] + 0.05 * np.cos(2 * np.pi * x)[None, None, :] * \
        np.sin(np.pi * y)[None, :, None] * np.ones((64, 1, 1))
    sol.initial(T)
    m0 = sol.diagnostics()["mass"]
    sol.step(200)
    d = sol.diagnostics()
    assert d["mass"] == pytest.approx(m0, rel=1e-12)
    m = sol.macros()
    np.testing.assert_allclose(m["T"][:, :, ::-1], m["T"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(m["u"][:, :, ::-1], -m["u"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_midplane_slices_match_oracle(axis):
    sol, ora = _run_pair(3, 20, 17, 12, (200,), rayleigh=1e6)
    a, b = sol.midplane_slice(axis), ora.midplane_slice(axis)
    for k in ("u", "v", "w", "T"):
        np.testing.assert_array_equal(a[k], b[k])
    for k in ("psi", "vort"):
        scale = np.abs(b[k]).max()
        np.testing.assert_allclose(a[k], b[k], rtol=0, atol=1e-13 * scale)


@pytest.mark.parametrize("case", ["rb_periodic", "side_heated"])
def test_psi_vorticity_2d_match_oracle(case):
    kw = dict(rayleigh=1e5, thot=1.0, tcold=0.0)
    kw.update(dict(vel_x=1, temp_x=2) if case == "rb_periodic" else
              dict(side_heated=1, vel_x=0, temp_x=0, temp_y=1))
    sol, ora = _run_pair(2, 48, 32, 1, (300,), **kw)
    a, b = sol.psi_vorticity_2d(), ora.psi_vorticity_2d()
    np.testing.assert_array_equal(a["psi"], b["psi"])
    np.testing.assert_array_equal(a["vort"], b["vort"])
    for k in ("psi_center_coarse", "psi_center_fine", "psi_abs_max", "x_at_max", "y_at_max"):
        assert a[k] == pytest.approx(b[k], rel=1e-12, abs=1e-15), (k, a[k], b[k])


@pytest.mark.parametrize("dims", [3, 2])
def test_legacy_thermal_scheme_bitwise(dims):
    """EnableLegacyThermalScheme: fixed Qk, Qnu, rest weight from paraA, legacy wall factor;
    bitwise against the oracle (3D in both kernel forms)."""
    if dims == 3:
        kw = dict(rayleigh=1e5, mach=0.1, legacy=1)
        sol, ora = _run_pair(3, 20, 16, 12, (1, 30), init="exact",
                             T0=np.random.default_rng(4).uniform(-0.3, 0.3, (12, 16, 20)), **kw)
        sol.set_kernel_variant(rblbm.KERNEL_ONE_PASS, 5)
        sol.step(20)
        ora.step(20)
        np.testing.assert_array_equal(sol.gather()["g"], ora.state()["g"])
        _compare_diag(sol, ora)
    else:
        sol, ora = _run_pair(2, 40, 21, 1, (1, 200), rayleigh=1e5, mach=0.1, thot=1.0, tcold=0.0,
                             vel_x=0, temp_x=1, legacy=1)
        np.testing.assert_array_equal(sol.gather()["f"], ora.state()["f"])
    with pytest.raises(rblbm.LbmError):
        rblbm.Solver(rblbm.Config(dims=3, nx=8, ny=8, nz=8, legacy=1,
                                  precision=rblbm.FP32_SHIFT))
