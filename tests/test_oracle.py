"""Pins the CPU oracle (oracle/rb_oracle.c) before it is trusted as the parity checker.

The reference ships no tests, fixtures or golden vectors for this path, and it cannot be built
in this image (its driver calls GNU Fortran's `ctime` extension, which flang's runtime does not
provide).  The oracle is therefore pinned by:
  * an independent matrix-form model of the same scheme (tests/mrt_model.py) on small,
    non-cubic grids, for every wall combination the reference's switches allow;
  * closed forms: the dimensionless groups, the least-squares fits against numpy.polyfit;
  * invariants: mass and heat conservation in closed cavities, the rest state as a fixed point,
    the conduction state reaching Nu = 1 on every Nusselt definition.
"""
import numpy as np
import pytest

import orc
from mrt_model import Model, basis_d2q9, basis_d3q19, E2, E3


def _rand_T0(cfg, seed, amp=0.3):
    rng = np.random.default_rng(seed)
    return rng.uniform(-amp, amp, size=cfg.shape)


def _pair(cfg, T0):
    o = orc.Oracle(cfg)
    o.init_exact(T0)
    m = Model(cfg.dims, cfg.nx, cfg.ny, cfg.nz, cfg.rayleigh, cfg.prandtl, cfg.mach, cfg.thot,
              cfg.tcold, cfg.vel_x, cfg.vel_z, cfg.temp_x, cfg.temp_y, cfg.temp_z,
              bool(cfg.use_g), bool(cfg.side_heated), bool(cfg.legacy))
    m.init(T0)
    return o, m


def _compare(o, m, tol):
    s = o.state()
    np.testing.assert_allclose(s["f"], m.f, rtol=0, atol=tol)
    np.testing.assert_allclose(s["g"], m.g, rtol=0, atol=tol)
    np.testing.assert_allclose(s["T"], m.T, rtol=0, atol=tol)
    np.testing.assert_allclose(s["rho"], m.rho, rtol=0, atol=tol)
    np.testing.assert_allclose(s["u"], m.u[..., 0], rtol=0, atol=tol)
    np.testing.assert_allclose(s["v"], m.u[..., 1], rtol=0, atol=tol)
    if o.cfg.dims == 3:
        np.testing.assert_allclose(s["w"], m.u[..., 2], rtol=0, atol=tol)
    np.testing.assert_allclose(s["bx"], m.B[..., 0], rtol=0, atol=tol)
    np.testing.assert_allclose(s["by"], m.B[..., 1], rtol=0, atol=tol)


# --------------------------------------------------------------------------------------------
# closed forms
# --------------------------------------------------------------------------------------------
@pytest.mark.parametrize("dims,n,ra", [(3, 128, 1e6), (3, 512, 1e8), (3, 1024, 1e10), (2, 129, 1e5)])
def test_params_match_physical_definitions(dims, n, ra):
    cfg = orc.Config(dims=dims, nx=2 * n - 1 if dims == 2 else n, ny=n, nz=n, rayleigh=ra)
    p = orc.params(cfg)
    q = Model(dims, cfg.nx, cfg.ny, cfg.nz, ra).p
    for k in ("tauf", "viscosity", "diffusivity", "gbeta", "taug", "Qk", "Sq"):
        assert p[k] == pytest.approx(q[k], rel=1e-12), k
    # free-fall velocity = Ma * c_s, and Ra, Pr recovered from the lattice transport coefficients
    assert p["velocity_unit"] == pytest.approx(0.1 / np.sqrt(3), rel=1e-12)
    assert p["gbeta"] * p["L"] ** 3 / (p["viscosity"] * p["diffusivity"]) == pytest.approx(ra, rel=1e-12)
    assert p["viscosity"] / p["diffusivity"] == pytest.approx(0.71, rel=1e-14)


@pytest.mark.parametrize("dims", [3, 2])
def test_legacy_thermal_params(dims):
    # paraA is the rest-weight parameter that makes kappa = cT^2 (1/Qk - 1/2) with the fixed
    # Qk = 3 - sqrt3 (the model derives the weights from that relation); outside the reference's
    # range (-6 or -4 < paraA < 1) the oracle refuses the configuration, as the reference stops.
    cfg = orc.Config(dims=dims, nx=12, ny=10, nz=8, rayleigh=1e5, mach=0.1, legacy=1)
    p = orc.params(cfg)
    q = Model(dims, cfg.nx, cfg.ny, cfg.nz, 1e5, legacy=True).p
    assert p["Qk"] == pytest.approx(q["Qk"], rel=1e-15) and p["Qnu"] == pytest.approx(q["Qnu"])
    qn = 6 if dims == 3 else 4
    wT1 = (qn + p["paraA"]) / (7 * qn if dims == 3 else 5 * qn)
    assert wT1 == pytest.approx(q["wT1"], rel=1e-12)
    with pytest.raises(ValueError):
        orc.Oracle(orc.Config(dims=dims, nx=12, ny=10, nz=8, rayleigh=1e5, mach=30.0, legacy=1))


def test_params_published_values():
    # tau_f at 512^3, Ra=1e8 and at 1024^3, Ra=1e10 (values quoted with the workload definitions)
    p = orc.params(orc.Config(nx=512, ny=512, nz=512, rayleigh=1e8))
    assert p["tauf"] == pytest.approx(0.507472, abs=5e-7)
    assert p["Qk"] == pytest.approx(1.94540, abs=5e-5)
    p = orc.params(orc.Config(nx=1024, ny=1024, nz=1024, rayleigh=1e10))
    assert p["tauf"] == pytest.approx(0.501494, abs=5e-7)


def test_moment_bases_orthogonal():
    for M in (basis_d3q19(E3), basis_d2q9(E2)):
        G = M @ M.T
        assert np.allclose(G, np.diag(np.diag(G)))


def test_fits_against_numpy_least_squares():
    rng = np.random.default_rng(3)
    for _ in range(20):
        y = np.sort(rng.uniform(0, 1, 4))
        t = rng.normal(size=4)
        y0 = rng.uniform(-0.5, 0.0)
        # model T = a + b (y - y0)^2, least squares; T_wall = a
        A = np.stack([np.ones(4), (y - y0) ** 2], axis=1)
        a = np.linalg.lstsq(A, t, rcond=None)[0][0]
        assert orc.fit_adiabatic_wall_T4(y0, y, t) == pytest.approx(a, rel=1e-10, abs=1e-12)
    for mode in (1, -1):
        for _ in range(20):
            y = np.sort(rng.uniform(0, 1, 5))
            c2 = -mode * rng.uniform(0.5, 2.0)
            yv = rng.uniform(y[1], y[3])
            f = c2 * (y - yv) ** 2 + 3.0 + 1e-3 * rng.normal(size=5)
            C, B, A0 = np.polyfit(y, f, 2)
            fs, ys = orc.fit_parabola_ls5(y, f, mode)
            yref = -B / (2 * C)
            if y.min() <= yref <= y.max():
                assert ys == pytest.approx(yref, rel=1e-9, abs=1e-12)
                assert fs == pytest.approx(A0 + B * yref + C * yref ** 2, rel=1e-9)
    # wrong-curvature parabola falls back to the best sample
    y = np.linspace(0, 1, 5)
    f = (y - 0.5) ** 2
    assert orc.fit_parabola_ls5(y, f, 1) == (f[0], y[0])


def test_noise_is_uniform_and_deterministic():
    vals = np.array([orc.noise(20261203, 0, i) for i in range(20000)])
    assert vals.min() >= -1.0 and vals.max() < 1.0
    assert abs(vals.mean()) < 0.02
    assert abs(vals.var() - 1 / 3) < 0.01
    assert orc.noise(1, 0, 5) == orc.noise(1, 0, 5)
    assert orc.noise(1, 0, 5) != orc.noise(1, 1, 5)


# --------------------------------------------------------------------------------------------
# oracle vs independent model
# --------------------------------------------------------------------------------------------
CASES_3D = {
    "rb_noslip": dict(),
    "periodic_xz": dict(vel_x=1, vel_z=1, temp_x=2, temp_z=2),
    "side_heated": dict(side_heated=1, temp_x=0, temp_y=1, temp_z=1),
    "periodic_x_walls_z": dict(vel_x=1, temp_x=2),
    "no_m1g": dict(use_g=0),
    "legacy_thermal": dict(legacy=1),
    "legacy_side_heated": dict(legacy=1, side_heated=1, temp_x=0, temp_y=1, temp_z=1),
}


@pytest.mark.parametrize("name", list(CASES_3D))
def test_oracle_matches_model_3d(name):
    cfg = orc.Config(dims=3, nx=7, ny=6, nz=5, rayleigh=1e5, mach=0.3, **CASES_3D[name])
    o, m = _pair(cfg, _rand_T0(cfg, 11))
    for n in (1, 2, 7):
        o.step(n)
        m.step(n)
        _compare(o, m, 2e-13)


CASES_2D = {
    "rb_periodic": dict(vel_x=1, temp_x=2),
    "rb_walls": dict(vel_x=0, temp_x=1),
    "side_heated": dict(side_heated=1, vel_x=0, temp_x=0, temp_y=1),
    "no_m1g": dict(vel_x=1, temp_x=2, use_g=0),
    "legacy_thermal": dict(vel_x=0, temp_x=1, legacy=1),
}


@pytest.mark.parametrize("name", list(CASES_2D))
def test_oracle_matches_model_2d(name):
    cfg = orc.Config(dims=2, nx=9, ny=6, rayleigh=1e5, mach=0.3, thot=1.0, tcold=0.0,
                     **CASES_2D[name])
    o, m = _pair(cfg, _rand_T0(cfg, 5) + 0.5)
    for n in (1, 2, 7):
        o.step(n)
        m.step(n)
        _compare(o, m, 2e-13)


# --------------------------------------------------------------------------------------------
# invariants
# --------------------------------------------------------------------------------------------
def test_mass_conserved_closed_cavity():
    cfg = orc.Config(dims=3, nx=8, ny=7, nz=6, rayleigh=1e5, mach=0.3)
    o = orc.Oracle(cfg)
    o.init_exact(_rand_T0(cfg, 2))
    m0 = o.diagnostics()["mass"]
    o.step(50)
    assert o.diagnostics()["mass"] == pytest.approx(m0, rel=1e-13)


def test_heat_conserved_adiabatic_box():
    cfg = orc.Config(dims=3, nx=8, ny=7, nz=6, rayleigh=1e5, mach=0.3, temp_y=orc.T_ADIABATIC)
    o = orc.Oracle(cfg)
    o.init_exact(_rand_T0(cfg, 4))
    h0 = o.diagnostics()["heat"]
    o.step(50)
    assert o.diagnostics()["heat"] == pytest.approx(h0, rel=1e-12, abs=1e-12)


def test_rest_state_is_fixed_point():
    cfg = orc.Config(dims=3, nx=6, ny=6, nz=6, temp_y=orc.T_ADIABATIC)
    o = orc.Oracle(cfg)
    o.init_exact(np.zeros(cfg.shape))          # T = Tref everywhere: no buoyancy
    s0 = o.state()
    o.step(20)
    s1 = o.state()
    np.testing.assert_allclose(s1["f"], s0["f"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(s1["g"], s0["g"], rtol=0, atol=1e-15)


@pytest.mark.parametrize("dims,periodic", [(2, 0), (2, 1), (3, 0), (3, 1)])
def test_conduction_state_gives_unit_nusselt(dims, periodic):
    # Far below the onset of convection the linear profile between the anti-bounce-back walls
    # is the steady state and every Nusselt definition reads 1.  The buoyancy-driven spurious
    # velocity of the discrete hydrostatic state scales with g*beta = Ma^2/(3L), so a small Mach
    # number (at fixed tau_f) leaves Nu - 1 at the 1e-8 level.
    cfg = orc.Config(dims=dims, nx=8, ny=12, nz=6, rayleigh=0.01, mach=0.001,
                     vel_x=periodic, temp_x=2 if periodic else 1)
    o = orc.Oracle(cfg)
    o.init_exact()
    o.step(12000)
    d = o.diagnostics()
    for k in ("nu_hot", "nu_cold", "nu_global", "nu_middle", "nu_vol"):
        assert d[k] == pytest.approx(1.0, abs=5e-8), (k, d[k])
    assert d["re_vol"] < 2e-6


def test_2d_rb_ra1e4_reproduces_published_nusselt():
    # Wang et al. (2013) Table 12, quoted in 2DRB/_paper_extract.txt:1123: 2D periodic RB at
    # Ra = 1e4, Pr = 0.71 has <Nu> = 2.6551312.  The scheme at Ma = 0.1 on 128 x 64 reaches a
    # steady roll pattern within 100 free-fall times; its Nusselt numbers agree within 1%.
    cfg = orc.Config(dims=2, nx=128, ny=64, rayleigh=1e4, vel_x=1, temp_x=2, thot=1.0, tcold=0.0)
    o = orc.Oracle(cfg)
    o.init_seeded(20261203, 1e-3)
    o.step(int(100 * orc.params(cfg)["time_unit"]))
    d = o.diagnostics()
    for k in ("nu_hot", "nu_cold", "nu_global", "nu_vol"):
        assert d[k] == pytest.approx(2.6551312, rel=0.01), (k, d[k])


def test_volume_diagnostics_closed_form():
    # A uniform flow U through the conduction profile T = Thot - (y/L) dT: every finite
    # difference of the Nusselt formulas is exact on it (dT/dy = -dT, wall rows included), so
    # each diagnostic reduces to a closed form in U.  The reload (reconstruct_macro_from_fg) adds
    # the half force 0.5 rho g beta (T - Tref) to the y momentum, which the closed forms include.
    cfg = orc.Config(dims=3, nx=8, ny=10, nz=6, rayleigh=1e5, mach=0.1)
    p = orc.params(cfg)
    U = np.array([0.01, -0.02, 0.005])
    y = (np.arange(cfg.ny) + 0.5) / cfg.ny
    T = (cfg.thot - y * (cfg.thot - cfg.tcold))[None, :, None] * np.ones(cfg.shape)
    m = Model(3, cfg.nx, cfg.ny, cfg.nz)
    f = m.feq(np.ones(cfg.shape), np.broadcast_to(U, cfg.shape + (3,)).copy())
    g = m.geq(T, np.zeros(cfg.shape + (3,)))
    o = orc.Oracle(cfg)
    o.init_exact()
    assert o.set_fg(f, g) == 0
    d = o.diagnostics()
    rho = f.sum(axis=-1)
    v = (rho * U[1] + 0.5 * rho * p["gbeta"] * (T - p["Tref"])) / rho
    u2 = U[0] ** 2 + v ** 2 + U[2] ** 2
    L, kappa, nu = p["L"], p["diffusivity"], p["viscosity"]
    assert d["nu_vol"] == pytest.approx(1 + L / kappa * np.mean(v * T), rel=1e-12)
    assert d["re_vol"] == pytest.approx(np.sqrt(np.mean(u2)) * L / nu, rel=1e-12)
    assert d["mass"] == pytest.approx(rho.sum(), rel=1e-14)
    dT = cfg.thot - cfg.tcold
    conv = L / kappa * v * (T - p["Tref"])
    assert d["nu_global"] == pytest.approx((np.mean(conv) + dT) / dT, rel=1e-12)
    assert d["nu_hot"] == pytest.approx(1.0, rel=1e-12)
    assert d["nu_cold"] == pytest.approx(1.0, rel=1e-12)
    jb, jt = cfg.ny // 2 - 1, cfg.ny // 2
    mid = 0.5 * (conv[:, jb, :] + conv[:, jt, :]).mean() / dT + 1.0
    assert d["nu_middle"] == pytest.approx(mid, rel=1e-12)


def _side_state(cfg, U, T):
    """f, g of a uniform flow U and temperature field T (Fortran-order arrays)."""
    m = Model(cfg.dims, cfg.nx, cfg.ny, cfg.nz)
    U = np.array(U, dtype=float)
    if cfg.dims == 2:
        U[2] = 0.0
    f = m.feq(np.ones(cfg.shape), np.broadcast_to(U, cfg.shape + (3,)).copy())
    g = m.geq(T, np.zeros(cfg.shape + (3,)))
    return f, g


@pytest.mark.parametrize("dims", [3, 2])
def test_side_heated_diagnostics_closed_form(dims):
    # Side-heated cell (hot wall x = 0, L = nx): a uniform flow U through the conduction profile
    # T = Thot - (x/L) dT.  The x differences are exact on it, so Nu_global, Nu_hot, Nu_cold are
    # closed forms, and Nu_vol uses u T (2DRB/3DRB.F90:1732-1741), not v T.  The reload's half
    # force acts on v only (gravity along y), so u = U_x exactly.
    cfg = orc.Config(dims=dims, nx=10, ny=9, nz=6, rayleigh=1e5, mach=0.1, thot=1.0, tcold=0.0,
                     side_heated=1, temp_x=orc.T_CONST, temp_y=orc.T_ADIABATIC)
    p = orc.params(cfg)
    assert p["L"] == cfg.nx
    U = np.array([0.01, -0.02, 0.005])
    x = (np.arange(cfg.nx) + 0.5) / cfg.nx
    T = (cfg.thot - x * (cfg.thot - cfg.tcold))[None, None, :] * np.ones(cfg.shape)
    f, g = _side_state(cfg, U, T)
    o = orc.Oracle(cfg)
    o.init_exact()
    assert o.set_fg(f, g) == 0
    d = o.diagnostics()
    L, kappa, dT = p["L"], p["diffusivity"], cfg.thot - cfg.tcold
    assert d["nu_vol"] == pytest.approx(1 + L / kappa * U[0] * T.mean(), rel=1e-12)
    conv = L / kappa * U[0] * (T - p["Tref"])
    assert d["nu_global"] == pytest.approx((conv.mean() + dT) / dT, rel=1e-12)
    assert d["nu_hot"] == pytest.approx(1.0, rel=1e-12)
    assert d["nu_cold"] == pytest.approx(1.0, rel=1e-12)
    il, ir = cfg.nx // 2 - 1, cfg.nx // 2
    mid = 0.5 * (conv[..., il] + conv[..., ir]).mean() / dT + 1.0
    assert d["nu_middle"] == pytest.approx(mid, rel=1e-12)
    # the hot-wall profile is flat (corner fits included): its extrema are 1
    assert d["nu_hot_max"] == pytest.approx(1.0, rel=1e-12)
    assert d["nu_hot_min"] == pytest.approx(1.0, rel=1e-12)


def test_side_heated_hot_wall_minimum_window():
    # The side-heated minimum is fitted on the five profile points jmin-4 .. jmin
    # (3DRB.F90:1975-1980), not on a window centred on jmin as for the maximum.  The first two
    # columns carry an upward parabola in y whose vertex y0 lies just above the grid point
    # nearest to it: a centred window contains y0 and returns it, the reference's window
    # jmin-4 .. jmin ends below y0, so the fit falls back to the grid point and its value.
    cfg = orc.Config(dims=3, nx=8, ny=16, nz=4, rayleigh=1e5, mach=0.1, thot=1.0, tcold=0.0,
                     side_heated=1, temp_x=orc.T_CONST, temp_y=orc.T_ADIABATIC)
    p = orc.params(cfg)
    L, dT = p["L"], cfg.thot - cfg.tcold
    yp = (np.arange(cfg.ny) + 0.5) / L
    jstar = 9
    y0 = yp[jstar] + 0.2 / L                      # nearest grid point: jstar
    T1 = 0.9 - 0.3 * (yp - y0) ** 2               # colder towards both corners
    T = np.broadcast_to(np.linspace(0.9, 0.1, cfg.nx)[None, None, :], cfg.shape).copy()
    T[:, :, 0] = T1[None, :]
    T[:, :, 1] = T1[None, :]
    f, g = _side_state(cfg, np.zeros(3), T)
    o = orc.Oracle(cfg)
    o.init_exact()
    assert o.set_fg(f, g) == 0
    d = o.diagnostics()
    nu_left = (8 * cfg.thot - 9 * T1 + T1) / (3 / L) / dT
    assert np.argmin(nu_left) == jstar
    assert d["nu_hot_min_pos"] == pytest.approx(yp[jstar], rel=1e-12)
    assert d["nu_hot_min"] == pytest.approx(nu_left[jstar], rel=1e-12)
    assert d["nu_hot"] == pytest.approx(nu_left.mean(), rel=1e-12)


def _velocity_state(cfg, u, v, w):
    """f of the given velocity field at rho = 1, g at T = Tref (no buoyancy in the reload)."""
    m = Model(cfg.dims, cfg.nx, cfg.ny, cfg.nz)
    U = np.stack([u, v, w], axis=-1)
    f = m.feq(np.ones(cfg.shape), U)
    Tref = 0.5 * (cfg.thot + cfg.tcold)
    g = m.geq(np.full(cfg.shape, Tref), np.zeros(cfg.shape + (3,)))
    return f, g


def test_midplane_maxima_3d_closed_form():
    # u varies along y only, v along z only, w along x only, each a parabola whose vertex lies
    # between grid points.  On its mid-plane every component is uniform along one in-plane axis,
    # so the reported maximum is the grid point nearest the vertex and the first index (1) on
    # the other axis (3DRB.F90:2399-2518); values are scaled by L / kappa.
    cfg = orc.Config(dims=3, nx=12, ny=10, nz=9, rayleigh=1e5, mach=0.1)
    p = orc.params(cfg)
    L, coef = p["L"], p["L"] / p["diffusivity"]
    xp, yp, zp = [(np.arange(n) + 0.5) / L for n in (cfg.nx, cfg.ny, cfg.nz)]
    prof = lambda s, s0: 0.02 - 0.05 * (s - s0) ** 2
    z3, y3, x3 = np.meshgrid(zp, yp, xp, indexing="ij")
    u, v, w = prof(y3, yp[6] + 0.01), prof(z3, zp[2] - 0.02), prof(x3, xp[4] + 0.03)
    zero = np.zeros(cfg.shape)
    o = orc.Oracle(cfg)
    o.init_exact()

    def mid(uu, vv, ww):        # one component at a time: the others stay exactly zero
        assert o.set_fg(*_velocity_state(cfg, uu, vv, ww)) == 0
        return o.midplane()

    m = mid(u, zero, zero)
    assert m["u_plane"] == pytest.approx(0.5 * cfg.nx / L, rel=1e-15)
    assert (m["u_pos1"], m["u_pos2"]) == (pytest.approx(yp[6]), pytest.approx(zp[0]))
    assert m["u_max"] == pytest.approx(coef * u[0, 6, 0], rel=1e-12)
    m = mid(zero, v, zero)
    assert (m["v_pos1"], m["v_pos2"]) == (pytest.approx(xp[0]), pytest.approx(zp[2]))
    assert m["v_max"] == pytest.approx(coef * v[2, 0, 0], rel=1e-12)
    m = mid(zero, zero, w)
    assert (m["w_pos1"], m["w_pos2"]) == (pytest.approx(xp[4]), pytest.approx(yp[0]))
    assert m["w_max"] == pytest.approx(coef * w[0, 0, 4], rel=1e-12)
    assert m["w_plane"] == pytest.approx(0.5 * cfg.nz / L, rel=1e-15)


@pytest.mark.parametrize("side", [0, 1])
def test_midplane_maxima_2d_parabola_fit(side):
    # 2D: the maxima come from a 5-point least-squares parabola, exact on parabolic profiles.
    # The Rayleigh-Benard cell samples u on x = 0.5 (2DRB.F90:2497-2520) and folds the fitted
    # position into the lower half (:2549-2550); the side-heated cell samples the middle
    # column and does not fold (:2081-2153).  v is sampled on the middle row in both.
    cfg = orc.Config(dims=2, nx=24, ny=20, rayleigh=1e5, mach=0.1, side_heated=side,
                     temp_x=0 if side else 1, temp_y=1 if side else 0)
    p = orc.params(cfg)
    L, coef = p["L"], p["L"] / p["diffusivity"]
    xp, yp = (np.arange(cfg.nx) + 0.5) / L, (np.arange(cfg.ny) + 0.5) / L
    y0, x0 = 0.73 * cfg.ny / L, 0.41 * cfg.nx / L
    yy, xx = np.meshgrid(yp, xp, indexing="ij")
    u = (0.03 - 0.4 * (yy - y0) ** 2)[None]
    v = (0.02 - 0.3 * (xx - x0) ** 2)[None]
    f, g = _velocity_state(cfg, u, v, np.zeros_like(u))
    o = orc.Oracle(cfg)
    o.init_exact()
    assert o.set_fg(f, g) == 0
    m = o.midplane()
    ylen = cfg.ny / L
    assert m["u_max"] == pytest.approx(0.03 * coef, rel=1e-9)
    assert m["u_pos1"] == pytest.approx(y0 if side else ylen - y0, rel=1e-9)
    assert m["v_max"] == pytest.approx(0.02 * coef, rel=1e-9)
    assert m["v_pos1"] == pytest.approx(x0, rel=1e-9)
    assert m["v_plane"] == pytest.approx(0.5 * (yp[cfg.ny // 2 - 1] + yp[cfg.ny // 2]))
    assert m["u_plane"] == pytest.approx(0.5 * (xp[11] + xp[12]) if side else 0.5)


def test_restart_reconstruction_roundtrip():
    cfg = orc.Config(dims=3, nx=7, ny=6, nz=5, rayleigh=1e5, mach=0.3)
    o = orc.Oracle(cfg)
    o.init_exact(_rand_T0(cfg, 9))
    o.step(5)
    s = o.state()
    r = orc.Oracle(cfg)
    r.init_exact(np.zeros(cfg.shape))
    assert r.set_fg(s["f"], s["g"]) == 0
    t = r.state()
    np.testing.assert_array_equal(t["T"], s["T"])
    np.testing.assert_array_equal(t["rho"], s["rho"])
    # The reload rebuilds u with the half force of the reloaded T (3DRB.F90:1428-1435), not
    # with the previous step's T as the in-run macro does.
    p = orc.params(cfg)
    jy = (s["f"] * np.array([e[1] for e in E3])).sum(axis=-1)
    v_expect = (jy + 0.5 * (s["rho"] * p["gbeta"] * (s["T"] - p["Tref"]))) / s["rho"]
    np.testing.assert_allclose(t["v"], v_expect, rtol=0, atol=1e-15)
    f = s["f"].copy()
    f[0, 0, 0, :] = -1.0
    assert r.set_fg(f, s["g"]) == -1


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_midplane_slice_stream_function_and_vorticity_closed_form(axis):
    # On each mid-plane the stream function integrates the first in-plane axis's velocity along
    # the second axis from the wall; with that velocity a*s^2 (s the wall distance along the
    # second axis) the first-cell rule, Simpson's rule and the quadratic midpoints are exact:
    # psi = (L/kappa) a s^3 / 3.  The plane-normal vorticity of the same field is exact too
    # (second-order differences): -(L/kappa) 2 a s for the x and z planes, +(L/kappa) 2 a s for
    # the y plane (omega_y = du/dz - dw/dx).  3DRB.F90:2721-3011.
    cfg = orc.Config(dims=3, nx=10, ny=12, nz=9, rayleigh=1e5, mach=0.1)
    p = orc.params(cfg)
    L, coef = p["L"], p["L"] / p["diffusivity"]
    a = 0.03
    zp, yp = (np.arange(cfg.nz) + 0.5) / L, (np.arange(cfg.ny) + 0.5) / L
    zero = np.zeros(cfg.shape)
    if axis == 0:      # (y, z): v(z)
        s = zp
        prof = np.broadcast_to((a * zp ** 2)[:, None, None], cfg.shape)
        u, v, w = zero, prof, zero
    elif axis == 1:    # (x, z): u(z)
        s = zp
        prof = np.broadcast_to((a * zp ** 2)[:, None, None], cfg.shape)
        u, v, w = prof, zero, zero
    else:              # (x, y): u(y)
        s = yp
        prof = np.broadcast_to((a * yp ** 2)[None, :, None], cfg.shape)
        u, v, w = prof, zero, zero
    o = orc.Oracle(cfg)
    o.init_exact()
    assert o.set_fg(*_velocity_state(cfg, u, v, w)) == 0
    sl = o.midplane_slice(axis)
    psi_exact = coef * a * s ** 3 / 3
    np.testing.assert_allclose(sl["psi"], np.broadcast_to(psi_exact[:, None], sl["psi"].shape),
                               rtol=1e-10, atol=1e-12 * coef)
    sign = 1.0 if axis == 1 else -1.0
    vort_exact = sign * coef * 2 * a * s
    np.testing.assert_allclose(sl["vort"], np.broadcast_to(vort_exact[:, None], sl["vort"].shape),
                               rtol=1e-9, atol=1e-9 * coef * a)
    comp = {0: "v", 1: "u", 2: "u"}[axis]
    np.testing.assert_allclose(sl[comp], np.broadcast_to(coef * a * s[:, None] ** 2, sl[comp].shape),
                               rtol=1e-12)


def test_spline_and_lagrange_against_scipy_numpy():
    # the natural cubic spline and the 4-point Lagrange interpolant of the 2D |psi| statistics
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(8)
    x = np.concatenate([[0.0], np.sort(rng.uniform(0.02, 0.98, 9)), [1.0]])
    y = rng.normal(size=x.size)
    xq = np.linspace(0.0, 1.0, 57)
    ref = CubicSpline(x, y, bc_type="natural")(xq)
    np.testing.assert_allclose(orc.spline_eval(x, y, xq), ref, rtol=1e-12, atol=1e-12)
    xk = np.sort(rng.uniform(0, 1, 4))
    fk = rng.normal(size=4)
    coeff = np.polyfit(xk, fk, 3)
    for xq1 in (0.3, 0.55, xk[1]):
        assert orc.lagrange4(xq1, xk, fk) == pytest.approx(np.polyval(coeff, xq1), rel=1e-10, abs=1e-12)


def test_psi_vorticity_2d_closed_form():
    # u = a y (Y - y) (zero on both walls y = 0, Y), v = 0: the 2D stream function
    # psi = (L/kappa) a (Y y^2 / 2 - y^3 / 3) and the vorticity -(L/kappa) a (Y - 2 y) are exact
    # for the first-cell rule, Simpson's rule with quadratic midpoints and the one-sided wall
    # differences through the wall value 0 (2DRB.F90:2657-2770).
    cfg = orc.Config(dims=2, nx=12, ny=10, rayleigh=1e5, mach=0.1, thot=1.0, tcold=0.0,
                     vel_x=1, temp_x=2)
    p = orc.params(cfg)
    L, coef = p["L"], p["L"] / p["diffusivity"]
    a = 0.02
    yp = (np.arange(cfg.ny) + 0.5) / L
    Y = cfg.ny / L
    u = np.broadcast_to((a * yp * (Y - yp))[None, :, None], cfg.shape)
    zero = np.zeros(cfg.shape)
    o = orc.Oracle(cfg)
    o.init_exact()
    assert o.set_fg(*_velocity_state(cfg, u, zero, zero)) == 0
    r = o.psi_vorticity_2d(stats=False)
    psi = coef * a * (Y * yp ** 2 / 2 - yp ** 3 / 3)
    np.testing.assert_allclose(r["psi"], np.broadcast_to(psi[:, None], r["psi"].shape), rtol=1e-10)
    vort = -coef * a * (Y - 2 * yp)
    np.testing.assert_allclose(r["vort"], np.broadcast_to(vort[:, None], r["vort"].shape),
                               rtol=1e-9, atol=1e-9 * coef * a)


# --------------------------------------------------------------------------------------------
# check (2DRB/3DRB.F90:1475-1528, 2DRB/2DRB.F90:1213-1260), the exact initial state
# (2DRB/3DRB.F90:517-600, 2DRB/2DRB.F90:540-608) and the kinetic-energy sum, each against values
# the chosen fields fix in closed form
# --------------------------------------------------------------------------------------------
def _load(o, cfg, U, T, rho=1.0):
    """Load f = f_eq(rho, U) node by node (U: (..., 3)), g = wT T; returns (u, v, w) the reload
    gives in closed form: U, with the half force 0.5 g beta (T - Tref) added to v."""
    m = Model(cfg.dims, cfg.nx, cfg.ny, cfg.nz)
    U = np.broadcast_to(np.asarray(U, dtype=float), cfg.shape + (3,)).copy()
    if cfg.dims == 2:
        U[..., 2] = 0.0
    T = np.broadcast_to(np.asarray(T, dtype=float), cfg.shape).copy()
    assert o.set_fg(m.feq(np.full(cfg.shape, rho), U), m.geq(T, np.zeros(cfg.shape + (3,)))) == 0
    p = orc.params(cfg)
    v = U[..., 1] + 0.5 * p["gbeta"] * (T - p["Tref"])
    return U[..., 0], v, U[..., 2]


@pytest.mark.parametrize("dims", [3, 2])
def test_check_relative_l2_velocity_and_l1_temperature(dims):
    cfg = orc.Config(dims=dims, nx=6, ny=5, nz=4, rayleigh=1e5, thot=1.0, tcold=0.0)
    tref = orc.params(cfg)["Tref"]
    o = orc.Oracle(cfg)
    o.init_exact()
    # first call: the previous fields are zero, so both errors are 1 exactly
    U1 = np.array([0.01, -0.02, 0.005])
    _load(o, cfg, U1, tref)
    eu, et = o.check()
    assert eu == pytest.approx(1.0, rel=1e-14) and et == pytest.approx(1.0, rel=1e-14)
    # second call: half the nodes (a checkerboard) move by dU, the temperature by eps on the
    # other half, at T = Tref + eps there the reload adds the half force to v
    k, j, i = np.indices(cfg.shape)
    odd = (i + j + k) % 2 == 1
    n_odd, n_even = int(odd.sum()), int((~odd).sum())
    dU = np.array([0.003, 0.004, -0.012])
    if dims == 2:
        U1 = U1 * np.array([1, 1, 0])
        dU = dU * np.array([1, 1, 0])
    eps = 0.25
    U2 = np.where(odd[..., None], U1 + dU, U1)
    T2 = np.where(odd, tref, tref + eps)
    u, v, w = _load(o, cfg, U2, T2)
    eu, et = o.check()
    du = np.where(odd[..., None], dU, np.zeros(3))
    dv_force = np.where(odd, 0.0, 0.5 * orc.params(cfg)["gbeta"] * eps)
    num = (du[..., 0] ** 2 + (du[..., 1] + dv_force) ** 2 + du[..., 2] ** 2).sum()
    den = (u ** 2 + v ** 2 + w ** 2).sum()
    assert eu == pytest.approx(np.sqrt(num) / np.sqrt(den), rel=1e-12)
    # L1 ratio: sum |T2 - T1| / sum |T2| = n_even eps / (n_odd Tref + n_even (Tref + eps))
    assert et == pytest.approx(n_even * eps / (n_odd * tref + n_even * (tref + eps)), rel=1e-13)
    # an L2 form of the temperature error would give a different number on these fields
    l2 = np.sqrt(n_even) * eps / np.sqrt(n_odd * tref ** 2 + n_even * (tref + eps) ** 2)
    assert abs(et - l2) > 1e-3
    # a third call on unchanged fields: zero change
    _load(o, cfg, U2, T2)
    eu, et = o.check()
    assert eu == 0.0 and et == 0.0


def test_check_zero_field_guard_3d_only():
    # 3D: with u = 0 and T = 0 the denominators vanish and the reference reports the plain norms
    # sqrt(sum |du|^2) and sum |dT| (2DRB/3DRB.F90:1513-1522); 2D has no guard
    # (2DRB/2DRB.F90:1244-1245), so the velocity ratio divides by zero
    for dims in (3, 2):
        cfg = orc.Config(dims=dims, nx=6, ny=5, nz=4, rayleigh=1e5, thot=0.5, tcold=-0.5)
        n = int(np.prod(cfg.shape))
        o = orc.Oracle(cfg)
        o.init_exact()
        U1 = np.array([0.01, -0.02, 0.005 if dims == 3 else 0.0])
        T1 = 0.125
        u, v, w = _load(o, cfg, U1, T1)
        o.check()
        _load(o, cfg, np.zeros(3), 0.0)   # T = Tref = 0: no half force, u = 0 everywhere
        with np.errstate(divide="ignore"):
            eu, et = o.check()
        if dims == 3:
            assert eu == pytest.approx(np.sqrt(((u ** 2 + v ** 2 + w ** 2)).sum()), rel=1e-13)
            assert eu == pytest.approx(np.sqrt(n * (U1[0] ** 2 + v.flat[0] ** 2 + U1[2] ** 2)), rel=1e-13)
            assert et == pytest.approx(n * T1, rel=1e-14)
        else:
            assert np.isinf(eu)
            assert np.isinf(et)


@pytest.mark.parametrize("dims,ra,perturbed", [(3, 1e5, False), (3, 1e4, True), (3, 1.0001e4, False),
                                               (3, 3e3, True), (2, 1e4, False), (2, 9999.0, True),
                                               (2, 1e5, False)])
def test_exact_initial_state_closed_form(dims, ra, perturbed):
    # T = Thot + (j - 1/2)/ny (Tcold - Thot) at the cell centres (L = ny for the RB cell); the
    # perturbation 1e-3 dT cos(2 pi x/X) sin(pi y/Y) (3D, Ra <= 1e4) or sin(2 pi x/X) sin(pi y/Y)
    # (2D, Ra < 1e4) with x = (i - 1/2)/L, X = nx/L, y = (j - 1/2)/L, Y = ny/L; rho = 1, u = 0,
    # f = w, g = wT T, B_prev = 0
    cfg = orc.Config(dims=dims, nx=10, ny=7, nz=5, rayleigh=ra, thot=1.5, tcold=-0.25)
    o = orc.Oracle(cfg)
    o.init_exact()
    s = o.state()
    k, j, i = np.indices(cfg.shape)
    dT = cfg.thot - cfg.tcold
    T = cfg.thot + (j + 0.5) / cfg.ny * (cfg.tcold - cfg.thot)
    if perturbed:
        trig = np.cos if dims == 3 else np.sin
        T = T + 1e-3 * dT * trig(2 * np.pi * (i + 0.5) / cfg.nx) * np.sin(np.pi * (j + 0.5) / cfg.ny)
    np.testing.assert_allclose(s["T"], T, rtol=0, atol=4e-16)
    if perturbed:   # the perturbation is really there (not a no-op within the tolerance)
        assert np.abs(s["T"] - (cfg.thot + (j + 0.5) / cfg.ny * (cfg.tcold - cfg.thot))).max() > 5e-4
    np.testing.assert_array_equal(s["rho"], 1.0)
    for key in ("u", "v", "w", "bx", "by", "bz"):
        np.testing.assert_array_equal(s[key], 0.0)
    wf = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12) if dims == 3 else \
        np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    wt = np.array([0.25] + [0.125] * 6) if dims == 3 else np.array([1 / 3] + [1 / 6] * 4)
    np.testing.assert_allclose(s["f"], np.broadcast_to(wf, s["f"].shape), rtol=1e-15, atol=0)
    np.testing.assert_allclose(s["g"], T[..., None] * wt, rtol=1e-15, atol=1e-17)


def test_exact_initial_state_side_heated_profile():
    # VerticalWallsConstT: the profile runs along x between the hot (x = 0) and cold walls, L = nx
    cfg = orc.Config(dims=3, nx=9, ny=6, nz=5, rayleigh=1e3, thot=1.0, tcold=0.0, side_heated=1,
                     temp_x=orc.T_CONST, temp_y=orc.T_ADIABATIC)
    o = orc.Oracle(cfg)
    o.init_exact()
    k, j, i = np.indices(cfg.shape)
    np.testing.assert_allclose(o.state()["T"], 1.0 - (i + 0.5) / cfg.nx, rtol=0, atol=4e-16)


@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("dims", [3, 2])
def test_kinetic_energy_sum_closed_form(dims, side):
    # kinetic = 0.5 sum rho |u|^2: a uniform flow at density rho0 through a uniform T (the reload
    # adds the half force to v), so the sum is 0.5 rho0 N (Ux^2 + v^2 + Uz^2); the RB and the
    # side-heated cells sum it in separate routines
    extra = dict(side_heated=1, temp_x=orc.T_CONST, temp_y=orc.T_ADIABATIC) if side else {}
    cfg = orc.Config(dims=dims, nx=7, ny=6, nz=5, rayleigh=1e5, thot=1.0, tcold=0.0, **extra)
    n = int(np.prod(cfg.shape))
    o = orc.Oracle(cfg)
    o.init_exact()
    rho0 = 1.2
    U = np.array([0.02, -0.01, 0.015])
    u, v, w = _load(o, cfg, U, 0.75, rho=rho0)
    d = o.diagnostics()
    uz = U[2] if dims == 3 else 0.0
    assert d["kinetic"] == pytest.approx(0.5 * rho0 * n * (U[0] ** 2 + v.flat[0] ** 2 + uz ** 2), rel=1e-12)
    assert d["mass"] == pytest.approx(rho0 * n, rel=1e-14)
    assert d["heat"] == pytest.approx(0.75 * n, rel=1e-14)
