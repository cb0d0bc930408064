"""In-place (A-A) population layout: one array per lattice, updated in place by both passes.

The layout changes only where populations are stored, never the arithmetic, so the FP64 form
must stay bitwise identical to the oracle and the reduced-precision forms bitwise identical to
their two-array runs, through the alternation of the two in-place states (odd and even step
counts), gather / macros / diagnostics in either state, restart, and the autotuner's layout
choice for LBM_LAYOUT_AUTO domains.
"""
import numpy as np
import pytest

import orc
import rblbm

pytestmark = pytest.mark.gpu


def _pair(nx, ny, nz, layout=rblbm.LAYOUT_AA, precision=rblbm.FP64, **kw):
    sol = rblbm.Solver(rblbm.Config(dims=3, nx=nx, ny=ny, nz=nz, precision=precision,
                                    layout=layout, **kw))
    ora = orc.Oracle(orc.Config(dims=3, nx=nx, ny=ny, nz=nz, **kw))
    return sol, ora


def _bitwise(sol, ora, label):
    s = ora.state()
    m = sol.macros()
    st = sol.gather()
    print(label, flush=True)
    for k in ("f", "g", "bx", "by", "bz"):
        np.testing.assert_array_equal(st[k], s[k], err_msg=f"{label} {k}")
    for k in ("rho", "u", "v", "w", "T"):
        np.testing.assert_array_equal(m[k], s[k], err_msg=f"{label} {k}")


@pytest.mark.parametrize("case", [
    dict(nx=24, ny=20, nz=16, rayleigh=1e6),                         # RB cell, adiabatic sides
    dict(nx=17, ny=29, nz=13, rayleigh=3e4, prandtl=7.0),            # odd sizes
    dict(nx=20, ny=16, nz=12, rayleigh=1e5, side_heated=1, temp_x=0, temp_y=1,
         thot=1.0, tcold=0.0),                                       # side-heated cell
    dict(nx=16, ny=12, nz=10, rayleigh=1e5, use_g=0),                # no M1G correction
    dict(nx=16, ny=12, nz=10, rayleigh=1e5, legacy=1),               # legacy thermal scheme
])
def test_aa_fp64_bitwise_against_oracle(case):
    kw = dict(case)
    nx, ny, nz = kw.pop("nx"), kw.pop("ny"), kw.pop("nz")
    sol, ora = _pair(nx, ny, nz, **kw)
    assert sol.layout == rblbm.LAYOUT_AA
    sol.initial_seeded(20261203, 1e-3)
    ora.init_seeded(20261203, 1e-3)
    done = 0
    for n in (1, 2, 5, 12):      # both in-place states, odd and even step counts
        sol.step(n - done)
        ora.step(n - done)
        done = n
        _bitwise(sol, ora, f"aa {case} step {n}")
    d, o = sol.diagnostics(), ora.diagnostics()
    for k in ("nu_vol", "nu_global", "re_vol", "nu_hot", "nu_cold"):
        assert d[k] == pytest.approx(o[k], rel=1e-12, abs=1e-14), k
    sol.close()


@pytest.mark.parametrize("precision", [rblbm.FP32_SHIFT, rblbm.FP16_SHIFT])
def test_aa_reduced_precision_equals_two_array(precision):
    out = {}
    for layout in (rblbm.LAYOUT_TWO_ARRAY, rblbm.LAYOUT_AA):
        sol = rblbm.Solver(rblbm.Config(dims=3, nx=40, ny=24, nz=20, rayleigh=1e7,
                                        precision=precision, layout=layout))
        assert sol.layout == layout
        sol.initial_seeded(7, 0.05)
        res = []
        for n in (1, 22):   # after 1 and 23 steps: both in-place states
            sol.step(n)
            res.append((sol.macros(), sol.gather(), sol.diagnostics()))
        out[layout] = res
        sol.close()
    for (m2, g2, d2), (ma, ga, da) in zip(out[rblbm.LAYOUT_TWO_ARRAY], out[rblbm.LAYOUT_AA]):
        for k in ("rho", "u", "v", "w", "T"):
            np.testing.assert_array_equal(ma[k], m2[k])
        for k in ("f", "g", "bx"):
            np.testing.assert_array_equal(ga[k], g2[k])
        assert da == d2


def test_aa_restart_continues_bitwise():
    # a state gathered from an in-place domain, scattered into another, runs on as the oracle
    nx, ny, nz = 18, 14, 12
    a, ora = _pair(nx, ny, nz, rayleigh=1e5, mach=0.3)
    a.initial_seeded(3, 1e-3)
    ora.init_seeded(3, 1e-3)
    a.step(9)
    ora.step(9)
    st = a.gather(with_bprev=False)
    a.close()
    b = rblbm.Solver(rblbm.Config(dims=3, nx=nx, ny=ny, nz=nz, rayleigh=1e5, mach=0.3,
                                  layout=rblbm.LAYOUT_AA))
    b.scatter(st["f"], st["g"])
    r = orc.Oracle(orc.Config(dims=3, nx=nx, ny=ny, nz=nz, rayleigh=1e5, mach=0.3))
    r.init_exact()
    assert r.set_fg(ora.state()["f"], ora.state()["g"]) == 0
    b.step(7)
    r.step(7)
    _bitwise(b, r, "aa restart")
    b.close()


def test_layout_auto_and_autotune(tmp_path):
    cfg = rblbm.Config(dims=3, nx=32, ny=24, nz=20, rayleigh=1e6, layout=rblbm.LAYOUT_AUTO)
    sol = rblbm.Solver(cfg)
    assert sol.layout == rblbm.LAYOUT_TWO_ARRAY     # it fits: two arrays
    sol.initial_seeded(11, 1e-3)
    cache = str(tmp_path / "tune.txt")
    sol.autotune(cache, steps=3)                    # may switch to in place before stepping
    lay = sol.layout
    assert lay in (rblbm.LAYOUT_TWO_ARRAY, rblbm.LAYOUT_AA)
    line = open(cache).read().split()
    assert line[0].endswith("_l2") and int(line[-1]) == lay
    ora = orc.Oracle(orc.Config(dims=3, nx=32, ny=24, nz=20, rayleigh=1e6))
    ora.init_seeded(11, 1e-3)
    sol.step(6)
    ora.step(6)
    _bitwise(sol, ora, f"auto layout {lay}")
    sol.close()
    # the cached choice is applied without timing to a new AUTO domain
    again = rblbm.Solver(cfg)
    again.initial_seeded(11, 1e-3)
    again.autotune(cache, steps=3)
    assert again.layout == lay
    again.close()


def test_aa_rejections():
    # periodic walls (a same-row periodic copy is read by two nodes), 2D, the one-pass form,
    # batches
    for kw in (dict(vel_x=rblbm.VEL_PERIODIC), dict(temp_z=rblbm.T_PERIODIC)):
        with pytest.raises(rblbm.LbmError) as e:
            rblbm.Solver(rblbm.Config(dims=3, nx=16, ny=12, nz=10, layout=rblbm.LAYOUT_AA, **kw))
        assert e.value.code == -5
    with pytest.raises(rblbm.LbmError):
        rblbm.Solver(rblbm.Config(dims=2, nx=16, ny=12, nz=1, layout=rblbm.LAYOUT_AA))
    auto2d = rblbm.Solver(rblbm.Config(dims=2, nx=16, ny=12, nz=1, layout=rblbm.LAYOUT_AUTO))
    assert auto2d.layout == rblbm.LAYOUT_TWO_ARRAY
    auto2d.close()
    sol = rblbm.Solver(rblbm.Config(dims=3, nx=16, ny=12, nz=10, layout=rblbm.LAYOUT_AA))
    with pytest.raises(rblbm.LbmError) as e:
        sol.set_kernel_variant(rblbm.KERNEL_ONE_PASS)
    assert e.value.code == -5
    sol.initial_seeded(1, 1e-3)
    sol.step_timed(0)      # the prologue alone: the plain f is gone in place
    with pytest.raises(rblbm.LbmError) as e:
        sol.macros()
    assert e.value.code == -4
    sol.step(1)
    sol.macros()
    sol.close()
    with pytest.raises(rblbm.LbmError):
        rblbm.Batch([rblbm.Config(dims=3, nx=16, ny=12, nz=10, layout=rblbm.LAYOUT_AA)])


def test_two_array_state_after_prologue_alone_is_the_plain_state():
    o = orc.Oracle(orc.Config(dims=3, nx=16, ny=12, nz=10, rayleigh=1e5))
    o.init_seeded(2, 1e-3)
    sol = rblbm.Solver(rblbm.Config(dims=3, nx=16, ny=12, nz=10, rayleigh=1e5))
    sol.initial_seeded(2, 1e-3)
    sol.step_timed(0)
    s, st, m = o.state(), sol.gather(with_bprev=False), sol.macros()
    np.testing.assert_array_equal(st["f"], s["f"])
    np.testing.assert_array_equal(st["g"], s["g"])
    np.testing.assert_array_equal(m["T"], s["T"])
    sol.close()
