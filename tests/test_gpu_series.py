"""Runs with periodic diagnostics replayed from a HIP graph (lbm_run_diagnostics and its batch
form) against the same steps run eagerly with lbm_step + lbm_diagnostics: every sample of every
replay must be identical (the graph holds the same kernels on the same buffers), and the state
after the run must be the eager state.  The graphs are replayed several times per call, calls
alternate with eager steps and synchronising reads, and the FP64 form is also checked against
the oracle."""
import numpy as np
import pytest

import orc
import rblbm

pytestmark = pytest.mark.gpu

DIAG_KEYS = ("nu_vol", "re_vol", "nu_global", "nu_hot", "nu_cold", "nu_middle", "nu_hot_max",
             "nu_hot_min", "mass", "heat", "kinetic")


def _eager(sol, nsteps, every):
    out = []
    for _ in range(nsteps // every):
        sol.step(every)
        out.append(sol.diagnostics())
    return out


def _same(a, b, label):
    assert len(a) == len(b), label
    for k, (x, y) in enumerate(zip(a, b)):
        for key in DIAG_KEYS:
            assert x[key] == y[key], (label, k, key, x[key], y[key])


@pytest.mark.parametrize("form", [
    dict(precision=rblbm.FP64, layout=rblbm.LAYOUT_TWO_ARRAY, variant=rblbm.KERNEL_TWO_PASS),
    dict(precision=rblbm.FP64, layout=rblbm.LAYOUT_AA, variant=rblbm.KERNEL_TWO_PASS),
    dict(precision=rblbm.FP64, layout=rblbm.LAYOUT_TWO_ARRAY, variant=rblbm.KERNEL_ONE_PASS),
    dict(precision=rblbm.FP16_SHIFT, layout=rblbm.LAYOUT_AA, variant=rblbm.KERNEL_TWO_PASS),
])
@pytest.mark.parametrize("every", [4, 3])     # even: one period per graph; odd: two
def test_run_diagnostics_equals_eager(form, every):
    cfg = rblbm.Config(dims=3, nx=28, ny=20, nz=18, rayleigh=1e6, precision=form["precision"],
                       layout=form["layout"])
    a, b = rblbm.Solver(cfg), rblbm.Solver(cfg)
    for s in (a, b):
        s.set_kernel_variant(form["variant"], 6)
        s.initial_seeded(12, 0.05)
    # graph runs alternate with eager steps and reads; 6 and 5 samples: several replays each
    ga = a.run_diagnostics(6 * every, every)
    _same(ga, _eager(b, 6 * every, every), f"{form} first call")
    for s in (a, b):
        s.step(5)
        s.macros()
    ga = a.run_diagnostics(5 * every, every)
    _same(ga, _eager(b, 5 * every, every), f"{form} second call")
    assert a.itc == b.itc == 11 * every + 5
    sa, sb = a.gather(), b.gather()
    for k in ("f", "g", "bx"):
        np.testing.assert_array_equal(sa[k], sb[k])
    if form["precision"] == rblbm.FP64:
        ora = orc.Oracle(orc.Config(dims=3, nx=28, ny=20, nz=18, rayleigh=1e6))
        ora.init_seeded(12, 0.05)
        ora.step(11 * every + 5)
        np.testing.assert_array_equal(sa["f"], ora.state()["f"])
        assert ga[-1]["nu_vol"] == pytest.approx(ora.diagnostics()["nu_vol"], rel=1e-12)
    a.close()
    b.close()


def test_run_diagnostics_2d_and_arguments():
    cfg = rblbm.Config(dims=2, nx=40, ny=21, nz=1, rayleigh=1e5, vel_x=1, temp_x=2, thot=1.0,
                       tcold=0.0)
    a, b = rblbm.Solver(cfg), rblbm.Solver(cfg)
    a.initial_seeded(3, 0.05)
    b.initial_seeded(3, 0.05)
    _same(a.run_diagnostics(50, 10), _eager(b, 50, 10), "2d")
    with pytest.raises(rblbm.LbmError):
        a.run_diagnostics(25, 10)          # not a multiple of the period
    with pytest.raises(rblbm.LbmError):
        a.run_diagnostics(10, 0)
    fresh = rblbm.Solver(cfg)
    with pytest.raises(rblbm.LbmError):   # no state yet
        fresh.run_diagnostics(10, 5)
    for s in (a, b, fresh):
        s.close()


@pytest.mark.parametrize("mode", [rblbm.BATCH_PER_DOMAIN, rblbm.BATCH_ONE_LAUNCH, rblbm.BATCH_GRAPH])
def test_batch_run_diagnostics_equals_eager(mode):
    shapes = [(24, 20, 16, 1e5), (32, 24, 20, 1e6), (17, 29, 13, 3e4)]
    cfgs = [rblbm.Config(dims=3, nx=a, ny=b, nz=c, rayleigh=ra) for a, b, c, ra in shapes]
    bt = rblbm.Batch(cfgs)
    bt.set_mode(mode)
    bt.initial_seeded(21, 0.05)
    ora = []
    for i, (a, b, c, ra) in enumerate(shapes):
        o = orc.Oracle(orc.Config(dims=3, nx=a, ny=b, nz=c, rayleigh=ra))
        o.init_seeded(21, 0.05, instance=i)
        ora.append(o)
    series = bt.run_diagnostics(7 * 4, 4)
    assert len(series) == 7 and all(len(s) == 3 for s in series)
    for k in range(7):
        for o in ora:
            o.step(4)
        for i, o in enumerate(ora):
            ref = o.diagnostics()
            for key in ("nu_vol", "re_vol", "nu_global", "nu_hot"):
                assert series[k][i][key] == pytest.approx(ref[key], rel=1e-12, abs=1e-14), (k, i, key)
    for d, o in zip(bt.domains, ora):
        np.testing.assert_array_equal(d.gather()["f"], o.state()["f"])
    bt.close()


@pytest.mark.parametrize("mode", [rblbm.BATCH_PER_DOMAIN, rblbm.BATCH_ONE_LAUNCH, rblbm.BATCH_GRAPH])
def test_batch_run_diagnostics_one_pass_every_step(mode):
    # every = 1 with one-pass domains: the first period is the single two-pass step from the
    # plain state, after which each domain's one-pass launch table still has to be built before
    # a period can be captured; samples and step counters must equal the oracle's
    shapes = [(24, 20, 16, 1e5), (32, 24, 20, 1e6)]
    cfgs = [rblbm.Config(dims=3, nx=a, ny=b, nz=c, rayleigh=ra) for a, b, c, ra in shapes]
    bt = rblbm.Batch(cfgs)
    bt.set_mode(mode)
    bt.set_kernel_variant(rblbm.KERNEL_ONE_PASS)
    bt.initial_seeded(8, 0.05)
    ora = []
    for i, (a, b, c, ra) in enumerate(shapes):
        o = orc.Oracle(orc.Config(dims=3, nx=a, ny=b, nz=c, rayleigh=ra))
        o.init_seeded(8, 0.05, instance=i)
        ora.append(o)
    series = bt.run_diagnostics(9, 1)
    assert len(series) == 9
    for k in range(9):
        for o in ora:
            o.step(1)
        for i, o in enumerate(ora):
            ref = o.diagnostics()
            for key in ("nu_vol", "re_vol", "nu_global"):
                assert series[k][i][key] == pytest.approx(ref[key], rel=1e-12, abs=1e-14), (k, i, key)
    for d, o in zip(bt.domains, ora):
        assert d.itc == 9
        np.testing.assert_array_equal(d.gather()["f"], o.state()["f"])
    bt.close()
