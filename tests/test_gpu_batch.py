"""Batched ensembles: independent cavities of different shapes and parameters advancing in one
table launch per pass (two-pass form) or per step (one-pass form), plain or HIP-graph replay,
or one launch per domain, against the oracle."""
import numpy as np
import pytest

import orc
import rblbm

pytestmark = pytest.mark.gpu

SHAPES = [(24, 20, 16, 1e5, 0.71), (32, 32, 32, 1e6, 7.0), (17, 29, 13, 3e4, 0.71),
          (40, 24, 20, 1e7, 7.0), (16, 16, 48, 1e4, 0.71)]


def _cfgs(prec):
    return [rblbm.Config(dims=3, nx=a, ny=b, nz=c, rayleigh=ra, prandtl=pr, precision=prec)
            for a, b, c, ra, pr in SHAPES]


def _oracles(seed, amp):
    out = []
    for i, (a, b, c, ra, pr) in enumerate(SHAPES):
        o = orc.Oracle(orc.Config(dims=3, nx=a, ny=b, nz=c, rayleigh=ra, prandtl=pr))
        o.init_seeded(seed, amp, instance=i)
        out.append(o)
    return out


VARIANTS = [rblbm.KERNEL_TWO_PASS, rblbm.KERNEL_ONE_PASS]


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("mode", [rblbm.BATCH_PER_DOMAIN, rblbm.BATCH_ONE_LAUNCH, rblbm.BATCH_GRAPH])
def test_batch_fp64_bitwise_against_oracle(mode, variant):
    b = rblbm.Batch(_cfgs(rblbm.FP64))
    b.set_mode(mode)
    b.set_kernel_variant(variant)
    b.initial_seeded(31, 0.05)
    ora = _oracles(31, 0.05)
    done = 0
    for n in (1, 2, 5, 12):
        b.step(n - done)
        for o in ora:
            o.step(n - done)
        done = n
        for d, o in zip(b.domains, ora):
            st, s = d.gather(), o.state()
            np.testing.assert_array_equal(st["f"], s["f"])
            np.testing.assert_array_equal(st["g"], s["g"])
            np.testing.assert_array_equal(st["bx"], s["bx"])
    for d, o in zip(b.domains, ora):
        assert d.diagnostics()["nu_vol"] == pytest.approx(o.diagnostics()["nu_vol"], rel=1e-12)
    assert b.nodes == sum(a * bb * c for a, bb, c, _, _ in SHAPES)
    b.close()


@pytest.mark.parametrize("variant", VARIANTS)
def test_batch_fp32_graph_tracks_oracle(variant):
    b = rblbm.Batch(_cfgs(rblbm.FP32_SHIFT))
    b.set_mode(rblbm.BATCH_GRAPH)
    b.set_kernel_variant(variant)
    b.initial_seeded(5, 0.05)
    ora = _oracles(5, 0.05)
    b.step(101)                     # first step two-pass, then 50 graph replays
    for o in ora:
        o.step(101)
    for d, o in zip(b.domains, ora):
        m, s = d.macros(), o.state()
        umax = max(np.abs(s[k]).max() for k in ("u", "v", "w"))
        du = max(np.abs(m[k] - s[k]).max() for k in ("u", "v", "w")) / umax
        assert du <= 2e-4
        assert np.abs(m["T"] - s["T"]).max() <= 1e-5
    b.close()


def test_batch_rejects_mixed_precision_and_owned_domain_destroy():
    cfgs = _cfgs(rblbm.FP64)
    cfgs[1].precision = rblbm.FP32_SHIFT
    with pytest.raises(rblbm.LbmError):
        rblbm.Batch(cfgs)
    b = rblbm.Batch(_cfgs(rblbm.FP64)[:2])
    lib = b.lib
    assert lib.lbm_destroy(b.domains[0]._h) == -4
    with pytest.raises(rblbm.LbmError):
        b.set_kernel_variant(7)
    b.close()


@pytest.mark.parametrize("variant", VARIANTS)
@pytest.mark.parametrize("prec", [rblbm.FP32_SHIFT, rblbm.FP16_SHIFT])
def test_batch_lowprec_table_launch_equals_per_domain(variant, prec):
    # the table launches run the same node arithmetic as the per-domain kernels: bitwise equal
    # (FP32 one-pass: the default 66x15 tile without prefetching, one launch for all domains)
    def run(mode):
        b = rblbm.Batch(_cfgs(prec))
        b.set_mode(mode)
        b.set_kernel_variant(variant)
        b.initial_seeded(9, 0.05)
        b.step(23)
        out = [d.macros() for d in b.domains]
        b.close()
        return out
    for m, r in zip(run(rblbm.BATCH_GRAPH), run(rblbm.BATCH_PER_DOMAIN)):
        for k in ("rho", "u", "v", "w", "T"):
            np.testing.assert_array_equal(m[k], r[k])


WALLS = [dict(),                                                        # RB cell, adiabatic sides
         dict(vel_x=1, vel_z=1),                                       # periodic flow in x, z
         dict(side_heated=1, temp_x=0, temp_y=1, thot=1.0, tcold=0.0),  # isothermal x walls
         dict(use_g=0)]                                                # no M1G correction


@pytest.mark.parametrize("prec", [rblbm.FP32_SHIFT, rblbm.FP16_SHIFT])
def test_lowprec_domain_kernels_equal_table_kernels_on_every_wall(prec):
    # A domain's own two-pass kernels load interior rows at precomputed offsets and redo the
    # x-wall nodes with the generic pulls; the batch table kernels use the generic pulls for
    # every node.  Both must agree bitwise for every kind of wall, including nodes on x walls.
    shapes = [(40, 20, 12), (33, 18, 14), (24, 16, 10), (20, 14, 12)]
    cfgs = [rblbm.Config(dims=3, nx=a, ny=b, nz=c, rayleigh=1e6, precision=prec, **w)
            for (a, b, c), w in zip(shapes, WALLS)]

    def run(mode):
        b = rblbm.Batch(cfgs)
        b.set_mode(mode)
        b.set_kernel_variant(rblbm.KERNEL_TWO_PASS)
        b.initial_seeded(17, 0.05)
        b.step(13)
        out = [(d.macros(), d.gather(with_bprev=False)) for d in b.domains]
        b.close()
        return out
    for n, ((m1, s1), (m2, s2)) in enumerate(zip(run(rblbm.BATCH_PER_DOMAIN),
                                                run(rblbm.BATCH_ONE_LAUNCH))):
        for k in ("rho", "u", "v", "w", "T"):
            np.testing.assert_array_equal(m1[k], m2[k], err_msg=f"{WALLS[n]} {k}")
        for k in ("f", "g"):
            np.testing.assert_array_equal(s1[k], s2[k], err_msg=f"{WALLS[n]} {k}")


@pytest.mark.parametrize("prec", [rblbm.FP32_SHIFT, rblbm.FP16_SHIFT, rblbm.FP64])
def test_wide_rows_domain_kernels_equal_table_kernels(prec):
    # rows of 160 and 300 nodes: the per-row kernels' last block of a row has idle threads
    # (192 threads for 160 nodes; 256 + 44 for 300); the table kernels' 256-node blocks do not
    # follow rows
    cfgs = [rblbm.Config(dims=3, nx=nx, ny=10, nz=8, rayleigh=1e6, precision=prec)
            for nx in (160, 300)]

    def run(mode):
        b = rblbm.Batch(cfgs)
        b.set_mode(mode)
        b.set_kernel_variant(rblbm.KERNEL_TWO_PASS)
        b.initial_seeded(23, 0.05)
        b.step(9)
        out = [d.gather(with_bprev=False) for d in b.domains]
        b.close()
        return out
    for s1, s2 in zip(run(rblbm.BATCH_PER_DOMAIN), run(rblbm.BATCH_ONE_LAUNCH)):
        for k in ("f", "g"):
            np.testing.assert_array_equal(s1[k], s2[k])


def test_batch_domain_refuses_state_changes_of_its_own():
    # stepping, initialising or scattering one domain of a batch would leave the batch's launch
    # tables describing stale buffers: refused, and the batch stays bitwise on the oracle
    b = rblbm.Batch(_cfgs(rblbm.FP64)[:2])
    b.set_mode(rblbm.BATCH_ONE_LAUNCH)
    b.initial_seeded(17, 0.05)
    ora = _oracles(17, 0.05)[:2]
    b.step(3)
    d = b.domains[1]
    st = d.gather()
    for call in (lambda: d.step(1), lambda: d.step_timed(1), lambda: d.initial_seeded(1, 0.01),
                 lambda: d.initial(), lambda: d.scatter(st["f"], st["g"]),
                 lambda: d.run_diagnostics(2, 1)):
        with pytest.raises(rblbm.LbmError) as e:
            call()
        assert e.value.code == -4
    b.step(4)
    for o in ora:
        o.step(7)
    for dd, o in zip(b.domains, ora):
        np.testing.assert_array_equal(dd.gather()["f"], o.state()["f"])
        np.testing.assert_array_equal(dd.gather()["g"], o.state()["g"])
    b.close()


def test_batch_autotune_per_shape_cache_and_bitwise_state(tmp_path):
    # lbm_batch_autotune tunes each distinct shape once (one cache line per shape), gives the
    # domains of one shape the same form, gives every domain domain 0's tile and prefetch (one
    # launch has one block shape), reads the cache on a second call, and leaves the results
    # bitwise those of the untuned batch (every form runs the same node arithmetic)
    shapes = [(24, 20, 16, 1e5, 0.71), (32, 32, 32, 1e6, 7.0), (24, 20, 16, 3e4, 7.0)]
    cfgs = [rblbm.Config(dims=3, nx=a, ny=b, nz=c, rayleigh=ra, prandtl=pr, precision=rblbm.FP32_SHIFT)
            for a, b, c, ra, pr in shapes]
    cache = str(tmp_path / "batch_tune.txt")
    tuned = rblbm.Batch(cfgs)
    tuned.initial_seeded(23, 0.05)
    tuned.autotune(cache, steps=2)
    lines = [l for l in open(cache).read().split("\n") if l]
    assert sorted(l.split()[0].split("_p")[0] for l in lines) == ["d3_24x20x16", "d3_32x32x32"]
    forms = [d.kernel_form() for d in tuned.domains]
    assert forms[0] == forms[2]
    assert all((f["tile"], f["prefetch"]) == (forms[0]["tile"], forms[0]["prefetch"]) for f in forms)
    tuned.autotune(cache, steps=2)                      # from the cache: nothing appended
    assert len([l for l in open(cache).read().split("\n") if l]) == len(lines)
    plain = rblbm.Batch(cfgs)
    plain.initial_seeded(23, 0.05)
    for b in (tuned, plain):
        b.set_mode(rblbm.BATCH_ONE_LAUNCH)
        b.step(9)
    for d1, d2 in zip(tuned.domains, plain.domains):
        s1, s2 = d1.gather(), d2.gather()
        for k in ("f", "g", "bx", "by", "bz"):
            np.testing.assert_array_equal(s1[k], s2[k])
    tuned.close()
    plain.close()


@pytest.mark.parametrize("prec", [rblbm.FP64, rblbm.FP32_SHIFT])
def test_batch_auto_table_form_by_total_size(prec):
    # under AUTO a one-launch step runs the one-pass table form once the batch as a whole has
    # 2^22 (FP64) or 2^24 (FP32-shift) nodes, though each cavity alone runs the two-pass form:
    # bitwise the per-domain steps (FP64 also the oracle's)
    shapes = [(128, 128, 128, 1e6, 0.71), (128, 128, 128, 3e5, 7.0)] if prec == rblbm.FP64 else \
             [(256, 256, 128, 1e6, 0.71), (256, 256, 128, 3e5, 7.0)]
    cfgs = [rblbm.Config(dims=3, nx=a, ny=b, nz=c, rayleigh=ra, prandtl=pr, precision=prec)
            for a, b, c, ra, pr in shapes]
    # a run switching modes mid-way passes between the one-pass table form and each cavity's
    # two-pass form in both directions
    runs = {"graph": [(rblbm.BATCH_GRAPH, 11)], "per_domain": [(rblbm.BATCH_PER_DOMAIN, 11)],
            "mixed": [(rblbm.BATCH_GRAPH, 5), (rblbm.BATCH_PER_DOMAIN, 3), (rblbm.BATCH_ONE_LAUNCH, 3)]}
    out = {}
    for name, plan in runs.items():
        b = rblbm.Batch(cfgs)
        assert all(d.kernel_form()["variant"] == rblbm.KERNEL_TWO_PASS for d in b.domains)
        b.initial_seeded(41, 0.05)
        for mode, n in plan:
            b.set_mode(mode)
            b.step(n)
        out[name] = [d.gather() for d in b.domains]
        b.close()
    for name in ("graph", "mixed"):
        for s1, s2 in zip(out[name], out["per_domain"]):
            for k in ("f", "g", "bx", "by", "bz"):
                np.testing.assert_array_equal(s1[k], s2[k], err_msg=f"{name} {k}")
