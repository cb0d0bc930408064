"""The one-pass step (tile + ring, z-marching) against the two-pass step and the oracle.

In the FP64 form both kernels keep the reference's arithmetic, so they must agree bit for bit
with each other and with the oracle, on grids that leave partial tiles and partial z-chunks.
"""
import numpy as np
import pytest

import orc
import rblbm

pytestmark = pytest.mark.gpu

CASES = {
    "rb_cavity": dict(),
    "side_heated": dict(side_heated=1, temp_x=0, temp_y=1, temp_z=1),
    "periodic_velocity_walls_heat": dict(vel_x=1, vel_z=1),
    "no_m1g": dict(use_g=0),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("kchunk", [5, 64, rblbm.KCHUNK_AUTO])
def test_one_pass_fp64_bitwise(name, kchunk):
    dims = dict(dims=3, nx=45, ny=37, nz=29, rayleigh=1e5, mach=0.3, **CASES[name])
    ora = orc.Oracle(orc.Config(**dims))
    one = rblbm.Solver(rblbm.Config(**dims))
    two = rblbm.Solver(rblbm.Config(**dims))
    one.set_kernel_variant(rblbm.KERNEL_ONE_PASS, kchunk)
    two.set_kernel_variant(rblbm.KERNEL_TWO_PASS)
    for s in (ora, ):
        s.init_seeded(4, 0.2)
    one.initial_seeded(4, 0.2)
    two.initial_seeded(4, 0.2)
    done = 0
    for n in (1, 2, 3, 25):
        ora.step(n - done)
        one.step(n - done)
        two.step(n - done)
        done = n
        a, b, s = one.gather(), two.gather(), ora.state()
        np.testing.assert_array_equal(a["f"], s["f"])
        np.testing.assert_array_equal(a["g"], s["g"])
        np.testing.assert_array_equal(a["bx"], s["bx"])
        np.testing.assert_array_equal(b["f"], s["f"])
        m = one.macros()
        np.testing.assert_array_equal(m["u"], s["u"])
        np.testing.assert_array_equal(m["T"], s["T"])


TILES = [rblbm.TILE_32X16, rblbm.TILE_16X16, rblbm.TILE_32X8, rblbm.TILE_32X32,
         rblbm.TILE_34X15, rblbm.TILE_66X15]
PREFETCH = [rblbm.PREFETCH_NONE, rblbm.PREFETCH_REGS, rblbm.PREFETCH_STAGED]


def _form_supported(tile, prefetch, prec):
    if tile in (rblbm.TILE_32X32, rblbm.TILE_66X15) and prec == rblbm.FP64:
        return False                                 # 128 registers per lane: FP64 does not fit
    if prefetch == rblbm.PREFETCH_STAGED:
        return prec == rblbm.FP32_SHIFT and tile in (rblbm.TILE_34X15, rblbm.TILE_66X15)
    return True


# FP16 runs the tiles of the default and wide shapes only (the others share its code)
TILE_CASES = [(tile, prefetch, prec)
              for prec in (rblbm.FP64, rblbm.FP32_SHIFT, rblbm.FP16_SHIFT)
              for prefetch in PREFETCH for tile in TILES
              if prec != rblbm.FP16_SHIFT or tile in (rblbm.TILE_34X15, rblbm.TILE_66X15)]


@pytest.mark.parametrize("tile, prefetch, prec", TILE_CASES)
def test_one_pass_tiles_and_prefetch(tile, prefetch, prec):
    """Every tile shape and prefetching mode gives the default kernel's results: bitwise
    against the oracle in FP64, bitwise against the two-pass form in FP32 and FP16 (the same
    per-node arithmetic, products fused only inside a source expression, and the exchanged
    thermal populations rounded as the storage rounds them).  Unsupported combinations are
    refused with LBM_ERR_UNSUPPORTED."""
    dims = dict(dims=3, nx=45, ny=37, nz=29, rayleigh=1e5, mach=0.3, precision=prec)
    sol = rblbm.Solver(rblbm.Config(**dims))
    sol.set_kernel_variant(rblbm.KERNEL_ONE_PASS, 7)
    if not _form_supported(tile, prefetch, prec):
        with pytest.raises(rblbm.LbmError) as e:
            sol.set_fused_tile(tile, prefetch)
        assert e.value.code == rblbm.ERR_UNSUPPORTED
        return
    sol.set_fused_tile(tile, prefetch)
    assert sol.kernel_form() == dict(variant=rblbm.KERNEL_ONE_PASS, kchunk=7, tile=tile,
                                     prefetch=prefetch)
    sol.initial_seeded(4, 0.2)
    sol.step(12)
    if prec == rblbm.FP64:
        ora = orc.Oracle(orc.Config(**{k: v for k, v in dims.items() if k != "precision"}))
        ora.init_seeded(4, 0.2)
        ora.step(12)
        s = ora.state()
    else:
        ref = rblbm.Solver(rblbm.Config(**dims))
        ref.set_kernel_variant(rblbm.KERNEL_TWO_PASS)
        ref.initial_seeded(4, 0.2)
        ref.step(12)
        s = ref.gather()
    a = sol.gather()
    for k in ("f", "g", "bx", "by", "bz"):
        np.testing.assert_array_equal(a[k], s[k])


def test_one_pass_switching_and_periodic_heat_rejected():
    dims = dict(dims=3, nx=20, ny=18, nz=16, rayleigh=1e6)
    ora = orc.Oracle(orc.Config(**dims))
    sol = rblbm.Solver(rblbm.Config(**dims))
    ora.init_seeded(8, 0.01)
    sol.initial_seeded(8, 0.01)
    for variant, n in ((rblbm.KERNEL_ONE_PASS, 7), (rblbm.KERNEL_TWO_PASS, 3),
                       (rblbm.KERNEL_ONE_PASS, 9)):
        sol.set_kernel_variant(variant, 4)
        sol.step(n)
        ora.step(n)
        np.testing.assert_array_equal(sol.gather()["f"], ora.state()["f"])
    per = rblbm.Solver(rblbm.Config(dims=3, nx=8, ny=8, nz=8, temp_x=2, vel_x=1))
    with pytest.raises(rblbm.LbmError):
        per.set_kernel_variant(rblbm.KERNEL_ONE_PASS)


def test_autotune_caches_choice_and_keeps_state(tmp_path):
    dims = dict(dims=3, nx=45, ny=37, nz=29, rayleigh=1e5)
    cache = str(tmp_path / "tune.txt")
    sol = rblbm.Solver(rblbm.Config(**dims))
    sol.initial_seeded(2, 0.01)
    ms = sol.autotune(cache, steps=3)
    assert ms > 0
    lines = open(cache).read().split("\n")
    assert lines[0].startswith("d3_45x37x29_p0")
    assert sol.autotune(cache, steps=3) == 0.0           # from the cache, no timing
    ora = orc.Oracle(orc.Config(**dims))
    ora.init_seeded(2, 0.01)
    sol.step(6)
    ora.step(6)
    assert sol.itc == 6
    np.testing.assert_array_equal(sol.gather()["f"], ora.state()["f"])


@pytest.mark.parametrize("prec", [rblbm.FP32_SHIFT, rblbm.FP16_SHIFT])
def test_one_pass_lowprec_matches_two_pass(prec):
    dims = dict(dims=3, nx=45, ny=37, nz=29, rayleigh=1e6, precision=prec)
    one = rblbm.Solver(rblbm.Config(**dims))
    two = rblbm.Solver(rblbm.Config(**dims))
    one.set_kernel_variant(rblbm.KERNEL_ONE_PASS, 6)
    two.set_kernel_variant(rblbm.KERNEL_TWO_PASS)
    one.initial_seeded(4, 0.05)
    two.initial_seeded(4, 0.05)
    one.step(40)
    two.step(40)
    a, b = one.macros(), two.macros()
    # the same arithmetic per node in both kernel forms: bitwise equal after 40 steps
    for k in ("rho", "u", "v", "w", "T"):
        np.testing.assert_array_equal(a[k], b[k])


@pytest.mark.parametrize("prec", [rblbm.FP64, rblbm.FP16_SHIFT])
@pytest.mark.parametrize("shape", [(96, 80, 40), (130, 70, 33)])
def test_one_pass_round_schedule_covers_every_tile(prec, shape):
    """Domains with several rounds of resident blocks and partial strips, bands and rounds: the
    block schedule visits every tile and z-chunk once (bitwise equal to the two-pass form)."""
    nx, ny, nz = shape
    dims = dict(dims=3, nx=nx, ny=ny, nz=nz, rayleigh=1e6, mach=0.2, precision=prec)
    one = rblbm.Solver(rblbm.Config(**dims))
    two = rblbm.Solver(rblbm.Config(**dims))
    one.set_kernel_variant(rblbm.KERNEL_ONE_PASS, 4)   # many z-chunks: several rounds
    one.set_fused_tile(rblbm.TILE_34X15, 1)
    two.set_kernel_variant(rblbm.KERNEL_TWO_PASS)
    one.initial_seeded(6, 0.1)
    two.initial_seeded(6, 0.1)
    one.step(9)
    two.step(9)
    a, b = one.gather(), two.gather()
    for k in ("f", "g", "bx", "by", "bz"):
        np.testing.assert_array_equal(a[k], b[k])


# the 66x15 tile is a reduced-precision tile
@pytest.mark.parametrize("prec, tile, prefetch", [
    (prec, tile, prefetch)
    for tile, prefetch in ((rblbm.TILE_34X15, rblbm.PREFETCH_REGS), (rblbm.TILE_34X15, rblbm.PREFETCH_NONE),
                           (rblbm.TILE_66X15, rblbm.PREFETCH_REGS))
    for prec in (rblbm.FP64, rblbm.FP32_SHIFT, rblbm.FP16_SHIFT)
    if not (tile == rblbm.TILE_66X15 and prec == rblbm.FP64)])
def test_one_pass_interior_waves(prec, tile, prefetch):
    """A grid on which whole waves hold interior nodes only, so that their pulls take the
    buffer-descriptor loads (34x15 tiles 1-2 cover x = 31..96 and tile row 1 covers
    y = 12..26; the 66x15 tile 1 covers x = 63..128), beside waves with wall and out-of-domain
    lanes: FP64 bitwise against the oracle, FP32 / FP16 bitwise against the two-pass form."""
    dims = dict(dims=3, nx=140, ny=44, nz=24, rayleigh=1e5, mach=0.3, precision=prec)
    sol = rblbm.Solver(rblbm.Config(**dims))
    sol.set_kernel_variant(rblbm.KERNEL_ONE_PASS, 8)
    sol.set_fused_tile(tile, prefetch)
    sol.initial_seeded(6, 0.2)
    sol.step(9)
    if prec == rblbm.FP64:
        ora = orc.Oracle(orc.Config(**{k: v for k, v in dims.items() if k != "precision"}))
        ora.init_seeded(6, 0.2)
        ora.step(9)
        s = ora.state()
    else:
        ref = rblbm.Solver(rblbm.Config(**dims))
        ref.set_kernel_variant(rblbm.KERNEL_TWO_PASS)
        ref.initial_seeded(6, 0.2)
        ref.step(9)
        s = ref.gather()
    a = sol.gather()
    for k in ("f", "g", "bx", "by", "bz"):
        np.testing.assert_array_equal(a[k], s[k])


@pytest.mark.parametrize("n, prec, variant, kchunk", [
    (128, rblbm.FP64, rblbm.KERNEL_TWO_PASS, 8), (192, rblbm.FP64, rblbm.KERNEL_ONE_PASS, 8),
    (256, rblbm.FP64, rblbm.KERNEL_ONE_PASS, 16), (192, rblbm.FP32_SHIFT, rblbm.KERNEL_TWO_PASS, 8),
    (256, rblbm.FP32_SHIFT, rblbm.KERNEL_ONE_PASS, 16), (256, rblbm.FP16_SHIFT, rblbm.KERNEL_TWO_PASS, 16)])
def test_auto_form_by_domain_size(n, prec, variant, kchunk):
    """AUTO runs the one-pass form for FP64 domains of 2^22 nodes or more and FP32-shift domains
    of 2^24 or more, with z-chunks of the largest power of two up to nz / 16 within 8..32."""
    sol = rblbm.Solver(rblbm.Config(dims=3, nx=n, ny=n, nz=n, rayleigh=1e6, precision=prec))
    f = sol.kernel_form()
    assert (f["variant"], f["kchunk"]) == (variant, kchunk)
    sol.close()
