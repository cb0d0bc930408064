"""lbm_step replays a HIP graph of 32 steps for runs of 128 steps or more on domains of up to
2^22 nodes.  The replayed run must leave exactly the state of the same steps launched one by
one (same kernels on the same buffers), keep the step counter, and leave every alternating
buffer and in-place state where the following eager steps and diagnostics expect it."""
import numpy as np
import pytest

import rblbm

pytestmark = pytest.mark.gpu

FORMS = {
    "fp64_two_array": dict(dims=3, nx=24, ny=20, nz=16, rayleigh=1e6),
    "fp64_in_place": dict(dims=3, nx=24, ny=20, nz=16, rayleigh=1e6, layout=rblbm.LAYOUT_AA),
    "fp16_two_array": dict(dims=3, nx=28, ny=18, nz=14, rayleigh=1e6, precision=rblbm.FP16_SHIFT),
    "fp32_in_place": dict(dims=3, nx=20, ny=16, nz=12, rayleigh=1e5, precision=rblbm.FP32_SHIFT,
                          layout=rblbm.LAYOUT_AA),
    "fp64_one_pass": dict(dims=3, nx=24, ny=20, nz=16, rayleigh=1e6),
    "d2_periodic": dict(dims=2, nx=40, ny=21, nz=1, rayleigh=1e5, vel_x=1, temp_x=2, thot=1.0,
                        tcold=0.0),
}


@pytest.mark.parametrize("name", list(FORMS))
def test_replayed_steps_equal_single_steps(name):
    cfg = rblbm.Config(**FORMS[name])
    a, b = rblbm.Solver(cfg), rblbm.Solver(cfg)
    for s in (a, b):
        if name == "fp64_one_pass":
            s.set_kernel_variant(rblbm.KERNEL_ONE_PASS, 8)
        s.initial_seeded(31, 0.05)
    a.step(203)                    # 1 eager step, 6 replays of 32, 10 eager steps
    for _ in range(203):
        b.step(1)
    assert a.itc == b.itc == 203
    for k, v in a.gather().items():
        np.testing.assert_array_equal(v, b.gather()[k], err_msg=f"{name} {k}")
    a.step(3)                      # eager steps after a replay find every buffer in place
    b.step(3)
    ma, mb = a.macros(), b.macros()
    for k in ("rho", "u", "v", "w", "T"):
        np.testing.assert_array_equal(ma[k], mb[k], err_msg=f"{name} {k}")
    assert a.diagnostics() == b.diagnostics()
    da = a.run_diagnostics(128, 64)
    db = b.run_diagnostics(128, 64)
    assert da == db and a.itc == b.itc == 334
    a.close()
    b.close()


def test_one_pass_diagnostics_every_step_right_after_initialisation():
    # the first period is the single two-pass step from the plain state; the one-pass launch
    # table is built by the next (eager) period before the capture
    cfg = rblbm.Config(dims=3, nx=24, ny=20, nz=16, rayleigh=1e6)
    a, b = rblbm.Solver(cfg), rblbm.Solver(cfg)
    for s in (a, b):
        s.set_kernel_variant(rblbm.KERNEL_ONE_PASS, 8)
        s.initial_seeded(5, 0.05)
    got = a.run_diagnostics(7, 1)
    ref = []
    for _ in range(7):
        b.step(1)
        ref.append(b.diagnostics())
    assert got == ref
    np.testing.assert_array_equal(a.gather()["f"], b.gather()["f"])
    a.close()
    b.close()
