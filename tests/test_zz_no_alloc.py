"""lbm_run_diagnostics_noalloc (series in the library's device scratch, nothing allocated) against
lbm_run_diagnostics, both through the C ABI: bitwise equal samples on two calls in a row (the
second with fewer samples, so the scratch still holds the first call's longer series), on a
second, smaller domain in another storage form, and a clear refusal beyond the scratch size."""
import ctypes as C

import numpy as np
import pytest

import rblbm

pytestmark = pytest.mark.gpu

FIELDS = [f for f, _ in rblbm._Diag._fields_]


def _noalloc(sol, nsteps, every):
    L = sol.lib
    L.lbm_run_diagnostics_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_int32,
                                              C.POINTER(rblbm._Diag)]
    ns = nsteps // every
    out = (rblbm._Diag * ns)()
    rc = L.lbm_run_diagnostics_noalloc(sol._h, nsteps, every, out)
    return rc, [{f: getattr(out[k], f) for f in FIELDS} for k in range(ns)]


def _pair(**kw):
    cfg = rblbm.Config(**kw)
    a, b = rblbm.Solver(cfg), rblbm.Solver(cfg)
    a.initial_seeded(41, 0.05)
    b.initial_seeded(41, 0.05)
    return a, b


def _same(x, y):
    assert len(x) == len(y)
    for k, (p, q) in enumerate(zip(x, y)):
        for f in FIELDS:
            assert p[f] == q[f] or (np.isnan(p[f]) and np.isnan(q[f])), (k, f, p[f], q[f])


@pytest.mark.parametrize("variant", [rblbm.KERNEL_TWO_PASS, rblbm.KERNEL_ONE_PASS])
def test_noalloc_series_equals_run_diagnostics(variant):
    a, b = _pair(dims=3, nx=48, ny=40, nz=36, rayleigh=1e6)
    for s in (a, b):
        s.set_kernel_variant(variant)
    rc, got = _noalloc(a, 40, 5)
    assert rc == 0, a.lib.lbm_last_error(a._h)
    _same(got, b.run_diagnostics(40, 5))
    # second call, fewer samples: the scratch still holds the first call's series beyond them
    rc, got = _noalloc(a, 12, 4)
    assert rc == 0
    _same(got, b.run_diagnostics(12, 4))
    assert a.itc == b.itc == 52
    np.testing.assert_array_equal(a.gather()["f"], b.gather()["f"])
    # a second, smaller domain in the FP32-shift form on the same device's scratch
    c, d = _pair(dims=3, nx=24, ny=20, nz=16, rayleigh=1e5, precision=rblbm.FP32_SHIFT)
    rc, got = _noalloc(c, 9, 3)
    assert rc == 0
    _same(got, d.run_diagnostics(9, 3))
    for s in (a, b, c, d):
        s.close()


def test_noalloc_2d_and_limits():
    a, b = _pair(dims=2, nx=40, ny=21, nz=1, rayleigh=1e5, vel_x=1, temp_x=2, thot=1.0, tcold=0.0)
    rc, got = _noalloc(a, 30, 10)
    assert rc == 0
    _same(got, b.run_diagnostics(30, 10))
    # beyond the scratch: 12,576 doubles a sample at 48 x 40 x 36, so 400 samples do not fit
    e, f = _pair(dims=3, nx=48, ny=40, nz=36, rayleigh=1e6)
    rc, _ = _noalloc(e, 400, 1)
    assert rc == -1
    msg = e.lib.lbm_last_error(e._h).decode()
    assert "scratch" in msg and "at most" in msg
    assert e.itc == 0                       # refused before any step
    rc, _ = _noalloc(e, 10, 3)              # not a multiple of the period
    assert rc == -1
    for s in (a, b, e, f):
        s.close()
