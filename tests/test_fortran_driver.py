"""The Fortran driver (iso_c_binding over the C ABI) runs the reference's program structure on
the GPU and reproduces the oracle: check() values, end-of-run Nusselt numbers, and the
backupData restart file (Fortran unformatted records, population index outermost)."""
import os
import re
import subprocess

import numpy as np
import pytest

import orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "-2d-3d-_amd", "lib", "rb_main")


def _read_records(path):
    out = []
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0
    while pos < len(data):
        n = int(np.frombuffer(data, "<i4", 1, pos)[0])
        out.append(np.frombuffer(data, "<f8", n // 8, pos + 4))
        assert int(np.frombuffer(data, "<i4", 1, pos + 4 + n)[0]) == n
        pos += n + 8
    return out


def _run(tmp_path, nml):
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built")
    (tmp_path / "rb.nml").write_text("&rb " + nml + " /\n")
    r = subprocess.run([EXE, "rb.nml"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("dims", [3, 2])
def test_driver_matches_oracle(tmp_path, dims):
    nx, ny, nz = (24, 20, 16) if dims == 3 else (40, 21, 1)
    extra = "" if dims == 3 else ", velX=1, tempX=2, Thot=1.0d0, Tcold=0.0d0, Rayleigh=1.0d5"
    out = _run(tmp_path, f"dims={dims}, nx={nx}, ny={ny}, nz={nz}, Rayleigh=1.0d6, itc_max=99, "
                         f"checkInterval=50, backupAtEnd=1, seed=77{extra}")
    kw = dict(dims=dims, nx=nx, ny=ny, nz=nz, rayleigh=1e6)
    if dims == 2:
        kw.update(vel_x=1, temp_x=2, thot=1.0, tcold=0.0, rayleigh=1e5)
    o = orc.Oracle(orc.Config(**kw))
    o.init_seeded(77, 1e-3)
    checks = [tuple(map(float, m.groups())) for m in
              re.finditer(r"^\s+(?:50|100)\s+(\S+)\s+(\S+)$", out, re.M)]
    assert len(checks) == 2, out
    for i, n in enumerate((50, 50)):
        o.step(n)
        eu, et = o.check()
        assert checks[i][0] == pytest.approx(eu, rel=1e-10)
        assert checks[i][1] == pytest.approx(et, rel=1e-10)
    d = o.diagnostics()
    vals = dict(re.findall(r"^(\w+)\s*(?:\(\w+\))?\s*=\s*(\S+)$", out, re.M))
    assert float(vals["NuVolAvg"]) == pytest.approx(d["nu_vol"], rel=1e-7)   # printed es16.8
    assert float(vals["Nu_global"]) == pytest.approx(d["nu_global"], rel=1e-7)
    # mid-plane maxima, printed as the reference's end-of-run report does
    mp = o.midplane()
    tag = "" if dims == 3 else r"\*"
    got = re.search(rf"^u_mid_max{tag} =\s+(\S+)", out, re.M)
    assert got, out
    assert float(got.group(1)) == pytest.approx(mp["u_max"], rel=1e-7)
    got = re.search(rf"^v_mid_max{tag} =\s+(\S+)", out, re.M)
    assert float(got.group(1)) == pytest.approx(mp["v_max"], rel=1e-7)
    recs = _read_records(tmp_path / "backupFile3D-100.bin")
    s = o.state()
    qf, qt = (19, 7) if dims == 3 else (9, 5)
    f = recs[0].reshape(qf, -1)
    g = recs[1].reshape(qt, -1)
    np.testing.assert_array_equal(f, s["f"].reshape(-1, qf).T)
    np.testing.assert_array_equal(g, s["g"].reshape(-1, qt).T)


def test_driver_restart_continues_bitwise(tmp_path):
    base = "dims=3, nx=16, ny=12, nz=10, Rayleigh=1.0d5, Mach=0.3d0, checkInterval=1000"
    _run(tmp_path, base + ", itc_max=39, backupAtEnd=1")              # 40 steps -> backup
    (tmp_path / "backupFile3D-40.bin").rename(tmp_path / "restart.bin")
    _run(tmp_path, base + ", itc_max=29, loadInitField=1, reloadFile='restart.bin', backupAtEnd=1")
    recs = _read_records(tmp_path / "backupFile3D-30.bin")
    o = orc.Oracle(orc.Config(dims=3, nx=16, ny=12, nz=10, rayleigh=1e5, mach=0.3))
    o.init_seeded(20261203, 1e-3)
    o.step(40)
    s = o.state()
    r = orc.Oracle(orc.Config(dims=3, nx=16, ny=12, nz=10, rayleigh=1e5, mach=0.3))
    r.init_exact()
    assert r.set_fg(s["f"], s["g"]) == 0
    r.step(30)
    t = r.state()
    np.testing.assert_array_equal(recs[0].reshape(19, -1), t["f"].reshape(-1, 19).T)
