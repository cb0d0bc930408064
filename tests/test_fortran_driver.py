"""The Fortran driver (iso_c_binding over the C ABI) runs the reference's program structure on
the GPU and reproduces the oracle: check() values, end-of-run Nusselt numbers, and the
backupData restart file (Fortran unformatted records, population index outermost)."""
import os
import re
import subprocess

import numpy as np
import pytest

import frecords
import orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "-2d-3d-_amd", "lib", "rb_main")


def _read_records(path, subrecords=None):
    """Logical records as float64 arrays (GNU Fortran's convention, tests/frecords.py); with
    subrecords, also checks that every record of more than that many bytes was split."""
    recs, subs = frecords.read_records(path, with_subrecords=True)
    if subrecords is not None:
        for r, lens in zip(recs, subs):
            assert max(lens) <= subrecords and len(lens) == max(1, -(-len(r) // subrecords))
    return [np.frombuffer(r, "<f8") for r in recs]


def _run(tmp_path, nml):
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built")
    (tmp_path / "binFile").mkdir(exist_ok=True)   # the 2D program's binFolderPrefix folder
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
    if dims == 2:   # the stream-function report of calc_psi_vort_and_output
        ps = o.psi_vorticity_2d()
        got = re.search(r"^max\(\|psi\|\) =\s+(\S+)\s+x\* =\s+(\S+)\s+y\* =\s+(\S+)", out, re.M)
        assert got, out
        assert float(got.group(1)) == pytest.approx(ps["psi_abs_max"], rel=1e-7)
        assert float(got.group(2)) == pytest.approx(ps["x_at_max"], rel=1e-7, abs=1e-12)
        got = re.search(r"^abs\(psi_center_fine\) =\s+(\S+)", out, re.M)
        assert float(got.group(1)) == pytest.approx(ps["psi_center_fine"], rel=1e-7)
    recs = _read_records(tmp_path / ("backupFile3D-100.bin" if dims == 3 else "backupFile-100.bin"))
    s = o.state()
    qf, qt = (19, 7) if dims == 3 else (9, 5)
    f = recs[0].reshape(qf, -1)
    g = recs[1].reshape(qt, -1)
    np.testing.assert_array_equal(f, s["f"].reshape(-1, qf).T)
    np.testing.assert_array_equal(g, s["g"].reshape(-1, qt).T)
    # steadyFlow end of run: output_binary of the final state (2DRB/3DRB.F90:1599-1627,
    # 2DRB/2DRB.F90:1357-1388): u, v(, w) scaled by L / kappa, then T, rho
    p = orc.params(orc.Config(**kw))
    scale = p["L"] / p["diffusivity"]
    name = "buoyancyCavity3DbinFile-000000000100.bin" if dims == 3 else "binFile/buoyancyCavity-100.bin"
    b = _read_records(tmp_path / name)
    keys = ("u", "v", "w", "T", "rho") if dims == 3 else ("u", "v", "T", "rho")
    assert len(b) == len(keys)
    for rec, key in zip(b, keys):
        want = s[key].reshape(-1) * (scale if key in ("u", "v", "w") else 1.0)
        np.testing.assert_array_equal(rec, want, err_msg=key)


@pytest.mark.parametrize("max_sub", [2147483639, 1000])
def test_driver_restart_continues_bitwise(tmp_path, max_sub):
    # maxSubrecord = 1000 bytes splits the 15 KiB-per-plane records into subrecords, as records
    # over 2 GiB split in a full-size restart file; the reload reads them back
    base = (f"dims=3, nx=16, ny=12, nz=10, Rayleigh=1.0d5, Mach=0.3d0, checkInterval=1000, "
            f"endOutputs=0, maxSubrecord={max_sub}")
    _run(tmp_path, base + ", itc_max=39, backupAtEnd=1")              # 40 steps -> backup
    recs = _read_records(tmp_path / "backupFile3D-40.bin", subrecords=max_sub)
    assert [len(r) for r in recs] == [19 * 16 * 12 * 10, 7 * 16 * 12 * 10]
    (tmp_path / "backupFile3D-40.bin").rename(tmp_path / "restart.bin")
    _run(tmp_path, base + ", itc_max=29, loadInitField=1, reloadFile='restart.bin', backupAtEnd=1")
    recs = _read_records(tmp_path / "backupFile3D-30.bin", subrecords=max_sub)
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


def _read_plt(path):
    """Parse a one-zone point-format Tecplot v101 binary file written by the driver."""
    data = open(path, "rb").read()
    assert data[:8] == b"#!TDV101"
    pos = 8
    def i32():
        nonlocal pos
        v = int(np.frombuffer(data, "<i4", 1, pos)[0])
        pos += 4
        return v
    def f32():
        nonlocal pos
        v = float(np.frombuffer(data, "<f4", 1, pos)[0])
        pos += 4
        return v
    def string():
        chars = []
        while True:
            c = i32()
            if c == 0:
                return "".join(chars)
            chars.append(chr(c))
    assert i32() == 1
    title = string()
    nv = i32()
    names = [string() for _ in range(nv)]
    assert f32() == 299.0
    assert string() == "ZONE 001"
    assert [i32() for _ in range(5)] == [-1, 0, 1, 0, 0]
    dims = [i32() for _ in range(3)]
    assert i32() == 0 and f32() == 357.0 and f32() == 299.0
    assert [i32() for _ in range(nv)] == [2] * nv
    assert i32() == 0 and i32() == -1
    npts = dims[0] * dims[1] * dims[2]
    vals = np.frombuffer(data, "<f8", nv * npts, pos).reshape(npts, nv)
    assert pos + 8 * nv * npts == len(data)
    return title, names, dims, vals


def test_driver_tecplot_files_match_oracle(tmp_path):
    # outputPltFile=1: the full field and, per mid-plane, the interpolated fields and the
    # stream function / vorticity, in the reference's Tecplot layout (2DRB/3DRB.F90:1676-1701,
    # :3018-3266), values equal to the oracle's state after the same steps.
    nx, ny, nz = 14, 12, 10
    _run(tmp_path, f"dims=3, nx={nx}, ny={ny}, nz={nz}, Rayleigh=1.0d6, itc_max=59, "
                   "checkInterval=1000, seed=5, outputPltFile=1, pltFolderPrefix='cav'")
    o = orc.Oracle(orc.Config(dims=3, nx=nx, ny=ny, nz=nz, rayleigh=1e6))
    o.init_seeded(5, 1e-3)
    o.step(60)
    s = o.state()
    L = float(ny)
    title, names, dims, vals = _read_plt(tmp_path / "cav-fullField-000000000060.plt")
    assert (title, names, dims) == ("full-field-3d", ["X", "Y", "Z", "U", "V", "W", "T"], [nx, ny, nz])
    for col, key in ((3, "u"), (4, "v"), (5, "w"), (6, "T")):
        np.testing.assert_array_equal(vals[:, col], s[key].reshape(-1))
    np.testing.assert_allclose(vals[:nx, 0], (np.arange(nx) + 0.5) / L)
    planes = {0: ("X", "Y", "Z", ny, nz), 1: ("Y", "X", "Z", nx, nz), 2: ("Z", "X", "Y", nx, ny)}
    for axis, (a, c1, c2, n1, n2) in planes.items():
        sl = o.midplane_slice(axis)
        title, names, dims, vals = _read_plt(tmp_path / f"cav-mid{a}-000000000060.plt")
        assert names == [c1, c2, "U_nd", "V_nd", "W_nd", "T"] and dims == [n1, n2, 1]
        for col, key in ((2, "u"), (3, "v"), (4, "w"), (5, "T")):
            np.testing.assert_allclose(vals[:, col], sl[key].reshape(-1), rtol=1e-13, atol=1e-13)
        title, names, dims, vals = _read_plt(tmp_path / f"cav-mid{a}psiVort-000000000060.plt")
        assert names == [c1, c2, "PSI_" + c1 + c2, "VORT_" + a]
        np.testing.assert_allclose(vals[:, 2], sl["psi"].reshape(-1), rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(vals[:, 3], sl["vort"].reshape(-1), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dims", [3, 2])
def test_driver_calnure_series(tmp_path, dims):
    # nuReInLoop=1: calNuRe every int(outputFrequency * timeUnit) steps (the reference's
    # commented-out in-loop call, 2DRB/3DRB.F90:249-250), served by lbm_run_diagnostics after
    # the first sample, plus the end-of-run call (:297); time column as calNuRe writes it
    nx, ny, nz = (16, 12, 10) if dims == 3 else (40, 21, 1)
    extra = "" if dims == 3 else ", velX=1, tempX=2, Thot=1.0d0, Tcold=0.0d0"
    kw = dict(dims=dims, nx=nx, ny=ny, nz=nz, rayleigh=1e5)
    if dims == 2:
        kw.update(vel_x=1, temp_x=2, thot=1.0, tcold=0.0)
    cfg = orc.Config(**kw)
    cadence = int(0.1 * orc.params(cfg)["time_unit"])
    itc_max = 5 * cadence + 7          # six samples in the loop span, then the final call
    _run(tmp_path, f"dims={dims}, nx={nx}, ny={ny}, nz={nz}, Rayleigh=1.0d5, itc_max={itc_max}, "
                   f"checkInterval=100000, seed=9, nuReInLoop=1, outputFrequency=0.1d0, "
                   f"reloadDimensionlessTime=3{extra}")
    suffix = "_3D" if dims == 3 else ""
    nu = np.loadtxt(tmp_path / f"Nu_VolAvg{suffix}.dat", ndmin=2)
    re = np.loadtxt(tmp_path / f"Re_VolAvg{suffix}.dat", ndmin=2)
    steps = itc_max + 1
    sample_at = [cadence * k for k in range(1, steps // cadence + 1)] + [steps]
    assert nu.shape == (len(sample_at), 2) and re.shape == nu.shape
    np.testing.assert_allclose(nu[:, 0], 3 + 0.1 * np.arange(1, len(sample_at) + 1))
    o = orc.Oracle(cfg)
    o.init_seeded(9, 1e-3)
    done = 0
    for k, n in enumerate(sample_at):
        o.step(n - done)
        done = n
        d = o.diagnostics()
        assert nu[k, 1] == pytest.approx(d["nu_vol"], rel=1e-12, abs=1e-14), (k, n)
        assert re[k, 1] == pytest.approx(d["re_vol"], rel=1e-12, abs=1e-14), (k, n)


@pytest.mark.parametrize("dims", [3, 2])
def test_driver_unsteady_flow_cadence(tmp_path, dims):
    # unsteadyFlow (2DRB/3DRB.F90:256-262, 2DRB/2DRB.F90:269-275): every outputFrequency
    # free-fall times output_binary with the output count as the file number, every
    # backupInterval free-fall times backupData with the same number; no check() calls
    nx, ny, nz = (16, 12, 10) if dims == 3 else (40, 21, 1)
    extra = "" if dims == 3 else ", velX=1, tempX=2, Thot=1.0d0, Tcold=0.0d0"
    kw = dict(dims=dims, nx=nx, ny=ny, nz=nz, rayleigh=1e5, mach=0.3)
    if dims == 2:
        kw.update(vel_x=1, temp_x=2, thot=1.0, tcold=0.0)
    cfg = orc.Config(**kw)
    p = orc.params(cfg)
    cadence = int(0.5 * p["time_unit"])
    itc_max = 5 * cadence + 3
    out = _run(tmp_path, f"dims={dims}, nx={nx}, ny={ny}, nz={nz}, Rayleigh=1.0d5, Mach=0.3d0, "
                         f"itc_max={itc_max}, seed=13, flowMode=1, outputBinFile=1, "
                         f"outputFrequency=0.5d0, backupInterval=1, checkInterval=7{extra}")
    assert not re.search(r"^\s+\d+\s+\S+E[-+]\d+\s+\S+E[-+]\d+$", out, re.M), out   # no check()
    o = orc.Oracle(cfg)
    o.init_seeded(13, 1e-3)
    scale = p["L"] / p["diffusivity"]
    keys = ("u", "v", "w", "T", "rho") if dims == 3 else ("u", "v", "T", "rho")
    qf = 19 if dims == 3 else 9
    for n in range(1, 6):                       # output points at itc = n * cadence
        o.step(cadence)
        s = o.state()
        name = (f"buoyancyCavity3DbinFile-{n:012d}.bin" if dims == 3 else f"binFile/buoyancyCavity-{n}.bin")
        b = _read_records(tmp_path / name)
        assert len(b) == len(keys)
        for rec, key in zip(b, keys):
            want = s[key].reshape(-1) * (scale if key in ("u", "v", "w") else 1.0)
            np.testing.assert_array_equal(rec, want, err_msg=f"{name} {key}")
        bak = tmp_path / (f"backupFile3D-{n}.bin" if dims == 3 else f"backupFile-{n}.bin")
        assert bak.exists() == (n % 2 == 0), (n, bak)   # backupInterval / outputFrequency = 2
        if n % 2 == 0:
            np.testing.assert_array_equal(_read_records(bak)[0].reshape(qf, -1), s["f"].reshape(-1, qf).T)
    assert not (tmp_path / (f"buoyancyCavity3DbinFile-{6:012d}.bin" if dims == 3 else "binFile/buoyancyCavity-6.bin")).exists()
