"""The driver's record module (rb_records.f90) against an independent reader / writer of GNU
Fortran's subrecord convention (tests/frecords.py), on the CPU: records longer than the
subrecord limit split into subrecords with sign-flagged markers, in both directions.  The limit
is a parameter, so small arrays exercise the splitting that records over 2 GiB need."""
import os
import struct
import subprocess

import numpy as np
import pytest

import frecords

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = os.environ.get("AMDFLANG", "/opt/rocm/bin/amdflang")
SIZES = [1, 5, 250, 1000, 0, 37]


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if not os.path.exists(FC):
        pytest.skip("no Fortran compiler")
    d = tmp_path_factory.mktemp("rec")
    src = [os.path.join(ROOT, "-2d-3d-_amd", "fortran", "rb_records.f90"),
           os.path.join(ROOT, "tests", "fortran", "records_check.f90")]
    subprocess.run([FC, "-O1", "-module-dir", str(d), *src, "-o", str(d / "records_check")],
                   check=True, capture_output=True)
    return str(d / "records_check")


def _expected(k, n):
    return np.arange(1, n + 1, dtype=np.float64) + 1000.0 * k


@pytest.mark.parametrize("max_sub", [frecords.MAX_SUBRECORD, 8, 24, 100, 7, 1])
def test_fortran_writer_against_python_reader(exe, tmp_path, max_sub):
    path = tmp_path / "w.bin"
    subprocess.run([exe, "write", str(path), str(max_sub)] + [str(n) for n in SIZES], check=True)
    recs, subs = frecords.read_records(path, with_subrecords=True)
    assert len(recs) == len(SIZES)
    for k, (rec, n, lens) in enumerate(zip(recs, SIZES, subs), start=1):
        np.testing.assert_array_equal(np.frombuffer(rec, "<f8"), _expected(k, n))
        # subrecords of max_sub bytes, the last one holding the rest
        nb = 8 * n
        want = [max_sub] * (nb // max_sub) + ([nb % max_sub] if nb % max_sub else [])
        assert lens == (want or [0]), (k, lens)


@pytest.mark.parametrize("max_sub", [frecords.MAX_SUBRECORD, 16, 9, 3])
def test_python_writer_against_fortran_reader(exe, tmp_path, max_sub):
    path = tmp_path / "p.bin"
    frecords.write_records(path, [_expected(k, n).tobytes() for k, n in enumerate(SIZES, start=1)],
                           max_sub=max_sub)
    out = subprocess.run([exe, "read", str(path)] + [str(n) for n in SIZES], check=True,
                         capture_output=True, text=True).stdout.split("\n")
    for k, n in enumerate(SIZES, start=1):
        kk, s = out[k - 1].split()
        assert int(kk) == k and float(s) == pytest.approx(_expected(k, n).sum(), rel=1e-15)


def test_python_reader_rejects_a_broken_marker(tmp_path):
    path = tmp_path / "b.bin"
    frecords.write_records(path, [b"x" * 20], max_sub=8)
    raw = bytearray(path.read_bytes())
    raw[4 + 8:4 + 8 + 4] = struct.pack("<i", -8)   # first subrecord's trailing marker negated
    path.write_bytes(bytes(raw))
    with pytest.raises(ValueError):
        frecords.read_records(path)
