"""The C-ABI library loads and exports every entry point include/lbm_rb.h declares (CPU only)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm_rb.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("lbm_create", "lbm_step", "lbm_macros", "lbm_diagnostics", "lbm_gather",
              "lbm_scatter", "lbm_destroy", "lbm_check"):
        assert n in names


def test_python_constants_match_the_header():
    """Every kernel-form, tile, prefetch, layout, precision and error constant of the ctypes
    binding has the value its #define gives."""
    import rblbm
    text = open(HEADER).read()
    defs = dict(re.findall(r"#define LBM_(\w+)\s+\(?(-?\d+)\)?", text))
    names = [n for n in defs if n.split("_")[0] in ("KERNEL", "TILE", "PREFETCH", "LAYOUT", "ERR",
                                                    "KCHUNK")]
    assert len(names) >= 20
    for n in names:
        assert getattr(rblbm, n) == int(defs[n]), n
    for n in ("FP64", "FP32_SHIFT", "FP16_SHIFT"):
        assert getattr(rblbm, n) == int(defs[n]), n


def test_library_exports_every_declared_symbol():
    import rblbm
    if not os.path.exists(rblbm.LIB_PATH):
        pytest.skip("librblbm.so not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", rblbm.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (lbm_\w+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing
    lib = rblbm.library()            # dlopen works without a GPU
    for n in declared():
        assert hasattr(lib, n)
    assert lib.lbm_error_string(-6).decode().startswith("non-positive")


def test_struct_layouts_match_the_python_mirrors(tmp_path):
    # every field of the ABI's structs at the same offset in include/lbm_rb.h (compiled by the
    # host C compiler) and in the ctypes mirrors the tests and tools use
    import ctypes
    import shutil

    import rblbm
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no host C compiler")
    mirrors = {"lbm_config": rblbm._Config, "lbm_params": rblbm._Params,
               "lbm_diag": rblbm._Diag, "lbm_midplane": rblbm._Mid}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "lbm_rb.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n"):
        if line:
            s, f, v = line.split()
            got[(s, f)] = int(v)
    for cname, py in mirrors.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)


def test_fortran_binding_types_match_the_header(tmp_path):
    # the bind(C) derived types of fortran/lbm_rb.f90 have the header's sizes and field offsets
    import shutil

    import rblbm
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    cc = shutil.which("gcc") or shutil.which("cc")
    if not fc or not cc:
        pytest.skip("amdflang or a host C compiler missing")
    fsrc = os.path.join(ROOT, "-2d-3d-_amd", "fortran", "lbm_rb.f90")
    types = {"lbm_config": rblbm._Config, "lbm_params": rblbm._Params, "lbm_diag": rblbm._Diag,
             "lbm_midplane": rblbm._Mid}
    prog = ["program layout", "  use iso_c_binding", "  use lbm_rb", "  implicit none"]
    for t in types:
        prog.append(f"  type({t}), target :: v_{t}")
    for t, py in types.items():
        prog.append(f"  print '(a,1x,a,1x,i0)', '{t}', 'sizeof', c_sizeof(v_{t})")
        for field, _ in py._fields_:
            prog.append(f"  print '(a,1x,a,1x,i0)', '{t}', '{field}', "
                        f"transfer(c_loc(v_{t}%{field}), 0_c_intptr_t) - transfer(c_loc(v_{t}), 0_c_intptr_t)")
    prog.append("end program layout")
    (tmp_path / "layout.f90").write_text("\n".join(prog) + "\n")
    subprocess.run([fc, "-module-dir", str(tmp_path), "-c", fsrc, "-o", str(tmp_path / "lbm_rb.o")],
                   check=True, cwd=tmp_path)
    # the program uses the module's types only: no procedure of it is linked
    subprocess.run([fc, "-I", str(tmp_path), str(tmp_path / "layout.f90"), "-o", str(tmp_path / "layout")],
                   check=True, cwd=tmp_path)
    out = subprocess.run([str(tmp_path / "layout")], capture_output=True, text=True, check=True).stdout
    got = {(s, f): int(v) for s, f, v in (l.split() for l in out.split("\n") if l.strip())}
    import ctypes
    for t, py in types.items():
        assert got[(t, "sizeof")] == ctypes.sizeof(py), t
        for field, _ in py._fields_:
            assert got[(t, field)] == getattr(py, field).offset, (t, field)


def test_missing_library_fails_loudly(tmp_path):
    """The product path has no fallback: a missing librblbm.so raises before any computation."""
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import rblbm\n"
            "try:\n"
            "    rblbm.library(%r)\n"
            "except FileNotFoundError as e:\n"
            "    print('raised:', e)\n"
            "else:\n"
            "    print('loaded')\n") % (os.path.join(ROOT, "-2d-3d-_amd"), str(tmp_path / "librblbm.so"))
    r = subprocess.run(["python3", "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("raised:") and "no fallback" in r.stdout
