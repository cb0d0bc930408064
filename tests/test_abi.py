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
