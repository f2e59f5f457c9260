"""The C-ABI library: it loads, exports every entry point include/ccs.h declares, and
reports errors through return codes (no compute calls: these tests run without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import PKG, ROOT

HEADER = os.path.join(ROOT, "include", "ccs.h")
LIB = os.path.join(PKG, "libccs.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|uint64_t|const char\*)\s+(ccs_\w+)\s*\(", text, re.M)))


def test_library_built():
    assert os.path.exists(LIB), "build first: python cunningham-chain-search_amd/build.py"


def test_exports_every_declared_symbol():
    names = declared()
    assert len(names) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    from ccs import _lib
    assert sorted(_lib.EXPORTS) == names


def test_loads_and_reports_errors_without_compute():
    from ccs import _lib
    lib = _lib.load()
    # invalid arguments are rejected by return code, not by exiting
    assert lib.ccs_shape_build(None, None, None, None) < 0
    assert lib.ccs_last_error()
    out = C.c_void_p()
    rc = lib.ccs_ctx_create(-1, C.byref(out))
    assert rc < 0 and not out.value
    assert lib.ccs_format_results(None, 0, 8, 18, None, 0, None) == 0
    n = C.c_int(-1)
    assert lib.ccs_device_count(C.byref(n)) == 0 and n.value >= 0


def test_struct_layouts():
    from ccs import _lib
    assert _lib.SURVIVOR_DTYPE.itemsize == 16          # v13 Survivor (v13.cu:207-212)
    assert _lib.JOB_SURVIVOR_DTYPE.itemsize == 8
    assert _lib.CHAIN_RESULT_DTYPE.itemsize == 32
    assert C.sizeof(_lib.Job) == 32
    assert C.sizeof(_lib.ShapeParams) == 80


def test_format_results_matches_reference_lines():
    import numpy as np
    from ccs import _lib, format_results
    res = np.zeros(3, _lib.CHAIN_RESULT_DTYPE)
    root = 0x57c4647809417bf1819285
    res[0] = (0, 18, 0, 7, root & ((1 << 64) - 1), root >> 64)
    res[1] = (1, 9, 2, 7 | 8, 0x1234567, 0)
    res[2] = (2, 3, 0, 7, 5, 0)                         # below the log threshold
    text = format_results(res, log_threshold=8, target=18)
    assert text == "TARGET CC18 0x57C4647809417BF1819285 87\nNON-ROOT CC9 0x1234567 25\n"
