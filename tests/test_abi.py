"""CPU checks of the drop-in boundary: the C-ABI library loads without a GPU, exports every symbol
declared in include/rerdmft_gpu.h, and the Python front-end binds exactly that set."""
import ctypes
import os
import re

import pytest

import rerdmft_amd as rr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "rerdmft_gpu.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(rr_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_bound_symbols():
    assert declared_symbols() == sorted(rr.EXPORTED)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(rr.LIB_PATH):
        pytest.skip("librerdmft_gpu.so not built")
    lib = ctypes.CDLL(rr.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_null_handle_is_rejected_without_a_gpu():
    if not os.path.exists(rr.LIB_PATH):
        pytest.skip("librerdmft_gpu.so not built")
    lib = rr.lib()
    assert lib.rr_destroy(None) == 1  # RR_ERR_ARG
    assert lib.rr_last_error(None) == b"null handle"


def test_header_compiles_as_c():
    import subprocess
    import tempfile
    src = '#include "rerdmft_gpu.h"\nint main(void) { rr_handle h = 0; (void)h; return 0; }\n'
    with tempfile.TemporaryDirectory() as d:
        f = os.path.join(d, "t.c")
        open(f, "w").write(src)
        r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", f, "-o",
                            os.path.join(d, "t.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
