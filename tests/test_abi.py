"""CPU checks of the C ABI library: it loads and exports every function the
header declares.  No compute calls (there is no GPU here)."""
import ctypes as C
import subprocess

import pytest

from cosmicatlas_amd import _abi


def test_library_exports_every_header_function():
    lib = _abi.load()
    names = _abi.header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(names) <= exported


def test_status_strings_and_size_query():
    lib = _abi.load()
    assert lib.cm_status_string(0) == b"ok"
    assert lib.cm_status_string(5) == b"over HBM budget"
    for N, ok in [(128, 1), (192, 1), (512, 1), (768, 1), (1024, 1), (100, 0), (8, 1), (6, 0), (2048, 0),
                  (72, 0)]:
        assert lib.cm_fft_supported(N) == ok, N


def test_null_context_is_rejected():
    lib = _abi.load()
    assert lib.cm_fft_r2c(None, 64, 0, None, None, 0.0) == _abi.CM_ERR_INVALID
    assert lib.cm_mesh_assign(None, None, None, None, 0, None, 0, 0) == _abi.CM_ERR_INVALID
    assert lib.cm_last_error(None) == b"null context"
