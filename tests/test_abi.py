"""The C-ABI library loads and exports every entry point include/ptbgpu.h declares."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ptbgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ptbgpu_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for want in ("ptbgpu_init", "ptbgpu_finalize", "ptbgpu_filter_sym", "ptbgpu_get_nel",
                 "ptbgpu_band_energy", "ptbgpu_idempotency"):
        assert want in syms


def test_library_exports_every_declared_symbol():
    from ptb_amd import _native
    lib = ctypes.CDLL(_native.lib_path("libptbgpu.so"))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"declared but not exported: {missing}"


def test_strerror_without_device():
    from ptb_amd import _native
    lib = ctypes.CDLL(_native.lib_path("libptbgpu.so"))
    lib.ptbgpu_strerror.restype = ctypes.c_char_p
    lib.ptbgpu_strerror.argtypes = [ctypes.c_int64]
    assert lib.ptbgpu_strerror(0) == b"ok"
    assert b"positive definite" in lib.ptbgpu_strerror(3)


def test_fortran_module_exports_drop_in():
    """The Fortran shim exports the reference's module procedure name."""
    import os
    import subprocess
    from ptb_amd import _native
    so = _native.lib_path("libptb_sm_gpu.so")
    if not os.path.exists(so):
        import pytest
        pytest.skip("Fortran shim not built (no flang)")
    syms = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True,
                          text=True).stdout
    assert "sm_stupid_simple" in syms.lower()
