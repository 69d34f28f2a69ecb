"""Loaders for the in-tree native libraries (built by ``ptb_amd.build``).

Every loader fails loudly when its library is missing: there is no CPU
fallback for the engine.
"""
from __future__ import annotations

import ctypes
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

_cache: dict = {}


def lib_path(name: str) -> str:
    return os.path.join(LIB_DIR, name)


def _load(name: str, mode=ctypes.RTLD_GLOBAL):
    if name in _cache:
        return _cache[name]
    path = lib_path(name)
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(or `python -m ptb_amd.build`) first")
    lib = ctypes.CDLL(path, mode=mode)
    _cache[name] = lib
    return lib


def surrogate_lib():
    lib = _load("libptbsurr.so")
    lib.ptbsurr_build.restype = ctypes.c_int64
    lib.ptbsurr_ndim.restype = ctypes.c_int64
    return lib


def engine_lib():
    """The HIP engine behind the C ABI declared in include/ptbgpu.h."""
    return _load("libptbgpu.so")
