"""(H, S) inputs for the density-build workloads.

Wraps ``csrc/surrogate.c`` (a Gaussian-overlap / extended-Hueckel stand-in
for PTB's Hamiltonian builder, which is not part of this engine) and returns
packed-upper FP64 arrays in the layout of the reference's ``Hvec``/``Svec``
(source/solve.f:149-165).  This module only produces inputs; the density build
itself never calls it.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .geometry import Molecule

# atom pairs farther apart than this (bohr) have all overlaps below 1e-12
RCUT_BOHR = 24.0


def build_hs(mol: Molecule, rcut: float = RCUT_BOHR):
    """Return packed-upper (H, S) float64 arrays of length ndim(ndim+1)/2."""
    lib = _native.surrogate_lib()
    nd = mol.ndim
    npk = nd * (nd + 1) // 2
    H = np.empty(npk, dtype=np.float64)
    S = np.empty(npk, dtype=np.float64)
    z = np.ascontiguousarray(mol.znum, dtype=np.int64)
    x = np.ascontiguousarray(mol.xyz, dtype=np.float64)
    got = lib.ptbsurr_build(ctypes.c_int64(mol.n), z.ctypes.data_as(ctypes.c_void_p),
                            x.ctypes.data_as(ctypes.c_void_p), ctypes.c_double(rcut),
                            H.ctypes.data_as(ctypes.c_void_p),
                            S.ctypes.data_as(ctypes.c_void_p))
    if got != nd:
        raise RuntimeError(f"surrogate build failed ({got} != {nd})")
    return H, S
