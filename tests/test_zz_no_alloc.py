"""ptbgpu_idempotency_noalloc, the allocation-free variant of ptbgpu_idempotency
(idempotent_packed, source/metrics.f90:316-360): same kernels, scratch in one
static device array.  Called through the C ABI on device inputs, it must give
exactly what the allocating entry point gives, on two calls in a row with the
second on a smaller matrix (the static scratch keeps the first call's
contents), and reject a size beyond its scratch."""
import ctypes

import numpy as np
import pytest

from conftest import small_system

pytestmark = pytest.mark.gpu


def _density(engine, size, seed):
    mol, H, S = small_system("water", size, seed)
    P = np.zeros_like(H)
    engine.sm_stupid_simple(mol.ndim, mol.nel, H, S, P, mol.n, mol.xyz, mol.aoat, 2, 2)
    return mol.ndim, P, S


def test_noalloc_matches_allocating_entry_point(engine):
    import torch
    lib = engine.lib
    lib.ptbgpu_idempotency_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p]
    out = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    for size, seed in (((3, 3, 2), 1), ((2, 2, 1), 0)):   # the second is smaller
        n, P, S = _density(engine, size, seed)
        ref = engine.idempotency_deviation(n, P, S)
        dP = torch.from_numpy(P).cuda()
        dS = torch.from_numpy(S).cuda()
        torch.cuda.synchronize()
        st = lib.ptbgpu_idempotency_noalloc(engine.ctx, n, dP.data_ptr(), dS.data_ptr(),
                                            out.data_ptr())
        assert st == 0, lib.ptbgpu_last_error(engine.ctx)
        engine.sync()
        got = out.item()
        print(f"\nndim {n}: noalloc {got!r} allocating {ref!r}")
        assert got == ref


def test_noalloc_rejects_size_beyond_scratch(engine):
    import torch
    lib = engine.lib
    lib.ptbgpu_idempotency_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p]
    dummy = torch.zeros(16, dtype=torch.float64, device="cuda")
    # the size check comes before any launch: nothing reads the small buffer
    st = lib.ptbgpu_idempotency_noalloc(engine.ctx, 60000, dummy.data_ptr(), dummy.data_ptr(),
                                        dummy.data_ptr())
    assert st == 1
    assert b"scratch" in lib.ptbgpu_last_error(engine.ctx)
