"""No read of an element a density build did not write.

With PTBGPU_POISON_ARENAS=1 set when the context is created, every
floating-point arena of the submatrix build (H, S/L, K, L^-1, the eigenvector
and work arenas, the fe projections, the mixed build's float copy, reflectors
and residuals) is filled with NaN before each use.  The kernels write only
the elements their readers look at (the triangular inverse, for example,
writes its diagonal blocks and the block above each, not the rest of its
arena, and the gather writes the overlap block's lower triangle only), so a
reader that strayed onto an unwritten element would carry a NaN
into the density.  The poisoned builds must equal the ordinary ones bit for
bit, single-molecule, batched, and with the divide-and-conquer merges over
the whole chip."""
import os

import numpy as np
import pytest

from conftest import small_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def poisoned():
    from ptb_amd.engine import Engine
    os.environ["PTBGPU_POISON_ARENAS"] = "1"
    try:
        eng = Engine(0)
    finally:
        del os.environ["PTBGPU_POISON_ARENAS"]
    assert eng.get_tuning("poison_arenas") == 1
    yield eng
    eng.close()


CASES = [("water", (4, 4, 4), 3, 2), ("water", (3, 3, 3), 1, 3), ("alkane", 12, 0, 2)]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind,size,seed,cols", CASES)
def test_poisoned_arenas_change_nothing(engine, poisoned, precision, kind, size, seed, cols):
    mol, H, S = small_system(kind, size, seed, filt=1e-4 if kind == "water" else 1e-5)
    n = mol.ndim
    out = []
    for eng in (engine, poisoned):
        saved = eng.get_tuning("mixed_min_m")
        eng.set_tuning("mixed_min_m", 0)  # every submatrix through the float path
        try:
            P = np.zeros_like(H)
            rep = eng.sm_stupid_simple(n, mol.nel, H, S, P, mol.n, mol.xyz, mol.aoat, cols, 2,
                                       precision=precision)
        finally:
            eng.set_tuning("mixed_min_m", saved)
        out.append((P, rep["mu"]))
    print(f"\n{kind}{size} cols={cols} {precision}: finite {np.isfinite(out[1][0]).all()}")
    assert np.isfinite(out[1][0]).all()
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])


def test_poisoned_batch_and_chip_wide_merges(engine, poisoned):
    """A two-molecule batch, and the merges from device scratch over the
    whole chip (knob dc_global_min), with poisoned arenas."""
    from ptb_amd import geometry, surrogate
    mols = geometry.mixed_batch(4, 2)
    items = []
    for mol in mols:
        Hm, Sm = surrogate.build_hs(mol)
        filt = 1e-4 if mol.name.startswith("h2o") else 1e-5
        engine.filter_matrix_sym(Sm, mol.ndim, filt)
        engine.filter_matrix_sym(Hm, mol.ndim, filt)
        items.append((mol, Hm, Sm))
    results = []
    for eng in (engine, poisoned):
        for v in (0, 64):
            eng.set_tuning("dc_global_min", v)
            batch = [(m.ndim, m.nel, Hm, Sm, np.zeros_like(Hm), m.n, m.xyz, m.aoat)
                     for m, Hm, Sm in items]
            try:
                eng.sm_stupid_simple_batch(batch, 2, 2, precision="mixed")
            finally:
                eng.set_tuning("dc_global_min", 0)
            results.append([b[4] for b in batch])
    for r in results[1:]:
        for a, b in zip(results[0], r):
            assert np.isfinite(b).all()
            assert np.array_equal(a, b)
