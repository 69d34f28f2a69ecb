"""Filter and positive-count parity (source/pgtb.f90:1532-1579): bit-exact."""
import numpy as np
import pytest

from conftest import small_system

pytestmark = pytest.mark.gpu


def _raw(kind="water", size=(4, 4, 4), seed=0):
    from ptb_amd import geometry, surrogate
    mol = geometry.water_box(size, 3.1, 0.2, seed) if kind == "water" else geometry.alkane(size)
    H, S = surrogate.build_hs(mol)
    return mol, H, S


@pytest.mark.parametrize("filt", [1e-4, 1e-5, 1e-6])
def test_filter_bit_exact_h2o64(engine, filt):
    from oracle import ptb_oracle as O
    mol, H, S = _raw()
    n = mol.ndim
    for A in (S, H):
        ref = A.copy()
        before = O.analyse_matrix_sym(ref, n)["nonzero"]
        O.filter_matrix_sym(ref, n, filt)
        after = O.analyse_matrix_sym(ref, n)["nonzero"]
        got = A.copy()
        c = engine.filter_matrix_sym(got, n, filt)
        assert got.tobytes() == ref.tobytes()
        assert c == (before, after)


def test_filter_edge_values(engine):
    from oracle import ptb_oracle as O
    f = 1e-4
    special = np.array([0.0, -0.0, f, -f, np.nextafter(f, 0), -np.nextafter(f, 0), np.nan,
                        -np.nan, np.inf, -np.inf, 5e-324, -5e-324, 1.0, -1.0, 2e-4, -2e-4])
    for n in (1, 2, 3, 5, 16, 33):
        npk = n * (n + 1) // 2
        rng = np.random.default_rng(n)
        A = special[rng.integers(0, len(special), npk)].copy()
        ref = A.copy()
        O.filter_matrix_sym(ref, n, f)
        got = A.copy()
        before, after = engine.filter_matrix_sym(got, n, f)
        assert got.tobytes() == ref.tobytes(), n
        assert before == O.analyse_matrix_sym(A, n)["nonzero"]
        assert after == O.analyse_matrix_sym(ref, n)["nonzero"]


def test_filter_device_unaligned(engine):
    """The _dev entry on a pointer that is 8 bytes off a 16-byte boundary."""
    import torch
    from oracle import ptb_oracle as O
    mol, H, S = _raw(size=(2, 2, 2))
    n = mol.ndim
    buf = torch.zeros(S.size + 1, dtype=torch.float64, device="cuda")
    view = buf[1:]
    view.copy_(torch.from_numpy(S))
    c = engine.filter_matrix_sym_dev(view, n, 1e-4)
    torch.cuda.synchronize()
    ref = S.copy()
    O.filter_matrix_sym(ref, n, 1e-4)
    assert view.cpu().numpy().tobytes() == ref.tobytes()
    assert c[1] == O.analyse_matrix_sym(ref, n)["nonzero"]


def test_analyse_matches(engine):
    from oracle import ptb_oracle as O
    mol, H, S = small_system(size=(2, 2, 2))
    for A in (H, S):
        a = engine.analyse_matrix_sym(A, mol.ndim)
        r = O.analyse_matrix_sym(A, mol.ndim)
        assert (a["nonzero"], a["average"]) == (r["nonzero"], r["average"])
