"""Dense spectral routines on the GPU: matrix_sign / matrix_root
(source/sm.f90:410-457) and the -check density solve2 with Fermi smearing
(source/pgtb.f90:1066-1170), against the oracle and closed forms."""
import numpy as np
import pytest
import scipy.linalg as sla

from conftest import small_system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [7, 60, 301])
def test_matrix_sign(engine, n):
    from oracle import ptb_oracle as O
    rng = np.random.default_rng(n)
    B = rng.normal(size=(n, n))
    A = B + B.T
    S = engine.matrix_sign(A)
    assert np.abs(S - O.matrix_sign(A)).max() < 1e-10
    assert np.abs(S @ S - np.eye(n)).max() < 1e-10          # sign^2 = I
    assert np.abs(S @ A - A @ S).max() < 1e-9               # commutes with A
    # only the upper triangle is read, as dsyevd('U')
    Ag = np.triu(A) + np.tril(rng.normal(size=(n, n)), -1)
    assert np.abs(engine.matrix_sign(Ag) - S).max() < 1e-10


@pytest.mark.parametrize("n", [97, 301])
def test_matrix_sign_degenerate_spectrum(engine, n):
    """Four eigenvalues, each repeated n/4 times: the divide-and-conquer merges
    deflate most of their poles, rotating pairs across the two children (the
    rotation path and the dense-column products).  Closed forms:
    sign(A) = Q sign(L) Q^T and A^{-1/2} = Q L^{-1/2} Q^T (A = Q L Q^T)."""
    rng = np.random.default_rng(7 + n)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    lam = np.array([-2.0, -1.0, 1.0, 3.0])[np.arange(n) % 4]
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)
    S = engine.matrix_sign(A)
    assert np.abs(S - (Q * np.sign(lam)) @ Q.T).max() < 1e-10
    assert np.abs(S @ S - np.eye(n)).max() < 1e-10
    P = A + 3.0 * np.eye(n)  # eigenvalues 1, 2, 4, 6
    R = engine.matrix_root(P)
    assert np.abs(R - (Q * (lam + 3.0) ** -0.5) @ Q.T).max() < 1e-10


@pytest.mark.parametrize("n", [5, 64, 250])
def test_matrix_root(engine, n):
    from oracle import ptb_oracle as O
    rng = np.random.default_rng(100 + n)
    B = rng.normal(size=(n, n))
    A = B @ B.T + n * np.eye(n)
    R = engine.matrix_root(A)
    assert np.abs(R - O.matrix_root(A)).max() < 1e-12
    assert np.abs(R @ A @ R - np.eye(n)).max() < 1e-10


@pytest.mark.parametrize("case", [("water", (2, 2, 2), 0), ("water", (3, 3, 2), 1),
                                  ("alkane", 8, 0)])
def test_solve2_matches_oracle(engine, case):
    from oracle import ptb_oracle as O
    mol, H, S = small_system(*case)
    n = mol.ndim
    for mode, shift in ((1, (0.0, 0.0)), (2, (-0.25, 0.1))):
        P, e, focc, U, fail = engine.solve2(mode, n, mol.nel, 0, 300.0, H, S, shift)
        Pr, er, fr = O.solve2(n, mol.nel, H, S, 300.0, 0, mode, shift)
        assert not fail
        assert np.abs(e - er).max() < 1e-9
        assert np.abs(focc - fr).max() < 1e-9
        assert np.abs(P - Pr).max() < 1e-8
        # U is S-orthonormal and diagonalises H
        Hd, Sd = O.blowsym(H, n), O.blowsym(S, n)
        assert np.abs(U.T @ Sd @ U - np.eye(n)).max() < 1e-9
        e0 = er if mode < 2 else (er - shift[0]) / (1.0 + shift[1])
        assert np.abs(U.T @ Hd @ U - np.diag(e0)).max() < 1e-8


def test_solve2_hot_smearing_and_open_shell(engine):
    """20000 K smears across the gap; an odd electron count gives alpha != beta."""
    from oracle import ptb_oracle as O
    mol, H, S = small_system("water", (2, 2, 1), 0)
    n = mol.ndim
    for nel, nopen, et in ((mol.nel, 0, 20000.0), (mol.nel - 1, 1, 300.0), (mol.nel, 2, 300.0)):
        P, e, focc, U, fail = engine.solve2(1, n, nel, nopen, et, H, S)
        Pr, er, fr = O.solve2(n, nel, H, S, et, nopen)
        assert np.abs(focc - fr).max() < 1e-9
        assert np.abs(P - Pr).max() < 1e-8
        assert abs(O.get_nel(n, P, S) - nel) < 1e-8
