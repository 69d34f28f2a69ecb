"""The eigensolver above 2048 rows: the multi-workgroup tridiagonalisation
(one launch per step of each column over the whole GPU) and the
divide-and-conquer merges whose vectors live in device scratch instead of LDS.

Checked through what the mathematics fixes, at sizes the oracle's dense
solves still finish in seconds: the -check solve2 (source/pgtb.f90:1066-1170)
at ndim 4140 (eigen-equation residual, S-orthonormality, the electron count
and P = U diag(focc) U^T), matrix_sign / matrix_root
(source/sm.f90:410-457) of matrices with a known spectrum (a degenerate one
forces the deflation paths of the large merges), and the submatrix build with
one 2944-row block against the oracle.  The inputs use filter 1e-5: at 1e-4
the filtered overlap of these boxes is indefinite (lowest eigenvalue about
-1e-4), which the last two tests use to check that a failed factorisation is
reported and leaves every result finite.
"""
import numpy as np
import pytest

from conftest import small_system

pytestmark = pytest.mark.gpu


def _mm(a, b):
    """FP64 product on the GPU (a checker, not the path under test)."""
    import torch
    return (torch.from_numpy(np.ascontiguousarray(a)).cuda() @
            torch.from_numpy(np.ascontiguousarray(b)).cuda()).cpu().numpy()


def test_solve2_at_4140(engine):
    """180 waters (ndim 4140): U is S-orthonormal and diagonalises H to the
    eigenvalues returned (ascending), the Fermi occupations sum to the
    electron count, P = U diag(focc) U^T, and tr(P S) = nel."""
    from oracle import ptb_oracle as O
    import time
    mol, H, S = small_system("water", (6, 6, 5), 2, filt=1e-5)
    n = mol.ndim
    assert n > 4096
    t0 = time.perf_counter()
    P, e, focc, U, fail = engine.solve2(1, n, mol.nel, 0, 300.0, H, S)
    dt = time.perf_counter() - t0
    print(f"\nsolve2 ndim {n}: {dt:.2f} s", flush=True)
    assert not fail
    assert np.all(np.diff(e) >= 0.0)
    Hd, Sd = O.blowsym(H, n), O.blowsym(S, n)
    SU = _mm(Sd, U)
    orth = np.abs(_mm(U.T, SU) - np.eye(n)).max()
    resid = np.abs(_mm(Hd, U) - SU * e).max() / max(1.0, np.abs(e).max())
    print(f"|U^T S U - I| {orth:.2e}  |H U - S U e| / |e| {resid:.2e}", flush=True)
    assert orth < 1e-8 and resid < 1e-8
    assert abs(focc.sum() - mol.nel) < 1e-9
    Pd = _mm(U * focc, U.T)
    assert np.abs(O.packsym(Pd) - P).max() < 1e-10
    assert abs(O.get_nel(n, P, S) - mol.nel) < 1e-8


def _known(n, lam, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = _mm(Q * lam, Q.T)
    return Q, 0.5 * (A + A.T)


@pytest.mark.parametrize("n,degenerate", [(3000, False), (4100, True)])
def test_matrix_sign_large(engine, n, degenerate):
    """sign(A) = Q sign(L) Q^T for A = Q L Q^T with a gap around zero; with
    four eigenvalues each repeated n/4 times most poles deflate in the
    merges (the device-scratch merges' rotation path)."""
    rng = np.random.default_rng(n)
    if degenerate:
        lam = np.array([-2.0, -1.0, 1.0, 3.0])[np.arange(n) % 4]
    else:
        lam = np.concatenate([-rng.uniform(1.0, 3.0, n // 2), rng.uniform(1.0, 3.0, n - n // 2)])
    Q, A = _known(n, lam, n + 1)
    S = engine.matrix_sign(A)
    ref = _mm(Q * np.sign(lam), Q.T)
    err = np.abs(S - ref).max()
    sq = np.abs(_mm(S, S) - np.eye(n)).max()
    print(f"\nmatrix_sign n {n}: |S - ref| {err:.2e}  |S^2 - I| {sq:.2e}")
    assert err < 1e-9 and sq < 1e-9


def test_matrix_root_large(engine):
    """A^-1/2 of an SPD matrix with eigenvalues in [1, 10] at n = 2500."""
    n = 2500
    lam = np.random.default_rng(5).uniform(1.0, 10.0, n)
    Q, A = _known(n, lam, 6)
    R = engine.matrix_root(A)
    err = np.abs(R - _mm(Q * lam ** -0.5, Q.T)).max()
    rar = np.abs(_mm(_mm(R, A), R) - np.eye(n)).max()
    print(f"\nmatrix_root n {n}: |R - ref| {err:.2e}  |R A R - I| {rar:.2e}")
    assert err < 1e-10 and rar < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_single_block_beyond_2048(engine, precision):
    """128 waters, all columns in one submatrix (submatrix_columns 3, m =
    2944): the FP64 build matches the oracle's dense solve to 1e-8 with the
    same mu; the mixed build (float storage in the multi-workgroup
    reduction) matches the FP64 one to 1e-6."""
    from oracle import ptb_oracle as O
    mol, H, S = small_system("water", (8, 4, 4), 0, filt=1e-5)
    n = mol.ndim
    P = np.full_like(H, np.nan)
    rep = engine.sm_stupid_simple(n, mol.nel, H, S, P, mol.n, mol.xyz, mol.aoat, 3, 2,
                                  precision=precision)
    assert rep["nsm"] == 1 and rep["max_m"] == n > 2048
    if precision == "fp64":
        res = O.sm_stupid_simple(n, mol.nel, H, S, mol.n, mol.xyz, mol.aoat, 3, 2)
        err = np.abs(P - res.P).max()
        print(f"\nm {n}: max|P - P_oracle| {err:.2e}, mu {rep['mu']} / {res.mu}")
        assert rep["mu"] == res.mu
        assert err < 1e-8
    else:
        P64 = np.full_like(H, np.nan)
        engine.sm_stupid_simple(n, mol.nel, H, S, P64, mol.n, mol.xyz, mol.aoat, 3, 2)
        err = np.abs(P - P64).max()
        print(f"\nm {n}: mixed max|P - P_fp64| {err:.2e}")
        assert err < 1e-6


def test_failed_factorisation_is_reported_and_finite(engine):
    """An indefinite overlap (filter 1e-4 on the 128- and 180-water boxes):
    the Cholesky meets non-positive pivots.  solve2 reports it (dsygvd
    INFO > 0), the submatrix build counts the block in not_spd and leaves its
    columns out of P; every returned value stays finite."""
    mol, H, S = small_system("water", (8, 4, 4), 0)
    n = mol.ndim
    P = np.full_like(H, np.nan)
    rep = engine.sm_stupid_simple(n, mol.nel, H, S, P, mol.n, mol.xyz, mol.aoat, 3, 2)
    print(f"\nm {n}: not_spd {rep['not_spd']}, max|P| {np.abs(P).max():.3e}")
    assert rep["not_spd"] == 1
    assert np.all(P == 0.0)
    mol, H, S = small_system("water", (6, 6, 5), 2)
    Pd, e, focc, U, fail = engine.solve2(1, mol.ndim, mol.nel, 0, 300.0, H, S)
    print(f"solve2 ndim {mol.ndim}: fail {fail}")
    assert fail
    assert np.isfinite(Pd).all() and np.isfinite(e).all() and np.isfinite(U).all()
