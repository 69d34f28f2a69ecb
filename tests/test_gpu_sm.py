"""Submatrix density build (source/sm.f90:34-409): GPU vs the oracle.

Bar (FP64 drop-in): max |P - P_oracle| <= 1e-8 per element, |dE_band| <= 1e-7 Eh,
|dN_el| <= 1e-10, identical chemical potential and bisection pass count.
"""
import numpy as np
import pytest

from conftest import small_system

pytestmark = pytest.mark.gpu

CASES = [
    ("water", (2, 2, 2), 0, 2),
    ("water", (2, 2, 2), 0, 1),
    ("water", (2, 2, 2), 0, 3),
    ("water", (3, 3, 3), 1, 2),
    ("alkane", 8, 0, 2),
    ("alkane", 8, 0, 3),
]


def _run(engine, kind, size, seed, cols):
    from oracle import ptb_oracle as O
    mol, H, S = small_system(kind, size, seed)
    n = mol.ndim
    res = O.sm_stupid_simple(n, mol.nel, H, S, mol.n, mol.xyz, mol.aoat, cols, 2)
    P = np.full_like(H, np.nan)
    rep = engine.sm_stupid_simple(n, mol.nel, H, S, P, mol.n, mol.xyz, mol.aoat, cols, 2)
    return mol, H, S, res, P, rep


@pytest.mark.parametrize("kind,size,seed,cols", CASES)
def test_density_parity(engine, kind, size, seed, cols):
    from oracle import ptb_oracle as O
    mol, H, S, res, P, rep = _run(engine, kind, size, seed, cols)
    n = mol.ndim
    print(f"\n{kind}{size} cols={cols} ndim={n} nsm={rep['nsm']} max_m={rep['max_m']} "
          f"mu={rep['mu']} ref_mu={res.mu} muiter={rep['muiter']}/{res.muiter}")
    m, eig, fe = engine.sm_spectra(rep["nsm"], rep["sum_m"])
    worst = 0.0
    for s in range(rep["nsm"]):
        assert m[s] == len(res.maps[s])
        err = np.abs(np.sort(eig[s]) - res.eigvals[s]).max()
        worst = max(worst, err)
    print(f"max eigenvalue error {worst:.3e}")
    assert worst < 1e-9
    assert rep["muiter"] == res.muiter
    assert rep["mu"] == res.mu
    err = np.abs(P - res.P).max()
    print(f"max|P-P_oracle| = {err:.3e}")
    assert err <= 1e-8
    e_gpu = O.band_structure(n, P, H)
    e_ref = O.band_structure(n, res.P, H)
    assert abs(e_gpu - e_ref) <= 1e-7
    assert abs(O.get_nel(n, P, S) - O.get_nel(n, res.P, S)) <= 1e-10
    # zeros outside each owner's map are exact zeros, as in the reference
    assert np.array_equal(P == 0.0, res.P == 0.0)


def test_fe_sums_to_owned_count(engine):
    mol, H, S, res, P, rep = _run(engine, "water", (2, 2, 2), 0, 2)
    m, eig, fe = engine.sm_spectra(rep["nsm"], rep["sum_m"])
    for s in range(rep["nsm"]):
        assert abs(fe[s].sum() - len(res.groups[s])) < 1e-9


def test_unsupported_modes_raise(engine):
    from ptb_amd.engine import PtbGpuError
    mol, H, S = small_system("water", (1, 1, 2), 0)
    P = np.zeros_like(H)
    with pytest.raises(PtbGpuError) as ex:
        engine.sm_stupid_simple(mol.ndim, mol.nel, H, S, P, mol.n, mol.xyz, mol.aoat, 2, 1)
    assert ex.value.status == 2
    with pytest.raises(PtbGpuError) as ex:
        engine.sm_stupid_simple(mol.ndim, mol.nel, H, S, P, mol.n, mol.xyz, mol.aoat, 5, 2)
    assert ex.value.status == 2


def splitmix_stream(count, seed=42):
    """The engine's own k-means generator (splitmix64 seeded with 42, uniform
    = top 53 bits / 2^53), restated to feed the oracle the same numbers."""
    mask = (1 << 64) - 1
    x, out = seed, []
    for _ in range(count):
        x = (x + 0x9E3779B97F4A7C15) & mask
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        out.append((z >> 11) * 2.0 ** -53)
    return np.array(out)


def _restart(stream):
    def uniform(_nsm):
        it = iter(stream)
        return lambda: next(it)
    return uniform


KMEANS_FIXTURES = ["sm_w16_kmeans.npz", "sm_c40_kmeans.npz"]


@pytest.mark.parametrize("fixture", KMEANS_FIXTURES)
def test_kmeans_columns_match_reference(engine, fixture):
    """submatrix_columns = 4 with the compiler's RANDOM_NUMBER stream: the
    reference's clusters, nsm choice and density (2 and 3 clusters)."""
    import os
    from conftest import ROOT
    d = np.load(os.path.join(ROOT, "tests", "golden", fixture))
    n = int(d["ndim"])
    P = np.zeros_like(d["P"])
    rep = engine.sm_stupid_simple(n, int(d["nel"]), d["H"].copy(), d["S"].copy(), P, int(d["n"]),
                                  d["xyz"], d["aoat"], 4, 2, kmeans_stream=d["stream"])
    err = np.abs(P - d["P"]).max()
    print(f"k-means nsm {rep['nsm']} mu {rep['mu']} muiter {rep['muiter']}: max|P-P_ref| {err:.3e}")
    assert rep["nsm"] == int(d["nsm"])
    assert rep["muiter"] == len(d["mu_trace"])
    assert rep["mu"] == d["mu_trace"][-1][1]
    assert err < 1e-8


def test_kmeans_columns_own_generator(engine):
    """Without a stream the engine draws from its own generator; the oracle fed
    the same numbers gives the same clusters and density."""
    from oracle import ptb_oracle as O
    mol, H, S = small_system("water", (8, 2, 1), 1)
    n = mol.ndim
    P = np.zeros_like(H)
    rep = engine.sm_stupid_simple(n, mol.nel, H, S, P, mol.n, mol.xyz, mol.aoat, 4, 2)
    res = O.sm_stupid_simple(n, mol.nel, H, S, mol.n, mol.xyz, mol.aoat, 4, 2,
                             uniform=_restart(splitmix_stream(64)))
    print(f"nsm {rep['nsm']} vs {len(res.groups)}, max|dP| {np.abs(P - res.P).max():.3e}")
    assert rep["nsm"] == len(res.groups)
    assert np.abs(P - res.P).max() < 1e-8


def test_kmeans_short_stream_is_an_error(engine):
    from ptb_amd.engine import PtbGpuError
    mol, H, S = small_system("water", (2, 2, 1), 0)
    P = np.zeros_like(H)
    with pytest.raises(PtbGpuError) as ex:
        engine.sm_stupid_simple(mol.ndim, mol.nel, H, S, P, mol.n, mol.xyz, mol.aoat, 4, 2,
                                kmeans_stream=np.array([0.5]))
    assert ex.value.status == 1
